// Projection preprocessing on gfx950: intensity -> line integral, Feldkamp weight,
// ramp-family row filter, and conversion into the backprojection layout.
//
// Reference behaviour restated (not translated):
//  - log forms: Ct/CtVolume.cpp:802-812 + 860-866 (CPU) and :892-904 (GPU);
//  - Feldkamp weight D/sqrt(D^2+u^2+v^2): Ct/CtVolume.cpp:814-830, Ct/cuda_ct.cu:102-131;
//  - row filter: real DFT of length W without padding, bins 0..W/2 scaled by the
//    window with r = k/(W/2+1), unnormalised real inverse: Ct/CtVolume.cpp:41-101,
//    832-858, 910-917 and Ct/cuda_ct.cu:48-100.
//
// Design: one workgroup filters two rows at once as the real and imaginary parts of
// one complex sequence (the response is real and even, so real rows stay real).
// The whole row pair lives in LDS; a power-of-two W runs a radix-2 DIT forward FFT
// (bit-reversed load), multiplies by the response and conjugates, then a DIF FFT
// whose bit-reversed output is read back in natural order.  Other W use a direct
// circular convolution with the precomputed kernel.
#include "ctg_internal.h"

#include <hip/hip_fp16.h>

#include <cmath>

namespace ctg {
namespace {

constexpr int kRowThreads = 256;
constexpr double kPi = 3.14159265358979323846;

template <typename T>
struct Cx {
    T re, im;
};

struct PreArgs {
    int W, H, log2W;
    long long nrows;  // total rows (nproj * H for images)
    float la, lb, lk;  // p = -log((I * la + lb) * lk)
    double dla, dlb, dlk;
    float fcd, upre, vpre;
    int do_log, do_weight, do_filter, pow2;
};

__device__ inline float dlog(float x) { return logf(x); }
__device__ inline double dlog(double x) { return log(x); }
__device__ inline float dsqrt(float x) { return sqrtf(x); }
__device__ inline double dsqrt(double x) { return sqrt(x); }

template <typename T, typename Tin>
__device__ inline T load_pixel(const Tin* __restrict__ raw, const float* __restrict__ lin, size_t idx,
                               const PreArgs& a, int r, int c) {
    T x;
    if (raw != nullptr) {
        T I = (T)raw[idx];
        if (a.do_log) {
            if constexpr (sizeof(T) == 4) {
                float y = fmaf((float)I, a.la, a.lb);
                x = -dlog(y * a.lk);
            } else {
                T y = I * (T)a.dla + (T)a.dlb;
                x = -dlog(y * (T)a.dlk);
            }
        } else {
            x = I;
        }
    } else {
        x = (T)lin[idx];
    }
    if (a.do_weight) {
        T u = (T)c - (T)a.upre;
        T v = (T)(-r) + (T)a.vpre;
        T D = (T)a.fcd;
        x = x * (D / dsqrt(D * D + u * u + v * v));
    }
    return x;
}

template <typename T>
__device__ void fft_dit(T* re, T* im, int W, int log2W, const Cx<T>* __restrict__ tw) {
    for (int lh = 0; lh < log2W; ++lh) {
        const int half = 1 << lh;
        const int sh = log2W - lh - 1;
        __syncthreads();
        for (int j = threadIdx.x; j < (W >> 1); j += blockDim.x) {
            int k = j & (half - 1);
            int i0 = ((j - k) << 1) + k, i1 = i0 + half;
            Cx<T> w = tw[k << sh];
            T br = re[i1], bi = im[i1];
            T xr = br * w.re - bi * w.im;
            T xi = br * w.im + bi * w.re;
            T ar = re[i0], ai = im[i0];
            re[i1] = ar - xr; im[i1] = ai - xi;
            re[i0] = ar + xr; im[i0] = ai + xi;
        }
    }
    __syncthreads();
}

template <typename T>
__device__ void fft_dif(T* re, T* im, int W, int log2W, const Cx<T>* __restrict__ tw) {
    for (int lh = log2W - 1; lh >= 0; --lh) {
        const int half = 1 << lh;
        const int sh = log2W - lh - 1;
        __syncthreads();
        for (int j = threadIdx.x; j < (W >> 1); j += blockDim.x) {
            int k = j & (half - 1);
            int i0 = ((j - k) << 1) + k, i1 = i0 + half;
            Cx<T> w = tw[k << sh];
            T ar = re[i0], ai = im[i0], br = re[i1], bi = im[i1];
            re[i0] = ar + br; im[i0] = ai + bi;
            T dr = ar - br, di = ai - bi;
            re[i1] = dr * w.re - di * w.im;
            im[i1] = dr * w.im + di * w.re;
        }
    }
    __syncthreads();
}

__device__ inline unsigned bitrev(unsigned n, int bits) { return __brev(n) >> (32 - bits); }

__device__ inline float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// One workgroup per row pair.  Output rows (T) row-major; for float output the
// per-image max |value| is folded into maxabs with an atomic max on the bits.
template <typename T, typename Tin>
__global__ __launch_bounds__(kRowThreads) void rows_kernel(PreArgs a, const Tin* __restrict__ raw,
                                                          const float* __restrict__ lin,
                                                          const T* __restrict__ resp,
                                                          const Cx<T>* __restrict__ tw,
                                                          const T* __restrict__ conv,
                                                          T* __restrict__ out,
                                                          unsigned* __restrict__ maxabs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* re = reinterpret_cast<T*>(smem_raw);
    T* im = re + a.W;
    const int W = a.W;
    const long long ra = 2ll * blockIdx.x;
    const long long rb = ra + 1;
    const bool hasb = rb < a.nrows;
    const int rin_a = (int)(ra % a.H), rin_b = (int)(rb % a.H);
    const size_t offa = (size_t)ra * W, offb = (size_t)rb * W;

    if (a.do_filter && a.pow2) {
        for (int n = threadIdx.x; n < W; n += blockDim.x) {
            T xa = load_pixel<T, Tin>(raw, lin, offa + n, a, rin_a, n);
            T xb = hasb ? load_pixel<T, Tin>(raw, lin, offb + n, a, rin_b, n) : (T)0;
            unsigned pos = bitrev((unsigned)n, a.log2W);
            re[pos] = xa;
            im[pos] = xb;
        }
        fft_dit<T>(re, im, W, a.log2W, tw);
        // multiply by the real even response and conjugate: IDFT(Z) = conj(DFT(conj Z))
        for (int k = threadIdx.x; k < W; k += blockDim.x) {
            T h = resp[k];
            re[k] = re[k] * h;
            im[k] = -im[k] * h;
        }
        fft_dif<T>(re, im, W, a.log2W, tw);
    } else if (a.do_filter) {
        T* g = im + W;
        for (int n = threadIdx.x; n < W; n += blockDim.x) {
            re[n] = load_pixel<T, Tin>(raw, lin, offa + n, a, rin_a, n);
            im[n] = hasb ? load_pixel<T, Tin>(raw, lin, offb + n, a, rin_b, n) : (T)0;
            g[n] = conv[n];
        }
        __syncthreads();
        // direct circular convolution; results go to registers, then back to LDS
        T ya[8], yb[8];
        int cnt = 0;
        for (int n = threadIdx.x; n < W && cnt < 8; n += blockDim.x, ++cnt) {
            T sa = 0, sb = 0;
            for (int m = 0; m <= n; ++m) { T gk = g[n - m]; sa += re[m] * gk; sb += im[m] * gk; }
            for (int m = n + 1; m < W; ++m) { T gk = g[n - m + W]; sa += re[m] * gk; sb += im[m] * gk; }
            ya[cnt] = sa;
            yb[cnt] = sb;
        }
        __syncthreads();
        cnt = 0;
        for (int n = threadIdx.x; n < W && cnt < 8; n += blockDim.x, ++cnt) {
            re[n] = ya[cnt];
            im[n] = -yb[cnt];  // read back with the same sign convention as the FFT path
        }
        __syncthreads();
    } else {
        for (int n = threadIdx.x; n < W; n += blockDim.x) {
            re[n] = load_pixel<T, Tin>(raw, lin, offa + n, a, rin_a, n);
            im[n] = hasb ? -load_pixel<T, Tin>(raw, lin, offb + n, a, rin_b, n) : (T)0;
        }
        __syncthreads();
    }

    const bool bitrev_out = a.do_filter && a.pow2;
    float mloc = 0.0f;
    for (int n = threadIdx.x; n < W; n += blockDim.x) {
        unsigned pos = bitrev_out ? bitrev((unsigned)n, a.log2W) : (unsigned)n;
        T ya = re[pos];
        T yb = -im[pos];
        out[offa + n] = ya;
        if (hasb) out[offb + n] = yb;
        if constexpr (sizeof(T) == 4) {
            mloc = fmaxf(mloc, fabsf((float)ya));
            if (hasb) mloc = fmaxf(mloc, fabsf((float)yb));
        }
    }
    if constexpr (sizeof(T) == 4) {
        if (maxabs != nullptr) {
            // rows a and b may belong to different images when H is odd
            const long long pa = ra / a.H, pb = hasb ? rb / a.H : pa;
            if (pa == pb) {
                float m = wave_max(mloc);
                if ((threadIdx.x & 63) == 0) atomicMax(&maxabs[pa], __float_as_uint(m));
            } else {
                float ma = 0.0f, mb = 0.0f;
                for (int n = threadIdx.x; n < W; n += blockDim.x) {
                    unsigned pos = bitrev_out ? bitrev((unsigned)n, a.log2W) : (unsigned)n;
                    ma = fmaxf(ma, fabsf((float)re[pos]));
                    mb = fmaxf(mb, fabsf((float)im[pos]));
                }
                ma = wave_max(ma);
                mb = wave_max(mb);
                if ((threadIdx.x & 63) == 0) {
                    atomicMax(&maxabs[pa], __float_as_uint(ma));
                    atomicMax(&maxabs[pb], __float_as_uint(mb));
                }
            }
        }
    }
}

// Row-major FP32 images -> backprojection layout.  Entry (u, v) of image p holds the
// bilinear coefficients of the cell whose top-left texel is (u, v):
//   A = q[v][u], dU = q[v][u+1] - A, dV = q[v+1][u] - A, dUV = q[v+1][u+1] - q[v+1][u] - dU
// (texels past the last column or row are 0), so that the interpolated value at
// (u + fu, v + fv) is A + fu*dU + fv*(dV + fu*dUV).  Entries are stored column-major
// (v fastest): a detector column is contiguous, and one voxel-update reads one entry.
__device__ inline uint32_t pack_f16(float a, float b) {
    __half ha = __float2half_rn(a), hb = __float2half_rn(b);
    return (uint32_t)__half_as_ushort(ha) | ((uint32_t)__half_as_ushort(hb) << 16);
}
__device__ inline uint32_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline uint32_t pack_bf16(float a, float b) { return f2bf(a) | (f2bf(b) << 16); }

template <int STORE>
__global__ __launch_bounds__(256) void layout_kernel(const float* __restrict__ rows, int W, int H,
                                                     const unsigned* __restrict__ maxabs,
                                                     void* __restrict__ out, float* __restrict__ scales) {
    __shared__ float tile[33][33];
    const int u0 = blockIdx.x * 32, v0 = blockIdx.y * 32;
    const size_t p = blockIdx.z;
    const float* img = rows + p * (size_t)W * H;
    for (int i = threadIdx.x; i < 33 * 33; i += 256) {
        int vl = i / 33, ul = i - vl * 33;
        int v = v0 + vl, u = u0 + ul;
        tile[vl][ul] = (v < H && u < W) ? img[(size_t)v * W + u] : 0.0f;
    }
    __syncthreads();
    float scale = 1.0f, inv = 1.0f;
    if (STORE != CTG_STORE_F32) {
        float m = __uint_as_float(maxabs[p]);
        if (m > 0.0f && m < INFINITY) {
            int e;
            frexpf(m, &e);
            scale = ldexpf(1.0f, e);
            inv = ldexpf(1.0f, -e);
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && scales != nullptr) scales[p] = scale;
    const int vl = threadIdx.x & 31;
    for (int ul = threadIdx.x >> 5; ul < 32; ul += 8) {
        int u = u0 + ul, v = v0 + vl;
        if (u < W && v < H) {
            const float A = tile[vl][ul] * inv, B = tile[vl][ul + 1] * inv;
            const float C = tile[vl + 1][ul] * inv, D = tile[vl + 1][ul + 1] * inv;
            const float dU = B - A, dV = C - A, dUV = (D - C) - dU;
            size_t idx = (p * (size_t)W + u) * H + v;
            if (STORE == CTG_STORE_F32) {
                reinterpret_cast<float4*>(out)[idx] = make_float4(A, dU, dV, dUV);
            } else if (STORE == CTG_STORE_F16) {
                reinterpret_cast<uint2*>(out)[idx] = make_uint2(pack_f16(A, dU), pack_f16(dV, dUV));
            } else {
                reinterpret_cast<uint2*>(out)[idx] = make_uint2(pack_bf16(A, dU), pack_bf16(dV, dUV));
            }
        }
    }
}

int ilog2_exact(int W) {
    int l = 0;
    while ((1 << l) < W) ++l;
    return (1 << l) == W ? l : -1;
}

void fill_log_params(PreArgs& a, ctg_pixel_type pixel, float base, ctg_log_form form) {
    const double B = base;
    const double S = pixel == CTG_PIXEL_U8 ? 255.0 : (pixel == CTG_PIXEL_U16 ? 65535.0 : 1.0);
    if (form == CTG_LOG_CPU_FORM) {
        // convertTo32bit: integer images I/S + 1e-6*B (float images unchanged), then * 1/(B+1e-6*B)
        a.dla = (pixel == CTG_PIXEL_F32) ? 1.0 : 1.0 / S;
        a.dlb = (pixel == CTG_PIXEL_F32) ? 0.0 : 0.000001 * B;
        a.dlk = 1.0 / (B + 0.000001 * B);
    } else {
        // convertTo(alpha = (1-1e-6)/(S*B), beta = 1e-6), log, * -1
        a.dla = (1.0 - 0.000001) / (S * B);
        a.dlb = 0.000001;
        a.dlk = 1.0;
    }
    a.la = (float)a.dla;
    a.lb = (float)a.dlb;
    a.lk = (float)a.dlk;
}

template <typename T>
ctg_status launch_rows(ctg_context* ctx, const void* raw, ctg_pixel_type pixel, const float* lin,
                       long long nrows, const PreArgs& a, T* out, unsigned* maxabs, hipStream_t st) {
    const FilterTables& ft = ctx->ftab;
    const T* resp = sizeof(T) == 4 ? (const T*)ft.resp_f.ptr : (const T*)ft.resp_d.ptr;
    const Cx<T>* tw = sizeof(T) == 4 ? (const Cx<T>*)ft.tw_f.ptr : (const Cx<T>*)ft.tw_d.ptr;
    const T* conv = sizeof(T) == 4 ? (const T*)ft.conv_f.ptr : (const T*)ft.conv_d.ptr;
    size_t lds = (size_t)a.W * sizeof(T) * ((a.do_filter && !a.pow2) ? 3 : 2);
    if (lds > 160 * 1024) return fail(CTG_ERROR_UNSUPPORTED, "detector width %d exceeds the LDS row buffer", a.W);
    if (a.do_filter && !a.pow2 && a.W > 8 * kRowThreads)
        return fail(CTG_ERROR_UNSUPPORTED, "non power-of-two detector width %d > %d", a.W, 8 * kRowThreads);
    dim3 grid((unsigned)((nrows + 1) / 2));
    if (raw == nullptr) {
        hipLaunchKernelGGL((rows_kernel<T, float>), grid, dim3(kRowThreads), lds, st, a, (const float*)nullptr,
                           lin, resp, tw, conv, out, maxabs);
    } else if (pixel == CTG_PIXEL_U8) {
        hipLaunchKernelGGL((rows_kernel<T, uint8_t>), grid, dim3(kRowThreads), lds, st, a, (const uint8_t*)raw,
                           lin, resp, tw, conv, out, maxabs);
    } else if (pixel == CTG_PIXEL_U16) {
        hipLaunchKernelGGL((rows_kernel<T, uint16_t>), grid, dim3(kRowThreads), lds, st, a,
                           (const uint16_t*)raw, lin, resp, tw, conv, out, maxabs);
    } else {
        hipLaunchKernelGGL((rows_kernel<T, float>), grid, dim3(kRowThreads), lds, st, a, (const float*)raw,
                           lin, resp, tw, conv, out, maxabs);
    }
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

}  // namespace

ctg_status launch_preprocess_rows(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int nproj,
                                  float base, ctg_log_form form, ctg_filter filter, float* d_out,
                                  unsigned* d_maxabs, bool do_log, bool do_weight, bool do_filter,
                                  hipStream_t st) {
    const ctg_geometry& g = ctx->geom;
    PreArgs a{};
    a.W = g.det_width;
    a.H = g.det_height;
    a.log2W = ilog2_exact(a.W);
    a.pow2 = a.log2W >= 0;
    a.nrows = (long long)nproj * a.H;
    fill_log_params(a, pixel, base, form);
    a.fcd = g.fcd;
    a.upre = g.u_pre;
    a.vpre = g.v_pre;
    a.do_log = do_log;
    a.do_weight = do_weight;
    a.do_filter = do_filter;
    if (do_filter) CTG_TRY(ensure_filter_tables(ctx, a.W, filter));
    return launch_rows<float>(ctx, do_log ? d_raw : nullptr, pixel, do_log ? nullptr : (const float*)d_raw,
                              a.nrows, a, d_out, d_maxabs, st);
}

ctg_status launch_filter_rows_only(ctg_context* ctx, const float* d_in, float* d_out, int64_t rows,
                                   ctg_filter filter, hipStream_t st) {
    const ctg_geometry& g = ctx->geom;
    PreArgs a{};
    a.W = g.det_width;
    a.H = g.det_height > 0 ? g.det_height : 1;
    a.log2W = ilog2_exact(a.W);
    a.pow2 = a.log2W >= 0;
    a.nrows = rows;
    a.do_filter = 1;
    CTG_TRY(ensure_filter_tables(ctx, a.W, filter));
    return launch_rows<float>(ctx, nullptr, CTG_PIXEL_F32, d_in, rows, a, d_out, nullptr, st);
}

ctg_status launch_preprocess_rows_f64(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int nproj,
                                      float base, ctg_filter filter, double* d_out, hipStream_t st) {
    const ctg_geometry& g = ctx->geom;
    PreArgs a{};
    a.W = g.det_width;
    a.H = g.det_height;
    a.log2W = ilog2_exact(a.W);
    a.pow2 = a.log2W >= 0;
    a.nrows = (long long)nproj * a.H;
    fill_log_params(a, pixel, base, CTG_LOG_CPU_FORM);
    a.fcd = g.fcd;
    a.upre = g.u_pre;
    a.vpre = g.v_pre;
    a.do_log = 1;
    a.do_weight = 1;
    a.do_filter = 1;
    CTG_TRY(ensure_filter_tables(ctx, a.W, filter));
    return launch_rows<double>(ctx, d_raw, pixel, nullptr, a.nrows, a, d_out, nullptr, st);
}

ctg_status launch_layout(const ctg_geometry& g, const float* d_rows, int nproj, ctg_store store,
                         const unsigned* d_maxabs, void* d_out, float* d_scales, hipStream_t st) {
    dim3 grid((unsigned)((g.det_width + 31) / 32), (unsigned)((g.det_height + 31) / 32), (unsigned)nproj);
    if (store == CTG_STORE_F32) {
        hipLaunchKernelGGL((layout_kernel<CTG_STORE_F32>), grid, dim3(256), 0, st, d_rows, g.det_width,
                           g.det_height, d_maxabs, d_out, d_scales);
    } else if (store == CTG_STORE_F16) {
        hipLaunchKernelGGL((layout_kernel<CTG_STORE_F16>), grid, dim3(256), 0, st, d_rows, g.det_width,
                           g.det_height, d_maxabs, d_out, d_scales);
    } else {
        hipLaunchKernelGGL((layout_kernel<CTG_STORE_BF16>), grid, dim3(256), 0, st, d_rows, g.det_width,
                           g.det_height, d_maxabs, d_out, d_scales);
    }
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

// Host-side tables in double: window H(k) with the float/double evaluation of
// Ct/CtVolume.cpp:868-885 (N = W/2+1), its real even extension, twiddles
// exp(-2 pi i m / W), and the circular-convolution kernel for other W.
ctg_status ensure_filter_tables(ctg_context* ctx, int W, int filter) {
    FilterTables& ft = ctx->ftab;
    if (ft.W == W && ft.filter == filter && ft.resp_f.ptr) return CTG_OK;
    const int nyq = W / 2 + 1;
    std::vector<double> h(nyq);
    const float N = (float)nyq;
    for (int k = 0; k < nyq; ++k) {
        float n = (float)k;
        float rl = n / N;
        float hv;
        if (filter == CTG_FILTER_SHEPP_LOGAN) {
            if (n == 0.0f) hv = 0.0f;
            else {
                double ang = (double)(rl * 0.5f) * kPi;
                hv = (float)((double)rl * std::sin(ang) / ang);
            }
        } else if (filter == CTG_FILTER_HANN) {
            hv = (float)((double)rl * (double)0.5f * (1.0 + std::cos(kPi * (double)rl)));
        } else {
            hv = rl;
        }
        h[k] = hv;
    }
    std::vector<double> resp(W), conv(W);
    for (int k = 0; k < W; ++k) resp[k] = h[k <= W / 2 ? k : W - k];
    std::vector<double> tw(2 * (size_t)(W / 2 > 0 ? W / 2 : 1));
    for (int m = 0; m < W / 2; ++m) {
        double ang = 2.0 * kPi * (double)m / (double)W;
        tw[2 * m] = std::cos(ang);
        tw[2 * m + 1] = -std::sin(ang);
    }
    if (ilog2_exact(W) < 0) {
        for (int m = 0; m < W; ++m) {
            double s = 0.0;
            for (int k = 0; k < W; ++k)
                s += resp[k] * std::cos(2.0 * kPi * (double)(((long long)k * m) % W) / (double)W);
            conv[m] = s;
        }
    }
    std::vector<float> resp_f(W), conv_f(W), tw_f(tw.size());
    for (int i = 0; i < W; ++i) { resp_f[i] = (float)resp[i]; conv_f[i] = (float)conv[i]; }
    for (size_t i = 0; i < tw.size(); ++i) tw_f[i] = (float)tw[i];
    CTG_TRY(ft.resp_f.reserve(sizeof(float) * W));
    CTG_TRY(ft.conv_f.reserve(sizeof(float) * W));
    CTG_TRY(ft.tw_f.reserve(sizeof(float) * tw_f.size()));
    CTG_TRY(ft.resp_d.reserve(sizeof(double) * W));
    CTG_TRY(ft.conv_d.reserve(sizeof(double) * W));
    CTG_TRY(ft.tw_d.reserve(sizeof(double) * tw.size()));
    CTG_HIP(hipMemcpy(ft.resp_f.ptr, resp_f.data(), sizeof(float) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.conv_f.ptr, conv_f.data(), sizeof(float) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.tw_f.ptr, tw_f.data(), sizeof(float) * tw_f.size(), hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.resp_d.ptr, resp.data(), sizeof(double) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.conv_d.ptr, conv.data(), sizeof(double) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.tw_d.ptr, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    ft.W = W;
    ft.filter = filter;
    return CTG_OK;
}

}  // namespace ctg
