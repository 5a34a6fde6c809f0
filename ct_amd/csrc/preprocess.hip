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
// Design: two rows are filtered at once as the real and imaginary parts of one complex
// sequence (the response is real and even, so real rows stay real).
//  - W = 256..2048 (FP32): one wave per row pair, values in registers, Stockham
//    radix-16/8/4 passes exchanging through the wave's LDS region (rows_fft_kernel).
//  - Other powers of two, and the FP64 yardstick: one workgroup per row pair, the pair
//    in LDS, a radix-2 DIT forward FFT (bit-reversed load), the response and a
//    conjugation, then a DIF FFT whose bit-reversed output is read back in natural order.
//  - Other W: the row's circular convolution with the filter kernel g (length W) is
//    the middle of a linear convolution with g extended periodically to 2W - 1 taps,
//    done in a power-of-two transform of length M >= 2W - 1 with the LDS radix-2 pair
//    (no wrap-around reaches the W outputs kept).
#include "ctg_internal.h"

#include <hip/hip_fp16.h>

#include <cmath>
#include <complex>
#include <vector>

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
    int M, log2M;  // transform length: W for powers of two, else the padded length
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
                                                          const Cx<T>* __restrict__ pad_resp,
                                                          T* __restrict__ out,
                                                          unsigned* __restrict__ maxabs) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* re = reinterpret_cast<T*>(smem_raw);
    T* im = re + a.M;
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
        // y[n] = sum_m x[m] g[(n - m) mod W] = (x * h)[n + W - 1] with h[j] = g[(j - W + 1)
        // mod W], j < 2W - 1: x zero-padded to M, times the M-point transform of h (over M,
        // precomputed), the same conjugate trick, outputs read at n + W - 1
        const int M = a.M;
        for (int n = threadIdx.x; n < M; n += blockDim.x) {
            T xa = (T)0, xb = (T)0;
            if (n < W) {
                xa = load_pixel<T, Tin>(raw, lin, offa + n, a, rin_a, n);
                if (hasb) xb = load_pixel<T, Tin>(raw, lin, offb + n, a, rin_b, n);
            }
            unsigned pos = bitrev((unsigned)n, a.log2M);
            re[pos] = xa;
            im[pos] = xb;
        }
        fft_dit<T>(re, im, M, a.log2M, tw);
        for (int k = threadIdx.x; k < M; k += blockDim.x) {
            const Cx<T> h = pad_resp[k];
            const T xr = re[k], xi = im[k];
            re[k] = xr * h.re - xi * h.im;
            im[k] = -(xr * h.im + xi * h.re);
        }
        fft_dif<T>(re, im, M, a.log2M, tw);
    } else {
        for (int n = threadIdx.x; n < W; n += blockDim.x) {
            re[n] = load_pixel<T, Tin>(raw, lin, offa + n, a, rin_a, n);
            im[n] = hasb ? -load_pixel<T, Tin>(raw, lin, offb + n, a, rin_b, n) : (T)0;
        }
        __syncthreads();
    }

    // filtered outputs sit in bit-reversed order (at n + W - 1 for padded widths)
    const int out_shift = a.pow2 ? 0 : W - 1;
    auto out_pos = [&](int n) -> unsigned {
        return a.do_filter ? bitrev((unsigned)(n + out_shift), a.log2M) : (unsigned)n;
    };
    float mloc = 0.0f;
    for (int n = threadIdx.x; n < W; n += blockDim.x) {
        unsigned pos = out_pos(n);
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
                    unsigned pos = out_pos(n);
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

// ---------------------------------------------------------------------------------
// Register-resident row filter for W = 256, 512, 1024, 2048 (FP32).
//
// One wave owns one row pair (rows a and b as the real and imaginary parts), each lane
// holds P = W/64 complex values, and the transform is a Stockham autosort FFT of 3-4
// passes with radix 16/8/4 butterflies evaluated in registers; between passes the wave
// exchanges values through its own LDS region (no workgroup barrier: a wave's LDS
// operations execute in order).  Pass i, radix R, stride NS: butterfly j reads
// x[j + r*W/R], multiplies by exp(-2 pi i (j mod NS) r / (NS R)) and writes the R-point
// DFT to y[(j / NS) NS R + j mod NS + r NS]; the output ends in natural order.  The
// filtered rows are the conjugate of the forward transform of conj(X * H), so the same
// forward passes run twice; the first pass reads the raw pixels from global memory and
// the last writes the filtered rows.
namespace fftw {

__host__ __device__ constexpr int passes(int N) { return N == 256 ? 4 : 3; }
__host__ __device__ constexpr int radix(int N, int i) {
    return N == 256 ? 4 : (N == 512 ? 8 : (i < 2 ? 16 : (N == 1024 ? 4 : 8)));
}
__host__ __device__ constexpr int stride(int N, int i) { return i == 0 ? 1 : stride(N, i - 1) * radix(N, i - 1); }
// LDS position of element i: one float2 of padding per 16 spreads the strided writes
__device__ inline int lds_pos(int i) { return i + (i >> 4); }

struct F2 {
    float x, y;
};
__device__ inline F2 add(F2 a, F2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline F2 sub(F2 a, F2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline F2 mul(F2 a, F2 w) { return {fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x)}; }
__device__ inline F2 mul_mi(F2 a) { return {a.y, -a.x}; }  // a * (-i)

// in-place forward DFTs (kernel exp(-2 pi i n k / R)), natural order in and out
__device__ inline void dft4(F2& a0, F2& a1, F2& a2, F2& a3) {
    const F2 t0 = add(a0, a2), t1 = sub(a0, a2), t2 = add(a1, a3), t3 = mul_mi(sub(a1, a3));
    a0 = add(t0, t2);
    a2 = sub(t0, t2);
    a1 = add(t1, t3);
    a3 = sub(t1, t3);
}
__device__ inline void dft8(F2* a) {
    constexpr float h = 0.70710678118654752440f;
    F2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    F2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = mul(o1, F2{h, -h});
    o2 = mul_mi(o2);
    o3 = mul(o3, F2{-h, -h});
    a[0] = add(e0, o0); a[4] = sub(e0, o0);
    a[1] = add(e1, o1); a[5] = sub(e1, o1);
    a[2] = add(e2, o2); a[6] = sub(e2, o2);
    a[3] = add(e3, o3); a[7] = sub(e3, o3);
}
__device__ inline void dft16(F2* a) {
    // 16 = 4 x 4: DFT4 over n_a of a[4 n_a + n_b], twiddle exp(-2 pi i n_b k_a / 16),
    // DFT4 over n_b; X[k_a + 4 k_b]
    constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
    F2 t[4][4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        t[nb][0] = a[nb]; t[nb][1] = a[4 + nb]; t[nb][2] = a[8 + nb]; t[nb][3] = a[12 + nb];
        dft4(t[nb][0], t[nb][1], t[nb][2], t[nb][3]);
    }
    t[1][1] = mul(t[1][1], F2{c1, -s1});
    t[1][2] = mul(t[1][2], F2{h, -h});
    t[1][3] = mul(t[1][3], F2{s1, -c1});
    t[2][1] = mul(t[2][1], F2{h, -h});
    t[2][2] = mul_mi(t[2][2]);
    t[2][3] = mul(t[2][3], F2{-h, -h});
    t[3][1] = mul(t[3][1], F2{s1, -c1});
    t[3][2] = mul(t[3][2], F2{-h, -h});
    t[3][3] = mul(t[3][3], F2{-c1, s1});
#pragma unroll
    for (int ka = 0; ka < 4; ++ka) {
        F2 y0 = t[0][ka], y1 = t[1][ka], y2 = t[2][ka], y3 = t[3][ka];
        dft4(y0, y1, y2, y3);
        a[ka] = y0; a[ka + 4] = y1; a[ka + 8] = y2; a[ka + 12] = y3;
    }
}
template <int R>
__device__ inline void dft(F2* a) {
    if constexpr (R == 4) dft4(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 8) dft8(a);
    else dft16(a);
}

// exp(-2 pi i k / N) for 0 <= k < N from the half table exp(-2 pi i m / N), m < N/2
template <int N>
__device__ inline F2 twiddle(const Cx<float>* __restrict__ tw, int k) {
    const Cx<float> t = tw[k & (N / 2 - 1)];
    return (k & (N / 2)) ? F2{-t.re, -t.im} : F2{t.re, t.im};
}

// orders the wave's LDS stores and loads around a cross-lane exchange
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int I>
__device__ inline void load_pass(const F2* __restrict__ buf, F2 (&v)[N / 64], int lane) {
    constexpr int R = radix(N, I), B = N / 64 / R;
#pragma unroll
    for (int b = 0; b < B; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = buf[lds_pos(lane + 64 * b + r * (N / R))];
}

template <int N, int I>
__device__ inline void store_pass(F2* __restrict__ buf, const F2 (&v)[N / 64], int lane) {
    constexpr int R = radix(N, I), NS = stride(N, I), B = N / 64 / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = lane + 64 * b;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) buf[lds_pos(base + r * NS)] = v[b * R + r];
    }
}

// Passes I.. of one forward transform.  On entry v holds pass I's inputs; on return
// v holds the last pass's outputs: value b*R + r is X[lane + 64 b + r * NS_last].
template <int N, int I>
__device__ inline void passes_from(F2 (&v)[N / 64], F2* __restrict__ buf, int lane,
                                   const Cx<float>* __restrict__ tw) {
    constexpr int R = radix(N, I), NS = stride(N, I), B = N / 64 / R;
    // an opaque zero offset per pass: the twiddles are then loaded where they are used
    // instead of being hoisted and kept live across both transforms
    int zero = 0;
    asm volatile("" : "+s"(zero));
    const Cx<float>* twp = tw + zero;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        if constexpr (NS > 1) {
            const int jm = (lane + 64 * b) % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) v[b * R + r] = mul(v[b * R + r], twiddle<N>(twp, jm * r * (N / (NS * R))));
        }
        dft<R>(v + b * R);
    }
    if constexpr (I + 1 < passes(N)) {
        store_pass<N, I>(buf, v, lane);
        wave_sync();
        load_pass<N, I + 1>(buf, v, lane);
        wave_sync();
        passes_from<N, I + 1>(v, buf, lane, tw);
    }
}

}  // namespace fftw

// row pairs (waves) per workgroup; LDS is (N + N/16) float2 per wave
__host__ __device__ constexpr int fft_waves(int N) { return N >= 2048 ? 2 : 4; }

template <int N, typename Tin>
__global__ __launch_bounds__(64 * fft_waves(N)) void rows_fft_kernel(PreArgs a, const Tin* __restrict__ raw,
                                                                  const float* __restrict__ lin,
                                                                  const float* __restrict__ resp,
                                                                  const Cx<float>* __restrict__ tw,
                                                                  float* __restrict__ out,
                                                                  unsigned* __restrict__ maxabs) {
    using namespace fftw;
    constexpr int P = N / 64;
    constexpr int LAST = passes(N) - 1;
    constexpr int R0 = radix(N, 0), RL = radix(N, LAST), NSL = stride(N, LAST);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    F2* buf = reinterpret_cast<F2*>(smem_raw) + (size_t)wave * (N + N / 16);
    const long long pair = (long long)blockIdx.x * fft_waves(N) + wave;
    const long long ra = 2 * pair, rb = ra + 1;
    if (ra >= a.nrows) return;  // waves share nothing, so a finished wave may leave
    const bool hasb = rb < a.nrows;
    const int rin_a = (int)(ra % a.H), rin_b = (int)(rb % a.H);
    const size_t offa = (size_t)ra * N, offb = (size_t)rb * N;

    F2 v[P];
    // pass 0 inputs straight from the images: x[lane + 64 b + r N/R0].  All loads are
    // issued first (a missing row b re-reads row a and is zeroed below); the uniform
    // log and weight flags are tested once, outside the per-pixel loops.
    const size_t offb_ld = hasb ? offb : offa;
    if (raw != nullptr) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int n = lane + 64 * (i / R0) + (i % R0) * (N / R0);
            v[i].x = (float)raw[offa + n];
            v[i].y = (float)raw[offb_ld + n];
        }
        if (a.do_log) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                v[i].x = -logf(fmaf(v[i].x, a.la, a.lb) * a.lk);
                v[i].y = -logf(fmaf(v[i].y, a.la, a.lb) * a.lk);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int n = lane + 64 * (i / R0) + (i % R0) * (N / R0);
            v[i].x = lin[offa + n];
            v[i].y = lin[offb_ld + n];
        }
    }
    if (a.do_weight) {
        // D / sqrt(D^2 + u^2 + v^2), u = column - uPre, v = vPre - row
        const float D = a.fcd;
        const float va = (float)(-rin_a) + a.vpre, vb = (float)(-rin_b) + a.vpre;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int n = lane + 64 * (i / R0) + (i % R0) * (N / R0);
            const float u = (float)n - a.upre;
            const float du = D * D + u * u;
            v[i].x = v[i].x * (D / sqrtf(du + va * va));
            v[i].y = v[i].y * (D / sqrtf(du + vb * vb));
        }
    }
    if (!hasb) {
#pragma unroll
        for (int i = 0; i < P; ++i) v[i].y = 0.0f;
    }
    passes_from<N, 0>(v, buf, lane, tw);
    // multiply by the real even response and conjugate; back to pass 0's distribution
#pragma unroll
    for (int b = 0; b < P / RL; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const int k = lane + 64 * b + r * NSL;
            const float h = resp[k];
            const F2 z = v[b * RL + r];
            buf[lds_pos(k)] = F2{z.x * h, -(z.y * h)};
        }
    wave_sync();
    load_pass<N, 0>(buf, v, lane);
    wave_sync();
    passes_from<N, 0>(v, buf, lane, tw);

    float mloc = 0.0f, ma = 0.0f, mb = 0.0f;
#pragma unroll
    for (int b = 0; b < P / RL; ++b)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
            const int k = lane + 64 * b + r * NSL;
            const float ya = v[b * RL + r].x, yb = -v[b * RL + r].y;
            out[offa + k] = ya;
            if (hasb) out[offb + k] = yb;
            ma = fmaxf(ma, fabsf(ya));
            mb = fmaxf(mb, fabsf(yb));
        }
    if (maxabs != nullptr) {
        const long long pa = ra / a.H, pb = hasb ? rb / a.H : pa;
        if (pa == pb) {
            mloc = wave_max(fmaxf(ma, hasb ? mb : 0.0f));
            if (lane == 0) atomicMax(&maxabs[pa], __float_as_uint(mloc));
        } else {
            ma = wave_max(ma);
            mb = wave_max(mb);
            if (lane == 0) {
                atomicMax(&maxabs[pa], __float_as_uint(ma));
                atomicMax(&maxabs[pb], __float_as_uint(mb));
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

// Transform length of the row filter: W itself for powers of two, otherwise the
// smallest power of two >= 2W - 1 (the padded linear convolution).
int filter_log2_length(int W) {
    const int l = ilog2_exact(W);
    if (l >= 0) return l;
    int m = 0;
    while ((1 << m) < 2 * W - 1) ++m;
    return m;
}

void set_transform(PreArgs& a) {
    a.log2W = ilog2_exact(a.W);
    a.pow2 = a.log2W >= 0;
    if (a.do_filter) {
        a.log2M = filter_log2_length(a.W);
        a.M = 1 << a.log2M;
    } else {
        a.log2M = a.log2W;
        a.M = a.W;
    }
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

template <int N>
ctg_status launch_rows_fft(const void* raw, ctg_pixel_type pixel, const float* lin, long long nrows,
                           const PreArgs& a, const float* resp, const Cx<float>* tw, float* out,
                           unsigned* maxabs, hipStream_t st) {
    constexpr int waves = fft_waves(N);
    const long long pairs = (nrows + 1) / 2;
    const long long blocks = (pairs + waves - 1) / waves;
    if (blocks > 0x7fffffffll) return fail(CTG_ERROR_UNSUPPORTED, "too many detector rows in one call");
    const size_t lds = (size_t)waves * (N + N / 16) * sizeof(fftw::F2);
    const dim3 grid((unsigned)blocks), block(64 * waves);
    if (raw == nullptr) {
        hipLaunchKernelGGL((rows_fft_kernel<N, float>), grid, block, lds, st, a, (const float*)nullptr, lin, resp,
                           tw, out, maxabs);
    } else if (pixel == CTG_PIXEL_U8) {
        hipLaunchKernelGGL((rows_fft_kernel<N, uint8_t>), grid, block, lds, st, a, (const uint8_t*)raw, lin, resp,
                           tw, out, maxabs);
    } else if (pixel == CTG_PIXEL_U16) {
        hipLaunchKernelGGL((rows_fft_kernel<N, uint16_t>), grid, block, lds, st, a, (const uint16_t*)raw, lin,
                           resp, tw, out, maxabs);
    } else {
        hipLaunchKernelGGL((rows_fft_kernel<N, float>), grid, block, lds, st, a, (const float*)raw, lin, resp, tw,
                           out, maxabs);
    }
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

template <typename T>
ctg_status launch_rows(ctg_context* ctx, const void* raw, ctg_pixel_type pixel, const float* lin,
                       long long nrows, const PreArgs& a, T* out, unsigned* maxabs, hipStream_t st) {
    const FilterTables& ft = ctx->ftab;
    const T* resp = sizeof(T) == 4 ? (const T*)ft.resp_f.ptr : (const T*)ft.resp_d.ptr;
    const Cx<T>* tw = sizeof(T) == 4 ? (const Cx<T>*)ft.tw_f.ptr : (const Cx<T>*)ft.tw_d.ptr;
    const Cx<T>* pad = sizeof(T) == 4 ? (const Cx<T>*)ft.pad_f.ptr : (const Cx<T>*)ft.pad_d.ptr;
    const size_t lds = (size_t)a.M * sizeof(T) * 2;
    if (lds > 160 * 1024)
        return fail(CTG_ERROR_UNSUPPORTED, "detector width %d (transform length %d) exceeds the LDS row buffer", a.W, a.M);
    if constexpr (sizeof(T) == 4) {
        if (a.do_filter && a.pow2 && a.W >= 256 && a.W <= 2048) {
            switch (a.W) {
                case 256: return launch_rows_fft<256>(raw, pixel, lin, nrows, a, resp, tw, out, maxabs, st);
                case 512: return launch_rows_fft<512>(raw, pixel, lin, nrows, a, resp, tw, out, maxabs, st);
                case 1024: return launch_rows_fft<1024>(raw, pixel, lin, nrows, a, resp, tw, out, maxabs, st);
                default: return launch_rows_fft<2048>(raw, pixel, lin, nrows, a, resp, tw, out, maxabs, st);
            }
        }
    }
    dim3 grid((unsigned)((nrows + 1) / 2));
    if (raw == nullptr) {
        hipLaunchKernelGGL((rows_kernel<T, float>), grid, dim3(kRowThreads), lds, st, a, (const float*)nullptr,
                           lin, resp, tw, pad, out, maxabs);
    } else if (pixel == CTG_PIXEL_U8) {
        hipLaunchKernelGGL((rows_kernel<T, uint8_t>), grid, dim3(kRowThreads), lds, st, a, (const uint8_t*)raw,
                           lin, resp, tw, pad, out, maxabs);
    } else if (pixel == CTG_PIXEL_U16) {
        hipLaunchKernelGGL((rows_kernel<T, uint16_t>), grid, dim3(kRowThreads), lds, st, a,
                           (const uint16_t*)raw, lin, resp, tw, pad, out, maxabs);
    } else {
        hipLaunchKernelGGL((rows_kernel<T, float>), grid, dim3(kRowThreads), lds, st, a, (const float*)raw,
                           lin, resp, tw, pad, out, maxabs);
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
    a.nrows = (long long)nproj * a.H;
    fill_log_params(a, pixel, base, form);
    a.fcd = g.fcd;
    a.upre = g.u_pre;
    a.vpre = g.v_pre;
    a.do_log = do_log;
    a.do_weight = do_weight;
    a.do_filter = do_filter;
    set_transform(a);
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
    a.nrows = rows;
    a.do_filter = 1;
    set_transform(a);
    CTG_TRY(ensure_filter_tables(ctx, a.W, filter));
    return launch_rows<float>(ctx, nullptr, CTG_PIXEL_F32, d_in, rows, a, d_out, nullptr, st);
}

ctg_status launch_preprocess_rows_f64(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int nproj,
                                      float base, ctg_filter filter, double* d_out, hipStream_t st) {
    const ctg_geometry& g = ctx->geom;
    PreArgs a{};
    a.W = g.det_width;
    a.H = g.det_height;
    a.nrows = (long long)nproj * a.H;
    fill_log_params(a, pixel, base, CTG_LOG_CPU_FORM);
    a.fcd = g.fcd;
    a.upre = g.u_pre;
    a.vpre = g.v_pre;
    a.do_log = 1;
    a.do_weight = 1;
    a.do_filter = 1;
    set_transform(a);
    CTG_TRY(ensure_filter_tables(ctx, a.W, filter));
    return launch_rows<double>(ctx, d_raw, pixel, nullptr, a.nrows, a, d_out, nullptr, st);
}

ctg_status launch_layout(const ctg_geometry& g, const float* d_rows, int nproj, ctg_store store,
                         const unsigned* d_maxabs, void* d_out, float* d_scales, hipStream_t st) {
    // one grid z index per projection; the grid's z extent is limited to 65535
    const size_t img = (size_t)g.det_width * g.det_height;
    const size_t ebytes = store == CTG_STORE_F32 ? 16 : 8;
    for (int p0 = 0; p0 < nproj; p0 += 65535) {
        const int np = std::min(65535, nproj - p0);
        dim3 grid((unsigned)((g.det_width + 31) / 32), (unsigned)((g.det_height + 31) / 32), (unsigned)np);
        const float* rows = d_rows + (size_t)p0 * img;
        const unsigned* mx = d_maxabs + p0;
        void* out = (char*)d_out + (size_t)p0 * img * ebytes;
        float* sc = d_scales ? d_scales + p0 : nullptr;
        if (store == CTG_STORE_F32) {
            hipLaunchKernelGGL((layout_kernel<CTG_STORE_F32>), grid, dim3(256), 0, st, rows, g.det_width,
                               g.det_height, mx, out, sc);
        } else if (store == CTG_STORE_F16) {
            hipLaunchKernelGGL((layout_kernel<CTG_STORE_F16>), grid, dim3(256), 0, st, rows, g.det_width,
                               g.det_height, mx, out, sc);
        } else {
            hipLaunchKernelGGL((layout_kernel<CTG_STORE_BF16>), grid, dim3(256), 0, st, rows, g.det_width,
                               g.det_height, mx, out, sc);
        }
        CTG_HIP(hipGetLastError());
    }
    return CTG_OK;
}

// Host-side tables in double: window H(k) with the float/double evaluation of
// Ct/CtVolume.cpp:868-885 (N = W/2+1) and its real even extension; twiddles
// exp(-2 pi i m / M) for the transform length M; for other W the circular-convolution
// kernel g[m] = sum_k H(k) cos(2 pi k m / W), extended periodically to 2W - 1 taps,
// and its M-point transform over M.
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
    std::vector<double> resp(W);
    for (int k = 0; k < W; ++k) resp[k] = h[k <= W / 2 ? k : W - k];
    const bool pow2 = ilog2_exact(W) >= 0;
    const int log2M = filter_log2_length(W), M = 1 << log2M;
    std::vector<double> tw(2 * (size_t)(M / 2 > 0 ? M / 2 : 1));
    for (int m = 0; m < M / 2; ++m) {
        double ang = 2.0 * kPi * (double)m / (double)M;
        tw[2 * m] = std::cos(ang);
        tw[2 * m + 1] = -std::sin(ang);
    }
    std::vector<double> pad(2 * (size_t)M, 0.0);  // complex, interleaved
    if (!pow2) {
        std::vector<double> cosw(W), g(W);
        for (int j = 0; j < W; ++j) cosw[j] = std::cos(2.0 * kPi * (double)j / (double)W);
        for (int m = 0; m < W; ++m) {
            double s = 0.0;
            for (int k = 0; k < W; ++k) s += resp[k] * cosw[(int)(((long long)k * m) % W)];
            g[m] = s;
        }
        std::vector<std::complex<double>> z(M);
        for (int j = 0; j < 2 * W - 1; ++j) z[j] = g[((j - (W - 1)) % W + W) % W];
        // iterative radix-2 transform exp(-2 pi i jk / M), over M
        for (int i = 1, r = 0; i < M; ++i) {
            int bit = M >> 1;
            for (; r & bit; bit >>= 1) r ^= bit;
            r ^= bit;
            if (i < r) std::swap(z[i], z[r]);
        }
        for (int len = 2; len <= M; len <<= 1) {
            for (int i0 = 0; i0 < M; i0 += len) {
                for (int k = 0; k < len / 2; ++k) {
                    const int t = k * (M / len);
                    const std::complex<double> w(tw[2 * t], tw[2 * t + 1]);
                    const std::complex<double> u = z[i0 + k], v = z[i0 + k + len / 2] * w;
                    z[i0 + k] = u + v;
                    z[i0 + k + len / 2] = u - v;
                }
            }
        }
        for (int k = 0; k < M; ++k) {
            pad[2 * (size_t)k] = z[k].real() / (double)M;
            pad[2 * (size_t)k + 1] = z[k].imag() / (double)M;
        }
    }
    std::vector<float> resp_f(W), tw_f(tw.size()), pad_f(pad.size());
    for (int i = 0; i < W; ++i) resp_f[i] = (float)resp[i];
    for (size_t i = 0; i < tw.size(); ++i) tw_f[i] = (float)tw[i];
    for (size_t i = 0; i < pad.size(); ++i) pad_f[i] = (float)pad[i];
    CTG_TRY(ft.resp_f.reserve(sizeof(float) * W));
    CTG_TRY(ft.tw_f.reserve(sizeof(float) * tw_f.size()));
    CTG_TRY(ft.pad_f.reserve(sizeof(float) * pad_f.size()));
    CTG_TRY(ft.resp_d.reserve(sizeof(double) * W));
    CTG_TRY(ft.tw_d.reserve(sizeof(double) * tw.size()));
    CTG_TRY(ft.pad_d.reserve(sizeof(double) * pad.size()));
    CTG_HIP(hipMemcpy(ft.resp_f.ptr, resp_f.data(), sizeof(float) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.tw_f.ptr, tw_f.data(), sizeof(float) * tw_f.size(), hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.pad_f.ptr, pad_f.data(), sizeof(float) * pad_f.size(), hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.resp_d.ptr, resp.data(), sizeof(double) * W, hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.tw_d.ptr, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    CTG_HIP(hipMemcpy(ft.pad_d.ptr, pad.data(), sizeof(double) * pad.size(), hipMemcpyHostToDevice));
    ft.W = W;
    ft.filter = filter;
    return CTG_OK;
}

}  // namespace ctg
