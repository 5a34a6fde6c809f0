// Voxel-driven FDK backprojection for gfx950.
//
// Per voxel and projection the arithmetic is the reference's (Ct/CtVolume.cpp:979-1039,
// Ct/cuda_ct.cu:234-283): cylinder test x^2+y^2 < R^2; t = -x sin + y cos + uOffset,
// s = x cos + y sin, U = FCD/(FCD-s), u = t U; window test on u; v = (z+h) U and window
// test on v; uIndex = u + uPre, vIndex = vPre - v; bilinear tap of the filtered
// projection; vol += U^2 * value.
//
// Organisation (MI355X-first, not the reference's one-launch-per-projection):
//  - one launch covers many projections; every lane owns one (x,y) voxel column and a
//    run of ZPL consecutive z, accumulating in registers across all projections, so the
//    volume slab is written once per launch (z-fastest, each lane stores ZPL*4 bytes);
//  - a workgroup (256 lanes = 16x16 columns) stages, per projection, the bounding box of
//    its voxel block's detector footprint into LDS (register prefetch of projection p+1
//    overlaps the compute of p).  The projection is stored as horizontal tap pairs
//    (q[v][u], q[v][u+1]) with v fastest, so one lane's four taps are two adjacent LDS
//    entries (one ds_read2);
//  - per z the lane does one fma for vIndex, floor/fract, one address add, the two-entry
//    LDS read, and 7 VALU for the interpolation and accumulation; the per-column geometry
//    (a division) is amortised over the z-run;
//  - boxes are computed by wave 0 for 64 projections at a time (one lane per projection);
//    a projection whose box exceeds the tile capacity falls back to reading the pairs
//    straight from global memory, so any geometry stays correct;
//  - workgroups are numbered so that each XCD (dispatch is round-robin over 8 XCDs)
//    receives a contiguous range of tiles and shares detector rows in its own L2.
#include "ctg_internal.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace ctg {
namespace {

constexpr int kTX = 16, kTY = 16, kThreads = 256;
constexpr int kMaxE = 10;  // staged entries per thread: tile capacity kMaxE * 256

struct BpArgs {
    int W, H, nx, ny, nzs, z0;
    float xpre, ypre, zpre, upre, vpre;
    float r2, fcd, uoff, ulo, uhi, vlo, vhi;
    int nproj, ntx, nty, nzb, cap, accumulate, vec4;
};

struct EF32 {
    using T = float2;
    __device__ static T zero() { return make_float2(0.0f, 0.0f); }
    __device__ static void taps(T e0, T e1, float& A, float& B, float& C, float& D) {
        A = e0.x; B = e0.y; C = e1.x; D = e1.y;
    }
};
struct EF16 {
    using T = uint32_t;
    __device__ static T zero() { return 0u; }
    __device__ static float lo(uint32_t e) { return __half2float(__ushort_as_half((unsigned short)(e & 0xffffu))); }
    __device__ static float hi(uint32_t e) { return __half2float(__ushort_as_half((unsigned short)(e >> 16))); }
    __device__ static void taps(T e0, T e1, float& A, float& B, float& C, float& D) {
        A = lo(e0); B = hi(e0); C = lo(e1); D = hi(e1);
    }
};
struct EBF16 {
    using T = uint32_t;
    __device__ static T zero() { return 0u; }
    __device__ static void taps(T e0, T e1, float& A, float& B, float& C, float& D) {
        A = __uint_as_float(e0 << 16); B = __uint_as_float(e0 & 0xffff0000u);
        C = __uint_as_float(e1 << 16); D = __uint_as_float(e1 & 0xffff0000u);
    }
};

__device__ inline float interp(float fu, float fv, float A, float B, float C, float D) {
    float top = fmaf(fu, B, fmaf(-fu, A, A));
    float bot = fmaf(fu, D, fmaf(-fu, C, C));
    return fmaf(fv, bot - top, top);
}

// box: x = first pair column, y = first row (may be -1: zero guard row), z = ncols | flag<<28,
// w = nrows.  flag 0: the block cannot hit the detector; 1: staged in LDS; 2: global reads.
__device__ inline int box_flag(int4 b) { return (unsigned)b.z >> 28; }
__device__ inline int box_nc(int4 b) { return b.z & 0x0fffffff; }

template <int ZPL, class E>
__global__ __launch_bounds__(kThreads) void bp_kernel(BpArgs a, const float4* __restrict__ pp,
                                                      const float* __restrict__ scales,
                                                      const typename E::T* __restrict__ proj,
                                                      float* __restrict__ vol) {
    using T = typename E::T;
    extern __shared__ __align__(16) unsigned char smem[];
    T* tile = reinterpret_cast<T*>(smem);
    __shared__ int4 boxes[2][64];

    // XCD-aware numbering: dispatch sends workgroup L to XCD L%8; give each XCD a
    // contiguous range of work items.
    const unsigned N = gridDim.x, L = blockIdx.x;
    const unsigned xcd = L & 7u, kk = L >> 3, q8 = N >> 3, r8 = N & 7u;
    const unsigned item = xcd * q8 + min(xcd, r8) + kk;
    const unsigned ntxy = (unsigned)(a.ntx * a.nty);
    const int zbi = (int)(item / ntxy);
    const int txy = (int)(item - (unsigned)zbi * ntxy);
    const int ty = txy / a.ntx, tx = txy - ty * a.ntx;
    const int tid = threadIdx.x;
    const int xi = tx * kTX + (tid & (kTX - 1));
    const int yi = ty * kTY + (tid / kTX);
    const int zb = zbi * ZPL;
    const int zcount = min(ZPL, a.nzs - zb);
    const bool colvol = xi < a.nx && yi < a.ny;
    const float x = (float)xi - a.xpre, y = (float)yi - a.ypre;
    const bool incyl = colvol && (x * x + y * y < a.r2);

    const float cxa = (float)(tx * kTX) - a.xpre;
    const float cxb = (float)min(tx * kTX + kTX - 1, a.nx - 1) - a.xpre;
    const float cya = (float)(ty * kTY) - a.ypre;
    const float cyb = (float)min(ty * kTY + kTY - 1, a.ny - 1) - a.ypre;
    const float zw0 = (float)(a.z0 + zb) - a.zpre;
    const float zw1 = zw0 + (float)(zcount - 1);

    float acc[ZPL];
#pragma unroll
    for (int k = 0; k < ZPL; ++k) acc[k] = 0.0f;

    // nearest point of the tile's xy rectangle to the axis
    const float ndx = cxa > 0.0f ? cxa : (cxb < 0.0f ? cxb : 0.0f);
    const float ndy = cya > 0.0f ? cya : (cyb < 0.0f ? cyb : 0.0f);
    const bool any_cyl = ndx * ndx + ndy * ndy < a.r2;

    if (any_cyl) {
        auto box_for = [&](int p) -> int4 {
            const float4 g = pp[p];
            float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
            bool bad = false;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float cx = (c & 1) ? cxb : cxa;
                const float cy = (c & 2) ? cyb : cya;
                float t = (-1.0f) * cx * g.x + cy * g.y;
                t += a.uoff;
                const float s = cx * g.y + cy * g.x;
                const float den = a.fcd - s;
                bad |= !(den > 0.0f);
                const float U = a.fcd / den;
                const float ui = t * U + a.upre;
                const float va = fmaf(-(zw0 + g.z), U, a.vpre);
                const float vb = fmaf(-(zw1 + g.z), U, a.vpre);
                umin = fminf(umin, ui); umax = fmaxf(umax, ui);
                vmin = fminf(vmin, fminf(va, vb)); vmax = fmaxf(vmax, fmaxf(va, vb));
            }
            int4 b = make_int4(0, 0, 2 << 28, 0);
            if (bad || !(umin == umin) || !(vmin == vmin)) return b;
            if (umax < -1.0f || umin > (float)a.W || vmax < -1.0f || vmin > (float)a.H) {
                b.z = 0;
                return b;
            }
            const int us = max((int)floorf(umin) - 1, 0);
            const int ue = min((int)floorf(umax) + 1, a.W - 1);
            const int vs = max((int)floorf(vmin) - 1, -1);
            const int ve = min((int)floorf(vmax) + 2, a.H - 1);
            const int nc = ue - us + 1, nr = ve - vs + 1;
            if (nc <= 0 || nr <= 1) { b.z = 0; return b; }
            b.x = us; b.y = vs; b.w = nr;
            b.z = nc | ((nc * nr <= a.cap ? 1 : 2) << 28);
            return b;
        };

        T regs[kMaxE];
        auto issue = [&](int p, int4 bx) {
            const int us = bx.x, vs = bx.y, nr = bx.w;
            const int ne = box_nc(bx) * nr;
            const T* src = proj + (size_t)p * a.W * a.H;
            int c = tid / nr, rr = tid - (tid / nr) * nr;
            const int qs = kThreads / nr, rs = kThreads - (kThreads / nr) * nr;
#pragma unroll
            for (int i = 0; i < kMaxE; ++i) {
                const int e = tid + i * kThreads;
                const int gv = vs + rr;
                T val = E::zero();
                if (e < ne && gv >= 0) val = src[(size_t)(us + c) * a.H + gv];
                regs[i] = val;
                c += qs;
                rr += rs;
                if (rr >= nr) { rr -= nr; ++c; }
            }
        };
        auto commit = [&](int4 bx) {
            const int ne = box_nc(bx) * bx.w;
#pragma unroll
            for (int i = 0; i < kMaxE; ++i) {
                const int e = tid + i * kThreads;
                if (e < ne) tile[e] = regs[i];
            }
        };

        auto compute = [&](int p, int4 bx) {
            const int flag = box_flag(bx);
            if (flag == 0) return;
            const float4 g = pp[p];
            const float sc = scales ? scales[p] : 1.0f;
            float t = (-1.0f) * x * g.x + y * g.y;
            t += a.uoff;
            const float s = x * g.y + y * g.x;
            const float U = a.fcd / (a.fcd - s);
            const float u = t * U;
            const bool cv = incyl && u >= a.ulo && u <= a.uhi;
            const float uIdx = u + a.upre;
            const int u0 = (int)uIdx;
            const float fu = uIdx - (float)u0;
            const float w = U * U * sc;
            if (flag == 1) {
                const int us = bx.x, vs = bx.y, nr = bx.w;
                const bool lane_full =
                    !cv || (zcount == ZPL && (zw0 + g.z) * U >= a.vlo && (zw1 + g.z) * U <= a.vhi);
                if (__all(lane_full)) {
                    // every active column keeps all ZPL voxels inside the detector window
                    const float Ue = cv ? U : 0.0f;
                    const float we = cv ? w : 0.0f;
                    const float fue = cv ? fu : 0.0f;
                    // inactive columns read entry 0 with zero weight
                    const int base = cv ? (u0 - us) * nr - vs : -(int)floorf(a.vpre);
                    const float vI0 = fmaf(-(zw0 + g.z), Ue, a.vpre);
#pragma unroll
                    for (int k = 0; k < ZPL; ++k) {
                        const float vI = fmaf(-(float)k, Ue, vI0);
                        const float fl = floorf(vI);
                        const float fv = vI - fl;
                        const int e = base + (int)fl;
                        const T e0 = tile[e], e1 = tile[e + 1];
                        float A, B, C, D;
                        E::taps(e0, e1, A, B, C, D);
                        acc[k] = fmaf(we, interp(fue, fv, A, B, C, D), acc[k]);
                    }
                } else {
                    const int base = (u0 - us) * nr - vs;
#pragma unroll
                    for (int k = 0; k < ZPL; ++k) {
                        const float zk = zw0 + (float)k;
                        const float v = (zk + g.z) * U;
                        if (cv && k < zcount && v >= a.vlo && v <= a.vhi) {
                            const float vI = (-1.0f) * v + a.vpre;
                            const int v0 = (int)vI;
                            const float fv = vI - (float)v0;
                            const int e = base + v0;
                            const T e0 = tile[e], e1 = tile[e + 1];
                            float A, B, C, D;
                            E::taps(e0, e1, A, B, C, D);
                            acc[k] = fmaf(w, interp(fu, fv, A, B, C, D), acc[k]);
                        }
                    }
                }
            } else {
                const T* src = proj + (size_t)p * a.W * a.H + (size_t)u0 * a.H;
#pragma unroll
                for (int k = 0; k < ZPL; ++k) {
                    const float zk = zw0 + (float)k;
                    const float v = (zk + g.z) * U;
                    if (cv && k < zcount && v >= a.vlo && v <= a.vhi) {
                        const float vI = (-1.0f) * v + a.vpre;
                        const int v0 = (int)vI;
                        const float fv = vI - (float)v0;
                        const T e0 = src[v0], e1 = src[v0 + 1];
                        float A, B, C, D;
                        E::taps(e0, e1, A, B, C, D);
                        acc[k] = fmaf(w, interp(fu, fv, A, B, C, D), acc[k]);
                    }
                }
            }
        };

        if (tid < 64) boxes[0][tid] = tid < a.nproj ? box_for(tid) : make_int4(0, 0, 0, 0);
        __syncthreads();
        {
            const int4 b0 = boxes[0][0];
            if (a.nproj > 0 && box_flag(b0) == 1) issue(0, b0);
        }
        for (int p = 0; p < a.nproj; ++p) {
            const int4 bx = boxes[(p >> 6) & 1][p & 63];
            __syncthreads();  // tile and the other box buffer are free
            if ((p & 63) == 0 && tid < 64) {
                const int pn = p + 64 + tid;
                boxes[((p >> 6) + 1) & 1][tid] = pn < a.nproj ? box_for(pn) : make_int4(0, 0, 0, 0);
            }
            if (box_flag(bx) == 1) commit(bx);
            __syncthreads();  // tile of p and the next box batch are visible
            if (p + 1 < a.nproj) {
                const int4 bn = boxes[((p + 1) >> 6) & 1][(p + 1) & 63];
                if (box_flag(bn) == 1) issue(p + 1, bn);
            }
            compute(p, bx);
        }
    }

    if (colvol) {
        float* dst = vol + ((size_t)xi * a.ny + yi) * a.nzs + zb;
        if (a.vec4 && zcount == ZPL) {
            float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
            for (int k = 0; k < ZPL; k += 4) {
                float4 v = make_float4(acc[k], acc[k + 1], acc[k + 2], acc[k + 3]);
                if (a.accumulate) {
                    const float4 o = d4[k >> 2];
                    v.x += o.x; v.y += o.y; v.z += o.z; v.w += o.w;
                }
                d4[k >> 2] = v;
            }
        } else {
#pragma unroll
            for (int k = 0; k < ZPL; ++k) {
                if (k < zcount) dst[k] = a.accumulate ? dst[k] + acc[k] : acc[k];
            }
        }
    }
}

// FP64 backprojection from row-major double images (the GPU error yardstick).
// One thread per voxel column and run of 16 z; the same masks as the FP32 path,
// evaluated in double.
__global__ __launch_bounds__(256) void bp_f64_kernel(BpArgs a, const double4* __restrict__ pp,
                                                     const double* __restrict__ rows,
                                                     double* __restrict__ vol) {
    constexpr int ZR = 16;
    const int xi = blockIdx.x * 16 + (threadIdx.x & 15);
    const int yi = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int zb = blockIdx.z * ZR;
    if (xi >= a.nx || yi >= a.ny) return;
    const int zc = min(ZR, a.nzs - zb);
    double acc[ZR];
#pragma unroll
    for (int k = 0; k < ZR; ++k) acc[k] = 0.0;
    const double x = (double)xi - a.xpre, y = (double)yi - a.ypre;
    const double r2 = (double)a.r2;
    if (x * x + y * y < r2) {
        for (int p = 0; p < a.nproj; ++p) {
            const double4 g = pp[p];
            const double sn = g.x, cs = g.y, h = g.z;
            const double t = -x * sn + y * cs + (double)a.uoff;
            const double s = x * cs + y * sn;
            const double U = (double)a.fcd / ((double)a.fcd - s);
            const double u = t * U;
            if (u < (double)a.ulo || u > (double)a.uhi) continue;
            const double w = U * U;
            const double uI = u + (double)a.upre;
            const int u0 = (int)uI;
            const double fu = uI - u0;
            const double* img = rows + (size_t)p * a.W * a.H;
#pragma unroll
            for (int k = 0; k < ZR; ++k) {
                if (k >= zc) break;
                const double v = ((double)(a.z0 + zb + k) - (double)a.zpre + h) * U;
                if (v < (double)a.vlo || v > (double)a.vhi) continue;
                const double vI = -v + (double)a.vpre;
                const int v0 = (int)vI;
                const double fv = vI - v0;
                const double* r0 = img + (size_t)v0 * a.W;
                const double* r1 = r0 + a.W;
                const double ta = (1.0 - fu) * r0[u0] + fu * r0[u0 + 1];
                const double tb = (1.0 - fu) * r1[u0] + fu * r1[u0 + 1];
                acc[k] += w * ((1.0 - fv) * ta + fv * tb);
            }
        }
    }
    double* dst = vol + ((size_t)xi * a.ny + yi) * a.nzs + zb;
    for (int k = 0; k < zc; ++k) dst[k] = acc[k];
}

BpArgs make_args(const ctg_geometry& g, int nproj, int z0, int nz, int accumulate) {
    BpArgs a{};
    a.W = g.det_width;
    a.H = g.det_height;
    a.nx = g.nx;
    a.ny = g.ny;
    a.nzs = nz;
    a.z0 = z0;
    a.xpre = g.x_pre;
    a.ypre = g.y_pre;
    a.zpre = g.z_pre;
    a.upre = g.u_pre;
    a.vpre = g.v_pre;
    a.r2 = g.radius * g.radius;
    a.fcd = g.fcd;
    a.uoff = g.u_offset;
    a.ulo = g.u_lo;
    a.uhi = g.u_hi;
    a.vlo = g.v_lo;
    a.vhi = g.v_hi;
    a.nproj = nproj;
    a.accumulate = accumulate;
    return a;
}

constexpr int kZPL = 32;

// Upper bound on the detector footprint (pair entries) of one 16x16xZPL block that
// intersects the cylinder, for any projection angle.
int footprint_bound(const ctg_geometry& g, int zpl) {
    const double diag = 15.0 * std::sqrt(2.0) + 0.01;
    const double Rt = (double)g.radius + diag + 1.0;
    const double fcd = g.fcd;
    if (fcd <= Rt + 1.0) return 0;
    const double Umax = fcd / (fcd - Rt);
    const double dU = fcd * diag / ((fcd - Rt) * (fcd - Rt));
    const double tmax = Rt + std::fabs((double)g.u_offset);
    const double urange = diag * Umax + tmax * dU;
    const double zmax = std::max((double)g.z_pre, (double)g.nz - g.z_pre) + 1.0;
    const double vrange = (zpl - 1) * Umax + zmax * dU;
    const int ncols = (int)std::ceil(urange) + 4;
    const int nrows = (int)std::ceil(vrange) + 5;
    return ncols * nrows;
}

}  // namespace

BpPlan plan_backproject(const ctg_geometry& g, ctg_store store, int nz) {
    BpPlan pl{};
    pl.zpl = kZPL;
    const int ntx = (g.nx + kTX - 1) / kTX, nty = (g.ny + kTY - 1) / kTY, nzb = (nz + kZPL - 1) / kZPL;
    pl.grid = ntx * nty * nzb;
    int cap = std::min(footprint_bound(g, kZPL), kMaxE * kThreads);
    if (cap < 2) cap = 2;
    pl.cap_entries = cap;
    pl.lds_bytes = cap * (store == CTG_STORE_F32 ? 8 : 4);
    return pl;
}

ctg_status launch_backproject(const BpLaunch& L, hipStream_t st) {
    const ctg_geometry& g = *L.g;
    if (L.nz <= 0 || g.nx <= 0 || g.ny <= 0) return CTG_OK;
    BpArgs a = make_args(g, L.nproj, L.z0, L.nz, L.accumulate);
    const BpPlan pl = plan_backproject(g, L.store, L.nz);
    a.ntx = (g.nx + kTX - 1) / kTX;
    a.nty = (g.ny + kTY - 1) / kTY;
    a.nzb = (L.nz + kZPL - 1) / kZPL;
    a.cap = pl.cap_entries;
    a.vec4 = (L.nz % 4 == 0 && ((uintptr_t)L.d_volume % 16) == 0) ? 1 : 0;
    const long long grid = (long long)a.ntx * a.nty * a.nzb;
    if (grid > 0x7fffffffll) return fail(CTG_ERROR_UNSUPPORTED, "backprojection grid too large");
    switch (L.store) {
        case CTG_STORE_F32:
            hipLaunchKernelGGL((bp_kernel<kZPL, EF32>), dim3((unsigned)grid), dim3(kThreads), pl.lds_bytes, st, a,
                               L.d_params, L.d_scales, (const float2*)L.d_filtered, L.d_volume);
            break;
        case CTG_STORE_F16:
            hipLaunchKernelGGL((bp_kernel<kZPL, EF16>), dim3((unsigned)grid), dim3(kThreads), pl.lds_bytes, st, a,
                               L.d_params, L.d_scales, (const uint32_t*)L.d_filtered, L.d_volume);
            break;
        case CTG_STORE_BF16:
            hipLaunchKernelGGL((bp_kernel<kZPL, EBF16>), dim3((unsigned)grid), dim3(kThreads), pl.lds_bytes, st, a,
                               L.d_params, L.d_scales, (const uint32_t*)L.d_filtered, L.d_volume);
            break;
        default:
            return fail(CTG_ERROR_INVALID_ARGUMENT, "unknown store type %d", (int)L.store);
    }
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

ctg_status launch_backproject_f64(const ctg_geometry& g, const double4* d_params, const double* d_rows,
                                  int nproj, double* d_volume, int z0, int nz, hipStream_t st) {
    if (nz <= 0) return CTG_OK;
    BpArgs a = make_args(g, nproj, z0, nz, 0);
    dim3 grid((unsigned)((g.nx + 15) / 16), (unsigned)((g.ny + 15) / 16), (unsigned)((nz + 15) / 16));
    hipLaunchKernelGGL(bp_f64_kernel, grid, dim3(256), 0, st, a, d_params, d_rows, d_volume);
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

}  // namespace ctg
