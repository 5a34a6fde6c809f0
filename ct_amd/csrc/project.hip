// Forward projector (exact adjoint of the backprojection), analytic phantom images,
// and volume index-order / byte-order conversion for the writer.
//
// The forward projector has no counterpart in the reference; it transposes the
// voxel-driven backprojection of Ct/CtVolume.cpp:979-1039 term by term: every voxel that
// the backprojection would read through taps (u0,v0),(u0+1,v0),(u0,v0+1),(u0+1,v0+1) with
// weight U^2 and bilinear weights instead adds U^2 * f * weight into those pixels, under
// the same cylinder and window masks.  A workgroup accumulates its block's splats for
// one projection in an LDS tile (ds_add_f32), then adds the tile to the detector image
// with global float atomics.
#include "ctg_internal.h"

#include <algorithm>
#include <cmath>

namespace ctg {
namespace {

constexpr int kTX = 16, kTY = 16, kThreads = 256, kZR = 16;
static_assert(kRowAnchor % kZR == 0, "a forward-projection run lies inside one row anchor block (det_row_index)");
constexpr int kFpCap = 4096;  // LDS accumulation tile capacity (floats)

struct FpArgs {
    int W, H, nx, ny, nzs, z0, nproj, ntx, nty;
    int zshift;  // z0 mod 16: runs start at grid z multiples of 16
    float xpre, ypre, zpre, upre, vpre;
    float r2, fcd, uoff, ulo, uhi, vlo, vhi;
};

__global__ __launch_bounds__(kThreads) void fp_kernel(FpArgs a, const float4* __restrict__ pp,
                                                      const float* __restrict__ vol,
                                                      float* __restrict__ proj) {
    __shared__ float acc[kFpCap];
    const int tile_id = blockIdx.x;
    const int ty = tile_id / a.ntx, tx = tile_id - ty * a.ntx;
    const int zb = blockIdx.y * kZR;
    const int tid = threadIdx.x;
    const int xi = tx * kTX + (tid & (kTX - 1));
    const int yi = ty * kTY + (tid / kTX);
    // run-local k is a slab voxel when zfirst <= k < zlast (slab index zb - zshift + k)
    const int zfirst = a.zshift - zb;
    const int zlast = a.nzs + a.zshift - zb;
    const bool colvol = xi < a.nx && yi < a.ny;
    const float x = (float)xi - a.xpre, y = (float)yi - a.ypre;
    const bool incyl = colvol && (x * x + y * y < a.r2);

    float f[kZR];
    {
        const float* col = vol + ((size_t)(colvol ? xi : 0) * a.ny + (colvol ? yi : 0)) * a.nzs;
#pragma unroll
        for (int k = 0; k < kZR; ++k) f[k] = (incyl && k >= zfirst && k < zlast) ? col[zb - a.zshift + k] : 0.0f;
    }
    const float cxa = (float)(tx * kTX) - a.xpre;
    const float cxb = (float)min(tx * kTX + kTX - 1, a.nx - 1) - a.xpre;
    const float cya = (float)(ty * kTY) - a.ypre;
    const float cyb = (float)min(ty * kTY + kTY - 1, a.ny - 1) - a.ypre;
    const float ndx = cxa > 0.0f ? cxa : (cxb < 0.0f ? cxb : 0.0f);
    const float ndy = cya > 0.0f ? cya : (cyb < 0.0f ? cyb : 0.0f);
    if (!(ndx * ndx + ndy * ndy < a.r2)) return;  // uniform: no voxel of the tile in the cylinder
    // detector rows come from the run's 32-slice anchor (det_row_index): k + koff past it
    const int zrun = a.z0 - a.zshift + zb;
    const float zw0 = (float)zrun - a.zpre;
    const float zw1 = zw0 + (float)(kZR - 1);
    const int koff = zrun & (kRowAnchor - 1);
    const float zwa = (float)(zrun - koff) - a.zpre;

    for (int p = 0; p < a.nproj; ++p) {
        const float4 g = pp[p];
        // footprint box of the block in pixels (same construction as the backprojection)
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
            const float va = (-1.0f) * ((zw0 + g.z) * U) + a.vpre;
            const float vb = (-1.0f) * ((zw1 + g.z) * U) + a.vpre;
            umin = fminf(umin, ui); umax = fmaxf(umax, ui);
            vmin = fminf(vmin, fminf(va, vb)); vmax = fmaxf(vmax, fmaxf(va, vb));
        }
        int us = 0, vs = 0, ncol = 0, nrow = 0;
        bool use_lds = false;
        if (!bad && umin == umin && vmin == vmin) {
            us = max((int)floorf(umin) - 1, 0);
            const int ue = min((int)floorf(umax) + 2, a.W - 1);
            vs = max((int)floorf(vmin) - 1, 0);
            const int ve = min((int)floorf(vmax) + 2, a.H - 1);
            ncol = ue - us + 1;
            nrow = ve - vs + 1;
            if (ncol <= 0 || nrow <= 0) continue;  // uniform: nothing of this block hits the detector
            use_lds = ncol * nrow <= kFpCap;
        }
        float* img = proj + (size_t)p * a.W * a.H;
        if (use_lds) {
            for (int e = tid; e < ncol * nrow; e += kThreads) acc[e] = 0.0f;
            __syncthreads();
        }
        float t = (-1.0f) * x * g.x + y * g.y;
        t += a.uoff;
        const float s = x * g.y + y * g.x;
        const float U = cone_magnification(a.fcd, s);
        const float u = t * U;
        const bool cv = incyl && u >= a.ulo && u <= a.uhi;
        if (cv) {
            const float uIdx = u + a.upre;
            const int u0 = (int)uIdx;
            const float fu = uIdx - (float)u0;
            const float w = U * U;
            const float vA = det_row_anchor(zwa, g.z, U, a.vpre);
#pragma unroll
            for (int k = 0; k < kZR; ++k) {
                const float zk = zw0 + (float)k;
                const float v = (zk + g.z) * U;
                if (k >= zfirst && k < zlast && v >= a.vlo && v <= a.vhi) {
                    const float vI = det_row_index((float)(k + koff), U, vA);
                    const int v0 = (int)vI;
                    const float fv = vI - (float)v0;
                    const float wf = w * f[k];
                    const float b0 = wf * (1.0f - fv), b1 = wf * fv;
                    const float c00 = b0 * (1.0f - fu), c10 = b0 * fu, c01 = b1 * (1.0f - fu), c11 = b1 * fu;
                    if (use_lds) {
                        const int e = (u0 - us) * nrow + (v0 - vs);
                        atomicAdd(&acc[e], c00);
                        atomicAdd(&acc[e + nrow], c10);
                        atomicAdd(&acc[e + 1], c01);
                        atomicAdd(&acc[e + nrow + 1], c11);
                    } else {
                        float* r0 = img + (size_t)v0 * a.W + u0;
                        atomicAdd(r0, c00);
                        atomicAdd(r0 + 1, c10);
                        atomicAdd(r0 + a.W, c01);
                        atomicAdd(r0 + a.W + 1, c11);
                    }
                }
            }
        }
        if (use_lds) {
            __syncthreads();
            for (int e = tid; e < ncol * nrow; e += kThreads) {
                const float val = acc[e];
                if (val != 0.0f) {
                    const int c = e / nrow, r = e - (e / nrow) * nrow;
                    atomicAdd(img + (size_t)(vs + r) * a.W + (us + c), val);
                }
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ phantom

struct EllDev {
    double c[3];
    double M[9];  // D^-1 R^T (rows)
    double mu;
};

__device__ inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ inline double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

// Poisson sample: inversion by sequential multiplication for small means, Hormann's
// transformed rejection (PTRS, 1993) otherwise.  The rejection loop is bounded; a draw
// that has not been accepted after 64 rounds returns the rounded mean.
__device__ double poisson(double lam, uint64_t& s) {
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double prod = uniform01(s);
        int k = 0;
        while (prod > L && k < 1000) { prod *= uniform01(s); ++k; }
        return (double)k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 64; ++it) {
        const double U = uniform01(s) - 0.5;
        const double V = uniform01(s);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0)) return k;
    }
    return floor(lam + 0.5);
}

__global__ __launch_bounds__(256) void phantom_kernel(int W, int H, int nproj, int p0, double fcd, double uoff,
                                                      double upre, double vpre, const double4* __restrict__ pp,
                                                      const EllDev* __restrict__ ell, int n_ell, double photons,
                                                      double full_scale, uint64_t seed,
                                                      uint16_t* __restrict__ img, float* __restrict__ lin) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    const int pl = blockIdx.z;
    if (c >= W) return;
    const double4 g = pp[p0 + pl];
    const double sn = g.x, cs = g.y, h = g.z;
    const double S[3] = {fcd * cs + uoff * sn, fcd * sn - uoff * cs, -h};
    const double u = (double)c - upre, v = vpre - (double)r;
    const double tq = u - uoff;
    const double Q[3] = {-tq * sn, tq * cs, v - h};
    double d[3] = {Q[0] - S[0], Q[1] - S[1], Q[2] - S[2]};
    const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= dn; d[1] /= dn; d[2] /= dn;
    double acc = 0.0;
    for (int e = 0; e < n_ell; ++e) {
        const EllDev E = ell[e];
        const double o[3] = {S[0] - E.c[0], S[1] - E.c[1], S[2] - E.c[2]};
        double q0[3], q1[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            q0[i] = E.M[3 * i] * o[0] + E.M[3 * i + 1] * o[1] + E.M[3 * i + 2] * o[2];
            q1[i] = E.M[3 * i] * d[0] + E.M[3 * i + 1] * d[1] + E.M[3 * i + 2] * d[2];
        }
        const double A = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
        const double B = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2];
        const double C = q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] - 1.0;
        const double disc = B * B - A * C;
        if (disc > 0.0) acc += E.mu * 2.0 * sqrt(disc) / A;
    }
    const size_t idx = ((size_t)pl * H + r) * W + c;
    if (lin) lin[idx] = (float)acc;
    if (img) {
        double val;
        if (photons > 0.0) {
            uint64_t st = seed ^ (0xd1b54a32d192ed03ull * ((uint64_t)(p0 + pl) * H * W + (uint64_t)r * W + c + 1));
            splitmix64(st);
            val = poisson(photons * exp(-acc), st);
        } else {
            val = floor(full_scale * exp(-acc) + 0.5);
        }
        val = fmin(fmax(val, 0.0), 65535.0);
        img[idx] = (uint16_t)val;
    }
}

// --------------------------------------------------------------- conversion

__device__ inline uint32_t maybe_swap(float v, int swap) {
    uint32_t b = __float_as_uint(v);
    return swap ? __builtin_bswap32(b) : b;
}

// Z_FASTEST [x][y][z] -> X_FASTEST [z][y][x] through a 32x32 LDS tile per (x,z) block.
__global__ __launch_bounds__(256) void to_xfastest_kernel(const float* __restrict__ in, int nx, int ny, int nz,
                                                          int swap, uint32_t* __restrict__ out) {
    __shared__ float tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tl = threadIdx.x & 31;
    for (int i = threadIdx.x >> 5; i < 32; i += 8) {
        const int x = x0 + i, z = z0 + tl;
        tile[i][tl] = (x < nx && z < nz) ? in[((size_t)x * ny + y) * nz + z] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x >> 5; i < 32; i += 8) {
        const int z = z0 + i, x = x0 + tl;
        if (x < nx && z < nz) out[((size_t)z * ny + y) * nx + x] = maybe_swap(tile[tl][i], swap);
    }
}

__global__ __launch_bounds__(256) void copy_swap_kernel(const float* __restrict__ in, size_t n, int swap,
                                                        uint32_t* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = maybe_swap(in[i], swap);
}

}  // namespace

ctg_status launch_forward_project(const ctg_geometry& g, const float4* d_params, const float* d_volume, int z0,
                                  int nz, int nproj, float* d_proj, hipStream_t st) {
    if (nz <= 0 || nproj <= 0) return CTG_OK;
    FpArgs a{};
    a.W = g.det_width; a.H = g.det_height; a.nx = g.nx; a.ny = g.ny; a.nzs = nz; a.z0 = z0; a.nproj = nproj;
    a.zshift = z0 & 15;
    a.ntx = (g.nx + kTX - 1) / kTX; a.nty = (g.ny + kTY - 1) / kTY;
    a.xpre = g.x_pre; a.ypre = g.y_pre; a.zpre = g.z_pre; a.upre = g.u_pre; a.vpre = g.v_pre;
    a.r2 = g.radius * g.radius; a.fcd = g.fcd; a.uoff = g.u_offset;
    a.ulo = g.u_lo; a.uhi = g.u_hi; a.vlo = g.v_lo; a.vhi = g.v_hi;
    dim3 grid((unsigned)(a.ntx * a.nty), (unsigned)((nz + a.zshift + kZR - 1) / kZR));
    hipLaunchKernelGGL(fp_kernel, grid, dim3(kThreads), 0, st, a, d_params, d_volume, d_proj);
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

ctg_status launch_phantom(const ctg_geometry& g, const double4* d_params, const double* h_ell, int n_ell,
                          int nproj, int p0, double photons, double full_scale, uint64_t seed, uint16_t* d_img,
                          float* d_lineint, DevBuf& ell_buf, hipStream_t st) {
    std::vector<EllDev> e(n_ell > 0 ? n_ell : 1);
    for (int i = 0; i < n_ell; ++i) {
        const double* E = h_ell + 10 * i;
        const double c1 = std::cos(E[6]), s1 = std::sin(E[6]), c2 = std::cos(E[7]), s2 = std::sin(E[7]);
        const double c3 = std::cos(E[8]), s3 = std::sin(E[8]);
        // R = Rz(phi) Ry(theta) Rz(psi), row-major; columns are the ellipsoid axes
        const double R[9] = {c1 * c2 * c3 - s1 * s3, -c1 * c2 * s3 - s1 * c3, c1 * s2,
                             s1 * c2 * c3 + c1 * s3, -s1 * c2 * s3 + c1 * c3, s1 * s2,
                             -s2 * c3,               s2 * s3,                c2};
        for (int a = 0; a < 3; ++a) {
            e[i].c[a] = E[a];
            for (int b = 0; b < 3; ++b) e[i].M[3 * a + b] = R[3 * b + a] / E[3 + a];
        }
        e[i].mu = E[9];
    }
    CTG_TRY(ell_buf.reserve(sizeof(EllDev) * e.size()));
    CTG_HIP(hipMemcpyAsync(ell_buf.ptr, e.data(), sizeof(EllDev) * e.size(), hipMemcpyHostToDevice, st));
    dim3 grid((unsigned)((g.det_width + 255) / 256), (unsigned)g.det_height, (unsigned)nproj);
    hipLaunchKernelGGL(phantom_kernel, grid, dim3(256), 0, st, g.det_width, g.det_height, nproj, p0, (double)g.fcd,
                       (double)g.u_offset, (double)g.u_pre, (double)g.v_pre, d_params, (const EllDev*)ell_buf.ptr,
                       n_ell, photons, full_scale, seed, d_img, d_lineint);
    CTG_HIP(hipGetLastError());
    // the host vector must outlive the asynchronous copy
    CTG_HIP(hipStreamSynchronize(st));
    return CTG_OK;
}

ctg_status launch_convert_volume(const ctg_geometry& g, const float* d_vol, int nz, ctg_index_order order,
                                 ctg_byte_order bo, void* d_out, hipStream_t st) {
    const int swap = bo == CTG_BIG_ENDIAN ? 1 : 0;
    if (order == CTG_Z_FASTEST) {
        const size_t n = (size_t)g.nx * g.ny * nz;
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535 * 8);
        hipLaunchKernelGGL(copy_swap_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, st, d_vol, n, swap,
                           (uint32_t*)d_out);
    } else {
        dim3 grid((unsigned)((nz + 31) / 32), (unsigned)((g.nx + 31) / 32), (unsigned)g.ny);
        hipLaunchKernelGGL(to_xfastest_kernel, grid, dim3(256), 0, st, d_vol, g.nx, g.ny, nz, swap,
                           (uint32_t*)d_out);
    }
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

}  // namespace ctg
