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
//  - projections arrive as bilinear-coefficient entries (A, dU, dV, dUV) stored
//    column-major, so one voxel-update is one LDS read of one entry and
//    val = A + fu*dU + fv*(dV + fu*dUV): 3 FMAs, plus 1 FMA into the accumulator;
//  - a workgroup (256 lanes = 16x16 columns) stages, per projection, the bounding box of
//    its voxel block's detector footprint into LDS with global->LDS DMA (buffer loads
//    to LDS), one detector column per DMA burst, double-buffered so that the DMA of
//    projection p+1 runs under the compute of p (the two buffers reach the compute as
//    __restrict__ pointers, without which the compiler waits for the DMA before the
//    first LDS read); a box's detector columns sit at an LDS pitch of +1 or -1 mod 16
//    entries, chosen per box against bank conflicts;
//  - per z the lane does one FMA for vIndex, v_cvt_flr/v_fract, one address op, the LDS
//    read and 4 FMAs, with the reads software-pipelined in small groups; the
//    per-column geometry (one reciprocal) is amortised over the run;
//  - boxes are computed by wave 0 for 64 projections at a time (one lane per projection);
//    a projection whose box exceeds the tile capacity reads entries from global memory
//    instead, so any geometry stays correct;
//  - workgroups are numbered so that each XCD (dispatch is round-robin over 8 XCDs)
//    receives a contiguous range of tiles, in 4x4 blocks of tiles, and shares detector
//    footprints in its own L2.
#include "ctg_internal.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace ctg {
namespace bp_detail {

constexpr int kTX = 16, kTY = 16, kThreads = 256, kWaves = kThreads / 64;
constexpr int kMaxTileBytes = 36 * 1024;  // per LDS buffer (two buffers per workgroup)
// Detector pixels added around a footprint box's corner bounds: the corner and voxel
// expressions round differently by about 1e-4 pixel at most.
constexpr float kBoxMargin = 0.05f;

struct BpArgs {
    int W, H, nx, ny, nzs, z0;
    float xpre, ypre, zpre, upre, vpre;
    float r2, fcd, uoff, ulo, uhi, vlo, vhi;
    int nproj, ntx, nty, nzb, cap, accumulate, vec4;
    int blk, ntxp, ntyp;     // tile blocks of blk x blk (0: rows of tiles); tile counts rounded up to blk
    int zshift;              // z0 mod ZPL: runs start at grid z multiples of ZPL, zshift voxels before the slab
    int dma16;               // 8-byte entries staged with 16-byte DMA (H even: rows start even)
    int pitch_mod, pitch_rem;  // LDS column pitch = rows rounded up to pitch_rem mod pitch_mod (a power of two)
    int pitch_adapt;           // per box, -pitch_rem mod pitch_mod where the footprint slants the other way
};

// Entry formats: FP32 (16 B), FP16 and BF16 (8 B: {A,dU} {dV,dUV} as 16-bit pairs).
struct EF32 {
    using T = float4;
    static constexpr int kBytes = 16;
    static constexpr int kMinWaves = 1;  // LDS limits FP32 tiles to 3 workgroups per CU anyway
    static constexpr int kGrp = 4;       // LDS reads per software-pipeline group
    __device__ static float eval(const T& e, float fu, float fv) {
        return fmaf(fv, fmaf(fu, e.w, e.z), fmaf(fu, e.y, e.x));
    }
};
struct EF16 {
    using T = uint2;
    static constexpr int kBytes = 8;
    static constexpr int kMinWaves = 4;  // VALU-bound: keep 4 waves per SIMD (at most 128 VGPRs)
    static constexpr int kGrp = 2;       // (groups of 4 need spills at 128 VGPRs: measured slower)
    __device__ static float lo(uint32_t e) { return __half2float(__ushort_as_half((unsigned short)(e & 0xffffu))); }
    __device__ static float hi(uint32_t e) { return __half2float(__ushort_as_half((unsigned short)(e >> 16))); }
    __device__ static float eval(const T& e, float fu, float fv) {
        return fmaf(fv, fmaf(fu, hi(e.y), lo(e.y)), fmaf(fu, hi(e.x), lo(e.x)));
    }
};
struct EBF16 {
    using T = uint2;
    static constexpr int kBytes = 8;
    static constexpr int kMinWaves = 4;
    static constexpr int kGrp = 2;
    __device__ static float lo(uint32_t e) { return __uint_as_float(e << 16); }
    __device__ static float hi(uint32_t e) { return __uint_as_float(e & 0xffff0000u); }
    __device__ static float eval(const T& e, float fu, float fv) {
        return fmaf(fv, fmaf(fu, hi(e.y), lo(e.y)), fmaf(fu, hi(e.x), lo(e.x)));
    }
};

// Global->LDS DMA through a buffer resource: B bytes (4 or 16) per lane from byte
// offset soff + voff of the resource into lds + lane*B.  The buffer form matters: an
// outstanding global_load_lds is modelled by the compiler as a flat access to memory
// and LDS, after which it waits for every outstanding LDS read (lgkmcnt(0)) at each LDS
// use and the read pipeline collapses; a buffer load to LDS is tracked by the memory
// counter only.  Offsets past the resource's size read zeros.  The builtins exist only
// in the device pass.
#if defined(__HIP_DEVICE_COMPILE__)
using dma_rsrc = __amdgpu_buffer_rsrc_t;
__device__ inline dma_rsrc make_dma_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, 0x00020000);
}
template <int B>
__device__ inline void lds_dma(dma_rsrc rs, unsigned voff, int soff, void* lds_wave_base) {
    auto* l = (__attribute__((address_space(3))) void*)lds_wave_base;
    if constexpr (B == 16) __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, l, 16, voff, soff, 0, 0);
    else __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, l, 4, voff, soff, 0, 0);
}
#else
struct dma_rsrc {};
__device__ inline dma_rsrc make_dma_rsrc(const void*, int) { return {}; }
template <int B>
__device__ inline void lds_dma(dma_rsrc, unsigned, int, void*) {}
#endif

__device__ inline int cvt_flr(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// box: x = first entry column, y = first row, w = rows, z = ncols | flags.
// bits 28-29: 0 the block cannot hit the detector, 1 staged in LDS, 2 global reads;
// bit 27: the fast (unmasked) path may be used.
__device__ inline int box_kind(int4 b) { return ((unsigned)b.z >> 28) & 3; }
__device__ inline bool box_fast(int4 b) { return (b.z >> 27) & 1; }
__device__ inline int box_nc(int4 b) { return b.z & 0x07ffffff; }
__device__ inline int box_nr(int4 b) { return b.w & 0xffff; }
__device__ inline int box_pitch(int4 b) { return (unsigned)b.w >> 16; }

// DMA: bytes per lane of the global->LDS staging (16, or 4 for 8-byte entries).  A
// compile-time width keeps the staging loop free of a run-time choice between the two
// DMA forms, which measured markedly slower for the 8-byte entries.
template <int ZPL, class E, int DMA>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(E::kMinWaves))) void bp_kernel(BpArgs a, const float4* __restrict__ pp,
                                                      const float* __restrict__ scales,
                                                      const void* __restrict__ proj_raw,
                                                      float* __restrict__ vol) {
    using T = typename E::T;
    const T* __restrict__ proj = reinterpret_cast<const T*>(proj_raw);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int4 boxes[2][64];
    // per projection: sin, cos, height, and the entry scale in w (written with the boxes)
    __shared__ float4 params[2][64];

    // XCD-aware numbering: dispatch sends workgroup L to XCD L%8; give each XCD a
    // contiguous range of work items.
    const unsigned N = gridDim.x, L = blockIdx.x;
    const unsigned xcd = L & 7u, kk = L >> 3, q8 = N >> 3, r8 = N & 7u;
    const unsigned item = xcd * q8 + min(xcd, r8) + kk;
    int zbi, tx, ty;
    if (a.blk > 0) {
        // consecutive items cover blk x blk squares of tiles, whose detector footprints
        // overlap along every ray direction (the workgroups an XCD runs together then
        // share their boxes in its L2); tiles of the rounded-up grid beyond it exit
        const unsigned ntxy = (unsigned)(a.ntxp * a.ntyp), bb = (unsigned)(a.blk * a.blk);
        zbi = (int)(item / ntxy);
        const unsigned t = item - (unsigned)zbi * ntxy, b = t / bb, w = t - b * bb;
        const unsigned sbx = (unsigned)(a.ntxp / a.blk), by = b / sbx;
        tx = (int)((b - by * sbx) * a.blk + w % a.blk);
        ty = (int)(by * a.blk + w / a.blk);
        if (tx >= a.ntx || ty >= a.nty) return;
    } else {
        const unsigned ntxy = (unsigned)(a.ntx * a.nty);
        zbi = (int)(item / ntxy);
        const int txy = (int)(item - (unsigned)zbi * ntxy);
        ty = txy / a.ntx;
        tx = txy - ty * a.ntx;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int xi = tx * kTX + (tid & (kTX - 1));
    const int yi = ty * kTY + (tid / kTX);
    // runs start at grid z indices that are multiples of ZPL (the slab begins zshift
    // voxels into the first run); run-local k is stored when zfirst <= k < zlast
    const int zb = zbi * ZPL;
    const int zfirst = a.zshift - zb;
    const int zlast = a.nzs + a.zshift - zb;
    const bool colvol = xi < a.nx && yi < a.ny;
    const float x = (float)xi - a.xpre, y = (float)yi - a.ypre;
    const bool incyl = colvol && (x * x + y * y < a.r2);

    const float cxa = (float)(tx * kTX) - a.xpre;
    const float cxb = (float)min(tx * kTX + kTX - 1, a.nx - 1) - a.xpre;
    const float cya = (float)(ty * kTY) - a.ypre;
    const float cyb = (float)min(ty * kTY + kTY - 1, a.ny - 1) - a.ypre;
    // the z-run always spans ZPL voxels (those outside the slab are computed, never
    // stored), so boxes and the fast-path test cover the full run.  Detector rows are
    // taken from the run's 32-slice anchor (det_row_index): k + koff slices past it.
    static_assert(ZPL == kRowAnchor || ZPL == kRowAnchor / 2, "a run lies inside one row anchor block");
    const int zrun = a.z0 - a.zshift + zb;
    const float zw0 = (float)zrun - a.zpre;
    const float zw1 = zw0 + (float)(ZPL - 1);
    const int koff = ZPL == kRowAnchor ? 0 : (zrun & (kRowAnchor - 1));
    const float zwa = (float)(zrun - koff) - a.zpre;

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
            // changes of uIdx and vIdx along x (the lanes of a 16-lane row), summed over
            // the tile's two y edges and the run's two z ends
            float du = 0.0f, dv = 0.0f;
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
                du += (c & 1) ? ui : -ui;
                dv += (c & 1) ? va + vb : -(va + vb);
            }
            int4 b = make_int4(0, 0, 2 << 28, 0);
            if (bad || !(umin == umin) || !(vmin == vmin)) return b;
            if (umax < -1.0f || umin > (float)a.W || vmax < -1.0f || vmin > (float)a.H) {
                b.z = 0;
                return b;
            }
            // A voxel reads the one entry at (floor(uIdx), floor(vIdx)).  uIdx and vIdx are
            // projective in the voxel's position, so over the tile and run they lie between
            // the corners' values; kBoxMargin covers the rounding of either expression.
            const int us0 = (int)floorf(umin - kBoxMargin), vs0 = (int)floorf(vmin - kBoxMargin);
            const int us = max(us0, 0);
            const int ue = min((int)floorf(umax + kBoxMargin), a.W - 1);
            const int vs = max(vs0, 0) & (a.dma16 ? ~1 : ~0);
            const int ve = min((int)floorf(vmax + kBoxMargin), a.H - 1);
            const int nc = ue - us + 1, nr = ve - vs + 1;
            if (nc <= 0 || nr <= 0) { b.z = 0; return b; }
            // LDS column pitch: the rows the DMA writes (an even count when 16-byte lanes
            // carry two 8-byte entries) rounded up to pitch_rem mod pitch_mod (a power of
            // two, so a mask: a division here would raise the kernel's register peak), so
            // that neighbouring detector columns start in different LDS bank groups
            const int wr = (DMA == 16 && E::kBytes == 8) ? (nr + 1) & ~1 : nr;
            int pitch = wr + ((a.pitch_rem - wr) & (a.pitch_mod - 1));
            // A 16-lane row of voxels reads entries whose column and row both drift along
            // x.  With the pitch at +1 mod 16 a lane's bank group is column + row, and two
            // lanes whose column step is undone by a row step share one; where the column
            // and row drift in opposite directions the pitch -1 mod 16 (row - column)
            // avoids that.  Taken only where the box still fits the buffer.
            if (a.pitch_adapt && du * dv < 0.0f) {
                const int pm = wr + ((a.pitch_mod - a.pitch_rem - wr) & (a.pitch_mod - 1));
                if (nc * pm <= a.cap) pitch = pm;
            }
            b.x = us; b.y = vs; b.w = nr | (pitch << 16);
            // the fused vIndex of the fast path may round a hair below 0 at the top edge
            const int fast = vs0 >= 0 ? 1 : 0;
            b.z = nc | (fast << 27) | ((nc * pitch <= a.cap ? 1 : 2) << 28);
            return b;
        };

        // Stage the box of projection p into LDS buffer `buf`: wave w copies detector
        // columns w, w+4, ...; a column's entries are contiguous in global memory and in
        // LDS.  All addresses are wave-uniform (SGPR arithmetic); each lane moves DMA bytes.
        const int wave_u = __builtin_amdgcn_readfirstlane(wave);
        auto stage = [&](unsigned char* dst, int p, int4 bx) {
            const int us = bx.x, vs = bx.y, nr = box_nr(bx), nc = box_nc(bx), pitch = box_pitch(bx);
            static_assert(DMA == 16 || (DMA == 4 && E::kBytes == 8), "16-byte entries stage 16 bytes per lane");
            constexpr bool wide = DMA == 16;
            constexpr int ub = DMA;  // bytes per lane per DMA
            const int units = wide ? (nr * E::kBytes + 15) / 16 : nr * (E::kBytes / 4);
            // one resource per projection image (its size is below 2^31 bytes, checked by
            // launch_backproject before the launch); the column offsets are wave-uniform
            const int img_bytes = a.W * a.H * E::kBytes;
            const dma_rsrc rs = make_dma_rsrc(reinterpret_cast<const char*>(proj) + (size_t)p * img_bytes, img_bytes);
            const int col0 = (us * a.H + vs) * E::kBytes;
            const unsigned lane_off = (unsigned)lane * ub;
            const int gstep = kWaves * a.H * E::kBytes, lstep = kWaves * pitch * E::kBytes;
            // A DMA instruction costs about the same whatever its lane mask, and the loop
            // around it is scalar work: every column takes `full` unmasked instructions,
            // then the columns' partial tails share one lane mask, set once per box.
            const int full = units >> 6, tail = units & 63;
            for (int j = 0; j < full; ++j) {
                int gcol = col0 + wave_u * a.H * E::kBytes + j * 64 * ub;
                unsigned char* lcol = dst + (size_t)wave_u * pitch * E::kBytes + (size_t)j * 64 * ub;
                for (int c = wave_u; c < nc; c += kWaves, gcol += gstep, lcol += lstep) lds_dma<ub>(rs, lane_off, gcol, lcol);
            }
            if (tail > 0 && lane < tail) {
                int gcol = col0 + wave_u * a.H * E::kBytes + full * 64 * ub;
                unsigned char* lcol = dst + (size_t)wave_u * pitch * E::kBytes + (size_t)full * 64 * ub;
                for (int c = wave_u; c < nc; c += kWaves, gcol += gstep, lcol += lstep) lds_dma<ub>(rs, lane_off, gcol, lcol);
            }
        };

        auto compute = [&](const T* tile, int p, int4 bx, float4 g, float sc) {
            const int kind = box_kind(bx);
            if (kind == 0) return;
            float t = (-1.0f) * x * g.x + y * g.y;
            t += a.uoff;
            const float s = x * g.y + y * g.x;
            const float U = cone_magnification(a.fcd, s);
            const float u = t * U;
            const bool cv = incyl && u >= a.ulo && u <= a.uhi;
            const float uIdx = u + a.upre;
            const int u0 = (int)uIdx;
            const float fu = uIdx - (float)u0;
            const float w = U * U * sc;
            if (kind == 1) {
                const int us = bx.x, vs = bx.y, nr = box_pitch(bx);
                const bool lane_full = !cv || ((zw0 + g.z) * U >= a.vlo && (zw1 + g.z) * U <= a.vhi);
                if (box_fast(bx) && __all(lane_full)) {
                    // every active column keeps all ZPL voxels inside the detector window;
                    // inactive columns read entry 0 with zero weight
                    const float Ue = cv ? U : 0.0f;
                    const float we = cv ? w : 0.0f;
                    const float fue = cv ? fu : 0.0f;
                    const int base = cv ? (u0 - us) * nr - vs : -(int)floorf(a.vpre);
                    const T* tb = tile + base;
                    const float vA = det_row_anchor(zwa, g.z, Ue, a.vpre);
                    // software pipeline in groups of kGrp: the LDS reads of group g+1 are
                    // issued before the arithmetic of group g, and scheduling barriers
                    // keep the compiler from trading those in-flight reads for hoisted
                    // row-index arithmetic
                    constexpr int kGrp = E::kGrp;
                    T e[2][kGrp];
                    float fv[2][kGrp];
                    auto fetch = [&](int g0, int slot) {
#pragma unroll
                        for (int j = 0; j < kGrp; ++j) {
                            const float vI = det_row_index((float)(g0 + j + koff), Ue, vA);
                            e[slot][j] = tb[cvt_flr(vI)];
                            fv[slot][j] = __builtin_amdgcn_fractf(vI);
                        }
                    };
                    fetch(0, 0);
#pragma unroll
                    for (int g0 = 0; g0 < ZPL; g0 += kGrp) {
                        const int slot = (g0 / kGrp) & 1;
                        if (g0 + kGrp < ZPL) fetch(g0 + kGrp, slot ^ 1);
#pragma unroll
                        for (int j = 0; j < kGrp; ++j)
                            acc[g0 + j] = fmaf(we, E::eval(e[slot][j], fue, fv[slot][j]), acc[g0 + j]);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                } else {
                    const T* tb = tile + ((u0 - us) * nr - vs);
                    const float vA = det_row_anchor(zwa, g.z, U, a.vpre);
#pragma unroll
                    for (int k = 0; k < ZPL; ++k) {
                        const float zk = zw0 + (float)k;
                        const float v = (zk + g.z) * U;
                        if (cv && v >= a.vlo && v <= a.vhi) {
                            const float vI = det_row_index((float)(k + koff), U, vA);
                            const int v0 = (int)vI;
                            const T e = tb[v0];
                            acc[k] = fmaf(w, E::eval(e, fu, vI - (float)v0), acc[k]);
                        }
                    }
                }
            } else {
                const T* src = proj + ((size_t)p * a.W + u0) * a.H;
                const float vA = det_row_anchor(zwa, g.z, U, a.vpre);
#pragma unroll
                for (int k = 0; k < ZPL; ++k) {
                    const float zk = zw0 + (float)k;
                    const float v = (zk + g.z) * U;
                    if (cv && v >= a.vlo && v <= a.vhi) {
                        const float vI = det_row_index((float)(k + koff), U, vA);
                        const int v0 = (int)vI;
                        const T e = src[v0];
                        acc[k] = fmaf(w, E::eval(e, fu, vI - (float)v0), acc[k]);
                    }
                }
            }
        };

        // Buffer p&1 holds projection p while it is computed.  The box of p+1 is always
        // produced at least one barrier before it is staged (batches are computed one
        // batch ahead by wave 0).
        // wave 0 fills one batch of 64 projections: footprint boxes and parameters
        auto fill_batch = [&](int slot, int pn) {
            boxes[slot][tid] = pn < a.nproj ? box_for(pn) : make_int4(0, 0, 0, 0);
            float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (pn < a.nproj) {
                g = pp[pn];
                g.w = scales ? scales[pn] : 1.0f;
            }
            params[slot][tid] = g;
        };
        if (tid < 64) fill_batch(0, tid);
        __syncthreads();
        auto buffer = [&](int p) { return smem + (size_t)(p & 1) * a.cap * E::kBytes; };
        if (a.nproj > 0) {
            const int4 b0 = boxes[0][0];
            if (box_kind(b0) == 1) stage(buffer(0), 0, b0);
        }
        // One projection step: stage projection p+1 into `nxt` (LDS DMA) and accumulate
        // projection p from `cur`.  The buffers are disjoint; declaring them __restrict__
        // lets the compiler issue the LDS reads of p while the DMA of p+1 is in flight
        // (otherwise it waits for all outstanding DMA before the first read).
        auto step = [&](const T* __restrict__ cur, unsigned char* __restrict__ nxt, int p, int4 bx, int4 bn,
                        float4 g, float sc) {
            if (box_kind(bn) == 1) stage(nxt, p + 1, bn);
            __builtin_amdgcn_sched_barrier(0);  // the DMA is issued before the compute
            compute(cur, p, bx, g, sc);
        };
        auto uniform_box = [](int4 b) {
            return make_int4(__builtin_amdgcn_readfirstlane(b.x), __builtin_amdgcn_readfirstlane(b.y),
                             __builtin_amdgcn_readfirstlane(b.z), __builtin_amdgcn_readfirstlane(b.w));
        };
        auto uniform_f = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
        auto uniform_f4 = [&](float4 v) {
            return make_float4(uniform_f(v.x), uniform_f(v.y), uniform_f(v.z), uniform_f(v.w));
        };
        // The box and parameters of projection p+1 are read from LDS during step p, so a
        // step starts without waiting for memory other than its own staging DMA.
        int4 bx = uniform_box(boxes[0][0]);
        float4 g = uniform_f4(params[0][0]);
        for (int p = 0; p < a.nproj; ++p) {
            // this wave's DMA for projection p has landed; the barrier publishes every
            // wave's and guarantees that nobody still reads buffer (p+1)&1 or the box
            // batch about to be overwritten
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            __syncthreads();
            if ((p & 63) == 0 && tid < 64) fill_batch(((p >> 6) + 1) & 1, p + 64 + tid);
            const int q = p + 1;
            const int4 bn = q < a.nproj ? uniform_box(boxes[(q >> 6) & 1][q & 63]) : make_int4(0, 0, 0, 0);
            const float4 gn = q < a.nproj ? uniform_f4(params[(q >> 6) & 1][q & 63]) : g;
            step(reinterpret_cast<const T*>(buffer(p)), buffer(p + 1), p, bx, bn, g, g.w);
            bx = bn;
            g = gn;
        }
    }

    if (colvol) {
        float* col = vol + ((size_t)xi * a.ny + yi) * a.nzs;
        if (a.vec4 && zfirst <= 0 && zlast >= ZPL) {
            float4* d4 = reinterpret_cast<float4*>(col + zb);
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
                if (k >= zfirst && k < zlast) {
                    float* d = col + (zb - a.zshift + k);
                    *d = a.accumulate ? *d + acc[k] : acc[k];
                }
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

// Largest footprint box (entries) over a sample of tiles and projections, with the
// same construction as box_for on the device.  Boxes that exceed the resulting
// capacity at run time take the global-memory path, so sampling only affects speed.
int sample_max_box(const ctg_geometry& g, const float4* hp, int nproj, int z0, int nz, int zpl, int dma16,
                   int pmod, int prem) {
    const int ntx = (g.nx + kTX - 1) / kTX, nty = (g.ny + kTY - 1) / kTY, nzb = (nz + zpl - 1) / zpl;
    // about 16 projections: boxes change slowly with the angle, and the capacity keeps a
    // margin above the largest sampled box
    const int pstep = std::max(1, nproj / 16);
    int best = 0;
    const int zblocks[3] = {0, nzb / 2, nzb - 1};
    for (int zi = 0; zi < 3; ++zi) {
        const int zb = zblocks[zi] * zpl;
        const float zw0 = (float)(z0 + zb) - g.z_pre, zw1 = zw0 + (float)(zpl - 1);
        for (int ty = 0; ty < nty; ++ty) {
            for (int tx = 0; tx < ntx; ++tx) {
                const float cxa = (float)(tx * kTX) - g.x_pre;
                const float cxb = (float)std::min(tx * kTX + kTX - 1, g.nx - 1) - g.x_pre;
                const float cya = (float)(ty * kTY) - g.y_pre;
                const float cyb = (float)std::min(ty * kTY + kTY - 1, g.ny - 1) - g.y_pre;
                const float ndx = cxa > 0.0f ? cxa : (cxb < 0.0f ? cxb : 0.0f);
                const float ndy = cya > 0.0f ? cya : (cyb < 0.0f ? cyb : 0.0f);
                if (!(ndx * ndx + ndy * ndy < g.radius * g.radius)) continue;
                for (int p = 0; p < nproj; p += pstep) {
                    const float4 q = hp[p];
                    float umin = INFINITY, umax = -INFINITY, vmin = INFINITY, vmax = -INFINITY;
                    bool bad = false;
                    for (int c = 0; c < 4; ++c) {
                        const float cx = (c & 1) ? cxb : cxa, cy = (c & 2) ? cyb : cya;
                        float t = (-1.0f) * cx * q.x + cy * q.y;
                        t += g.u_offset;
                        const float s = cx * q.y + cy * q.x;
                        const float den = g.fcd - s;
                        if (!(den > 0.0f)) bad = true;
                        const float U = g.fcd / den;
                        const float ui = t * U + g.u_pre;
                        const float va = std::fma(-(zw0 + q.z), U, g.v_pre);
                        const float vb = std::fma(-(zw1 + q.z), U, g.v_pre);
                        umin = std::min(umin, ui); umax = std::max(umax, ui);
                        vmin = std::min(vmin, std::min(va, vb)); vmax = std::max(vmax, std::max(va, vb));
                    }
                    if (bad) continue;
                    const int us = std::max((int)std::floor(umin - kBoxMargin), 0);
                    const int ue = std::min((int)std::floor(umax + kBoxMargin), g.det_width - 1);
                    const int vs = std::max((int)std::floor(vmin - kBoxMargin), 0) & (dma16 ? ~1 : ~0);
                    const int ve = std::min((int)std::floor(vmax + kBoxMargin), g.det_height - 1);
                    if (ue >= us && ve >= vs) {
                        const int nr = ve - vs + 1;
                        const int wr = dma16 ? (nr + 1) & ~1 : nr;
                        const int pitch = wr + ((prem - wr) & (pmod - 1));
                        best = std::max(best, (ue - us + 1) * pitch);
                    }
                }
            }
        }
    }
    return best;
}

// sample_max_box behind a small per-thread cache: callers that backproject the same
// projections into the same slab again (iterative methods, slab by slab) pay the host
// sampling once.  The key hashes the geometry, the projection parameters and the slab;
// a collision can only mis-size the LDS tile, which costs speed, not correctness
// (boxes beyond the capacity read global memory).
int cached_max_box(const ctg_geometry& g, const float4* hp, int nproj, int z0, int nz, int zpl, int dma16, int pmod,
                   int prem) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    const int keys[7] = {nproj, z0, nz, zpl, dma16, pmod, prem};
    mix(&g, sizeof(g));
    mix(keys, sizeof(keys));
    mix(hp, sizeof(float4) * (size_t)nproj);
    struct Entry { uint64_t key; int cap; bool used; };
    static thread_local Entry cache[16] = {};
    static thread_local int next = 0;
    for (const Entry& e : cache)
        if (e.used && e.key == h) return e.cap;
    const int cap = sample_max_box(g, hp, nproj, z0, nz, zpl, dma16, pmod, prem);
    cache[next] = {h, cap, true};
    next = (next + 1) % 16;
    return cap;
}

template <int ZPL>
ctg_status launch_zpl(const BpLaunch& L, BpArgs a, hipStream_t st) {
    const ctg_geometry& g = *L.g;
    const int ebytes = L.store == CTG_STORE_F32 ? 16 : 8;
    a.zshift = L.z0 & (ZPL - 1);
    // whole runs store 16-byte vectors when the slab starts on a run boundary
    a.vec4 = (a.zshift == 0 && L.nz % 4 == 0 && ((uintptr_t)L.d_volume % 16) == 0) ? 1 : 0;
    a.nzb = (L.nz + a.zshift + ZPL - 1) / ZPL;
    // 8-byte entries are staged 16 bytes (two rows) per lane when the detector height is
    // even, so that the first row of every box can be even and each lane's 16 bytes are
    // aligned in global memory: one DMA instruction per column of up to 128 rows instead
    // of two beyond 32.  A DMA instruction costs about the same whatever it moves, and the
    // LDS side needs only 8-byte alignment, so the odd pitch stays.  CTG_BP_DMA16=0
    // selects 4 bytes per lane (tuning).
    const char* dma_env = std::getenv("CTG_BP_DMA16");
    a.dma16 = (ebytes == 8 && g.det_height % 2 == 0 && !(dma_env && std::atoi(dma_env) == 0)) ? 1 : 0;
    // column pitch: rows rounded up to 1 mod 16.  CTG_BP_PITCH="mod,rem" overrides (tuning).
    a.pitch_mod = 16;
    a.pitch_rem = 1;
    // per-box choice of +1 or -1 mod 16 for 16-byte entries (FP32 179.4 -> 174.8 ms at
    // 1024^3 x 720; no change for 8-byte entries, which keep +1).  CTG_BP_PITCH_ADAPT=0/1
    // overrides (tuning).
    a.pitch_adapt = ebytes == 16 ? 1 : 0;
    if (const char* env = std::getenv("CTG_BP_PITCH_ADAPT")) a.pitch_adapt = std::atoi(env) != 0;
    if (const char* env = std::getenv("CTG_BP_PITCH")) {
        int m = 0, r = 0;
        if (std::sscanf(env, "%d,%d", &m, &r) == 2 && m >= 1 && (m & (m - 1)) == 0 && r >= 0) {
            a.pitch_mod = m;
            a.pitch_rem = r % m;
        }
    }
    int cap = L.h_params ? cached_max_box(g, L.h_params, L.nproj, L.z0 - a.zshift, L.nz + a.zshift, ZPL, a.dma16,
                                          a.pitch_mod, a.pitch_rem)
                         : 0;
    cap = cap + cap / 32 + 16;
    cap = std::min(cap, kMaxTileBytes / ebytes);
    cap &= ~7;  // keeps the second buffer and every column start 16-byte aligned
    a.cap = cap;
    const size_t lds = (size_t)2 * cap * ebytes;
    // workgroups in 4 x 4 blocks of tiles (measured against rows of tiles and 8 x 8 and
    // 16 x 16 blocks); CTG_BP_BLOCK=n sets the block edge, 0 for rows (tuning)
    a.blk = 4;
    if (const char* e = std::getenv("CTG_BP_BLOCK")) a.blk = std::max(0, std::min(64, std::atoi(e)));
    a.ntxp = a.blk > 0 ? (a.ntx + a.blk - 1) / a.blk * a.blk : a.ntx;
    a.ntyp = a.blk > 0 ? (a.nty + a.blk - 1) / a.blk * a.blk : a.nty;
    const long long grid = (long long)a.ntxp * a.ntyp * a.nzb;
    if (grid > 0x7fffffffll) return fail(CTG_ERROR_UNSUPPORTED, "backprojection grid too large");
#define CTG_BP_LAUNCH(E, DMA)                                                                           \
    hipLaunchKernelGGL((bp_kernel<ZPL, E, DMA>), dim3((unsigned)grid), dim3(kThreads), lds, st, a, L.d_params, \
                       L.d_scales, L.d_filtered, L.d_volume)
    switch (L.store) {
        case CTG_STORE_F32:
            CTG_BP_LAUNCH(EF32, 16);
            break;
        case CTG_STORE_F16:
            if (a.dma16) CTG_BP_LAUNCH(EF16, 16);
            else CTG_BP_LAUNCH(EF16, 4);
            break;
        case CTG_STORE_BF16:
            if (a.dma16) CTG_BP_LAUNCH(EBF16, 16);
            else CTG_BP_LAUNCH(EBF16, 4);
            break;
        default:
            return fail(CTG_ERROR_INVALID_ARGUMENT, "unknown store type %d", (int)L.store);
    }
#undef CTG_BP_LAUNCH
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

}  // namespace bp_detail

using namespace bp_detail;

ctg_status launch_backproject(const BpLaunch& L, hipStream_t st) {
    const ctg_geometry& g = *L.g;
    if (L.nz <= 0 || g.nx <= 0 || g.ny <= 0) return CTG_OK;
    if (((uintptr_t)L.d_filtered % 16) != 0)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "filtered projections must be 16-byte aligned");
    // The staging DMA addresses one projection through a buffer resource whose size and
    // column offsets are 32-bit: a projection's entries must stay below 2^31 bytes.
    const long long img_bytes = (long long)g.det_width * g.det_height * (L.store == CTG_STORE_F32 ? 16 : 8);
    if (img_bytes > 0x7fffffffll)
        return fail(CTG_ERROR_UNSUPPORTED,
                    "detector %dx%d: %lld bytes of backprojection entries per projection exceed the 2^31-byte limit",
                    g.det_width, g.det_height, img_bytes);
    BpArgs a = make_args(g, L.nproj, L.z0, L.nz, L.accumulate);
    a.ntx = (g.nx + kTX - 1) / kTX;
    a.nty = (g.ny + kTY - 1) / kTY;
    // z per lane: 32 by default; CTG_BP_ZPL=16 selects the shorter run
    const char* env = std::getenv("CTG_BP_ZPL");
    const int zpl = (env && std::atoi(env) == 16) ? 16 : 32;
    return zpl == 16 ? launch_zpl<16>(L, a, st) : launch_zpl<32>(L, a, st);
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
