// Forward projector (exact adjoint of the backprojection), analytic phantom images,
// and volume index-order / byte-order conversion for the writer.
//
// The forward projector has no counterpart in the reference; it transposes the
// voxel-driven backprojection of Ct/CtVolume.cpp:979-1039 term by term: every voxel that
// the backprojection would read through taps (u0,v0),(u0+1,v0),(u0,v0+1),(u0+1,v0+1) with
// weight U^2 and bilinear weights instead adds U^2 * f * weight into those pixels, under
// the same cylinder and window masks.  It is computed as a gather (fp_gather_kernel):
// each detector pixel sums the taps that land on it.
#include "ctg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace ctg {
namespace {

struct FpArgs {
    int W, H, nx, ny, nzs, z0, nproj;
    int zstride;  // slices per voxel column of the volume buffer (nzs of them from z0 are projected)
    float xpre, ypre, zpre, upre, vpre;
    float r2, fcd, uoff, ulo, uhi, vlo, vhi;
};

// Gather form of the projector.  No global or LDS float atomics: scattering splats
// costs LDS float atomics (about 30x an LDS read here) or global ones that contend
// across workgroups.  A workgroup owns one detector column c of one projection.
//  - The voxel columns that can feed column c are exactly those whose ray lands within
//    [c - 1, c + 1) on the detector (u0 = c takes weight 1 - fu, u0 = c - 1 takes fu).
//    Along a line of constant outer coordinate u is monotonic in the inner one, so the
//    two rays from the source through u = c - 1 and u = c + 1 (plus 0.05 pixel) bound
//    the candidates up to rounding; a 0.01-voxel margin covers that.
//  - The per-column work (cylinder and window tests, U, the tap weight) is done once per
//    candidate: the lanes fill a compacted list of candidates in LDS, in walk order.
//    The v-window test (v = (z + h) U inside [vLo, vHi]) is monotonic in z for U > 0, so
//    the fill also turns it, intersected with the slab, into an exact interval of z
//    (found by evaluating the backprojection's own expression at the boundaries);
//    candidates whose interval is empty are dropped.
//  - The waves then walk items of up to four consecutive 64-voxel z chunks of a
//    candidate (one voxel per lane and chunk), dealt round-robin, with the
//    backprojection's anchored row index, and add the two row taps of each voxel into
//    the wave's own row accumulators in LDS.  Chunks start at multiples of 32 in grid z,
//    so a lane's row anchor and its offset from it are fixed per lane.  Rows of voxels k
//    slices apart differ by at least k*U - 1, so lanes are split by residue modulo
//    ceil(2/U), rounded up to 2 or 4 (which divide the chunk length), between two row
//    copies per wave and, for 4, two passes: taps that meet never share a row, and the
//    read-modify-writes need no atomics.
//  - At the end the four waves' rows are summed into the image.
// Every (voxel, tap) pair is added exactly once, and the order of all additions is fixed:
// the result does not depend on scheduling.
constexpr int kGThreads = 256, kGWaves = kGThreads / 64;
constexpr int kSlots = 512;  // candidate slots per pass (16 B each)
constexpr int kMaxRowsFp = 4096;  // detector rows held in LDS per wave
constexpr int kFpZBits = 14;      // z interval bounds in a candidate record (relative to the run base)
constexpr int kFpMaxSlab = (1 << kFpZBits) - 2 * kRowAnchor;  // slices per launch

// Raw buffer resource over `bytes` bytes of global memory: a 4-byte load at an offset at
// or past the end returns zero.  The builtins exist only in the device pass.
#if defined(__HIP_DEVICE_COMPILE__)
using vbuf = __amdgpu_buffer_rsrc_t;
__device__ inline vbuf make_vbuf(const float* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, bytes, 0x00020000);
}
__device__ inline float vbuf_load(vbuf rs, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, (int)off, 0, 0));
}
// a + b in a vector register, opaque to the compiler, so that a constant part of the
// sum is not moved into a load's immediate offset
__device__ inline unsigned vadd_sv(int a, unsigned b) {
    unsigned r;
    asm("v_add_u32 %0, %1, %2" : "=v"(r) : "s"(a), "v"(b));
    return r;
}
#else
__device__ inline unsigned vadd_sv(int a, unsigned b) { return (unsigned)a + b; }
struct vbuf {};
__device__ inline vbuf make_vbuf(const float*, int) { return {}; }
__device__ inline float vbuf_load(vbuf, unsigned) { return 0.0f; }
#endif

__global__ __launch_bounds__(kGThreads) void fp_gather_kernel(FpArgs a, const float4* __restrict__ pp,
                                                              const float* __restrict__ vol,
                                                              float* __restrict__ proj) {
    __shared__ int4 slots[kSlots];
    __shared__ int wave_cnt[kGWaves];
    // kGWaves x 2 copies of (H + 4) rows: row r at index r + 2; rows -2, -1, H, H + 1
    // absorb taps off the image.  Even lanes of a wave add into its first copy, odd
    // lanes into its second.
    extern __shared__ float rows_lds[];
    const int c = blockIdx.x;
    const int p = blockIdx.y;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int H = a.H, pitch = a.H + 4;
    float* myrows = rows_lds + (size_t)(2 * wv + (lane & 1)) * pitch;
    for (int i = tid; i < 2 * kGWaves * pitch; i += kGThreads) rows_lds[i] = 0.0f;
    __syncthreads();

    const float4 g = pp[p];
    const float sn = g.x, cs = g.y, h = g.z;
    // source and detector points in world xy (voxel units; the virtual detector passes
    // through the axis, lateral coordinate u)
    const float Sx = a.fcd * cs + a.uoff * sn, Sy = a.fcd * sn - a.uoff * cs;
    const float ulo = (float)c - 1.05f - a.upre, uhi = (float)c + 1.05f - a.upre;
    const float q0x = -(ulo - a.uoff) * sn, q0y = (ulo - a.uoff) * cs;
    const float q1x = -(uhi - a.uoff) * sn, q1y = (uhi - a.uoff) * cs;
    const float umid = (float)c - a.upre - a.uoff;
    const bool xmajor = fabsf(-umid * sn - Sx) >= fabsf(umid * cs - Sy);
    const int n_out = xmajor ? a.nx : a.ny, n_in = xmajor ? a.ny : a.nx;
    const float pre_out = xmajor ? a.xpre : a.ypre, pre_in = xmajor ? a.ypre : a.xpre;
    const float So = xmajor ? Sx : Sy, Si = xmajor ? Sy : Sx;
    const float k0 = ((xmajor ? q0y : q0x) - Si) / ((xmajor ? q0x : q0y) - So);
    const float k1 = ((xmajor ? q1y : q1x) - Si) / ((xmajor ? q1x : q1y) - So);
    const float R = sqrtf(a.r2);
    const int o_lo = max(0, (int)floorf(pre_out - R) - 1), o_hi = min(n_out - 1, (int)ceilf(pre_out + R) + 1);
    const int zend = a.z0 + a.nzs - 1;  // last grid z index of the slab
    const int zrb = a.z0 & ~(kRowAnchor - 1);  // run base: chunks start at zrb + 64 k

    // lateral candidates of depth step o: [jl, jh]
    auto jrange = [&](int o, int& jl, int& jh) {
        const float ow = (float)o - pre_out;
        const float i0 = Si + (ow - So) * k0, i1 = Si + (ow - So) * k1;
        jl = max(0, (int)ceilf(fminf(i0, i1) + pre_in - 0.01f));
        jh = min(n_in - 1, (int)floorf(fmaxf(i0, i1) + pre_in + 0.01f));
    };
    // slots per depth step: the wedge width |i1 - i0| is linear in the depth and the
    // source lies outside the volume, so it peaks at an end of the depth range; the
    // count of integers in a window of width w + 0.02 is at most floor(w) + 2
    auto width = [&](int o) {
        const float ow = (float)o - pre_out;
        return fabsf((ow - So) * (k1 - k0));
    };
    const int jm = (int)fmaxf(width(o_lo), width(o_hi)) + 2;
    const int jm_pass = min(jm, kSlots);
    const int per_pass = max(1, kSlots / jm_pass);

    for (int ob = o_lo; ob <= o_hi; ob += per_pass) {
        const int no = min(per_pass, o_hi - ob + 1);
        for (int jb = 0; jb < jm; jb += jm_pass) {
            const int nj = min(jm_pass, jm - jb);
            const int nslot = no * nj;
            // fill: slot (o - ob) * nj + jj covers lateral candidate jl + jb + jj of depth o;
            // the candidates that pass the tests are stored compacted, in slot order
            // (a ballot prefix within each wave, wave totals through LDS)
            int ncand = 0;
            for (int sbase = 0; sbase < nslot; sbase += kGThreads) {
                const int sidx = sbase + tid;
                const int oi = sidx / nj, jj = sidx - oi * nj;
                const int o = ob + oi;
                int jl = 0, jh = -1;
                if (sidx < nslot) jrange(o, jl, jh);
                const int j = jl + jb + jj;
                int4 rec = make_int4(-1, 0, 0, 0);
                if (sidx < nslot && j <= jh) {
                    const int xi = xmajor ? o : j, yi = xmajor ? j : o;
                    // per-voxel arithmetic identical to the backprojection's
                    const float x = (float)xi - a.xpre, y = (float)yi - a.ypre;
                    if (x * x + y * y < a.r2) {
                        float t = (-1.0f) * x * g.x + y * g.y;
                        t += a.uoff;
                        const float s = x * g.y + y * g.x;
                        const float U = cone_magnification(a.fcd, s);
                        const float u = t * U;
                        if (u >= a.ulo && u <= a.uhi) {
                            const float uIdx = u + a.upre;
                            const int u0 = (int)uIdx;
                            if (u0 == c || u0 == c - 1) {
                                const float fu = uIdx - (float)u0;
                                const float wc = U * U * (u0 == c ? 1.0f - fu : fu);
                                // z interval of the window test v = (zk + h) U in [vLo, vHi]:
                                // v is non-decreasing in z (U > 0, monotone rounding), so the
                                // passing z form an interval; estimate its ends, then settle
                                // them with the backprojection's own expression
                                // (the estimate is within about 1e-2 of the true end for grids
                                // of up to 65536 slices, so its ceiling or floor is off by at
                                // most one)
                                auto v_of = [&](int zi) { const float zk = (float)zi - a.zpre; return (zk + h) * U; };
                                const float iU = __builtin_amdgcn_rcpf(U);
                                const float zlo_f = fminf(fmaxf(a.vlo * iU - h + a.zpre, (float)(a.z0 - 4)), (float)(zend + 4));
                                const float zhi_f = fminf(fmaxf(a.vhi * iU - h + a.zpre, (float)(a.z0 - 4)), (float)(zend + 4));
                                int lo = (int)ceilf(zlo_f), hi = (int)floorf(zhi_f);
#pragma unroll
                                for (int it = 0; it < 2; ++it) {
                                    if (lo > a.z0 && v_of(lo - 1) >= a.vlo) --lo;
                                    if (hi < zend && v_of(hi + 1) <= a.vhi) ++hi;
                                }
#pragma unroll
                                for (int it = 0; it < 2; ++it) {
                                    if (lo <= zend && !(v_of(lo) >= a.vlo)) ++lo;
                                    if (hi >= a.z0 && !(v_of(hi) <= a.vhi)) --hi;
                                }
                                lo = max(lo, a.z0);
                                hi = min(hi, zend);
                                if (lo <= hi) {
                                    // ceil(2/U) passes, rounded up to a power of two (a
                                    // multiple of ceil(2/U) passes is as safe); U > 1/1.9
                                    // keeps it at most 4
                                    const int np0 = (int)ceilf(2.0f * __builtin_amdgcn_rcpf(U) + 1e-3f);
                                    const int npass = np0 <= 1 ? 1 : (np0 == 2 ? 2 : 4);
                                    rec = make_int4(xi * a.ny + yi, __float_as_int(U), __float_as_int(wc),
                                                    (lo - zrb) | ((hi - zrb) << kFpZBits) | (npass << (2 * kFpZBits)));
                                }
                            }
                        }
                    }
                }
                const unsigned long long m = __ballot(rec.x >= 0);
                const int below = __popcll(m & ((1ull << lane) - 1ull));
                if (lane == 0) wave_cnt[wv] = __popcll(m);
                __syncthreads();
                int base = ncand;
                for (int w = 0; w < wv; ++w) base += wave_cnt[w];
                int total = 0;
                for (int w = 0; w < kGWaves; ++w) total += wave_cnt[w];
                if (rec.x >= 0) slots[base + below] = rec;
                ncand += total;
                __syncthreads();  // wave_cnt is rewritten by the next round
            }
            // walk: a voxel's taps go to rows v0, v0 + 1 of the wave's own accumulators.
            // Chunks start at zrb + 64 k (multiples of 32 in grid z), so lane l's voxel
            // lies (l & 31) slices past its row anchor, which is (l & 32) slices past the
            // chunk start.  Lanes outside the candidate's z interval load zero and send
            // their (zero) taps to the absorbing rows H, H + 1: their own rows could meet
            // an active lane's, because truncation folds (-1, 1) onto row 0.  A lane whose
            // voxel is zero may do the same, as its taps are zero too.
            const float lane_k = (float)(lane & 31), anchor_off = (float)(lane & 32);
            const float* colbase = vol + (zrb - a.z0);  // voxel (column, z) at column * zstride + z - zrb
            auto tap = [&](float zcw, float xv, float U, float wc, float& ta, float& tb) -> int {
                const float vA = det_row_anchor(zcw + anchor_off, h, U, a.vpre);
                const float vI = det_row_index(lane_k, U, vA);
                const int v0 = (int)vI;
                const float fv = vI - (float)v0;
                const float val = wc * xv;
                tb = fv * val;
                ta = val - tb;
                return (xv != 0.0f ? v0 : H) + 2;  // in-window rows are 0 .. H - 2 (checked on the host)
            };
            // a candidate record: column, U, weight, and z interval [lo, hi] relative to
            // zrb with the residue pass count in the top bits.  The column's voxels in
            // [lo, hi] are read through a buffer resource covering exactly them: a load
            // outside it returns zero, which is the mask of the window and the slab.
            struct Item {
                vbuf rs;
                float U, wc;
                int lo, npass;
            };
            // past the last candidate an item still issues its loads, over an empty
            // resource (they return zero without touching memory): every slot then has
            // the same loads in flight and the walk waits for exactly its own
            auto item_of = [&](int ci, bool valid) {
                const int4 r = slots[valid ? ci : 0];
                Item it;
                const int w = __builtin_amdgcn_readfirstlane(r.w);
                it.lo = w & ((1 << kFpZBits) - 1);
                const int hi = (w >> kFpZBits) & ((1 << kFpZBits) - 1);
                it.rs = make_vbuf(colbase + (size_t)__builtin_amdgcn_readfirstlane(r.x) * a.zstride + it.lo,
                                  valid ? (hi - it.lo + 1) * (int)sizeof(float) : 0);
                it.U = __int_as_float(__builtin_amdgcn_readfirstlane(r.y));
                it.wc = __int_as_float(__builtin_amdgcn_readfirstlane(r.z));
                it.npass = (unsigned)w >> (2 * kFpZBits);
                return it;
            };
            // this lane's voxel of the chunk starting at zc (relative to zrb).  The byte
            // offset is formed in one register per chunk: lanes before the interval get
            // an offset that wraps to a huge value, which the resource's range check
            // rejects.  With the chunk's distance in the load's immediate offset instead,
            // a wrapped register offset plus the immediate was measured to read zeros
            // where the sum lands at the start of the resource.
            const unsigned lane4 = (unsigned)lane * (unsigned)sizeof(float);
            auto load = [&](const Item& it, int zc) -> float {
                return vbuf_load(it.rs, vadd_sv((zc - it.lo) * (int)sizeof(float), lane4));
            };
            auto zworld = [&](int zc) { return (float)(zrb + zc) - a.zpre; };
            const int nzr = zend - zrb + 1;  // slices from the run base to the slab end
            if (ncand == 0) {
                __syncthreads();
                continue;
            }
            // Items are (candidate, group of G consecutive 64-voxel chunks), in walk order,
            // dealt to the four waves round-robin: item q = ci * ngrp + j goes to wave
            // q mod 4.  With n = ceil(2/U) rounded up to 1, 2 or 4, voxels n or more slices
            // apart have rows at least two apart.  Lanes of one parity add into one of the
            // wave's two row copies, and for n = 4 in two passes by (lane >> 1) & 1: lanes
            // l and l' of any two chunks of an item (chunks are 64 slices apart) meet in a
            // copy and pass only if their z differ by a nonzero multiple of 2 (n <= 2) or
            // 4, so a pass's read-modify-writes touch distinct rows.  Each wave walks its
            // items through S slots: the voxels of the next S - 1 items are in flight
            // while the current one is processed (the slots are named, not rotated, so that
            // no in-flight register is copied).
            const int nch = (nzr + 63) / 64;
            auto walk = [&](auto g_const) {
                constexpr int G = decltype(g_const)::value;
                constexpr int S = G == 1 ? 3 : 2;
                const int ngrp = (nch + G - 1) / G;
                const int dci = kGWaves / ngrp, dj = kGWaves - dci * ngrp;
                int ci_f = 0, j_f = wv;  // next item to fetch
                while (j_f >= ngrp) { j_f -= ngrp; ++ci_f; }
                struct Slot { Item it; int zc; float x[G]; };
                auto next = [&](Slot& s) -> bool {
                    const bool valid = ci_f < ncand;
                    s.it = item_of(ci_f, valid);
                    s.zc = 64 * G * j_f;
#pragma unroll
                    for (int k = 0; k < G; ++k) s.x[k] = load(s.it, s.zc + 64 * k);
                    ci_f += dci;
                    j_f += dj;
                    if (j_f >= ngrp) { j_f -= ngrp; ++ci_f; }
                    __builtin_amdgcn_sched_barrier(0);  // keep the slots' loads in walk order
                    return valid;
                };
                auto process = [&](const Slot& s) {
                    float ta[G], tb[G];
                    int ix[G];
                    // world z of the chunks: exact half-integers, so chunk k's is chunk 0's + 64 k
                    const float zw0 = zworld(s.zc);
#pragma unroll
                    for (int k = 0; k < G; ++k) ix[k] = tap(zw0 + (float)(64 * k), s.x[k], s.it.U, s.it.wc, ta[k], tb[k]);
                    // loads before stores: the 2G rows of a pass are distinct
                    auto rmw = [&]() {
                        float r0[G], r1[G];
#pragma unroll
                        for (int k = 0; k < G; ++k) { r0[k] = myrows[ix[k]]; r1[k] = myrows[ix[k] + 1]; }
#pragma unroll
                        for (int k = 0; k < G; ++k) { myrows[ix[k]] = r0[k] + ta[k]; myrows[ix[k] + 1] = r1[k] + tb[k]; }
                    };
                    if (s.it.npass <= 2) {
                        rmw();  // the lane parity picks the copy: one pass
                    } else {
                        const int pid = (lane >> 1) & 1;
                        for (int ps = 0; ps < 2; ++ps) {
                            if (pid == ps) rmw();
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                };
                if constexpr (S == 3) {
                    Slot sa, sb, sc;
                    bool ha = next(sa), hb = next(sb), hc = next(sc);
                    while (ha) {
                        process(sa);
                        ha = next(sa);
                        if (!hb) break;
                        process(sb);
                        hb = next(sb);
                        if (!hc) break;
                        process(sc);
                        hc = next(sc);
                    }
                } else {
                    Slot sa, sb;
                    bool ha = next(sa), hb = next(sb);
                    while (ha) {
                        process(sa);
                        ha = next(sa);
                        if (!hb) break;
                        process(sb);
                        hb = next(sb);
                    }
                }
            };
            if (nch >= 4) walk(std::integral_constant<int, 4>{});
            else if (nch == 3) walk(std::integral_constant<int, 3>{});
            else if (nch == 2) walk(std::integral_constant<int, 2>{});
            else walk(std::integral_constant<int, 1>{});
            __syncthreads();
        }
    }
    // both row copies of the four waves, summed in a fixed order, into the image
    for (int r = tid; r < H; r += kGThreads) {
        float sum = rows_lds[r + 2];
        for (int w = 1; w < 2 * kGWaves; ++w) sum += rows_lds[(size_t)w * pitch + r + 2];
        proj[((size_t)p * H + r) * a.W + c] += sum;
    }
}

// -------------------------------------------------------- entry-space projector
//
// The backprojection evaluates a voxel's value from ONE coefficient entry (A, dU, dV, dUV)
// at (u0, v0): val = A + fu dU + fv (dV + fu dUV), weighted by w = U^2.  Its transpose
// therefore adds (w x, w x fu, w x fv, w x fu fv) into one entry gradient per (voxel,
// projection), and the transpose of the layout (A = q00, dU = q10 - q00, dV = q01 - q00,
// dUV = q11 - q10 - q01 + q00) turns entry gradients into pixels:
//     q00 += gA - gdU - gdV + gdUV,  q01 += gdV - gdUV,  q10 += gdU - gdUV,  q11 += gdUV.
// Every voxel is visited once per projection (by the workgroup of its own u0), with one
// LDS read and one write of 16 bytes, instead of twice with two row taps each.
//
// fp_entry_kernel<G>: a workgroup owns entry column c of one projection within one
// window of 64 G slices (grid z [wz0, wz0 + 64 G), wz0 - zrb a multiple of 64).
//  - Candidates are the voxel columns with u0 == c that pass the cylinder and u window
//    tests, between the two fan rays through u = c - 0.05 and u = c + 1.05 (plus a
//    0.01-voxel margin).  The lanes test the wedge's (depth, lateral) slots and compact
//    the survivors (column, U, fu) into LDS; then one lane per survivor settles its z
//    interval of the v-window test (exactly, with the backprojection's expression),
//    intersected with the slab and the window, and writes the final record.
//  - A candidate is one item of G chunks of 64 slices: lane l takes slices
//    wz0 + 64 k + l, whose row anchor is (l & 32) slices past the chunk start and whose
//    offset from it is l & 31, as in the backprojection.  The voxels come through a
//    buffer resource covering the candidate's interval (zero outside it, which is the
//    window, slab and v-window mask).
//  - Each wave accumulates into two private copies of the window's entry rows (even and
//    odd lanes).  Lanes two slices apart have rows at least 2U - 1e-3 > 1 apart (U > 2/3
//    whenever the cylinder radius is below half the FCD, checked on the host), and chunks
//    are 64 slices apart, so no two lanes of one copy touch one entry in an instruction
//    and the read-modify-writes need no atomics.  The host bounds the rows that any slice
//    of the window can reach for every candidate U and height, [rbase, rbase + span), so
//    lanes outside their interval (adding zero) stay inside their own copy.
//  - At the end the eight copies are summed in a fixed order and turned into pixel
//    partials: own column c (q00, q01) and the next column (q10, q11), added into the
//    workgroup's own float2 column ps[p][c][*] of a column-major buffer.  Windows run one
//    after another on one stream; fp_finish_kernel then adds ps[p][c][r].x +
//    ps[p][c-1][r].y into the row-major image.
// The order of every addition is fixed: the result does not depend on scheduling.
constexpr int kESlots = 256;
constexpr int kEMaxSpan = 256;  // entry rows per copy

struct FpWin {
    int wz0;     // first grid slice of the window (a multiple of 64 past the run base)
    int rbase;   // first entry row any voxel of the window can reach
    int span;    // rows per copy
    int stride;  // entries between copies: span rounded up to 8 mod 16 (see the kernel)
};

__device__ inline int flr_i32(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

template <int G>
__global__ __launch_bounds__(kGThreads) void fp_entry_kernel(FpArgs a, FpWin win, const float4* __restrict__ pp,
                                                             const float* __restrict__ vol,
                                                             float2* __restrict__ ps) {
    // column, U, w fu and the z interval relative to the run base (w = U^2 is recomputed
    // from U): one 16-byte LDS read per item
    __shared__ int4 slots[kESlots];
    __shared__ int wave_cnt[kGWaves];
    // [wave][parity][row], copies `stride` entries apart (a multiple of 16 entries unless
    // the host is told otherwise).  On gfx950 a ds_write_b128 occupies the LDS for about
    // 13.5 cycles per wave instruction whatever its lane mask and a conflict-free
    // ds_read_b128 about 4 (16 lanes per cycle); of the copy offsets 0, 2, 5, 6, 8 and 13
    // entries modulo 16, 0 was the fastest on the 512^3 x 360 workload.
    extern __shared__ float4 acc_lds[];
    const int c = blockIdx.x;
    const int p = blockIdx.y;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int span = win.span, stride = win.stride;
    for (int i = tid; i < 2 * kGWaves * stride; i += kGThreads) acc_lds[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    const float4 g = pp[p];
    const float sn = g.x, cs = g.y, h = g.z;
    const float Sx = a.fcd * cs + a.uoff * sn, Sy = a.fcd * sn - a.uoff * cs;
    const float ulo = (float)c - 0.05f - a.upre, uhi = (float)c + 1.05f - a.upre;
    const float q0x = -(ulo - a.uoff) * sn, q0y = (ulo - a.uoff) * cs;
    const float q1x = -(uhi - a.uoff) * sn, q1y = (uhi - a.uoff) * cs;
    const float umid = (float)c + 0.5f - a.upre - a.uoff;
    const bool xmajor = fabsf(-umid * sn - Sx) >= fabsf(umid * cs - Sy);
    const int n_out = xmajor ? a.nx : a.ny, n_in = xmajor ? a.ny : a.nx;
    const float pre_out = xmajor ? a.xpre : a.ypre, pre_in = xmajor ? a.ypre : a.xpre;
    const float So = xmajor ? Sx : Sy, Si = xmajor ? Sy : Sx;
    const float k0 = ((xmajor ? q0y : q0x) - Si) / ((xmajor ? q0x : q0y) - So);
    const float k1 = ((xmajor ? q1y : q1x) - Si) / ((xmajor ? q1x : q1y) - So);
    const float R = sqrtf(a.r2);
    const int o_lo = max(0, (int)floorf(pre_out - R) - 1), o_hi = min(n_out - 1, (int)ceilf(pre_out + R) + 1);
    const int zend = a.z0 + a.nzs - 1;
    const int zrb = a.z0 & ~(kRowAnchor - 1);
    // the slices of this window inside the slab
    const int wlo = max(a.z0, win.wz0), whi = min(zend, win.wz0 + 64 * G - 1);

    auto jrange = [&](int o, int& jl, int& jh) {
        const float ow = (float)o - pre_out;
        const float i0 = Si + (ow - So) * k0, i1 = Si + (ow - So) * k1;
        jl = max(0, (int)ceilf(fminf(i0, i1) + pre_in - 0.01f));
        jh = min(n_in - 1, (int)floorf(fmaxf(i0, i1) + pre_in + 0.01f));
    };
    // lateral slots per depth step: the wedge width |i1 - i0| is linear in the depth and
    // the source lies outside the volume, so it peaks at an end of the depth range; a
    // window of width w + 0.02 holds at most floor(w) + 2 integers
    auto width = [&](int o) {
        const float ow = (float)o - pre_out;
        return fabsf((ow - So) * (k1 - k0));
    };
    const int jm = (int)fmaxf(width(o_lo), width(o_hi)) + 2;
    const int nslot = (o_hi - o_lo + 1) * jm;

    // per lane and chunk: world z of the slice's row anchor plus the height (the
    // backprojection's zwa + h), and the slice's offset from the anchor
    float zh[G];
#pragma unroll
    for (int k = 0; k < G; ++k) zh[k] = ((float)(win.wz0 + 64 * k + (lane & 32)) - a.zpre) + h;
    const float lane_k = (float)(lane & 31);
    // LDS byte address of entry row v0 in this lane's copy: v0 * 16 + lane_base
    int lane_base = (wv * 2 + (lane & 1)) * stride * 16 - win.rbase * 16;
    // opaque to the optimiser, so that v0 * 16 + lane_base stays one shift-and-add per
    // voxel instead of being re-associated around rbase
    asm volatile("" : "+v"(lane_base));
    char* acc_bytes = reinterpret_cast<char*>(acc_lds);
    const float* colbase = vol + (zrb - a.z0);
    const int zc = win.wz0 - zrb;  // the window's first slice relative to the run base
    const unsigned lane4 = (unsigned)lane * (unsigned)sizeof(float);
    __syncthreads();

    int ncand = 0;
    for (int sbase = 0; sbase < nslot && wlo <= whi; sbase += kGThreads) {
        // fill, step 1: the cheap tests for 256 slots, survivors compacted (ballot prefix
        // within each wave, wave totals through LDS)
        const int sidx = sbase + tid;
        int4 rec = make_int4(-1, 0, 0, 0);
        if (sidx < nslot) {
            const int oi = sidx / jm, jj = sidx - oi * jm;
            const int o = o_lo + oi;
            int jl, jh;
            jrange(o, jl, jh);
            const int j = jl + jj;
            if (j <= jh) {
                const int xi = xmajor ? o : j, yi = xmajor ? j : o;
                // per-voxel arithmetic identical to the backprojection's
                const float x = (float)xi - a.xpre, y = (float)yi - a.ypre;
                if (x * x + y * y < a.r2) {
                    float t = (-1.0f) * x * g.x + y * g.y;
                    t += a.uoff;
                    const float s = x * g.y + y * g.x;
                    const float U = cone_magnification(a.fcd, s);
                    const float u = t * U;
                    if (u >= a.ulo && u <= a.uhi) {
                        const float uIdx = u + a.upre;
                        const int u0 = (int)uIdx;
                        if (u0 == c) rec = make_int4(xi * a.ny + yi, __float_as_int(U), __float_as_int(uIdx - (float)u0), 0);
                    }
                }
            }
        }
        {
            const unsigned long long m = __ballot(rec.x >= 0);
            const int below = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wave_cnt[wv] = __popcll(m);
            __syncthreads();
            int base = ncand;
            for (int w = 0; w < wv; ++w) base += wave_cnt[w];
            int total = 0;
            for (int w = 0; w < kGWaves; ++w) total += wave_cnt[w];
            if (rec.x >= 0) slots[base + below] = rec;
            __syncthreads();  // the slots are complete; wave_cnt is rewritten below
            // fill, step 2: one lane per new survivor settles its z interval and writes its
            // final record (candidates with an empty interval keep an empty one)
            const int ci = ncand + tid;
            if (tid < total) {
                const int4 r = slots[ci];
                const float U = __int_as_float(r.y), fu = __int_as_float(r.z);
                auto v_of = [&](int zi) { const float zk = (float)zi - a.zpre; return (zk + h) * U; };
                int lo = wlo, hi = whi;
                if (!(v_of(wlo) >= a.vlo && v_of(whi) <= a.vhi)) {
                    // the passing z of v = (zk + h) U in [vLo, vHi] form an interval (v is
                    // non-decreasing in z): estimate its ends, then settle them with the
                    // backprojection's own expression
                    const float iU = __builtin_amdgcn_rcpf(U);
                    const float zlo_f = fminf(fmaxf(a.vlo * iU - h + a.zpre, (float)(wlo - 4)), (float)(whi + 4));
                    const float zhi_f = fminf(fmaxf(a.vhi * iU - h + a.zpre, (float)(wlo - 4)), (float)(whi + 4));
                    lo = (int)ceilf(zlo_f);
                    hi = (int)floorf(zhi_f);
#pragma unroll
                    for (int it = 0; it < 2; ++it) {
                        if (lo > wlo && v_of(lo - 1) >= a.vlo) --lo;
                        if (hi < whi && v_of(hi + 1) <= a.vhi) ++hi;
                    }
#pragma unroll
                    for (int it = 0; it < 2; ++it) {
                        if (lo <= whi && !(v_of(lo) >= a.vlo)) ++lo;
                        if (hi >= wlo && !(v_of(hi) <= a.vhi)) --hi;
                    }
                    lo = max(lo, wlo);
                    hi = min(hi, whi);
                }
                const float w = U * U;
                // an empty interval [lo, hi] with hi < lo becomes a zero-length resource
                slots[ci] = make_int4(r.x, r.y, __float_as_int(w * fu),
                                      (lo - zrb) | (max(hi, lo - 1) - zrb + 1) << kFpZBits);
            }
            ncand += total;
            __syncthreads();  // records complete; wave_cnt free again
        }
        // walk when another fill round might not fit, and after the last round
        if (ncand > kESlots - kGThreads || sbase + kGThreads >= nslot) {
            struct Item {
                vbuf rs;
                float U, w, wfu;
                int lo;
            };
            // past the last candidate an item still issues its loads, over an empty
            // resource (they return zero without touching memory)
            auto item_of = [&](int ci, bool valid) {
                const int4 r = slots[valid ? ci : 0];
                const int zb = __builtin_amdgcn_readfirstlane(r.w);
                Item it;
                it.lo = zb & ((1 << kFpZBits) - 1);
                const int hi1 = (zb >> kFpZBits) & ((1 << kFpZBits) - 1);  // one past the last slice
                it.rs = make_vbuf(colbase + (size_t)__builtin_amdgcn_readfirstlane(r.x) * a.zstride + it.lo,
                                  valid ? (hi1 - it.lo) * (int)sizeof(float) : 0);
                // the same value in every lane: kept in vector registers (operands of the
                // per-voxel arithmetic), where moving them to scalars would cost a
                // readfirstlane each
                it.U = __int_as_float(r.y);
                it.w = it.U * it.U;  // as in the fill: the weight is U^2
                it.wfu = __int_as_float(r.z);
                return it;
            };
            // Candidate ci goes to wave ci mod 4.  Three named slots keep the next two
            // candidates' voxels in flight while one is accumulated.
            struct Slot { Item it; float x[G]; };
            int ci_f = wv;
            auto next = [&](Slot& sl) -> bool {
                const bool valid = ci_f < ncand;
                sl.it = item_of(ci_f, valid);
#pragma unroll
                for (int k = 0; k < G; ++k)
                    sl.x[k] = vbuf_load(sl.it.rs, vadd_sv((zc + 64 * k - sl.it.lo) * (int)sizeof(float), lane4));
                ci_f += kGWaves;
                __builtin_amdgcn_sched_barrier(0);  // keep the slots' loads in walk order
                return valid;
            };
            // The G chunks of one item touch distinct entries of each copy (same-parity
            // lanes of the item are two or more slices apart, so their rows differ), so
            // all G reads are issued before the first write: one LDS round trip per item
            // instead of one per chunk.
            auto process = [&](const Slot& sl) {
                float4* e[G];
                float4 v[G];
                float fv[G];
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const float vA = fmaf(-zh[k], sl.it.U, a.vpre);  // det_row_anchor(zwa, h, U, vpre)
                    const float vI = det_row_index(lane_k, sl.it.U, vA);
                    const int v0 = flr_i32(vI);
                    fv[k] = __builtin_amdgcn_fractf(vI);
                    e[k] = reinterpret_cast<float4*>(acc_bytes + (v0 * 16 + lane_base));
                    v[k] = *e[k];
                }
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const float t = sl.it.w * sl.x[k], tu = sl.it.wfu * sl.x[k];
                    v[k].x += t;
                    v[k].y += tu;
                    v[k].z = fmaf(t, fv[k], v[k].z);
                    v[k].w = fmaf(tu, fv[k], v[k].w);
                    *e[k] = v[k];
                }
            };
            Slot sa, sb, sc;
            bool ha = next(sa), hb = next(sb), hc = next(sc);
            while (ha) {
                process(sa);
                ha = next(sa);
                if (!hb) break;
                process(sb);
                hb = next(sb);
                if (!hc) break;
                process(sc);
                hc = next(sc);
            }
            ncand = 0;
            __syncthreads();  // every wave is done with the slots
        }
    }
    // sum the eight copies per entry row (waves in order, even before odd) and turn the
    // entry gradients into pixel partials: own column (q00 at row v, q01 at v + 1) and
    // the next column (q10 at v, q11 at v + 1); pixel row rbase + r takes entry rows r
    // and r - 1
    float2* col = ps + ((size_t)p * a.W + c) * a.H;
    auto entry = [&](int r) {
        float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (r >= 0 && r < span) {
            for (int w = 0; w < kGWaves; ++w)
                for (int par = 0; par < 2; ++par) {
                    const float4 e = acc_lds[(w * 2 + par) * stride + r];
                    sum.x += e.x; sum.y += e.y; sum.z += e.z; sum.w += e.w;
                }
        }
        return sum;
    };
    for (int r = tid; r <= span; r += kGThreads) {
        const int row = win.rbase + r;
        if (row < 0 || row >= a.H) continue;
        const float4 e0 = entry(r), e1 = entry(r - 1);  // entries at rows v = row and v = row - 1
        const float own = ((e0.x - e0.y) - e0.z + e0.w) + (e1.z - e1.w);
        const float side = (e0.y - e0.w) + e1.w;
        float2 o = col[row];
        o.x += own;
        o.y += side;
        col[row] = o;
    }
}

// img[p][r][c] += ps[p][c][r].x + ps[p][c - 1][r].y: column-major partials into the
// row-major images, through a 32 x 33 tile per (row, column) block.
__global__ __launch_bounds__(256) void fp_finish_kernel(const float2* __restrict__ ps, int W, int H,
                                                        float* __restrict__ img) {
    __shared__ float2 tile[33][33];  // [column - c0 + 1][row - r0]
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const size_t p = blockIdx.z;
    const int tl = threadIdx.x & 31;
    for (int i = threadIdx.x >> 5; i < 33; i += 8) {
        const int cc = c0 - 1 + i, r = r0 + tl;
        tile[i][tl] = (cc >= 0 && cc < W && r < H) ? ps[(p * W + cc) * H + r] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    for (int i = threadIdx.x >> 5; i < 32; i += 8) {
        const int r = r0 + i, cc = c0 + tl;
        if (cc < W && r < H) img[(p * H + r) * W + cc] += tile[tl + 1][i].x + tile[tl][i].y;
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

namespace {

// Entry rows that the slices [wz0, wz0 + nsl) of any voxel column can reach: vIndex =
// vPre - (z + h) U over every world z of the window, every height of the launch's
// projections and every magnification of the cylinder, with two rows of margin for the
// rounding of the kernel's anchored expression.
void window_rows(const ctg_geometry& g, double hmin, double hmax, double Umin, double Umax, int wz0, int nsl,
                 int& rbase, int& span) {
    const double ylo = (double)wz0 - g.z_pre + hmin, yhi = (double)(wz0 + nsl - 1) - g.z_pre + hmax;
    const double c[4] = {ylo * Umin, ylo * Umax, yhi * Umin, yhi * Umax};
    const double mn = std::min(std::min(c[0], c[1]), std::min(c[2], c[3]));
    const double mx = std::max(std::max(c[0], c[1]), std::max(c[2], c[3]));
    rbase = (int)std::floor(g.v_pre - mx) - 2;
    const int rtop = (int)std::floor(g.v_pre - mn) + 2;
    span = rtop - rbase + 1;
}

}  // namespace

ctg_status launch_forward_project(const ctg_geometry& g, const float4* d_params, const float4* h_params,
                                  const float* d_volume, int z0, int nz, int nproj, float* d_proj, DevBuf& work,
                                  hipStream_t st) {
    if (nz <= 0 || nproj <= 0) return CTG_OK;
    FpArgs a{};
    a.W = g.det_width; a.H = g.det_height; a.nx = g.nx; a.ny = g.ny; a.nzs = nz; a.z0 = z0; a.nproj = nproj;
    a.xpre = g.x_pre; a.ypre = g.y_pre; a.zpre = g.z_pre; a.upre = g.u_pre; a.vpre = g.v_pre;
    a.r2 = g.radius * g.radius; a.fcd = g.fcd; a.uoff = g.u_offset;
    a.ulo = g.u_lo; a.uhi = g.u_hi; a.vlo = g.v_lo; a.vhi = g.v_hi;
    a.zstride = nz;
    if (a.H > kMaxRowsFp) return fail(CTG_ERROR_UNSUPPORTED, "forward projection of more than %d detector rows", kMaxRowsFp);
    // the candidate z intervals are settled from an estimate accurate to about 1e-2
    // slices for grids of up to 65536 slices
    if (g.nz > 65536 || std::fabs(g.z_pre) > 65536.0f)
        return fail(CTG_ERROR_UNSUPPORTED, "forward projection of more than 65536 slices");
    // U = FCD / (FCD - s) with |s| < radius stays in (1/1.9, 10): at most 4 residue passes
    if (!(g.radius < 0.9f * g.fcd))
        return fail(CTG_ERROR_UNSUPPORTED, "forward projection needs the reconstruction cylinder within 0.9 FCD");
    // a voxel inside the v window reads rows v0, v0 + 1 with 0 <= v0 <= H - 2 (vIndex =
    // vPre - v within about 1e-3 of rounding); the kernel indexes its row accumulators
    // with v0 unclamped
    if (!(g.v_pre - g.v_hi >= -0.5f && g.v_pre - g.v_lo <= (float)g.det_height - 1.05f))
        return fail(CTG_ERROR_UNSUPPORTED, "forward projection needs the v window inside the detector rows");
    dim3 grid((unsigned)a.W, (unsigned)nproj);
    if (grid.y > 65535u) return fail(CTG_ERROR_UNSUPPORTED, "forward projection grid too large");

    // Entry-space form (fp_entry_kernel) when the windows' rows fit the LDS copies: it
    // needs U > 2/3 (radius below half the FCD) and host parameters for the row bounds.
    int G = 0;
    double hmin = 0.0, hmax = 0.0;
    const double R = g.radius, F = g.fcd;
    const double Umin = F / (F + R) * (1.0 - 1e-5), Umax = F / (F - R) * (1.0 + 1e-5);
    const bool entry_ok = h_params != nullptr && R < 0.5 * F && !std::getenv("CTG_FP_GATHER");
    if (entry_ok) {
        hmin = hmax = h_params[0].z;
        for (int i = 1; i < nproj; ++i) {
            hmin = std::min(hmin, (double)h_params[i].z);
            hmax = std::max(hmax, (double)h_params[i].z);
        }
        const int zrb = z0 & ~(kRowAnchor - 1);
        // CTG_FP_MAXSPAN overrides the rows per copy the window height may need (tuning;
        // at most 400, which keeps the eight copies within 54 KB of LDS)
        int max_span = kEMaxSpan;
        if (const char* e = std::getenv("CTG_FP_MAXSPAN")) max_span = std::max(16, std::min(400, std::atoi(e)));
        for (int cand : {4, 2, 1}) {
            bool fits = true;
            for (int w0 = zrb; w0 < z0 + nz && fits; w0 += 64 * cand) {
                int rb, sp;
                window_rows(g, hmin, hmax, Umin, Umax, w0, 64 * cand, rb, sp);
                fits = sp <= max_span;
            }
            if (fits) {
                G = cand;
                break;
            }
        }
    }
    if (G > 0 && nz <= kFpMaxSlab) {
        const size_t cells = (size_t)nproj * a.W * a.H;
        CTG_TRY(work.reserve(cells * sizeof(float2)));
        float2* ps = work.as<float2>();
        CTG_TRY(fill_u32(ps, 0u, cells * 2, st));
        const int zrb = z0 & ~(kRowAnchor - 1);
        for (int w0 = zrb; w0 < z0 + nz; w0 += 64 * G) {
            FpWin win{};
            win.wz0 = w0;
            window_rows(g, hmin, hmax, Umin, Umax, w0, 64 * G, win.rbase, win.span);
            // copies a multiple of 16 entries apart (CTG_FP_SPAD=r sets the remainder
            // modulo 16 instead, tuning)
            int pad = 0;
            if (const char* e = std::getenv("CTG_FP_SPAD")) pad = std::atoi(e) & 15;
            win.stride = win.span + ((pad - win.span) & 15);
            const size_t lds = sizeof(float4) * 2 * kGWaves * (size_t)win.stride;
            if (G == 4) hipLaunchKernelGGL(fp_entry_kernel<4>, grid, dim3(kGThreads), lds, st, a, win, d_params, d_volume, ps);
            else if (G == 2) hipLaunchKernelGGL(fp_entry_kernel<2>, grid, dim3(kGThreads), lds, st, a, win, d_params, d_volume, ps);
            else hipLaunchKernelGGL(fp_entry_kernel<1>, grid, dim3(kGThreads), lds, st, a, win, d_params, d_volume, ps);
            CTG_HIP(hipGetLastError());
        }
        dim3 fg((unsigned)((a.W + 31) / 32), (unsigned)((a.H + 31) / 32), (unsigned)nproj);
        hipLaunchKernelGGL(fp_finish_kernel, fg, dim3(256), 0, st, ps, a.W, a.H, d_proj);
        CTG_HIP(hipGetLastError());
        return CTG_OK;
    }

    const size_t lds = sizeof(float) * (size_t)(2 * kGWaves) * (a.H + 4);
    // the candidate records hold z relative to the launch's run base in kFpZBits bits
    for (int s = 0; s < nz; s += kFpMaxSlab) {
        a.z0 = z0 + s;
        a.nzs = std::min(kFpMaxSlab, nz - s);
        hipLaunchKernelGGL(fp_gather_kernel, grid, dim3(kGThreads), lds, st, a, d_params, d_volume + s, d_proj);
        CTG_HIP(hipGetLastError());
    }
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
