// isa_rates: measures the instruction rates that bound the library's kernels on gfx950:
// LDS reads and writes (the entry-space forward projector's 16-byte read-modify-writes,
// the backprojection's 16-byte entry reads) and VALU instructions: those of the
// backprojection's per-voxel arithmetic and the alternatives weighed for it (FP32
// add/mul, integer add/and, shifts, conversions).  Not used by the library.
//
// LDS: every workgroup (256 lanes, 64 KB of LDS: 16 KB per wave) issues `iters` x 8 LDS
// instructions of one kind per wave at addresses inside its own wave's region; the
// kernel is launched twice and the second launch is timed with HIP events.  The result
// is printed as device clock cycles per wave instruction per CU:
//     ms * clock * CUs / (wave instructions),
// with the clock and CU count from the device attributes.  The arithmetic around each
// access (address, one add) is a few VALU instructions per access, so a figure near 6
// or below for reads is bounded by VALU issue, not by the LDS.
//
// VALU: 8 workgroups of 256 lanes per CU (8 waves per SIMD) run 8 independent register
// chains per lane of one instruction kind (inline assembly on registers only, so the
// compiler neither folds nor reorders them), 512 per loop iteration; the result is
// device cycles per wave64 instruction per SIMD.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define LR_CHECK(x)                                                                    \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

namespace {

constexpr int kEntries = 1024;  // 16-byte entries per wave region
constexpr int kBlocks = 2048, kIters = 1000;

enum Kind { kRead128 = 0, kWrite128 = 1, kAdd32 = 2, kWrite32 = 3 };

// entry (16-byte units, or dwords for the 32-bit kinds) of lane l; mode 1 spaces lanes
// 4 entries (64 bytes) apart
__device__ constexpr int lane_entry(int mode, int l) { return mode == 1 ? 4 * l : l; }

// active lanes: mode 2 one lane in four, mode 3 lanes 0-7 only, otherwise all
__device__ inline bool lane_active(int mode, int l) {
    return mode == 2 ? (l & 3) == 0 : mode == 3 ? l < 8 : true;
}

template <int KIND, int MODE>
__global__ __launch_bounds__(256) void lds_kernel(int iters, float* sink) {
    extern __shared__ float4 region[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float4* base4 = region + wv * kEntries;
    float* base1 = reinterpret_cast<float*>(base4);
    for (int i = lane; i < kEntries; i += 64) base4[i] = make_float4((float)i, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    const int e = lane_entry(MODE, lane);
    // a run-time zero the compiler cannot fold, so that every access forms its address
    volatile int zero_src = 0;
    const int z = zero_src;
    float acc = 0.0f;
    if (lane_active(MODE, lane)) {
        float4 v = make_float4((float)lane, 1.0f, 2.0f, 3.0f);
        for (int it = 0; it < iters; ++it) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int idx = (e + z * (k + it)) & (kEntries - 1);
                if constexpr (KIND == kRead128) {
                    const float4 r = base4[idx];
                    acc += (r.x + r.y) + (r.z + r.w);
                } else if constexpr (KIND == kWrite128) {
                    base4[idx] = v;
                    v.x += 1.0f;
                } else if constexpr (KIND == kAdd32) {
                    atomicAdd(&base1[(idx + 1024 * (k & 3)) & (4 * kEntries - 1)], v.x);
                } else {
                    base1[(idx + 1024 * (k & 3)) & (4 * kEntries - 1)] = v.x;
                    v.x += 1.0f;
                }
            }
        }
    }
    __syncthreads();
    if (acc == -1.0f || base4[lane].x == -1.0f) sink[threadIdx.x] = acc;
}

enum Valu {
    kFma = 0, kFmaMix = 1, kFract = 2, kCvtFlr = 3, kLshlAdd = 4, kAddU32 = 5,
    kLshl = 6, kMadU24 = 7, kAddF32 = 8, kMulF32 = 9, kCvtF32I32 = 10, kAndB32 = 11
};

template <int KIND>
__global__ __launch_bounds__(256) void valu_kernel(int iters, float a, float b, float* sink) {
    float x[8];
    int y[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        x[i] = (float)threadIdx.x * 0.001f + (float)i;
        y[i] = (int)threadIdx.x + i;
    }
    const int ah = __float_as_int(a) & 0xffff;  // a 16-bit pattern for the mixed-precision form
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int s = 0; s < 64; ++s) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if constexpr (KIND == kFma) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(x[i]) : "v"(a), "v"(b));
                else if constexpr (KIND == kFmaMix)
                    asm volatile("v_fma_mix_f32 %0, %0, %1, %2 op_sel_hi:[0,1,0]" : "+v"(x[i]) : "v"(ah), "v"(b));
                else if constexpr (KIND == kFract) asm volatile("v_fract_f32 %0, %0" : "+v"(x[i]));
                else if constexpr (KIND == kCvtFlr) asm volatile("v_cvt_flr_i32_f32 %0, %1" : "=v"(y[i]) : "v"(x[i]));
                else if constexpr (KIND == kLshlAdd) asm volatile("v_lshl_add_u32 %0, %0, 4, %1" : "+v"(y[i]) : "v"(ah));
                else if constexpr (KIND == kLshl) asm volatile("v_lshlrev_b32 %0, 3, %0" : "+v"(y[i]));
                else if constexpr (KIND == kMadU24) asm volatile("v_mad_u32_u24 %0, %0, 8, %1" : "+v"(y[i]) : "v"(ah));
                else if constexpr (KIND == kAddF32) asm volatile("v_add_f32 %0, %0, %1" : "+v"(x[i]) : "v"(b));
                else if constexpr (KIND == kMulF32) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(x[i]) : "v"(a));
                else if constexpr (KIND == kCvtF32I32) asm volatile("v_cvt_f32_i32 %0, %1" : "=v"(x[i]) : "v"(y[i]));
                else if constexpr (KIND == kAndB32) asm volatile("v_and_b32 %0, %0, %1" : "+v"(y[i]) : "v"(ah));
                else asm volatile("v_add_u32 %0, %0, %1" : "+v"(y[i]) : "v"(ah));
            }
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += x[i] + (float)y[i];
    if (s == -1.0f) sink[threadIdx.x] = s;
}

}  // namespace

int main() {
    int dev = 0, clock_khz = 0, cus = 0;
    LR_CHECK(hipGetDevice(&dev));
    LR_CHECK(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, dev));
    LR_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    float* sink = nullptr;
    LR_CHECK(hipMalloc(&sink, 256 * sizeof(float)));
    hipEvent_t t0, t1;
    LR_CHECK(hipEventCreate(&t0));
    LR_CHECK(hipEventCreate(&t1));
    const size_t lds = (size_t)4 * kEntries * 16;
    const double instr = (double)kBlocks * 4 * kIters * 8;
    std::printf("clock %.0f MHz, %d CUs; cycles per wave instruction per CU\n", clock_khz / 1e3, cus);
    using Kernel = void (*)(int, float*);
    struct Case { Kernel k; const char* what; };
    const Case cases[] = {
        {lds_kernel<kRead128, 0>, "ds_read_b128, lanes on consecutive 16-byte entries"},
        {lds_kernel<kRead128, 1>, "ds_read_b128, lanes 64 bytes apart (4-way bank conflicts)"},
        {lds_kernel<kWrite128, 0>, "ds_write_b128, lanes on consecutive 16-byte entries"},
        {lds_kernel<kWrite128, 1>, "ds_write_b128, lanes 64 bytes apart"},
        {lds_kernel<kWrite128, 2>, "ds_write_b128, one lane in four active"},
        {lds_kernel<kWrite128, 3>, "ds_write_b128, lanes 0-7 active"},
        {lds_kernel<kWrite32, 0>, "ds_write_b32, consecutive dwords"},
        {lds_kernel<kAdd32, 0>, "ds_add_f32 (no return), consecutive dwords"},
    };
    for (const Case& c : cases) {
        for (int rep = 0; rep < 2; ++rep) {
            if (rep == 1) LR_CHECK(hipEventRecord(t0));
            hipLaunchKernelGGL(c.k, dim3(kBlocks), dim3(256), lds, 0, kIters, sink);
            LR_CHECK(hipGetLastError());
        }
        LR_CHECK(hipEventRecord(t1));
        LR_CHECK(hipEventSynchronize(t1));
        float ms = 0.0f;
        LR_CHECK(hipEventElapsedTime(&ms, t0, t1));
        std::printf("%6.2f  %s\n", ms * 1e-3 * clock_khz * 1e3 * cus / instr, c.what);
        std::fflush(stdout);
    }
    const int vblocks = 8 * cus, viters = 200;
    const double vinstr = (double)vblocks * 4 * viters * 512;
    std::printf("cycles per wave64 instruction per SIMD (8 waves per SIMD)\n");
    using VKernel = void (*)(int, float, float, float*);
    struct VCase { VKernel k; const char* what; };
    const VCase vcases[] = {
        {valu_kernel<kFma>, "v_fma_f32"},           {valu_kernel<kFmaMix>, "v_fma_mix_f32 (one 16-bit operand)"},
        {valu_kernel<kFract>, "v_fract_f32"},       {valu_kernel<kCvtFlr>, "v_cvt_flr_i32_f32"},
        {valu_kernel<kLshlAdd>, "v_lshl_add_u32"}, {valu_kernel<kAddU32>, "v_add_u32"},
        {valu_kernel<kLshl>, "v_lshlrev_b32"},      {valu_kernel<kMadU24>, "v_mad_u32_u24"},
        {valu_kernel<kAddF32>, "v_add_f32"},        {valu_kernel<kMulF32>, "v_mul_f32"},
        {valu_kernel<kCvtF32I32>, "v_cvt_f32_i32"}, {valu_kernel<kAndB32>, "v_and_b32"},
    };
    for (const VCase& c : vcases) {
        for (int rep = 0; rep < 2; ++rep) {
            if (rep == 1) LR_CHECK(hipEventRecord(t0));
            hipLaunchKernelGGL(c.k, dim3(vblocks), dim3(256), 0, 0, viters, 0.999f, 0.001f, sink);
            LR_CHECK(hipGetLastError());
        }
        LR_CHECK(hipEventRecord(t1));
        LR_CHECK(hipEventSynchronize(t1));
        float ms = 0.0f;
        LR_CHECK(hipEventElapsedTime(&ms, t0, t1));
        std::printf("%6.2f  %s\n", ms * 1e-3 * clock_khz * 1e3 * cus * 4 / vinstr, c.what);
        std::fflush(stdout);
    }
    LR_CHECK(hipFree(sink));
    return 0;
}
