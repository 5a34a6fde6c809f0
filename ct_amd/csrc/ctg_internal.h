// Internal declarations shared by the libctg translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ctg.h"

namespace ctg {

// Records the message for ctg_last_error() on this thread and returns `s`.
ctg_status fail(ctg_status s, const char* fmt, ...);
void clear_error();

#define CTG_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            return ::ctg::fail(e_ == hipErrorOutOfMemory ? CTG_ERROR_OUT_OF_MEMORY      \
                                                         : CTG_ERROR_DEVICE,            \
                               "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                               __FILE__, __LINE__);                                     \
        }                                                                               \
    } while (0)

#define CTG_TRY(expr)                       \
    do {                                    \
        ctg_status s_ = (expr);             \
        if (s_ != CTG_OK) return s_;        \
    } while (0)

// Cone-beam magnification U = FCD / (FCD - s) of a voxel at depth s along the central
// ray (Ct/cuda_ct.cu:250, Ct/CtVolume.cpp:1004).  v_rcp_f32 (1 ulp) replaces the IEEE
// division sequence, as the reference's fast-math kernel did.  The backprojection and its
// adjoint (the forward projector) both take U from here: identical U means identical
// window masks and bilinear weights, so the pair stays an exact transpose.
__device__ inline float cone_magnification(float fcd, float s) { return fcd * __builtin_amdgcn_rcpf(fcd - s); }

// Detector row coordinate vIndex = vPre - (z + h) U of a voxel (Ct/CtVolume.cpp:1008,
// 1455), evaluated from an anchor every 32 grid slices: for grid z index zi, with
// za = zi rounded down to a multiple of 32 and zwa its world z,
//     vIndex = fma(-(zi - za), U, fma(-(zwa + h), U, vPre)).
// The anchor costs one FMA per (voxel column, projection, 32 slices), the voxel one FMA.
// At |vIndex| ~ 500 one float ulp is 3e-5 pixel, so different evaluation orders place a
// voxel differently by that much; the backprojection and the forward projector both
// use exactly this expression, whatever their run lengths and slab boundaries, so their
// bilinear weights agree bit for bit and each is the exact transpose of the other.
constexpr int kRowAnchor = 32;
__device__ inline float det_row_anchor(float zwa, float h, float U, float vpre) { return fmaf(-(zwa + h), U, vpre); }
__device__ inline float det_row_index(float k, float U, float anchor) { return fmaf(-k, U, anchor); }

// Grow-only device buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    ctg_status reserve(size_t n);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Grow-only pinned host buffer owned by a context (hipHostMalloc).
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    ctg_status reserve(size_t n);
    void release();
};

// Cached filter response for one (W, filter): full-length real even response
// Htilde(k) (W values), twiddles of the transform length M (M/2 complex), and for
// widths that are not powers of two the M-point transform of the periodically extended
// filter kernel (M complex); float and double copies.
struct FilterTables {
    int W = -1;
    int filter = -1;
    DevBuf resp_f, tw_f, pad_f;  // pad: transform of the padded kernel (widths that are not powers of two)
    DevBuf resp_d, tw_d, pad_d;
};

}  // namespace ctg

struct ctg_context {
    int device = 0;
    ctg_geometry geom{};
    bool has_geom = false;
    std::vector<float4> proj_host;   // sin, cos, height, 0 (rounded from FP64)
    std::vector<double4> proj_host64;
    ctg::DevBuf proj_dev;            // float4 per projection
    ctg::DevBuf proj_dev64;          // double4 per projection
    int nproj_table = 0;
    std::atomic<bool> cancel{false};
    uint64_t spare_mib = 200;        // CtVolume::gpuSpareMemory default (Ct/CtVolume.h:242)
    uint64_t budget_bytes = 0;       // cap on what a reconstruction plans with (0 = free memory)
    ctg::FilterTables ftab;
    ctg::DevBuf rows_f32;            // row-major FP32 work images
    ctg::DevBuf rows_f64;
    ctg::DevBuf maxabs;              // per-projection max |q| (float bits)
    ctg::DevBuf scales_tmp;          // per-projection scales of internally produced layouts
    ctg::DevBuf filtered;            // backprojection-layout projections (pipeline)
    ctg::DevBuf raw_stage;           // raw images on the device (pipeline)
    ctg::DevBuf slab[2];             // volume slabs (pipeline)
    ctg::DevBuf slab_conv;           // index-order converted slab (pipeline)
    ctg::HostBuf pin_up[2];          // pinned staging of raw images (pipeline)
    ctg::HostBuf pin_dl[2];          // pinned download chunks (pipeline)
    ctg::DevBuf ell;                 // phantom ellipsoids
    ctg::DevBuf fp_work;             // column partials of the forward projector
    unsigned* noalloc_scratch = nullptr;  // the device's static max-abs counters (ctg_layout_projections_noalloc)
};

namespace ctg {

// Fill helpers used instead of hipMemsetAsync so that stream capture stays valid.
ctg_status fill_u32(void* d, uint32_t value, size_t count, hipStream_t st);
// Per-image max |x| of nproj row-major FP32 images, as float bits, max-combined into
// d_maxabs (which the caller clears).
ctg_status launch_maxabs(const float* d_images, size_t per_image, int nproj, unsigned* d_maxabs, hipStream_t st);

ctg_status ensure_filter_tables(ctg_context* ctx, int W, int filter);

// ---- preprocess.hip
ctg_status launch_preprocess_rows(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel,
                                  int nproj, float base, ctg_log_form form, ctg_filter filter,
                                  float* d_out, unsigned* d_maxabs, bool do_log, bool do_weight,
                                  bool do_filter, hipStream_t st);
ctg_status launch_filter_rows_only(ctg_context* ctx, const float* d_in, float* d_out, int64_t rows,
                                   ctg_filter filter, hipStream_t st);
ctg_status launch_preprocess_rows_f64(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel,
                                      int nproj, float base, ctg_filter filter, double* d_out,
                                      hipStream_t st);
ctg_status launch_layout(const ctg_geometry& g, const float* d_rows, int nproj, ctg_store store,
                         const unsigned* d_maxabs, void* d_out, float* d_scales, hipStream_t st);

// ---- backproject.hip
struct BpLaunch {
    const ctg_geometry* g;
    const float4* d_params;  // nproj entries: sin, cos, height, unused
    const float4* h_params;  // host copy of the same entries (tile sizing)
    const float* d_scales;   // nproj per-projection scales, or nullptr for 1
    const void* d_filtered;  // BP layout, nproj projections
    ctg_store store;
    int nproj;
    float* d_volume;
    int z0, nz;
    int accumulate;
};
ctg_status launch_backproject(const BpLaunch& L, hipStream_t st);
ctg_status launch_backproject_f64(const ctg_geometry& g, const double4* d_params, const double* d_rows,
                                  int nproj, double* d_volume, int z0, int nz, hipStream_t st);

// ---- project.hip
// Adds the forward projection of the slab into d_proj (row-major nproj x H x W images).
// h_params: host copy of the projections' parameters (row bounds of the entry-space
// form; null selects the gather form); work: the context buffer for its column partials.
ctg_status launch_forward_project(const ctg_geometry& g, const float4* d_params, const float4* h_params,
                                  const float* d_volume, int z0, int nz, int nproj, float* d_proj, DevBuf& work,
                                  hipStream_t st);
ctg_status launch_phantom(const ctg_geometry& g, const double4* d_params, const double* h_ell,
                          int n_ell, int nproj, int p0, double photons, double full_scale,
                          uint64_t seed, uint16_t* d_img, float* d_lineint, DevBuf& ell_buf,
                          hipStream_t st);
ctg_status launch_convert_volume(const ctg_geometry& g, const float* d_vol, int nz,
                                 ctg_index_order order, ctg_byte_order bo, void* d_out,
                                 hipStream_t st);

}  // namespace ctg
