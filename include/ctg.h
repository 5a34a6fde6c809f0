/*
 * ctg.h -- C ABI of the MI355X (gfx950) FDK reconstruction library libctg.so.
 *
 * The entry points replace Ct's device layer (namespace ct::cuda, Ct/cuda_ct.h:16-71,
 * plus CtVolume::FftFilter, Ct/CtVolume.h:137-155) and the GPU half of the engine
 * (CtVolume::cudaReconstructionCore / launchCudaThreads, Ct/CtVolume.cpp:1056-1397).
 * Each declaration cites the reference interface it stands in for.
 *
 * Conventions
 *  - Every call returns a ctg_status; no C++ exception crosses this boundary.
 *    On failure, ctg_last_error() returns a message for the calling thread
 *    (the reference printed to stdout and set CtVolume::lastErrorMessage,
 *    Ct/CtVolume.cpp:1097-1099, 1150, 1211).
 *  - Pointers named d_* are device pointers (hipMalloc or any allocator on the
 *    context's device); h_* are host pointers.  `stream` is a hipStream_t passed
 *    as void* (NULL = the null stream).  Stream-ordered calls return after the
 *    launch; errors found later surface on the next synchronising call.
 *  - All lengths are in elements unless named *_bytes.  Detector images are
 *    row-major W x H (row r, column c at r*W + c), as Ct loads them.
 *  - Volumes on the device are Z_FASTEST slabs [x][y][z] (Common/Volume.h:260)
 *    holding the z-range [z0, z0+nz) of the reconstructed grid.
 *  - A context is used from one host thread at a time; ctg_cancel may be called
 *    from any thread (CtVolume::stop, Ct/CtVolume.cpp:723-726).
 */
#ifndef CTG_H
#define CTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ctg_status {
    CTG_OK = 0,
    CTG_ERROR_INVALID_ARGUMENT = 1,
    CTG_ERROR_DEVICE = 2,             /* HIP runtime / launch failure ("Kernel launch ERROR", Ct/cuda_ct.cu:95-99) */
    CTG_ERROR_OUT_OF_MEMORY = 3,      /* device or host allocation failed (Ct/cuda_ct.cu:138-144) */
    CTG_ERROR_INSUFFICIENT_VRAM = 4,  /* no slab fits (getMaxChunkSize()==0, Ct/CtVolume.cpp:1120-1126) */
    CTG_ERROR_INTERRUPTED = 5,        /* ctg_cancel observed (CompletionStatus::interrupted, Common/types.h:15) */
    CTG_ERROR_UNSUPPORTED = 6,
    CTG_ERROR_IO = 7,
    CTG_ERROR_NO_DEVICE = 8           /* cudaErrorNoDevice analogue (Ct/CtVolume.cpp:119-134) */
} ctg_status;

/* Ct/CtVolume.h:50-54 (same order: passed as int to the kernel, Ct/CtVolume.cpp:913). */
typedef enum ctg_filter { CTG_FILTER_RAMLAK = 0, CTG_FILTER_SHEPP_LOGAN = 1, CTG_FILTER_HANN = 2 } ctg_filter;
/* Image depths accepted by the loader (Ct/CtVolume.cpp:247): 8U, 16U, 32F. */
typedef enum ctg_pixel_type { CTG_PIXEL_U8 = 0, CTG_PIXEL_U16 = 1, CTG_PIXEL_F32 = 2 } ctg_pixel_type;
/* Intensity -> line integral: CPU form (Ct/CtVolume.cpp:802-812, 860-866) or
 * GPU form (Ct/CtVolume.cpp:892-904).  They agree to float rounding. */
typedef enum ctg_log_form { CTG_LOG_CPU_FORM = 0, CTG_LOG_GPU_FORM = 1 } ctg_log_form;
/* Storage of filtered projections; interpolation and accumulation are FP32 in all cases. */
typedef enum ctg_store { CTG_STORE_F32 = 0, CTG_STORE_F16 = 1, CTG_STORE_BF16 = 2 } ctg_store;
/* Common/Volume.h:29-32 */
typedef enum ctg_index_order { CTG_X_FASTEST = 0, CTG_Z_FASTEST = 1 } ctg_index_order;
/* QDataStream::ByteOrder values (BigEndian = 0, LittleEndian = 1), Common/Volume.h:490. */
typedef enum ctg_byte_order { CTG_BIG_ENDIAN = 0, CTG_LITTLE_ENDIAN = 1 } ctg_byte_order;

/* Reconstruction geometry in detector-pixel units: the CtVolume members the
 * GPU kernel takes as arguments (Ct/cuda_ct.cu:294-314, filled at
 * Ct/CtVolume.cpp:1065-1071, 1245-1265). */
typedef struct ctg_geometry {
    int32_t det_width, det_height;   /* imageWidth, imageHeight */
    int32_t nx, ny, nz;              /* reconstructed grid (xMax, yMax, zMax) */
    float fcd;                       /* source-to-axis distance / pixel size */
    float u_offset;                  /* rotation axis offset / pixel size */
    float radius;                    /* reconstruction cylinder radius (voxels) */
    float x_pre, y_pre, z_pre;       /* world = index - pre */
    float u_pre, v_pre;              /* uIndex = u + u_pre, vIndex = v_pre - v */
    float u_lo, u_hi, v_lo, v_hi;    /* accepted detector window on u and v */
} ctg_geometry;

typedef struct ctg_device_info {
    char name[256];
    uint64_t total_bytes, free_bytes;
    int32_t compute_units;
    int32_t memory_bus_width_bits;
    int32_t memory_clock_khz;
    int32_t clock_khz;
} ctg_device_info;

typedef struct ctg_context ctg_context;

/* Progress callback: fraction in [0,1].  Stands in for the Qt signal
 * cudaThreadProgressUpdate / reconstructionProgress (Ct/CtVolume.h:270-277). */
typedef void (*ctg_progress_fn)(double fraction, void* user);

/* ---------------------------------------------------------------- devices */

/* CtVolume::getCudaDeviceCount (Ct/CtVolume.cpp:119-134). */
ctg_status ctg_device_count(int32_t* count);
/* ct::cuda::getFreeMemory/getTotalMemory/getMultiprocessorCnt/getMemoryBusWidth/
 * getMemoryClockRate/getDeviceName (Ct/cuda_ct.cu:8-46). */
ctg_status ctg_get_device_info(int32_t device, ctg_device_info* info);
/* Message of the last failing call on this thread ("" if none). */
const char* ctg_last_error(void);
/* Library version string. */
const char* ctg_version(void);

/* --------------------------------------------------------------- geometry */

/* CtVolume semantics: cylinder radius (Ct/CtVolume.cpp:370-373), grid 2R x 2R x H
 * (:311-316), setVolumeBounds clamping (:487-500, including its yTo/zTo clamp
 * against xFrom), ROI rounding and precomputes (:1399-1424), detector window
 * (:924-928).  bounds = {xFrom,xTo,yFrom,yTo,zFrom,zTo} in [0,1] or NULL.
 * roi_out (6 entries, may be NULL) = voxel ranges {xFrom,xTo,...}. */
ctg_status ctg_geometry_ctvolume(uint32_t det_width, uint32_t det_height, float fcd_px,
                                 float u_offset_px, const float* bounds,
                                 ctg_geometry* out, uint32_t* roi_out);
/* Explicit nx x ny x nz grid centred on the axis (same model, radius as above). */
ctg_status ctg_geometry_grid(uint32_t det_width, uint32_t det_height, float fcd_px,
                             float u_offset_px, int32_t nx, int32_t ny, int32_t nz,
                             ctg_geometry* out);

/* ---------------------------------------------------------------- context */

/* Binds a device (cudaSetDevice in cudaReconstructionCore, Ct/CtVolume.cpp:1060).
 * A context keeps the working buffers of its calls (filter tables, staged and
 * filtered projections, volume slabs, pinned host staging) for the next call;
 * ctg_destroy releases them. */
ctg_status ctg_create(int32_t device, ctg_context** out);
ctg_status ctg_destroy(ctg_context* ctx);
ctg_status ctg_set_geometry(ctg_context* ctx, const ctg_geometry* geometry);
/* Per-projection angle (degrees, after correctAngleDirection) and height
 * offset (pixels, after makeHeightOffsetRelative).  sin/cos are taken in FP64
 * on the host, as the GPU path did (Ct/CtVolume.cpp:1200-1202). */
ctg_status ctg_set_projections(ctg_context* ctx, const double* angles_deg,
                               const double* heights_px, int32_t count);
/* CtVolume::stop (Ct/CtVolume.cpp:723-726): long calls return CTG_ERROR_INTERRUPTED
 * at their next projection/slab boundary.  ctg_reset_cancel clears the flag. */
void ctg_cancel(ctg_context* ctx);
void ctg_reset_cancel(ctg_context* ctx);
/* VRAM to leave free when sizing slabs, in MiB (setGpuSpareMemory, Ct/CtVolume.cpp:522-524). */
ctg_status ctg_set_spare_memory(ctg_context* ctx, uint64_t mib);
/* Upper bound, in bytes, on the device memory ctg_reconstruct plans with (0 = the free
 * memory minus the spare memory).  For sharing a GPU with other work; no counterpart in
 * the reference, whose only control is the spare memory. */
ctg_status ctg_set_memory_budget(ctg_context* ctx, uint64_t bytes);

/* ----------------------------------------------------- stream-ordered ops */

/* Bytes of the backprojection layout for nproj projections in a store type (0 for a
 * null geometry, a negative count or an unknown store). */
size_t ctg_filtered_bytes(const ctg_geometry* geometry, ctg_store store, int32_t nproj);

/* Fused preprocessing of nproj raw images: intensity -> line integral, Feldkamp
 * weight, ramp-family row filter (real DFT of length W without padding, bins
 * 0..W/2 times the window, unnormalised inverse), then conversion into the
 * backprojection layout in the chosen store with per-projection power-of-two
 * scales written to d_scales (nproj floats, 1 for F32; may be null for F32 only).
 * The filter runs a transform of length
 * M = W for powers of two, else the power of two >= 2W - 1 (padded); M up to 16384
 * fits the LDS row buffer (8192 for the FP64 yardstick, ctg_fdk_f64), and wider
 * rows return CTG_ERROR_UNSUPPORTED.
 * Replaces cudaPreprocessImage (Ct/CtVolume.cpp:887-921): convertTo/log/multiply,
 * applyFeldkampWeightFiltering (Ct/cuda_ct.cu:119-131), FftFilter::applyForward/
 * applyInverse (Ct/CtVolume.cpp:75-101) and applyFrequencyFiltering (Ct/cuda_ct.cu:87-100). */
ctg_status ctg_preprocess(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel,
                          int32_t nproj, float base_intensity, ctg_log_form log_form,
                          ctg_filter filter, ctg_store store, void* d_filtered,
                          float* d_scales, void* stream);

/* Feldkamp weight in place on nproj line-integral images
 * (applyFeldkampWeightFiltering, Ct/cuda_ct.cu:119-131). */
ctg_status ctg_feldkamp_weight(ctg_context* ctx, float* d_images, int32_t nproj, void* stream);

/* Ramp-family row filter alone, rows of length det_width, d_in may equal d_out
 * (FftFilter + applyFrequencyFiltering, Ct/CtVolume.cpp:910-917).  The operator is
 * symmetric, so it is also its own adjoint. */
ctg_status ctg_filter_rows(ctg_context* ctx, const float* d_in, float* d_out, int64_t rows,
                           ctg_filter filter, void* stream);

/* Row-major FP32 detector images -> backprojection layout in a store type, with
 * per-projection power-of-two scales in d_scales (may be null for F32 only; no
 * counterpart: the reference kept FP32). */
ctg_status ctg_layout_projections(ctg_context* ctx, const float* d_images, int32_t nproj,
                                  ctg_store store, void* d_filtered, float* d_scales,
                                  void* stream);

/* Maximum projection count of one ctg_layout_projections_noalloc call. */
#define CTG_NOALLOC_MAX_PROJ 65536

/* ctg_layout_projections without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc), e.g. to lay out the
 * images of an A^T application that is captured whole (ctg_backproject allocates
 * nothing).  The per-projection max-abs counters that ctg_layout_projections keeps in
 * context memory live here in a library-owned static device buffer of
 * CTG_NOALLOC_MAX_PROJ entries (256 KiB, held on every GPU the library runs on until
 * the process ends); its address on the context's device is resolved by ctg_create.
 * nproj above CTG_NOALLOC_MAX_PROJ returns CTG_ERROR_INVALID_ARGUMENT.  The counters
 * are cleared by a fill kernel in every call; calls that overlap on one GPU (other
 * streams, other contexts, other host threads) share the buffer and race.  Results
 * are identical to ctg_layout_projections'. */
ctg_status ctg_layout_projections_noalloc(ctg_context* ctx, const float* d_images, int32_t nproj,
                                          ctg_store store, void* d_filtered, float* d_scales,
                                          void* stream);

/* Voxel-driven backprojection of projections [p0, p0+nproj) (geometry indices
 * from ctg_set_projections; d_filtered holds them in order starting with p0)
 * into the slab d_volume = z-range [z0, z0+nz).  accumulate = 0 overwrites the
 * slab, 1 adds to it (nproj = 0 then zeroes, or leaves, the slab).  d_scales: the
 * scales ctg_preprocess wrote, required for F16/BF16 (null means 1 for F32).
 * Replaces startReconstruction/reconstructionKernel (Ct/cuda_ct.cu:201-352) with one
 * launch for many projections. */
ctg_status ctg_backproject(ctg_context* ctx, const void* d_filtered, ctg_store store,
                           const float* d_scales, int32_t p0, int32_t nproj,
                           float* d_volume, int32_t z0, int32_t nz, int32_t accumulate,
                           void* stream);

/* Exact adjoint of ctg_backproject with store F32 and unit scales: splats each
 * voxel of the slab into the four bilinear taps with the same masks and weights.
 * d_proj: nproj row-major W x H images, accumulated into when accumulate = 1.
 * No counterpart in the reference (added for differentiation). */
ctg_status ctg_forward_project(ctg_context* ctx, const float* d_volume, int32_t z0, int32_t nz,
                               int32_t p0, int32_t nproj, float* d_proj, int32_t accumulate,
                               void* stream);

/* Backward of the linear FDK operator F = B . H . Wf (B backprojection, H row
 * filter, Wf Feldkamp weight) with respect to line integrals:
 * d_grad_lineint = Wf . H . B^T (d_grad_volume), nproj row-major W x H images
 * (overwritten; the filter and weight run in place there, no other workspace).
 * No counterpart in the reference. */
ctg_status ctg_fdk_backward(ctg_context* ctx, const float* d_grad_volume, int32_t z0, int32_t nz,
                            int32_t p0, int32_t nproj, ctg_filter filter, float* d_grad_lineint,
                            void* stream);

/* FP64 FDK on the GPU (log in the CPU form, weight, filter and backprojection in
 * double): the error yardstick for the FP16/BF16 store.  d_volume: double slab. */
ctg_status ctg_fdk_f64(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int32_t nproj,
                       float base_intensity, ctg_filter filter, double* d_volume,
                       int32_t z0, int32_t nz, void* stream);

/* Z_FASTEST slab -> the requested index order and byte order, as raw bytes
 * (the element order of Volume::saveToBinaryFile, Common/Volume.h:492-556).
 * For X_FASTEST the slab must hold the whole z-range (z0 = 0, nz = grid nz) or the
 * caller places slabs itself; d_out receives nx*ny*nz*4 bytes. */
ctg_status ctg_convert_volume(ctg_context* ctx, const float* d_volume, int32_t nz,
                              ctg_index_order order, ctg_byte_order byte_order,
                              void* d_out, void* stream);

/* ------------------------------------------------------- whole pipeline */

/* Full reconstruction from host images (nproj x H x W, pixel type as given) into
 * a host volume in the requested index order (nx*ny*nz floats).  Plans z-slabs
 * from free VRAM minus the spare memory, keeps filtered projections resident
 * when they fit and re-filters per slab otherwise, overlaps the upload of raw
 * images with preprocessing and backprojection, and each slab's download with the
 * next slab's backprojection (see ctg_plan_reconstruct), on separate streams.
 * Progress is reported per slab.  Replaces launchCudaThreads/cudaReconstructionCore
 * (Ct/CtVolume.cpp:1056-1397). */
ctg_status ctg_reconstruct(ctg_context* ctx, const void* h_raw, ctg_pixel_type pixel,
                           int32_t nproj, float base_intensity, ctg_log_form log_form,
                           ctg_filter filter, ctg_store store, float* h_volume,
                           ctg_index_order order, ctg_progress_fn progress, void* user);

/* The plan ctg_reconstruct would follow for these inputs under the current memory
 * limits: *resident = 1 when every filtered projection stays in device memory for all
 * slabs (else each slab re-derives them batch by batch from the uploaded images), and
 * *slab_nz = z-slices per slab.  A resident volume of 256 or more slices that fits
 * whole is still cut into four slabs when a second slab buffer fits, so that each
 * slab's download runs under the next slab's backprojection (for Z_FASTEST output the
 * cut runs along x, into twice as many ranges of whole 16-column tiles, which the
 * host volume holds contiguously).  Fails with CTG_ERROR_INSUFFICIENT_VRAM, as the
 * reconstruction would, when no slab fits (getMaxChunkSize() == 0,
 * Ct/CtVolume.cpp:1120-1126); the reference's counterpart for callers sizing a job is
 * getRequiredMemoryUpperBound (Ct/CtVolume.cpp:469-480). */
ctg_status ctg_plan_reconstruct(ctg_context* ctx, ctg_pixel_type pixel, int32_t nproj, ctg_store store,
                                int32_t* resident, int32_t* slab_nz);

/* ------------------------------------------------------ synthetic inputs */

/* Analytic cone-beam line integrals of ellipsoids in the same geometry model
 * (ell: n_ell x 10 doubles {cx,cy,cz,a,b,c,phi,theta,psi,mu}, voxel units, ZYZ
 * Euler angles), then intensities: photons > 0 draws Poisson counts with mean
 * photons*exp(-p) (counter-based generator seeded by `seed`), photons <= 0 gives
 * round(full_scale*exp(-p)).  Output: nproj x H x W uint16, clamped to 65535.
 * No counterpart in the reference (which reads images from disk). */
ctg_status ctg_phantom_images(ctg_context* ctx, const double* ell, int32_t n_ell,
                              int32_t p0, int32_t nproj, double photons, double full_scale,
                              uint64_t seed, uint16_t* d_images, float* d_lineint, void* stream);

#ifdef __cplusplus
}
#endif
#endif
