// C ABI of libctg: context, geometry, stream-ordered operations and the z-slab
// reconstruction pipeline.  See include/ctg.h for the contract of each entry point.
#include "ctg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

namespace ctg {

namespace {
thread_local std::string g_last_error;

__global__ void fill_u32_kernel(uint32_t* d, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        d[i] = v;
}

// Per-image max |x| of row-major FP32 images (bits of a non-negative float order like
// unsigned integers).
__global__ void maxabs_kernel(const float* __restrict__ img, size_t per_image, unsigned* __restrict__ out) {
    const size_t p = blockIdx.y;
    const float* src = img + p * per_image;
    float m = 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < per_image; i += (size_t)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(src[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&out[p], __float_as_uint(m));
}
}  // namespace

ctg_status fail(ctg_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

void clear_error() { g_last_error.clear(); }

ctg_status DevBuf::reserve(size_t n) {
    if (n <= bytes && ptr) return CTG_OK;
    release();
    if (n == 0) return CTG_OK;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return fail(CTG_ERROR_OUT_OF_MEMORY, "device allocation of %zu bytes failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return CTG_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

ctg_status HostBuf::reserve(size_t n) {
    if (n <= bytes && ptr) return CTG_OK;
    release();
    if (n == 0) return CTG_OK;
    hipError_t e = hipHostMalloc(&ptr, n, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return fail(CTG_ERROR_OUT_OF_MEMORY, "pinned host allocation of %zu bytes failed: %s", n, hipGetErrorString(e));
    }
    bytes = n;
    return CTG_OK;
}

void HostBuf::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

ctg_status fill_u32(void* d, uint32_t value, size_t count, hipStream_t st) {
    if (count == 0) return CTG_OK;
    const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, st, (uint32_t*)d, value, count);
    CTG_HIP(hipGetLastError());
    return CTG_OK;
}

ctg_status launch_maxabs(const float* d_images, size_t per_image, int nproj, unsigned* d_maxabs, hipStream_t st) {
    const unsigned bx = (unsigned)std::min<size_t>((per_image + 255) / 256, 64);
    // one grid row per image; the grid's y extent is limited to 65535
    for (int p0 = 0; p0 < nproj; p0 += 65535) {
        const int np = std::min(65535, nproj - p0);
        hipLaunchKernelGGL(maxabs_kernel, dim3(bx, (unsigned)np), dim3(256), 0, st, d_images + (size_t)p0 * per_image,
                           per_image, d_maxabs + p0);
        CTG_HIP(hipGetLastError());
    }
    return CTG_OK;
}

}  // namespace ctg

using namespace ctg;

// Max-abs counters of ctg_layout_projections_noalloc: a static device array, one copy
// per GPU.  In host code `scratch` names a host-side stand-in, not the device array, so
// kernels and fills get the per-device address that hipGetSymbolAddress returns,
// resolved once per device with that device current (noalloc_scratch_for).
constexpr int kNoAllocMaxProj = CTG_NOALLOC_MAX_PROJ;
static __device__ __align__(256) unsigned scratch[kNoAllocMaxProj];

namespace {

size_t pixel_bytes(ctg_pixel_type p) { return p == CTG_PIXEL_U8 ? 1 : (p == CTG_PIXEL_U16 ? 2 : 4); }
size_t store_bytes(ctg_store s) { return s == CTG_STORE_F32 ? 16 : 8; }

// Cylinder radius, Ct/CtVolume.cpp:370-373 (unsigned halving, float FCD^2, truncation).
uint32_t cylinder_radius(uint32_t W, float fcd) {
    const double r = (double)(W / 2u);
    const float f2 = fcd * fcd;
    return (uint32_t)std::sqrt(((double)f2 * r * r) / ((double)f2 + r * r));
}

// Detector window, Ct/CtVolume.cpp:924-928 with matToImageU/V (:1459-1466).
void detector_window(ctg_geometry* g) {
    const uint32_t W = (uint32_t)g->det_width, H = (uint32_t)g->det_height;
    g->u_pre = (float)((double)(float)W / 2.0);
    g->v_pre = (float)((double)(float)H / 2.0);
    g->u_lo = 0.0f - g->u_pre;
    g->u_hi = (float)((double)(W - 1u) - 0.1) - g->u_pre;
    g->v_lo = (-1.0f) * (float)((double)(H - 1u) - 0.1) + g->v_pre;
    g->v_hi = (-1.0f) * 0.0f + g->v_pre;
}

ctg_status check_ctx(ctg_context* ctx, bool need_geom = true) {
    if (!ctx) return fail(CTG_ERROR_INVALID_ARGUMENT, "null context");
    if (need_geom && !ctx->has_geom) return fail(CTG_ERROR_INVALID_ARGUMENT, "geometry not set");
    CTG_HIP(hipSetDevice(ctx->device));
    return CTG_OK;
}

ctg_status check_range(ctg_context* ctx, int p0, int nproj) {
    if (nproj < 0 || p0 < 0 || p0 + nproj > ctx->nproj_table)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "projection range [%d,%d) outside the %d set projections", p0,
                    p0 + nproj, ctx->nproj_table);
    return CTG_OK;
}

ctg_status check_slab(ctg_context* ctx, int z0, int nz) {
    if (z0 < 0 || nz < 0 || z0 + nz > ctx->geom.nz)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "slab [%d,%d) outside the grid z-range [0,%d)", z0, z0 + nz,
                    ctx->geom.nz);
    return CTG_OK;
}

bool bad_pixel(ctg_pixel_type v) { return (int)v < 0 || (int)v > 2; }
bool bad_filter(ctg_filter v) { return (int)v < 0 || (int)v > 2; }
bool bad_store(ctg_store v) { return (int)v < 0 || (int)v > 2; }
bool bad_log_form(ctg_log_form v) { return (int)v < 0 || (int)v > 1; }

// FP16/BF16 entries are stored divided by a per-projection power of two: without the
// scale array the backprojection would silently use 1
ctg_status check_scales(ctg_store store, const float* d_scales, int nproj) {
    if (nproj > 0 && store != CTG_STORE_F32 && !d_scales)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "FP16/BF16 stores need a scale array");
    return CTG_OK;
}

// Preprocess nproj raw images into the BP layout in batches that bound the FP32 work buffer.
ctg_status preprocess_batched(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int nproj, float base,
                              ctg_log_form form, ctg_filter filter, ctg_store store, void* d_filtered,
                              float* d_scales, hipStream_t st) {
    const ctg_geometry& g = ctx->geom;
    const size_t img = (size_t)g.det_width * g.det_height;
    const size_t budget = (size_t)1 << 30;  // 1 GiB of FP32 work images per batch
    int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nproj, budget / (img * 4)));
    CTG_TRY(ctx->rows_f32.reserve(img * 4 * batch));
    CTG_TRY(ctx->maxabs.reserve(sizeof(unsigned) * std::max(nproj, 1)));
    CTG_TRY(fill_u32(ctx->maxabs.ptr, 0u, (size_t)nproj, st));
    for (int b0 = 0; b0 < nproj; b0 += batch) {
        if (ctx->cancel.load()) return fail(CTG_ERROR_INTERRUPTED, "interrupted");
        const int nb = std::min(batch, nproj - b0);
        const char* raw = (const char*)d_raw + (size_t)b0 * img * pixel_bytes(pixel);
        CTG_TRY(launch_preprocess_rows(ctx, raw, pixel, nb, base, form, filter, ctx->rows_f32.as<float>(),
                                       ctx->maxabs.as<unsigned>() + b0, true, true, true, st));
        char* out = (char*)d_filtered + (size_t)b0 * img * store_bytes(store);
        CTG_TRY(launch_layout(g, ctx->rows_f32.as<float>(), nb, store, ctx->maxabs.as<unsigned>() + b0, out,
                              d_scales ? d_scales + b0 : nullptr, st));
    }
    return CTG_OK;
}

// Host copy split over worker threads (a single thread copies pageable memory at a
// fraction of what the pinned upload that follows can take).
void parallel_memcpy(void* dst, const void* src, size_t n) {
    const size_t chunk = (size_t)4 << 20;  // at least 4 MiB per thread
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::min<size_t>(std::min<size_t>(8, hw), (n + chunk - 1) / chunk);
    if (workers <= 1) {
        std::memcpy(dst, src, n);
        return;
    }
    const size_t per = (n + workers - 1) / workers;
    std::vector<std::thread> pool;
    for (size_t w = 1; w < workers; ++w) {
        const size_t o = w * per;
        if (o >= n) break;
        pool.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, std::min(per, n - o)); });
    }
    std::memcpy(dst, src, std::min(per, n));
    for (auto& t : pool) t.join();
}

// Runs fn(i0, i1) over [0, n) split across up to 8 host threads.
template <class F>
void parallel_ranges(size_t n, size_t min_per_thread, F fn) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t workers = std::max<size_t>(1, std::min<size_t>(std::min<size_t>(8, hw), n / std::max<size_t>(1, min_per_thread)));
    const size_t per = (n + workers - 1) / workers;
    std::vector<std::thread> pool;
    for (size_t w = 1; w < workers; ++w) {
        const size_t a = w * per, b = std::min(n, a + per);
        if (a < b) pool.emplace_back([=, &fn] { fn(a, b); });
    }
    fn(0, std::min(n, per));
    for (auto& t : pool) t.join();
}

std::mutex g_scratch_mu;
std::vector<unsigned*> g_scratch_ptr;  // indexed by device

// Address of `scratch` on `device`, which must be the current device.
ctg_status noalloc_scratch_for(int device, unsigned** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        CTG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<unsigned*>(p);
    }
    *out = g_scratch_ptr[device];
    return CTG_OK;
}

}  // namespace

extern "C" {

const char* ctg_last_error(void) { return g_last_error.c_str(); }

const char* ctg_version(void) { return "ctg 0.1 (gfx950)"; }

ctg_status ctg_device_count(int32_t* count) {
    if (!count) return fail(CTG_ERROR_INVALID_ARGUMENT, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return CTG_OK;
    }
    *count = n;
    return CTG_OK;
}

ctg_status ctg_get_device_info(int32_t device, ctg_device_info* info) {
    if (!info) return fail(CTG_ERROR_INVALID_ARGUMENT, "null info");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return fail(CTG_ERROR_NO_DEVICE, "device %d not available", device);
    }
    hipDeviceProp_t prop;
    CTG_HIP(hipGetDeviceProperties(&prop, device));
    memset(info, 0, sizeof(*info));
    snprintf(info->name, sizeof(info->name), "%s (%s)", prop.name, prop.gcnArchName);
    CTG_HIP(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    CTG_HIP(hipMemGetInfo(&fr, &tot));
    info->free_bytes = fr;
    info->total_bytes = tot;
    info->compute_units = prop.multiProcessorCount;
    info->memory_bus_width_bits = prop.memoryBusWidth;
    info->memory_clock_khz = prop.memoryClockRate;
    info->clock_khz = prop.clockRate;
    return CTG_OK;
}

ctg_status ctg_geometry_ctvolume(uint32_t W, uint32_t H, float fcd_px, float u_offset_px, const float* bounds,
                                 ctg_geometry* out, uint32_t* roi_out) {
    clear_error();
    if (!out) return fail(CTG_ERROR_INVALID_ARGUMENT, "null geometry");
    if (W < 2 || H < 2) return fail(CTG_ERROR_INVALID_ARGUMENT, "detector %ux%u too small", W, H);
    // sinogramFromImages rejects imageWidth >= FCD (Ct/CtVolume.cpp:294-301)
    if (!((float)W < fcd_px)) return fail(CTG_ERROR_INVALID_ARGUMENT, "the FCD must be greater than the image width");
    const uint32_t R = cylinder_radius(W, fcd_px);
    const uint32_t xSize = 2u * R, ySize = 2u * R, zSize = H;
    float xf = 0, xt = 1, yf = 0, yt = 1, zf = 0, zt = 1;
    if (bounds && !(bounds[0] == bounds[1] || bounds[2] == bounds[3] || bounds[4] == bounds[5])) {
        // setVolumeBounds, Ct/CtVolume.cpp:493-498 (yTo and zTo are clamped against xFrom there)
        xf = std::max(0.0f, std::min(1.0f, bounds[0]));
        xt = std::max(xf, std::min(1.0f, bounds[1]));
        yf = std::max(0.0f, std::min(1.0f, bounds[2]));
        yt = std::max(xf, std::min(1.0f, bounds[3]));
        zf = std::max(0.0f, std::min(1.0f, bounds[4]));
        zt = std::max(xf, std::min(1.0f, bounds[5]));
    }
    // updateBoundaries, Ct/CtVolume.cpp:1401-1423
    const uint32_t xFrom = (uint32_t)((double)(xf * (float)xSize) + 0.5);
    const uint32_t xTo = (uint32_t)((double)(xt * (float)xSize) + 0.5);
    const uint32_t yFrom = (uint32_t)((double)(yf * (float)ySize) + 0.5);
    const uint32_t yTo = (uint32_t)((double)(yt * (float)ySize) + 0.5);
    const uint32_t zFrom = (uint32_t)((double)(zf * (float)zSize) + 0.5);
    const uint32_t zTo = (uint32_t)((double)(zt * (float)zSize) + 0.5);
    if (roi_out) {
        roi_out[0] = xFrom; roi_out[1] = xTo; roi_out[2] = yFrom;
        roi_out[3] = yTo; roi_out[4] = zFrom; roi_out[5] = zTo;
    }
    if (xTo < xFrom || yTo < yFrom || zTo < zFrom)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "volume bounds give a negative extent");
    memset(out, 0, sizeof(*out));
    out->det_width = (int32_t)W;
    out->det_height = (int32_t)H;
    out->nx = (int32_t)(xTo - xFrom);
    out->ny = (int32_t)(yTo - yFrom);
    out->nz = (int32_t)(zTo - zFrom);
    out->fcd = fcd_px;
    out->u_offset = u_offset_px;
    out->radius = (float)R;
    out->x_pre = (float)(((double)(float)xSize / 2.0) - (double)(float)xFrom);
    out->y_pre = (float)(((double)(float)ySize / 2.0) - (double)(float)yFrom);
    out->z_pre = (float)(((double)(float)zSize / 2.0) - (double)(float)zFrom);
    detector_window(out);
    return CTG_OK;
}

ctg_status ctg_geometry_grid(uint32_t W, uint32_t H, float fcd_px, float u_offset_px, int32_t nx, int32_t ny,
                             int32_t nz, ctg_geometry* out) {
    clear_error();
    if (!out) return fail(CTG_ERROR_INVALID_ARGUMENT, "null geometry");
    if (W < 2 || H < 2 || nx < 0 || ny < 0 || nz < 0) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad sizes");
    if (!((float)W < fcd_px)) return fail(CTG_ERROR_INVALID_ARGUMENT, "the FCD must be greater than the image width");
    memset(out, 0, sizeof(*out));
    out->det_width = (int32_t)W;
    out->det_height = (int32_t)H;
    out->nx = nx;
    out->ny = ny;
    out->nz = nz;
    out->fcd = fcd_px;
    out->u_offset = u_offset_px;
    out->radius = (float)cylinder_radius(W, fcd_px);
    out->x_pre = (float)((double)(float)nx / 2.0);
    out->y_pre = (float)((double)(float)ny / 2.0);
    out->z_pre = (float)((double)(float)nz / 2.0);
    detector_window(out);
    return CTG_OK;
}

ctg_status ctg_create(int32_t device, ctg_context** out) {
    clear_error();
    if (!out) return fail(CTG_ERROR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(CTG_ERROR_NO_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(CTG_ERROR_NO_DEVICE, "device %d not available (%d devices)", device, n);
    CTG_HIP(hipSetDevice(device));
    unsigned* sp = nullptr;
    CTG_TRY(noalloc_scratch_for(device, &sp));
    ctg_context* c = new (std::nothrow) ctg_context();
    if (!c) return fail(CTG_ERROR_OUT_OF_MEMORY, "host allocation failed");
    c->device = device;
    c->noalloc_scratch = sp;
    *out = c;
    return CTG_OK;
}

ctg_status ctg_destroy(ctg_context* ctx) {
    if (!ctx) return CTG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    ctx->proj_dev.release();
    ctx->proj_dev64.release();
    ctx->ftab.resp_f.release(); ctx->ftab.tw_f.release(); ctx->ftab.pad_f.release();
    ctx->ftab.resp_d.release(); ctx->ftab.tw_d.release(); ctx->ftab.pad_d.release();
    ctx->rows_f32.release();
    ctx->rows_f64.release();
    ctx->maxabs.release();
    ctx->scales_tmp.release();
    ctx->filtered.release();
    ctx->raw_stage.release();
    for (int i = 0; i < 2; ++i) {
        ctx->slab[i].release();
        ctx->pin_up[i].release();
        ctx->pin_dl[i].release();
    }
    ctx->slab_conv.release();
    ctx->ell.release();
    ctx->fp_work.release();
    delete ctx;
    return CTG_OK;
}

ctg_status ctg_set_geometry(ctg_context* ctx, const ctg_geometry* g) {
    clear_error();
    CTG_TRY(check_ctx(ctx, false));
    if (!g) return fail(CTG_ERROR_INVALID_ARGUMENT, "null geometry");
    if (g->det_width < 2 || g->det_height < 2 || g->nx < 0 || g->ny < 0 || g->nz < 0 || !(g->fcd > 0.0f))
        return fail(CTG_ERROR_INVALID_ARGUMENT, "invalid geometry");
    ctx->geom = *g;
    ctx->has_geom = true;
    return CTG_OK;
}

ctg_status ctg_set_projections(ctg_context* ctx, const double* angles_deg, const double* heights_px, int32_t count) {
    clear_error();
    CTG_TRY(check_ctx(ctx, false));
    if (count < 0 || (count > 0 && !angles_deg)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad projection table");
    ctx->proj_host.resize(count);
    ctx->proj_host64.resize(count);
    for (int i = 0; i < count; ++i) {
        // Ct/CtVolume.cpp:1200-1202: angle_rad = angle/180*pi, sin/cos in double
        const double a = (angles_deg[i] / 180.0) * 3.14159265358979323846;
        const double sn = std::sin(a), cs = std::cos(a);
        const double h = heights_px ? heights_px[i] : 0.0;
        ctx->proj_host[i] = make_float4((float)sn, (float)cs, (float)h, 0.0f);
        ctx->proj_host64[i] = make_double4(sn, cs, h, 0.0);
    }
    CTG_TRY(ctx->proj_dev.reserve(sizeof(float4) * std::max(count, 1)));
    CTG_TRY(ctx->proj_dev64.reserve(sizeof(double4) * std::max(count, 1)));
    if (count > 0) {
        CTG_HIP(hipMemcpy(ctx->proj_dev.ptr, ctx->proj_host.data(), sizeof(float4) * count, hipMemcpyHostToDevice));
        CTG_HIP(hipMemcpy(ctx->proj_dev64.ptr, ctx->proj_host64.data(), sizeof(double4) * count,
                          hipMemcpyHostToDevice));
    }
    ctx->nproj_table = count;
    return CTG_OK;
}

void ctg_cancel(ctg_context* ctx) {
    if (ctx) ctx->cancel.store(true);
}

void ctg_reset_cancel(ctg_context* ctx) {
    if (ctx) ctx->cancel.store(false);
}

ctg_status ctg_set_spare_memory(ctg_context* ctx, uint64_t mib) {
    if (!ctx) return fail(CTG_ERROR_INVALID_ARGUMENT, "null context");
    ctx->spare_mib = mib;
    return CTG_OK;
}

ctg_status ctg_set_memory_budget(ctg_context* ctx, uint64_t bytes) {
    if (!ctx) return fail(CTG_ERROR_INVALID_ARGUMENT, "null context");
    ctx->budget_bytes = bytes;
    return CTG_OK;
}

size_t ctg_filtered_bytes(const ctg_geometry* g, ctg_store store, int32_t nproj) {
    if (!g || nproj < 0 || bad_store(store)) return 0;
    return (size_t)g->det_width * g->det_height * store_bytes(store) * (size_t)nproj;
}

ctg_status ctg_preprocess(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int32_t nproj,
                          float base_intensity, ctg_log_form log_form, ctg_filter filter, ctg_store store,
                          void* d_filtered, float* d_scales, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (nproj < 0 || (nproj > 0 && (!d_raw || !d_filtered)))
        return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (base_intensity == 0.0f)  // Ct/CtVolume.cpp:177-181
        return fail(CTG_ERROR_INVALID_ARGUMENT, "the base intensity must not be equal to zero");
    if (bad_pixel(pixel) || bad_filter(filter) || bad_store(store) || bad_log_form(log_form))
        return fail(CTG_ERROR_INVALID_ARGUMENT, "bad enum value");
    CTG_TRY(check_scales(store, d_scales, nproj));
    if (nproj == 0) return CTG_OK;
    return preprocess_batched(ctx, d_raw, pixel, nproj, base_intensity, log_form, filter, store, d_filtered,
                              d_scales, (hipStream_t)stream);
}

ctg_status ctg_feldkamp_weight(ctg_context* ctx, float* d_images, int32_t nproj, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (nproj < 0 || (nproj > 0 && !d_images)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (nproj == 0) return CTG_OK;
    return launch_preprocess_rows(ctx, d_images, CTG_PIXEL_F32, nproj, 1.0f, CTG_LOG_CPU_FORM, CTG_FILTER_RAMLAK,
                                  d_images, nullptr, false, true, false, (hipStream_t)stream);
}

ctg_status ctg_filter_rows(ctg_context* ctx, const float* d_in, float* d_out, int64_t rows, ctg_filter filter,
                           void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (bad_filter(filter)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad filter");
    if (rows < 0 || (rows > 0 && (!d_in || !d_out))) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (rows == 0) return CTG_OK;
    return launch_filter_rows_only(ctx, d_in, d_out, rows, filter, (hipStream_t)stream);
}

ctg_status ctg_layout_projections(ctg_context* ctx, const float* d_images, int32_t nproj, ctg_store store,
                                  void* d_filtered, float* d_scales, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (bad_store(store)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad store");
    if (nproj < 0 || (nproj > 0 && (!d_images || !d_filtered))) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    CTG_TRY(check_scales(store, d_scales, nproj));
    if (nproj == 0) return CTG_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t img = (size_t)ctx->geom.det_width * ctx->geom.det_height;
    CTG_TRY(ctx->maxabs.reserve(sizeof(unsigned) * nproj));
    CTG_TRY(fill_u32(ctx->maxabs.ptr, 0u, (size_t)nproj, st));
    if (store != CTG_STORE_F32) CTG_TRY(launch_maxabs(d_images, img, nproj, ctx->maxabs.as<unsigned>(), st));
    return launch_layout(ctx->geom, d_images, nproj, store, ctx->maxabs.as<unsigned>(), d_filtered, d_scales, st);
}

ctg_status ctg_layout_projections_noalloc(ctg_context* ctx, const float* d_images, int32_t nproj, ctg_store store,
                                          void* d_filtered, float* d_scales, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (bad_store(store)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad store");
    if (nproj < 0 || (nproj > 0 && (!d_images || !d_filtered))) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (nproj > kNoAllocMaxProj)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "%d projections exceed the %d of the allocation-free layout", nproj,
                    kNoAllocMaxProj);
    CTG_TRY(check_scales(store, d_scales, nproj));
    if (nproj == 0) return CTG_OK;
    if (!ctx->noalloc_scratch) return fail(CTG_ERROR_DEVICE, "the static max-abs buffer was not resolved");
    hipStream_t st = (hipStream_t)stream;
    const size_t img = (size_t)ctx->geom.det_width * ctx->geom.det_height;
    unsigned* mx = ctx->noalloc_scratch;  // the device address of `scratch`, never the host stand-in
    CTG_TRY(fill_u32(mx, 0u, (size_t)nproj, st));
    if (store != CTG_STORE_F32) CTG_TRY(launch_maxabs(d_images, img, nproj, mx, st));
    return launch_layout(ctx->geom, d_images, nproj, store, mx, d_filtered, d_scales, st);
}

ctg_status ctg_backproject(ctg_context* ctx, const void* d_filtered, ctg_store store, const float* d_scales,
                           int32_t p0, int32_t nproj, float* d_volume, int32_t z0, int32_t nz, int32_t accumulate,
                           void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, p0, nproj));
    CTG_TRY(check_slab(ctx, z0, nz));
    if (bad_store(store)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad store");
    if (!d_volume || (nproj > 0 && !d_filtered)) return fail(CTG_ERROR_INVALID_ARGUMENT, "null buffers");
    CTG_TRY(check_scales(store, d_scales, nproj));
    if (ctx->cancel.load()) return fail(CTG_ERROR_INTERRUPTED, "interrupted");
    if (nproj == 0) {  // no projection to stage: an overwrite leaves a zero slab
        if (accumulate) return CTG_OK;
        return fill_u32(d_volume, 0u, (size_t)ctx->geom.nx * ctx->geom.ny * nz, (hipStream_t)stream);
    }
    BpLaunch L{};
    L.g = &ctx->geom;
    L.d_params = ctx->proj_dev.as<float4>() + p0;
    L.h_params = ctx->proj_host.data() + p0;
    L.d_scales = d_scales;
    L.d_filtered = d_filtered;
    L.store = store;
    L.nproj = nproj;
    L.d_volume = d_volume;
    L.z0 = z0;
    L.nz = nz;
    L.accumulate = accumulate ? 1 : 0;
    return launch_backproject(L, (hipStream_t)stream);
}

ctg_status ctg_forward_project(ctg_context* ctx, const float* d_volume, int32_t z0, int32_t nz, int32_t p0,
                               int32_t nproj, float* d_proj, int32_t accumulate, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, p0, nproj));
    CTG_TRY(check_slab(ctx, z0, nz));
    if (!d_volume || (nproj > 0 && !d_proj)) return fail(CTG_ERROR_INVALID_ARGUMENT, "null buffers");
    if (nproj == 0) return CTG_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)nproj * ctx->geom.det_width * ctx->geom.det_height;
    if (!accumulate) CTG_TRY(fill_u32(d_proj, 0u, n, st));
    return launch_forward_project(ctx->geom, ctx->proj_dev.as<float4>() + p0, ctx->proj_host.data() + p0, d_volume, z0,
                                  nz, nproj, d_proj, ctx->fp_work, st);
}

ctg_status ctg_fdk_backward(ctg_context* ctx, const float* d_grad_volume, int32_t z0, int32_t nz, int32_t p0,
                            int32_t nproj, ctg_filter filter, float* d_grad_lineint, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, p0, nproj));
    CTG_TRY(check_slab(ctx, z0, nz));
    if (bad_filter(filter)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad filter");
    if (nproj > 0 && (!d_grad_volume || !d_grad_lineint)) return fail(CTG_ERROR_INVALID_ARGUMENT, "null buffers");
    if (nproj == 0) return CTG_OK;
    hipStream_t st = (hipStream_t)stream;
    const ctg_geometry& g = ctx->geom;
    const size_t n = (size_t)nproj * g.det_width * g.det_height;
    CTG_TRY(fill_u32(d_grad_lineint, 0u, n, st));
    CTG_TRY(launch_forward_project(g, ctx->proj_dev.as<float4>() + p0, ctx->proj_host.data() + p0, d_grad_volume, z0,
                                   nz, nproj, d_grad_lineint, ctx->fp_work, st));
    CTG_TRY(launch_filter_rows_only(ctx, d_grad_lineint, d_grad_lineint, (int64_t)nproj * g.det_height, filter, st));
    return launch_preprocess_rows(ctx, d_grad_lineint, CTG_PIXEL_F32, nproj, 1.0f, CTG_LOG_CPU_FORM, filter,
                                  d_grad_lineint, nullptr, false, true, false, st);
}

ctg_status ctg_fdk_f64(ctg_context* ctx, const void* d_raw, ctg_pixel_type pixel, int32_t nproj, float base_intensity,
                       ctg_filter filter, double* d_volume, int32_t z0, int32_t nz, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, 0, nproj));
    CTG_TRY(check_slab(ctx, z0, nz));
    if (base_intensity == 0.0f) return fail(CTG_ERROR_INVALID_ARGUMENT, "the base intensity must not be equal to zero");
    if (bad_pixel(pixel) || bad_filter(filter)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad enum value");
    if ((nz > 0 && !d_volume) || (nproj > 0 && !d_raw)) return fail(CTG_ERROR_INVALID_ARGUMENT, "null buffers");
    hipStream_t st = (hipStream_t)stream;
    const ctg_geometry& g = ctx->geom;
    const size_t img = (size_t)g.det_width * g.det_height;
    CTG_TRY(ctx->rows_f64.reserve(img * 8 * std::max(nproj, 1)));
    if (nproj > 0)
        CTG_TRY(launch_preprocess_rows_f64(ctx, d_raw, pixel, nproj, base_intensity, filter,
                                           ctx->rows_f64.as<double>(), st));
    return launch_backproject_f64(g, ctx->proj_dev64.as<double4>(), ctx->rows_f64.as<double>(), nproj, d_volume, z0,
                                  nz, st);
}

ctg_status ctg_convert_volume(ctg_context* ctx, const float* d_volume, int32_t nz, ctg_index_order order,
                              ctg_byte_order byte_order, void* d_out, void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    if (!d_volume || !d_out || nz < 0) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (nz > ctx->geom.nz) return fail(CTG_ERROR_INVALID_ARGUMENT, "slab of %d slices exceeds the grid's %d", nz, ctx->geom.nz);
    if ((int)order < 0 || (int)order > 1 || (int)byte_order < 0 || (int)byte_order > 1)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "bad enum value");
    if (nz == 0) return CTG_OK;
    return launch_convert_volume(ctx->geom, d_volume, nz, order, byte_order, d_out, (hipStream_t)stream);
}

ctg_status ctg_phantom_images(ctg_context* ctx, const double* ell, int32_t n_ell, int32_t p0, int32_t nproj,
                              double photons, double full_scale, uint64_t seed, uint16_t* d_images, float* d_lineint,
                              void* stream) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, p0, nproj));
    if (n_ell < 0 || (n_ell > 0 && !ell)) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad ellipsoid list");
    if (nproj == 0) return CTG_OK;
    return launch_phantom(ctx->geom, ctx->proj_dev64.as<double4>(), ell, n_ell, nproj, p0, photons, full_scale, seed,
                          d_images, d_lineint, ctx->ell, (hipStream_t)stream);
}

// Slab pipeline.  Raw images go host -> pinned staging -> device in batches on an upload
// stream while the compute stream preprocesses the previous batch into the resident
// backprojection layout; then every z-slab is backprojected from all projections,
// converted to the requested index order and copied out.  When the filtered projections
// do not fit next to a useful slab, they are recomputed per slab from re-uploaded images.
namespace {

struct ReconPlan {
    int batch = 1;       // projections uploaded and preprocessed per step
    bool resident = true;
    int slab_nz = 0;
    bool double_buffer = false;  // two device slabs: slab s downloads under slab s+1
};

// Memory plan of ctg_reconstruct: the working set (staging and FP32 rows of one batch),
// the filtered projections (all of them when they fit next to a useful slab, else one
// batch, recomputed per slab), and the largest slab (z-slices of the volume plus its
// index-order conversion) that fits the rest.  Available memory is the device's free
// memory minus the spare memory, capped by the budget; buffers the context already
// holds count as available because they are reused.
ctg_status plan_reconstruct(ctg_context* ctx, ctg_pixel_type pixel, int nproj, ctg_store store, ReconPlan& plan) {
    const ctg_geometry& g = ctx->geom;
    const size_t img = (size_t)g.det_width * g.det_height;
    const size_t pb = pixel_bytes(pixel);
    const size_t col = (size_t)g.nx * g.ny;
    size_t fr = 0, tot = 0;
    CTG_HIP(hipMemGetInfo(&fr, &tot));
    const long long spare = (long long)ctx->spare_mib * 1024ll * 1024ll;
    long long avail = (long long)fr - spare;
    avail += (long long)ctx->filtered.bytes + (long long)ctx->raw_stage.bytes + (long long)ctx->rows_f32.bytes +
             (long long)ctx->slab[0].bytes + (long long)ctx->slab[1].bytes + (long long)ctx->slab_conv.bytes;
    if (ctx->budget_bytes > 0) avail = std::min<long long>(avail, (long long)ctx->budget_bytes);
    plan.batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nproj, ((size_t)256 << 20) / (img * 4)));
    const long long work = (long long)img * plan.batch * (4 + 2 * pb) + (long long)nproj * 8;
    const long long filt = (long long)ctg_filtered_bytes(&g, store, nproj);
    const long long slab_col = (long long)col * 4 * 2;  // slab + converted copy, per z
    plan.resident = avail - work - filt >= slab_col * std::min(g.nz, 32);
    const long long for_slabs =
        avail - work - (plan.resident ? filt : (long long)ctg_filtered_bytes(&g, store, plan.batch));
    if (for_slabs < slab_col) {
        return fail(CTG_ERROR_INSUFFICIENT_VRAM,
                    "The VRAM of the GPU is insufficient to run a reconstruction (free %zu MiB, spare %llu MiB).",
                    fr >> 20, (unsigned long long)ctx->spare_mib);
    }
    plan.slab_nz = (int)std::min<long long>(g.nz, for_slabs / slab_col);
    // split volumes use slabs of whole 32-slice runs (the backprojection's run length)
    if (plan.slab_nz < g.nz && plan.slab_nz >= 32) plan.slab_nz -= plan.slab_nz % 32;
    // A resident volume of 256 or more slices that fits whole is cut into about four
    // slabs of whole runs when two slab buffers and the converted copy fit: the
    // download of each slab then runs under the backprojection of the next, and only
    // the last slab's download is left at the end.
    if (plan.resident && plan.slab_nz == g.nz && g.nz >= 256) {
        const int part = (g.nz + 3) / 4;
        const int pipe_nz = (part + 31) / 32 * 32;
        if ((long long)col * 4 * 3 * pipe_nz <= for_slabs) {
            plan.slab_nz = pipe_nz;
            plan.double_buffer = true;
        }
    }
    return CTG_OK;
}

}  // namespace

ctg_status ctg_plan_reconstruct(ctg_context* ctx, ctg_pixel_type pixel, int32_t nproj, ctg_store store,
                                int32_t* resident, int32_t* slab_nz) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, 0, nproj));
    if ((int)pixel < 0 || (int)pixel > 2 || (int)store < 0 || (int)store > 2 || nproj <= 0)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "bad pixel type, store or projection count");
    ReconPlan plan;
    CTG_TRY(plan_reconstruct(ctx, pixel, nproj, store, plan));
    if (resident) *resident = plan.resident ? 1 : 0;
    if (slab_nz) *slab_nz = plan.slab_nz;
    return CTG_OK;
}

ctg_status ctg_reconstruct(ctg_context* ctx, const void* h_raw, ctg_pixel_type pixel, int32_t nproj,
                           float base_intensity, ctg_log_form log_form, ctg_filter filter, ctg_store store,
                           float* h_volume, ctg_index_order order, ctg_progress_fn progress, void* user) {
    clear_error();
    CTG_TRY(check_ctx(ctx));
    CTG_TRY(check_range(ctx, 0, nproj));
    if (!h_raw || !h_volume || nproj <= 0) return fail(CTG_ERROR_INVALID_ARGUMENT, "bad buffers");
    if (base_intensity == 0.0f) return fail(CTG_ERROR_INVALID_ARGUMENT, "the base intensity must not be equal to zero");
    if (bad_pixel(pixel) || bad_store(store) || bad_filter(filter) || bad_log_form(log_form) || (int)order < 0 ||
        (int)order > 1)
        return fail(CTG_ERROR_INVALID_ARGUMENT, "bad pixel type, store, filter, log form or index order");
    const ctg_geometry& g = ctx->geom;
    const size_t img = (size_t)g.det_width * g.det_height;
    const size_t pb = pixel_bytes(pixel);
    const size_t col = (size_t)g.nx * g.ny;
    if (col == 0 || g.nz == 0) return CTG_OK;

    ReconPlan plan;
    CTG_TRY(plan_reconstruct(ctx, pixel, nproj, store, plan));
    const int batch = plan.batch;
    const bool resident = plan.resident;
    const int slab_nz = plan.slab_nz;
    const bool dbl = plan.double_buffer;
    const long long filt = (long long)ctg_filtered_bytes(&g, store, nproj);

    // Streams: uploads (s_up), preprocessing (s_pre), backprojection (s_cmp) and
    // downloads (s_dl).  Staging buffer i (pinned host and device raw) carries batch b
    // with i = b & 1: the host refills pinned[i] once upload b-2 is done, the upload of
    // b waits until preprocessing b-2 has read raw_stage[i].  With two slab buffers,
    // slab s is backprojected into d_slab[s & 1] while slab s-1 is downloaded.
    hipStream_t s_up = nullptr, s_pre = nullptr, s_cmp = nullptr, s_dl = nullptr;
    hipEvent_t ev_up[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr}, ev_dl[2] = {nullptr, nullptr},
               ev_slab[2] = {nullptr, nullptr}, ev_bp = nullptr;
    // pinned staging and device slabs are the context's, kept for the next call
    void* pinned[2] = {nullptr, nullptr};
    void* pinned_dl[2] = {nullptr, nullptr};
    float* d_slab[2] = {nullptr, nullptr};
    void* d_conv = nullptr;
    ctg_status st = CTG_OK;
    auto cleanup = [&]() {
        if (s_up) (void)hipStreamSynchronize(s_up);
        if (s_pre) (void)hipStreamSynchronize(s_pre);
        if (s_cmp) (void)hipStreamSynchronize(s_cmp);
        if (s_dl) (void)hipStreamSynchronize(s_dl);
        for (int i = 0; i < 2; ++i) {
            if (ev_up[i]) (void)hipEventDestroy(ev_up[i]);
            if (ev_done[i]) (void)hipEventDestroy(ev_done[i]);
            if (ev_dl[i]) (void)hipEventDestroy(ev_dl[i]);
            if (ev_slab[i]) (void)hipEventDestroy(ev_slab[i]);
        }
        if (ev_bp) (void)hipEventDestroy(ev_bp);
        if (s_up) (void)hipStreamDestroy(s_up);
        if (s_pre) (void)hipStreamDestroy(s_pre);
        if (s_cmp) (void)hipStreamDestroy(s_cmp);
        if (s_dl) (void)hipStreamDestroy(s_dl);
    };
#define PIPE_HIP(call)                                                                                      \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            st = fail(e_ == hipErrorOutOfMemory ? CTG_ERROR_OUT_OF_MEMORY : CTG_ERROR_DEVICE, "%s failed: %s", \
                      #call, hipGetErrorString(e_));                                                        \
            cleanup();                                                                                      \
            return st;                                                                                      \
        }                                                                                                   \
    } while (0)
#define PIPE_TRY(expr)            \
    do {                          \
        st = (expr);              \
        if (st != CTG_OK) {       \
            cleanup();            \
            return st;            \
        }                         \
    } while (0)

    // pinned staging: one batch of raw images for the uploads, 32 MiB chunks for the
    // downloads
    const size_t stage_bytes = img * pb * batch;
    const size_t dl_bytes = (size_t)32 << 20;
    PIPE_HIP(hipStreamCreateWithFlags(&s_up, hipStreamNonBlocking));
    PIPE_HIP(hipStreamCreateWithFlags(&s_pre, hipStreamNonBlocking));
    PIPE_HIP(hipStreamCreateWithFlags(&s_cmp, hipStreamNonBlocking));
    PIPE_HIP(hipStreamCreateWithFlags(&s_dl, hipStreamNonBlocking));
    PIPE_HIP(hipEventCreateWithFlags(&ev_bp, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) {
        PIPE_HIP(hipEventCreateWithFlags(&ev_up[i], hipEventDisableTiming));
        PIPE_HIP(hipEventCreateWithFlags(&ev_done[i], hipEventDisableTiming));
        PIPE_HIP(hipEventCreateWithFlags(&ev_dl[i], hipEventDisableTiming));
        PIPE_HIP(hipEventCreateWithFlags(&ev_slab[i], hipEventDisableTiming));
        PIPE_TRY(ctx->pin_up[i].reserve(stage_bytes));
        PIPE_TRY(ctx->pin_dl[i].reserve(dl_bytes));
        pinned[i] = ctx->pin_up[i].ptr;
        pinned_dl[i] = ctx->pin_dl[i].ptr;
    }
    PIPE_TRY(ctx->raw_stage.reserve(img * pb * batch * 2));
    // resident: every projection's entries stay in HBM for all slabs; otherwise one
    // batch of entries is re-derived from its uploaded images for every slab
    PIPE_TRY(ctx->filtered.reserve((size_t)(resident ? filt : (long long)ctg_filtered_bytes(&g, store, batch))));
    PIPE_TRY(ctx->scales_tmp.reserve(sizeof(float) * nproj));
    // Z_FASTEST host volumes ([x][y][z]) take a cut planned along z as twice as many x
    // ranges (whole 16-column tiles) instead: an x range is contiguous in the host
    // volume, so its download is one sequential copy rather than a strided one per
    // column.  Every range's download runs under the next range's backprojection except
    // the last.  1024^3 from 720 x 1024^2 is planned as four z-slabs, so eight ranges:
    // 206.6 ms host to host, against 228.8 / 212.7 / 212.9 ms for 2 / 4 / 16 ranges
    // (CTG_RECON_XPARTS=n sets the count, tuning).  The backprojection of a range is the
    // one of the whole grid restricted to it (a grid of its columns with x_pre shifted by
    // its start): same voxels, same tiles, same results.
    const int nslabs_z = (g.nz + slab_nz - 1) / slab_nz;
    int part_x = g.nx;
    if (dbl && order == CTG_Z_FASTEST) {
        int parts = 2 * nslabs_z;
        if (const char* e = std::getenv("CTG_RECON_XPARTS")) parts = std::max(1, std::atoi(e));
        part_x = ((g.nx + parts - 1) / parts + 15) / 16 * 16;
        if ((size_t)part_x * g.ny * g.nz > col * (size_t)slab_nz) part_x = g.nx;  // no larger than the planned slab
    }
    const bool xcut = part_x < g.nx;
    const size_t slab_floats = xcut ? (size_t)part_x * g.ny * g.nz : col * (size_t)slab_nz;
    for (int i = 0; i < (dbl ? 2 : 1); ++i) {
        PIPE_TRY(ctx->slab[i].reserve(slab_floats * 4));
        d_slab[i] = ctx->slab[i].as<float>();
    }
    int cur = 0;                     // slab buffer being backprojected
    ctg_geometry gcur = g;           // grid of the slab being backprojected
    if (order == CTG_X_FASTEST) {
        PIPE_TRY(ctx->slab_conv.reserve(col * 4 * (size_t)slab_nz));
        d_conv = ctx->slab_conv.ptr;
    }
    const int nb_total = (nproj + batch - 1) / batch;

    // Upload batch b through pinned staging and preprocess it on s_pre into `out`
    // (entries of projections b*batch..).  Returns after the host copy; the GPU work is
    // queued.  ev_done[b & 1] marks the end of its preprocessing.
    auto stage_batch = [&](int b, char* out) -> ctg_status {
        if (ctx->cancel.load()) return fail(CTG_ERROR_INTERRUPTED, "interrupted");
        const int i = b & 1;
        const int b0 = b * batch, nb = std::min(batch, nproj - b0);
        CTG_HIP(hipEventSynchronize(ev_up[i]));  // upload b-2 has read pinned[i]
        parallel_memcpy(pinned[i], (const char*)h_raw + (size_t)b0 * img * pb, (size_t)nb * img * pb);
        char* dst = (char*)ctx->raw_stage.ptr + (size_t)i * img * pb * batch;
        CTG_HIP(hipStreamWaitEvent(s_up, ev_done[i], 0));  // preprocessing b-2 has read raw_stage[i]
        CTG_HIP(hipMemcpyAsync(dst, pinned[i], (size_t)nb * img * pb, hipMemcpyHostToDevice, s_up));
        CTG_HIP(hipEventRecord(ev_up[i], s_up));
        CTG_HIP(hipStreamWaitEvent(s_pre, ev_up[i], 0));
        CTG_TRY(preprocess_batched(ctx, dst, pixel, nb, base_intensity, log_form, filter, store, out,
                                   ctx->scales_tmp.as<float>() + b0, s_pre));
        CTG_HIP(hipEventRecord(ev_done[i], s_pre));
        return CTG_OK;
    };
    // backprojection on s_cmp of projections [p0, p0+np) once their preprocessing (in
    // stream order on s_pre, up to the last recorded batch) has finished
    auto backproject = [&](int p0, int np, int z0, int nz, int accumulate, int last_batch) -> ctg_status {
        if (last_batch >= 0) CTG_HIP(hipStreamWaitEvent(s_cmp, ev_done[last_batch & 1], 0));
        BpLaunch L{};
        L.g = &gcur;
        L.d_params = ctx->proj_dev.as<float4>() + p0;
        L.h_params = ctx->proj_host.data() + p0;
        L.d_scales = ctx->scales_tmp.as<float>() + p0;
        L.d_filtered = (const char*)ctx->filtered.ptr + (resident ? (size_t)p0 * img * store_bytes(store) : 0);
        L.store = store;
        L.nproj = np;
        L.d_volume = d_slab[cur];
        L.z0 = z0;
        L.nz = nz;
        L.accumulate = accumulate;
        CTG_TRY(launch_backproject(L, s_cmp));
        CTG_HIP(hipEventRecord(ev_bp, s_cmp));
        return CTG_OK;
    };
    // Slab buffer `buf` -> host volume on s_dl, through the pinned download buffers in
    // chunks: the DMA of chunk k+1 runs while host threads place chunk k (contiguous for
    // X_FASTEST; for Z_FASTEST, columns of nz values at a stride of nz_total).  It
    // starts once the slab's backprojection (ev_slab[buf] on s_cmp) has finished and
    // has finished itself when this returns.
    auto download = [&](int z0, int nz, int x0, int nxs, int buf) -> ctg_status {
        CTG_HIP(hipStreamWaitEvent(s_dl, ev_slab[buf], 0));
        const float* src = d_slab[buf];
        if (order == CTG_X_FASTEST) {
            CTG_TRY(launch_convert_volume(g, d_slab[buf], nz, CTG_X_FASTEST, CTG_LITTLE_ENDIAN, d_conv, s_dl));
            src = (const float*)d_conv;
        }
        // an x range of a Z_FASTEST volume, or any X_FASTEST slab: one contiguous block
        const bool contiguous = xcut || order == CTG_X_FASTEST;
        const size_t block0 = xcut ? (size_t)x0 * g.ny * g.nz : (size_t)z0 * col;
        const size_t unit = contiguous ? 1 : (size_t)nz;  // floats per unit: one value, or a column
        const size_t n_units = xcut ? (size_t)nxs * g.ny * g.nz : (contiguous ? col * (size_t)nz : col);
        const size_t per_chunk = std::max<size_t>(1, dl_bytes / (unit * 4));
        if (per_chunk * unit * 4 > dl_bytes) return fail(CTG_ERROR_UNSUPPORTED, "volume column exceeds the download chunk");
        const size_t nchunks = (n_units + per_chunk - 1) / per_chunk;
        auto issue = [&](size_t k) -> ctg_status {
            const size_t u0 = k * per_chunk, u1 = std::min(n_units, u0 + per_chunk);
            CTG_HIP(hipMemcpyAsync(pinned_dl[k & 1], src + u0 * unit, (u1 - u0) * unit * 4, hipMemcpyDeviceToHost,
                                   s_dl));
            CTG_HIP(hipEventRecord(ev_dl[k & 1], s_dl));
            return CTG_OK;
        };
        CTG_TRY(issue(0));
        for (size_t k = 0; k < nchunks; ++k) {
            if (k + 1 < nchunks) CTG_TRY(issue(k + 1));
            CTG_HIP(hipEventSynchronize(ev_dl[k & 1]));
            const size_t u0 = k * per_chunk, u1 = std::min(n_units, u0 + per_chunk);
            const float* stage = (const float*)pinned_dl[k & 1];
            if (contiguous) {
                parallel_memcpy(h_volume + block0 + u0, stage, (u1 - u0) * 4);
            } else if (nz == g.nz) {
                parallel_memcpy(h_volume + u0 * (size_t)g.nz, stage, (u1 - u0) * unit * 4);
            } else {
                parallel_ranges(u1 - u0, 4096, [&](size_t a, size_t b) {
                    for (size_t c = a; c < b; ++c)
                        std::memcpy(h_volume + (u0 + c) * (size_t)g.nz + z0, stage + c * unit, unit * 4);
                });
            }
        }
        return CTG_OK;
    };

    const int nslabs = xcut ? (g.nx + part_x - 1) / part_x : nslabs_z;
    // slab s: z range [z0, z0 + nz) of x range [x0, x0 + nxs)
    auto slab_of = [&](int s, int& z0, int& nz, int& x0, int& nxs) {
        if (xcut) {
            z0 = 0; nz = g.nz;
            x0 = s * part_x; nxs = std::min(part_x, g.nx - x0);
        } else {
            z0 = s * slab_nz; nz = std::min(slab_nz, g.nz - z0);
            x0 = 0; nxs = g.nx;
        }
    };
    for (int s = 0; s < nslabs; ++s) {
        if (ctx->cancel.load()) {
            st = fail(CTG_ERROR_INTERRUPTED, "interrupted");
            cleanup();
            return st;
        }
        int z0, nz, x0, nxs;
        slab_of(s, z0, nz, x0, nxs);
        gcur = g;
        gcur.nx = nxs;
        gcur.x_pre = g.x_pre - (float)x0;
        if (resident) {
            // Every slab is backprojected in the same groups of batches (up to four), so
            // each voxel sees the same sequence of launches whatever the slab plan.  In
            // the first slab a group is backprojected as soon as it is preprocessed, and
            // the uploads and preprocessing of the next group run under it.
            const int groups = std::min(nb_total, 4);
            int bdone = 0;
            for (int grp = 0; grp < groups; ++grp) {
                const int bend = (int)((long long)nb_total * (grp + 1) / groups);
                if (s == 0) {
                    for (int b = bdone; b < bend; ++b)
                        PIPE_TRY(stage_batch(b, (char*)ctx->filtered.ptr + (size_t)b * batch * img * store_bytes(store)));
                }
                const int p0 = bdone * batch, p1 = std::min(nproj, bend * batch);
                PIPE_TRY(backproject(p0, p1 - p0, z0, nz, grp > 0 ? 1 : 0, s == 0 ? bend - 1 : -1));
                bdone = bend;
            }
        } else {
            // the single batch buffer is rewritten only after the previous batch's
            // backprojection has read it
            for (int b = 0; b < nb_total; ++b) {
                const int b0 = b * batch, nb = std::min(batch, nproj - b0);
                PIPE_HIP(hipStreamWaitEvent(s_pre, ev_bp, 0));
                PIPE_TRY(stage_batch(b, (char*)ctx->filtered.ptr));
                PIPE_TRY(backproject(b0, nb, z0, nz, b > 0 ? 1 : 0, b));
            }
        }
        PIPE_HIP(hipEventRecord(ev_slab[cur], s_cmp));
        if (!dbl) {
            PIPE_TRY(download(z0, nz, x0, nxs, cur));
            if (progress) progress((double)(s + 1) / nslabs, user);
        } else {
            // the previous slab downloads while this one is backprojected; the buffer it
            // used is rewritten (slab s+1) only after that download has returned
            if (s > 0) {
                int zp, nzp, xp, nxp;
                slab_of(s - 1, zp, nzp, xp, nxp);
                PIPE_TRY(download(zp, nzp, xp, nxp, cur ^ 1));
                if (progress) progress((double)s / nslabs, user);
            }
            cur ^= 1;
        }
    }
    if (dbl) {
        int zl, nzl, xl, nxl;
        slab_of(nslabs - 1, zl, nzl, xl, nxl);
        PIPE_TRY(download(zl, nzl, xl, nxl, cur ^ 1));
        if (progress) progress(1.0, user);
    }
    cleanup();
#undef PIPE_HIP
#undef PIPE_TRY
    return CTG_OK;
}

}  // extern "C"
