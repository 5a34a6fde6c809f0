/*
 * oracle.c -- CPU restatement of Ct's FDK hot path (TEST INFRASTRUCTURE ONLY;
 * see oracle.h).  Plain C99 + OpenMP, built by oracle/Makefile.
 *
 * What pins each function (checks live in tests/test_oracle.py):
 *   orc_cylinder_radius / orc_volume_geometry / orc_grid_geometry
 *       closed forms for R at every configured detector; ROI rounding
 *       floor(f*N+0.5) including the reference's yTo/zTo clamp against xFrom;
 *       precomputes against hand-evaluated values.
 *   orc_log_cpu / orc_log_gpu
 *       closed form: I = round(S*B*exp(-p)) returns p to the 1e-6 epsilon and
 *       16-bit quantisation; both forms agree to float rounding.
 *   orc_feldkamp
 *       closed form FCD/sqrt(FCD^2+u^2+v^2); centre pixel weight 1; symmetry.
 *   orc_filter_window
 *       hand-evaluated H(k) at chosen k for each window; H(0) = 0.
 *   orc_filter_rows (FFT / circular-convolution paths)
 *       against orc_filter_rows_direct_f64 (explicit DFT + C2R definition),
 *       impulse response = unnormalised inverse DFT of H, DC row -> 0,
 *       linearity, shift equivariance.
 *   orc_filter_rows_direct_f64
 *       numpy.fft.rfft/irfft(*W) on random rows (independent library routine).
 *   orc_backproject_f32 / _f64
 *       8^3 brute force written independently in the test from the A.3 model;
 *       single projection at theta=0 of a constant detector: value = U^2 in
 *       closed form; cylinder mask exact zeros; linearity; f32 vs f64.
 *   orc_forward_project_f64
 *       dot-product adjoint test against orc_backproject_f64 (<Ax,y>=<x,A^T y>).
 *   orc_phantom_project
 *       centred sphere: chord 2*sqrt(r^2-d^2) per pixel in closed form.
 *   orc_fdk_f32 / _f64 / orc_preprocess_f32
 *       compositions of the pinned pieces; centred uniform-sphere phantom
 *       reconstructs to a flat interior (sanity, not parity).
 * The reference itself cannot be built here (OpenCV, cuFFT and Qt absent), so
 * no function is pinned against reference binaries; cv::dft/cv::log/cuFFT
 * rounding is unpinned and far below the 1e-4 parity tolerance.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }

int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#ifdef _OPENMP
#define ORC_NTHREADS num_threads(orc_get_threads())
#else
#define ORC_NTHREADS
#endif

/* ------------------------------------------------------------------ geometry */

/* Ct/CtVolume.cpp:370-373: double radius = imageWidth / 2 (unsigned integer
 * division); sqrt((FCD*FCD * r*r) / (FCD*FCD + r*r)) truncated to unsigned.
 * FCD is a float member, so FCD*FCD is a float product. */
uint32_t orc_cylinder_radius(uint32_t det_width, float fcd_px) {
    double r = (double)(det_width / 2u);
    float f2 = fcd_px * fcd_px;
    return (uint32_t)sqrt(((double)f2 * r * r) / ((double)f2 + r * r));
}

/* Detector window, Ct/CtVolume.cpp:924-928 (matToImageU/V, :1459-1466). */
static void set_detector_window(orc_geometry* g) {
    uint32_t W = (uint32_t)g->det_width, H = (uint32_t)g->det_height;
    g->u_pre = (float)((double)(float)W / 2.0);
    g->v_pre = (float)((double)(float)H / 2.0);
    g->u_lo = 0.0f - g->u_pre;
    g->u_hi = (float)((double)(W - 1u) - 0.1) - g->u_pre;
    g->v_lo = (-1.0f) * (float)((double)(H - 1u) - 0.1) + g->v_pre;
    g->v_hi = (-1.0f) * 0.0f + g->v_pre;
}

int orc_volume_geometry(uint32_t W, uint32_t H, float fcd_px, float u_offset_px,
                        const float* bounds, orc_geometry* g, uint32_t* roi_out,
                        float* bounds_eff) {
    /* Ct/CtVolume.cpp:311-316 */
    uint32_t R = orc_cylinder_radius(W, fcd_px);
    uint32_t xSize = R * 2u, ySize = R * 2u, zSize = H;
    float xf = 0, xt = 1, yf = 0, yt = 1, zf = 0, zt = 1;
    /* Ct/CtVolume.cpp:487-500 (the yTo/zTo clamp against xFrom is kept) */
    if (bounds && !(bounds[0] == bounds[1] || bounds[2] == bounds[3] || bounds[4] == bounds[5])) {
        xf = fmaxf(0.0f, fminf(1.0f, bounds[0]));
        xt = fmaxf(xf, fminf(1.0f, bounds[1]));
        yf = fmaxf(0.0f, fminf(1.0f, bounds[2]));
        yt = fmaxf(xf, fminf(1.0f, bounds[3]));
        zf = fmaxf(0.0f, fminf(1.0f, bounds[4]));
        zt = fmaxf(xf, fminf(1.0f, bounds[5]));
    }
    /* Ct/CtVolume.cpp:1399-1424 */
    uint32_t xFrom = (uint32_t)((double)(xf * (float)xSize) + 0.5);
    uint32_t xTo = (uint32_t)((double)(xt * (float)xSize) + 0.5);
    uint32_t yFrom = (uint32_t)((double)(yf * (float)ySize) + 0.5);
    uint32_t yTo = (uint32_t)((double)(yt * (float)ySize) + 0.5);
    uint32_t zFrom = (uint32_t)((double)(zf * (float)zSize) + 0.5);
    uint32_t zTo = (uint32_t)((double)(zt * (float)zSize) + 0.5);
    if (roi_out) {
        roi_out[0] = xFrom; roi_out[1] = xTo; roi_out[2] = yFrom;
        roi_out[3] = yTo; roi_out[4] = zFrom; roi_out[5] = zTo;
    }
    if (bounds_eff) {
        bounds_eff[0] = xf; bounds_eff[1] = xt; bounds_eff[2] = yf;
        bounds_eff[3] = yt; bounds_eff[4] = zf; bounds_eff[5] = zt;
    }
    memset(g, 0, sizeof(*g));
    g->det_width = (int32_t)W;
    g->det_height = (int32_t)H;
    g->fcd = fcd_px;
    g->u_offset = u_offset_px;
    g->radius = (float)R;
    if (xTo < xFrom || yTo < yFrom || zTo < zFrom) return -1;
    g->nx = (int32_t)(xTo - xFrom);
    g->ny = (int32_t)(yTo - yFrom);
    g->nz = (int32_t)(zTo - zFrom);
    g->x_pre = (float)(((double)(float)xSize / 2.0) - (double)(float)xFrom);
    g->y_pre = (float)(((double)(float)ySize / 2.0) - (double)(float)yFrom);
    g->z_pre = (float)(((double)(float)zSize / 2.0) - (double)(float)zFrom);
    set_detector_window(g);
    return 0;
}

void orc_grid_geometry(uint32_t W, uint32_t H, float fcd_px, float u_offset_px,
                       int32_t nx, int32_t ny, int32_t nz, orc_geometry* g) {
    memset(g, 0, sizeof(*g));
    g->det_width = (int32_t)W;
    g->det_height = (int32_t)H;
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->fcd = fcd_px;
    g->u_offset = u_offset_px;
    g->radius = (float)orc_cylinder_radius(W, fcd_px);
    g->x_pre = (float)((double)(float)nx / 2.0);
    g->y_pre = (float)((double)(float)ny / 2.0);
    g->z_pre = (float)((double)(float)nz / 2.0);
    set_detector_window(g);
}

/* ------------------------------------------------------------ preprocessing */

static double pixel_value(const void* img, int depth, size_t i) {
    if (depth == ORC_DEPTH_U8) return (double)((const uint8_t*)img)[i];
    if (depth == ORC_DEPTH_U16) return (double)((const uint16_t*)img)[i];
    return (double)((const float*)img)[i];
}

static double depth_scale(int depth) {
    return depth == ORC_DEPTH_U8 ? 255.0 : (depth == ORC_DEPTH_U16 ? 65535.0 : 1.0);
}

/* CPU form.  convertTo32bit (Ct/CtVolume.cpp:802-812): integer images become
 * I/S + 1e-6*B, float images are left as they are; then applyLogScaling
 * (:860-866): multiply by 1/(B + 1e-6*B), natural log, negate. */
void orc_log_cpu_f64(const void* img, int depth, size_t n, double B, double* out) {
    double S = depth_scale(depth);
    double eps = (depth == ORC_DEPTH_F32) ? 0.0 : 0.000001 * B;
    double k = 1.0 / (B + 0.000001 * B);
    #pragma omp parallel for ORC_NTHREADS
    for (long long i = 0; i < (long long)n; ++i) {
        double x = pixel_value(img, depth, (size_t)i);
        double y = (depth == ORC_DEPTH_F32) ? x : x / S + eps;
        out[i] = -log(y * k);
    }
}

void orc_log_cpu(const void* img, int depth, size_t n, double B, float* out) {
    double S = depth_scale(depth);
    double eps = (depth == ORC_DEPTH_F32) ? 0.0 : 0.000001 * B;
    double k = 1.0 / (B + 0.000001 * B);
    #pragma omp parallel for ORC_NTHREADS
    for (long long i = 0; i < (long long)n; ++i) {
        double x = pixel_value(img, depth, (size_t)i);
        float y = (depth == ORC_DEPTH_F32) ? (float)x : (float)(x / S + eps);
        out[i] = (float)(-log((double)y * k));
    }
}

/* GPU form (Ct/CtVolume.cpp:892-904): convertTo with alpha = (1-1e-6)/(S*B),
 * beta = 1e-6 (S = 1 for float images), then log and multiply by -1. */
void orc_log_gpu(const void* img, int depth, size_t n, double B, float* out) {
    double S = depth_scale(depth);
    double alpha = (1.0 - 0.000001) / (S * B);
    #pragma omp parallel for ORC_NTHREADS
    for (long long i = 0; i < (long long)n; ++i) {
        double x = pixel_value(img, depth, (size_t)i);
        float y = (float)(x * alpha + 0.000001);
        out[i] = (float)(-log((double)y));
    }
}

/* Ct/CtVolume.cpp:814-830: p[r][c] *= D / sqrt(D^2 + u^2 + v^2) with
 * u = c - uPre, v = -r + vPre (matToImageU/V). */
void orc_feldkamp(float* img, int32_t W, int32_t H, int32_t nimg, float fcd, float u_pre, float v_pre) {
    #pragma omp parallel for ORC_NTHREADS
    for (long long rr = 0; rr < (long long)nimg * H; ++rr) {
        int r = (int)(rr % H);
        float* ptr = img + (size_t)rr * W;
        float v = (-1.0f) * (float)r + v_pre;
        for (int c = 0; c < W; ++c) {
            float u = (float)c - u_pre;
            double wgt = (double)fcd / sqrt((double)fcd * fcd + (double)u * u + (double)v * v);
            ptr[c] = (float)(ptr[c] * wgt);
        }
    }
}

void orc_feldkamp_f64(double* img, int32_t W, int32_t H, int32_t nimg, double fcd, double u_pre, double v_pre) {
    #pragma omp parallel for ORC_NTHREADS
    for (long long rr = 0; rr < (long long)nimg * H; ++rr) {
        int r = (int)(rr % H);
        double* ptr = img + (size_t)rr * W;
        double v = -(double)r + v_pre;
        for (int c = 0; c < W; ++c) {
            double u = (double)c - u_pre;
            ptr[c] *= fcd / sqrt(fcd * fcd + u * u + v * v);
        }
    }
}

/* Ct/CtVolume.cpp:868-885 with N = W/2 + 1 (:835).  The float/double mix of
 * the reference expressions is kept (rl is a float, M_PI a double). */
void orc_filter_window(int filter, int32_t W, float* h) {
    int nyq = W / 2 + 1;
    float N = (float)nyq;
    for (int k = 0; k < nyq; ++k) {
        float n = (float)k;
        float rl = n / N;
        if (filter == ORC_SHEPP_LOGAN) {
            if (n == 0.0f) h[k] = 0.0f;
            else {
                double a = (double)(rl * 0.5f) * M_PI;
                h[k] = (float)((double)rl * sin(a) / a);
            }
        } else if (filter == ORC_HANN) {
            h[k] = (float)((double)rl * (double)0.5f * (1.0 + cos(M_PI * (double)rl)));
        } else {
            h[k] = rl;
        }
    }
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* Iterative radix-2 complex FFT, sign = -1 forward, +1 inverse (unscaled). */
static void fft_pow2(double* re, double* im, int n, int sign) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                int j = i + half;
                double xr = re[j] * wr - im[j] * wi;
                double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr; im[j] = im[i] - xi;
                re[i] += xr; im[i] += xi;
            }
        }
    }
}

/* Full-length real even response Htilde(k), k = 0..W-1. */
static void full_response(int filter, int W, double* Hf) {
    float* h = (float*)malloc(sizeof(float) * (size_t)(W / 2 + 1));
    orc_filter_window(filter, W, h);
    for (int k = 0; k < W; ++k) {
        int kk = k <= W / 2 ? k : W - k;
        Hf[k] = (double)h[kk];
    }
    free(h);
}

/* Filter rows of doubles.  Two real rows share one complex FFT (Htilde is real
 * and even, so real rows stay real); non power-of-two W uses the circular
 * convolution with g[m] = sum_k Htilde(k) cos(2 pi k m / W). */
void orc_filter_rows_f64(double* img, int32_t W, int64_t rows, int filter) {
    double* Hf = (double*)malloc(sizeof(double) * (size_t)W);
    full_response(filter, W, Hf);
    if (is_pow2(W)) {
        #pragma omp parallel ORC_NTHREADS
        {
            double* re = (double*)malloc(sizeof(double) * (size_t)W);
            double* im = (double*)malloc(sizeof(double) * (size_t)W);
            #pragma omp for schedule(static)
            for (long long pr = 0; pr < (rows + 1) / 2; ++pr) {
                double* a = img + (size_t)(2 * pr) * W;
                double* b = (2 * pr + 1 < rows) ? a + W : NULL;
                for (int i = 0; i < W; ++i) { re[i] = a[i]; im[i] = b ? b[i] : 0.0; }
                fft_pow2(re, im, W, -1);
                for (int k = 0; k < W; ++k) { re[k] *= Hf[k]; im[k] *= Hf[k]; }
                fft_pow2(re, im, W, +1);
                for (int i = 0; i < W; ++i) { a[i] = re[i]; if (b) b[i] = im[i]; }
            }
            free(re); free(im);
        }
    } else {
        double* gk = (double*)malloc(sizeof(double) * (size_t)W);
        for (int m = 0; m < W; ++m) {
            double s = 0.0;
            for (int k = 0; k < W; ++k) s += Hf[k] * cos(2.0 * M_PI * (double)(((long long)k * m) % W) / (double)W);
            gk[m] = s;
        }
        #pragma omp parallel ORC_NTHREADS
        {
            double* x = (double*)malloc(sizeof(double) * (size_t)W);
            #pragma omp for schedule(static)
            for (long long r = 0; r < rows; ++r) {
                double* row = img + (size_t)r * W;
                memcpy(x, row, sizeof(double) * (size_t)W);
                for (int n = 0; n < W; ++n) {
                    double s = 0.0;
                    for (int m = 0; m < W; ++m) {
                        int d = n - m; if (d < 0) d += W;
                        s += x[m] * gk[d];
                    }
                    row[n] = s;
                }
            }
            free(x);
        }
        free(gk);
    }
    free(Hf);
}

void orc_filter_rows(float* img, int32_t W, int64_t rows, int filter) {
    const int64_t chunk = 4096;
    double* buf = (double*)malloc(sizeof(double) * (size_t)W * (size_t)chunk);
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        int64_t n = rows - r0 < chunk ? rows - r0 : chunk;
        for (int64_t i = 0; i < n * W; ++i) buf[i] = (double)img[r0 * W + i];
        orc_filter_rows_f64(buf, W, n, filter);
        for (int64_t i = 0; i < n * W; ++i) img[r0 * W + i] = (float)buf[i];
    }
    free(buf);
}

/* Definition-level evaluation: X_k = sum_n x_n e^{-2 pi i k n / W} for
 * k = 0..W/2, Y_k = H(k) X_k, then the C2R inverse that reads only the
 * Hermitian half (imaginary parts of Y_0 and, for even W, Y_{W/2} ignored),
 * unnormalised. */
void orc_filter_rows_direct_f64(double* img, int32_t W, int64_t rows, int filter) {
    int nyq = W / 2 + 1;
    float* h = (float*)malloc(sizeof(float) * (size_t)nyq);
    orc_filter_window(filter, W, h);
    double* Xr = (double*)malloc(sizeof(double) * (size_t)nyq);
    double* Xi = (double*)malloc(sizeof(double) * (size_t)nyq);
    double* x = (double*)malloc(sizeof(double) * (size_t)W);
    for (int64_t r = 0; r < rows; ++r) {
        double* row = img + (size_t)r * W;
        memcpy(x, row, sizeof(double) * (size_t)W);
        for (int k = 0; k < nyq; ++k) {
            double sr = 0, si = 0;
            for (int n = 0; n < W; ++n) {
                double a = -2.0 * M_PI * (double)(((long long)k * n) % W) / (double)W;
                sr += x[n] * cos(a); si += x[n] * sin(a);
            }
            Xr[k] = sr * h[k]; Xi[k] = si * h[k];
        }
        for (int n = 0; n < W; ++n) {
            double s = Xr[0];
            int kmax = (W % 2 == 0) ? W / 2 - 1 : W / 2;
            for (int k = 1; k <= kmax; ++k) {
                double a = 2.0 * M_PI * (double)(((long long)k * n) % W) / (double)W;
                s += 2.0 * (Xr[k] * cos(a) - Xi[k] * sin(a));
            }
            if (W % 2 == 0) s += Xr[W / 2] * ((n % 2) ? -1.0 : 1.0);
            row[n] = s;
        }
    }
    free(h); free(Xr); free(Xi); free(x);
}

/* ----------------------------------------------------------- backprojection */

/* Ct/CtVolume.cpp:1048-1054: the two u-lerps are float expressions stored in
 * doubles, the v-lerp is evaluated in double and returned as float. */
static inline float bilerp_ref(float u, float v, float u0v0, float u1v0, float u0v1, float u1v1) {
    double v0 = (1.0f - u) * u0v0 + u * u1v0;
    double v1 = (1.0f - u) * u0v1 + u * u1v1;
    return (float)((1.0f - v) * v0 + v * v1);
}

void orc_backproject_f32(const orc_geometry* g, const float* filtered,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, int32_t step, float* vol) {
    const int W = g->det_width, H = g->det_height;
    const long nx = g->nx, ny = g->ny, nz = g->nz;
    const float uLo = g->u_lo, uHi = g->u_hi, vLo = g->v_lo, vHi = g->v_hi;
    const float yLo = 0.0f - g->y_pre, yHi = (float)ny - g->y_pre;   /* :930-931 */
    const float zLo = 0.0f - g->z_pre, zHi = (float)nz - g->z_pre;   /* :932-933 */
    const float FCD = g->fcd, uOffset = g->u_offset;
    const float radius = g->radius;
    const float radiusSquared = radius * radius;                     /* :939-940 */
    const float uPre = g->u_pre, vPre = g->v_pre, xPre = g->x_pre;
    if (step < 1) step = 1;
    for (int p = 0; p < P; p += step) {                              /* :945 */
        float angle_rad = (float)((angles_deg[p] / 180.0) * M_PI);   /* :953-955 */
        float sine = (float)sin((double)angle_rad);
        float cosine = (float)cos((double)angle_rad);
        const float* image = filtered + (size_t)p * W * H;
        float heightOffset = (float)heights_px[p];                   /* :972 */
        #pragma omp parallel for schedule(dynamic) ORC_NTHREADS
        for (long xIndex = 0; xIndex < nx; ++xIndex) {               /* :975-976 */
            float x = (float)xIndex - xPre;
            float* volumePtr = vol + (size_t)xIndex * ny * nz;
            for (float y = yLo; y < yHi; ++y) {                      /* :979 */
                if ((x * x + y * y) >= radiusSquared) { volumePtr += nz; continue; }
                float t = (-1) * x * sine + y * cosine;              /* :990-995 */
                t += uOffset;
                float s = x * cosine + y * sine;
                float rdw = FCD / (FCD - s);
                float u = t * rdw;
                if (u < uLo || u > uHi) { volumePtr += nz; continue; }
                float w = rdw * rdw;                                 /* :1004 */
                for (float z = zLo; z < zHi; ++z, ++volumePtr) {     /* :1006 */
                    float v = (z + heightOffset) * rdw;
                    if (v >= vLo && v <= vHi) {
                        float uIndex = u + uPre;                     /* :1013-1014 */
                        float vIndex = (-1) * v + vPre;
                        int u0 = (int)uIndex, u1 = u0 + 1;
                        int v0 = (int)vIndex, v1 = v0 + 1;
                        const float* row0 = image + (size_t)v0 * W;
                        const float* row1 = image + (size_t)v1 * W;
                        float value = w * bilerp_ref(uIndex - (float)u0, vIndex - (float)v0,
                                                     row0[u0], row0[u1], row1[u0], row1[u1]);
                        (*volumePtr) += value;                       /* :1038 */
                    }
                }
            }
        }
    }
}

/* Same loop nest and masks in double precision (no float rounding anywhere
 * after the geometry, which is promoted from the stored float values). */
void orc_backproject_f64(const orc_geometry* g, const double* filtered,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, double* vol) {
    const int W = g->det_width, H = g->det_height;
    const long nx = g->nx, ny = g->ny, nz = g->nz;
    const double uLo = g->u_lo, uHi = g->u_hi, vLo = g->v_lo, vHi = g->v_hi;
    const double FCD = g->fcd, uOffset = g->u_offset;
    const double r2 = (double)g->radius * (double)g->radius;
    for (int p = 0; p < P; ++p) {
        double ang = angles_deg[p] / 180.0 * M_PI;
        double sine = sin(ang), cosine = cos(ang);
        const double* image = filtered + (size_t)p * W * H;
        double h = heights_px[p];
        #pragma omp parallel for schedule(dynamic) ORC_NTHREADS
        for (long xi = 0; xi < nx; ++xi) {
            double x = (double)xi - g->x_pre;
            for (long yi = 0; yi < ny; ++yi) {
                double y = (double)yi - g->y_pre;
                if (x * x + y * y >= r2) continue;
                double t = -x * sine + y * cosine + uOffset;
                double s = x * cosine + y * sine;
                double U = FCD / (FCD - s);
                double u = t * U;
                if (u < uLo || u > uHi) continue;
                double w = U * U;
                double uIndex = u + g->u_pre;
                int u0 = (int)uIndex;
                double fu = uIndex - u0;
                double* col = vol + ((size_t)xi * ny + yi) * nz;
                for (long zi = 0; zi < nz; ++zi) {
                    double v = ((double)zi - g->z_pre + h) * U;
                    if (v < vLo || v > vHi) continue;
                    double vIndex = -v + g->v_pre;
                    int v0 = (int)vIndex;
                    double fv = vIndex - v0;
                    const double* r0 = image + (size_t)v0 * W;
                    const double* r1 = r0 + W;
                    double a = (1.0 - fu) * r0[u0] + fu * r0[u0 + 1];
                    double b = (1.0 - fu) * r1[u0] + fu * r1[u0 + 1];
                    col[zi] += w * ((1.0 - fv) * a + fv * b);
                }
            }
        }
    }
}

void orc_forward_project_f64(const orc_geometry* g, const double* vol,
                             const double* angles_deg, const double* heights_px,
                             int32_t P, double* proj) {
    const int W = g->det_width, H = g->det_height;
    const long nx = g->nx, ny = g->ny, nz = g->nz;
    const double uLo = g->u_lo, uHi = g->u_hi, vLo = g->v_lo, vHi = g->v_hi;
    const double FCD = g->fcd, uOffset = g->u_offset;
    const double r2 = (double)g->radius * (double)g->radius;
    #pragma omp parallel for schedule(dynamic) ORC_NTHREADS
    for (int p = 0; p < P; ++p) {
        double ang = angles_deg[p] / 180.0 * M_PI;
        double sine = sin(ang), cosine = cos(ang);
        double* image = proj + (size_t)p * W * H;
        double h = heights_px[p];
        for (long xi = 0; xi < nx; ++xi) {
            double x = (double)xi - g->x_pre;
            for (long yi = 0; yi < ny; ++yi) {
                double y = (double)yi - g->y_pre;
                if (x * x + y * y >= r2) continue;
                double t = -x * sine + y * cosine + uOffset;
                double s = x * cosine + y * sine;
                double U = FCD / (FCD - s);
                double u = t * U;
                if (u < uLo || u > uHi) continue;
                double w = U * U;
                double uIndex = u + g->u_pre;
                int u0 = (int)uIndex;
                double fu = uIndex - u0;
                const double* col = vol + ((size_t)xi * ny + yi) * nz;
                for (long zi = 0; zi < nz; ++zi) {
                    double v = ((double)zi - g->z_pre + h) * U;
                    if (v < vLo || v > vHi) continue;
                    double vIndex = -v + g->v_pre;
                    int v0 = (int)vIndex;
                    double fv = vIndex - v0;
                    double f = w * col[zi];
                    double* r0 = image + (size_t)v0 * W;
                    double* r1 = r0 + W;
                    r0[u0] += (1.0 - fv) * (1.0 - fu) * f;
                    r0[u0 + 1] += (1.0 - fv) * fu * f;
                    r1[u0] += fv * (1.0 - fu) * f;
                    r1[u0 + 1] += fv * fu * f;
                }
            }
        }
    }
}

/* ------------------------------------------------------------- whole chains */

void orc_preprocess_f32(const orc_geometry* g, const void* raw, int depth, double B,
                        int log_gpu_form, int filter, int32_t P, float* out) {
    size_t n = (size_t)P * g->det_width * g->det_height;
    if (log_gpu_form) orc_log_gpu(raw, depth, n, B, out);
    else orc_log_cpu(raw, depth, n, B, out);
    orc_feldkamp(out, g->det_width, g->det_height, P, g->fcd, g->u_pre, g->v_pre);
    orc_filter_rows(out, g->det_width, (int64_t)P * g->det_height, filter);
}

void orc_fdk_f32(const orc_geometry* g, const void* raw, int depth, double B,
                 int log_gpu_form, int filter, const double* angles_deg,
                 const double* heights_px, int32_t P, float* vol) {
    size_t n = (size_t)P * g->det_width * g->det_height;
    float* q = (float*)malloc(sizeof(float) * n);
    orc_preprocess_f32(g, raw, depth, B, log_gpu_form, filter, P, q);
    orc_backproject_f32(g, q, angles_deg, heights_px, P, 1, vol);
    free(q);
}

void orc_fdk_f64(const orc_geometry* g, const void* raw, int depth, double B,
                 int filter, const double* angles_deg, const double* heights_px,
                 int32_t P, double* vol) {
    size_t n = (size_t)P * g->det_width * g->det_height;
    double* q = (double*)malloc(sizeof(double) * n);
    orc_log_cpu_f64(raw, depth, n, B, q);
    orc_feldkamp_f64(q, g->det_width, g->det_height, P, g->fcd, g->u_pre, g->v_pre);
    orc_filter_rows_f64(q, g->det_width, (int64_t)P * g->det_height, filter);
    orc_backproject_f64(g, q, angles_deg, heights_px, P, vol);
    free(q);
}

/* ------------------------------------------------------------------ phantom */

/* Rotation Rz(phi) Ry(theta) Rz(psi); columns are the ellipsoid axes. */
static void zyz(double phi, double th, double psi, double R[9]) {
    double c1 = cos(phi), s1 = sin(phi), c2 = cos(th), s2 = sin(th), c3 = cos(psi), s3 = sin(psi);
    R[0] = c1 * c2 * c3 - s1 * s3; R[1] = -c1 * c2 * s3 - s1 * c3; R[2] = c1 * s2;
    R[3] = s1 * c2 * c3 + c1 * s3; R[4] = -s1 * c2 * s3 + c1 * c3; R[5] = s1 * s2;
    R[6] = -s2 * c3;               R[7] = s2 * s3;                R[8] = c2;
}

/* Source at (s = fcd, t = -uOffset, z = -h); detector pixel (c, r) is the
 * point (s = 0, t = u - uOffset, z = v - h) with u = c - uPre, v = vPre - r,
 * which is the inverse of the mapping used by reconstructionCore
 * (u = (t + uOffset) * U, v = (z + h) * U, Ct/CtVolume.cpp:990-1014). */
void orc_phantom_project(const orc_geometry* g, const double* ell, int32_t n_ell,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, double* out) {
    const int W = g->det_width, H = g->det_height;
    double* rot = (double*)malloc(sizeof(double) * 9 * (size_t)(n_ell > 0 ? n_ell : 1));
    for (int e = 0; e < n_ell; ++e) zyz(ell[e * 10 + 6], ell[e * 10 + 7], ell[e * 10 + 8], rot + 9 * e);
    #pragma omp parallel for schedule(dynamic) ORC_NTHREADS
    for (long long pr = 0; pr < (long long)P * H; ++pr) {
        int p = (int)(pr / H), r = (int)(pr % H);
        double ang = angles_deg[p] / 180.0 * M_PI, cs = cos(ang), sn = sin(ang);
        double uo = g->u_offset, h = heights_px[p];
        double S[3] = {g->fcd * cs + uo * sn, g->fcd * sn - uo * cs, -h};
        double v = (double)g->v_pre - r;
        for (int c = 0; c < W; ++c) {
            double u = (double)c - g->u_pre;
            double tq = u - uo;
            double Q[3] = {-tq * sn, tq * cs, v - h};
            double d[3] = {Q[0] - S[0], Q[1] - S[1], Q[2] - S[2]};
            double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            d[0] /= dn; d[1] /= dn; d[2] /= dn;
            double acc = 0.0;
            for (int e = 0; e < n_ell; ++e) {
                const double* E = ell + e * 10;
                const double* Rm = rot + 9 * e;
                double o[3] = {S[0] - E[0], S[1] - E[1], S[2] - E[2]};
                double q0[3], q1[3];
                for (int a = 0; a < 3; ++a) {
                    q0[a] = (Rm[a] * o[0] + Rm[3 + a] * o[1] + Rm[6 + a] * o[2]) / E[3 + a];
                    q1[a] = (Rm[a] * d[0] + Rm[3 + a] * d[1] + Rm[6 + a] * d[2]) / E[3 + a];
                }
                double A = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
                double Bq = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2];
                double C = q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] - 1.0;
                double disc = Bq * Bq - A * C;
                if (disc > 0.0) acc += E[9] * 2.0 * sqrt(disc) / A;
            }
            out[(size_t)pr * W + c] = acc;
        }
    }
    free(rot);
}
