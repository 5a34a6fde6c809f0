/*
 * oracle.h -- CPU restatement of Ct's FDK hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * (or as the timed CPU baseline).  The product (ct_amd/libctg.so) never links,
 * loads or calls anything in this directory.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the bastian-weber/Ct source tree).  Pinning status of each function (what
 * independent check holds it) is listed in the header comment of oracle.c.
 */
#ifndef CT_ORACLE_H
#define CT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry in detector-pixel units, as CtVolume holds it after loading
 * (Ct/CtVolume.h:226-267).  Field order is identical to ctg_geometry so the
 * same bytes can be handed to both sides in tests. */
typedef struct orc_geometry {
    int32_t det_width, det_height;   /* imageWidth, imageHeight */
    int32_t nx, ny, nz;              /* xMax, yMax, zMax (reconstructed grid) */
    float fcd;                       /* FCD / pixelSize */
    float u_offset;                  /* uOffset / pixelSize */
    float radius;                    /* getReconstructionCylinderRadius() */
    float x_pre, y_pre, z_pre;       /* x/y/zPrecomputed */
    float u_pre, v_pre;              /* u/vPrecomputed */
    float u_lo, u_hi, v_lo, v_hi;    /* imageLower/UpperBoundU/V */
} orc_geometry;

enum { ORC_DEPTH_U8 = 0, ORC_DEPTH_U16 = 1, ORC_DEPTH_F32 = 2 };
enum { ORC_RAMLAK = 0, ORC_SHEPP_LOGAN = 1, ORC_HANN = 2 };

/* Ct/CtVolume.cpp:370-373 */
uint32_t orc_cylinder_radius(uint32_t det_width, float fcd_px);

/* Ct/CtVolume.cpp:311-316 + 487-500 + 1399-1424 + 924-928.
 * bounds = {xFrom,xTo,yFrom,yTo,zFrom,zTo} as handed to setVolumeBounds
 * (NULL = never called).  roi_out (may be NULL) receives
 * {xFrom,xTo,yFrom,yTo,zFrom,zTo} in voxels; bounds_eff (may be NULL) the
 * clamped float bounds.  Returns 0, or -1 if the ROI would wrap (yTo<yFrom or
 * zTo<zFrom through the reference's clamp, Ct/CtVolume.cpp:496,498). */
int orc_volume_geometry(uint32_t det_width, uint32_t det_height, float fcd_px,
                        float u_offset_px, const float* bounds,
                        orc_geometry* g, uint32_t* roi_out, float* bounds_eff);

/* Same model with an explicit nx x ny x nz grid centred on the axis. */
void orc_grid_geometry(uint32_t det_width, uint32_t det_height, float fcd_px,
                       float u_offset_px, int32_t nx, int32_t ny, int32_t nz,
                       orc_geometry* g);

/* Intensity -> line integral.  CPU form: Ct/CtVolume.cpp:802-812 + 860-866.
 * GPU form: Ct/CtVolume.cpp:892-904.  n pixels. */
void orc_log_cpu(const void* img, int depth, size_t n, double base_intensity, float* out);
void orc_log_gpu(const void* img, int depth, size_t n, double base_intensity, float* out);
void orc_log_cpu_f64(const void* img, int depth, size_t n, double base_intensity, double* out);

/* Feldkamp weight, in place on nimg images of W x H.  Ct/CtVolume.cpp:814-830. */
void orc_feldkamp(float* img, int32_t W, int32_t H, int32_t nimg, float fcd, float u_pre, float v_pre);
void orc_feldkamp_f64(double* img, int32_t W, int32_t H, int32_t nimg, double fcd, double u_pre, double v_pre);

/* Window H(k), k = 0..W/2 (W/2+1 values).  Ct/CtVolume.cpp:835-853, 868-885. */
void orc_filter_window(int filter, int32_t W, float* h_out);

/* Ramp-family row filter, in place on `rows` rows of length W:
 * real DFT (no padding), bins 0..W/2 times H(k), unnormalised real inverse.
 * Ct/CtVolume.cpp:832-858.  Power-of-two W uses an FFT, other W a direct
 * circular convolution; both in double. */
void orc_filter_rows(float* img, int32_t W, int64_t rows, int filter);
void orc_filter_rows_f64(double* img, int32_t W, int64_t rows, int filter);
/* Direct O(W^2) evaluation of the same map via an explicit DFT (pins the FFT). */
void orc_filter_rows_direct_f64(double* img, int32_t W, int64_t rows, int filter);

/* Voxel-driven backprojection, faithful FP32 form of
 * CtVolume::reconstructionCore (Ct/CtVolume.cpp:923-1054).
 * filtered: P images, row-major W x H.  angles in degrees as stored after
 * correctAngleDirection, heights in pixels after makeHeightOffsetRelative.
 * vol: Z_FASTEST [x][y][z] (Common/Volume.h:260), nx*ny*nz, accumulated into. */
void orc_backproject_f32(const orc_geometry* g, const float* filtered,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, int32_t step, float* vol);
/* Same loop nest in double precision throughout (the FP64 yardstick). */
void orc_backproject_f64(const orc_geometry* g, const double* filtered,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, double* vol);
/* Exact adjoint of orc_backproject_f64: splats every voxel into the four
 * bilinear taps with the same masks and weights.  proj: P x H x W, accumulated. */
void orc_forward_project_f64(const orc_geometry* g, const double* vol,
                             const double* angles_deg, const double* heights_px,
                             int32_t P, double* proj);

/* Whole CPU FDK from raw images (depth as above): log (CPU or GPU form),
 * Feldkamp weight, row filter, backprojection.  Ct/CtVolume.cpp:787-800 + 923. */
void orc_fdk_f32(const orc_geometry* g, const void* raw, int depth, double base_intensity,
                 int log_gpu_form, int filter, const double* angles_deg,
                 const double* heights_px, int32_t P, float* vol);
void orc_fdk_f64(const orc_geometry* g, const void* raw, int depth, double base_intensity,
                 int filter, const double* angles_deg, const double* heights_px,
                 int32_t P, double* vol);
/* Preprocessing only (log + weight + filter), FP32 out, P images row-major. */
void orc_preprocess_f32(const orc_geometry* g, const void* raw, int depth, double base_intensity,
                        int log_gpu_form, int filter, int32_t P, float* out);

/* Analytic cone-beam line integrals of a set of ellipsoids in the reference's
 * geometry model (source at s = fcd, virtual detector through the axis).
 * ell: n_ell x 10 doubles {cx,cy,cz, a,b,c, phi,theta,psi, mu} in voxel units,
 * ZYZ Euler angles.  out: P x H x W line integrals (double). */
void orc_phantom_project(const orc_geometry* g, const double* ell, int32_t n_ell,
                         const double* angles_deg, const double* heights_px,
                         int32_t P, double* out);

/* Threads used by the OpenMP loops (0 = OpenMP default). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
