/*
 * xr.h — C ABI of the MI355X cone-beam reconstruction library (libxr.so).
 *
 * This is the drop-in boundary for the reference's reconstruction hot path.
 * The reference has no FFI; its interface is the C++ host API in
 * reconstruct/include/reconstruct.cuh:14-73, direction/include/pca.cuh:9-12 and
 * reconstruct/include/params.h:37.  Each entry point below names the reference
 * function or kernel it replaces.  The reference-compatible C++ API
 * (include/xr/reconstruct.hpp) is a thin header over these functions.
 *
 * Conventions
 *  - Every function returns XR_OK (0) or a negative XR_ERR_* code; the message
 *    of the last failure on a context is available from xr_last_error().
 *  - Plain pointers and sizes only.  "host" buffers are ordinary CPU memory;
 *    "device" buffers are HIP device memory of the context's device.
 *  - Volumes at the host boundary use the reference layout (volume.h:65-66):
 *    x fastest, then y, then z; channel k at k*N^3.  Device-level entry points
 *    take volumes in the library's column layout (z fastest, then x, then y;
 *    convert with xr_volume_to_columns / xr_volume_from_columns).
 *  - Sinograms use the reference layout u + U*(v + V*view), one block of
 *    U*V*P floats per condition (ir.cu:1518, reconstruct.cu:816).
 *  - A context owns one HIP stream (the caller's, see xr_context_set_stream)
 *    and is not thread-safe; use one context per host thread.
 */
#ifndef XR_H
#define XR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XR_OK 0
#define XR_ERR_INVALID (-1)     /* bad argument / unsupported size */
#define XR_ERR_HIP (-2)         /* a HIP runtime call or kernel launch failed */
#define XR_ERR_ALLOC (-3)       /* device or host allocation failed */
#define XR_ERR_IO (-4)          /* settings / raw file could not be read */
#define XR_ERR_NOT_FOUND (-5)   /* settings tag or key missing */

/* Projector conventions (parity-critical, see the reference files cited). */
#define XR_CONV_IR 0   /* source R(0,+sod,0), voxel centre x-(N+1)/2 (ir.cu:1874-1884) */
#define XR_CONV_FDK 1  /* source R(0,-sod,0), voxel centre x-(N-1)/2 (fdk.cu:125-135) */
/* The gradient grid of IR::gradReconstruct: a plan of side N (the reference's
 * geom.voxel + 1) with the source at R(0,-sod,0) and voxel centre x - N/2
 * (forwardProjGrad / backwardProjGrad, ir.cu:447-462).  Its splat is
 * vox_size*ratio*mu (ir.cu:489-491), XR_OUT_RAW included, with no chord
 * normalisation: XR_OUT_NORMALISED equals XR_OUT_RAW.  Scalar volumes only. */
#define XR_CONV_GRAD 2

/* Rotate (reconstruct.cuh:19-24): CW -> view n at +2 pi n/P, CCW -> -2 pi n/P. */
#define XR_CW 0
#define XR_CCW 1

/* Method (reconstruct.cuh:14-18), same numbering. */
#define XR_METHOD_XTT 0
#define XR_METHOD_MLEM 1
#define XR_METHOD_ART 2
#define XR_METHOD_FDK 3

/* FDK filters: the reference's kernel 1/(pi^2 d (1-4m^2)) (reconstruct.cu:1694-1699)
 * or Ram-Lak in the same scaling convention, applied to cosine-weighted rows;
 * or the Hilbert filter of FDK::hilbertReconstruct (fdk.cu:298-375: R2C FFT of
 * each raw row, every half-spectrum bin times -i/(U/2+1), unnormalised C2R),
 * a circular convolution with U taps. */
#define XR_FILTER_REFERENCE 0
#define XR_FILTER_RAMLAK 1
#define XR_FILTER_HILBERT 2

/* Forward-projection output modes (xr_project). */
#define XR_OUT_RAW 0         /* sum_c c*ratio*mu, un-normalised splat (ir.cu:1518-1525) */
#define XR_OUT_NORMALISED 1  /* ... * L/(factor+1e-5) (ir.cu:1602) */
#define XR_OUT_RATIO 2       /* MLEM ratio y/p clamped at 2 (projRatio, ir.cu:1616-1635) */
#define XR_OUT_RESIDUAL 3    /* ART residual y-p (projSubtract, ir.cu:1656-1666) */

/* Mirrors Geometry (reconstruct/include/geometry.h:8-25): same field order, the
 * detector width and height split (the reference uses `detect` for both).
 * vox_size is explicit; the reference derives sod*det_size/sdd (geometry.h:13),
 * which xr_default_vox_size() returns. */
typedef struct xr_geometry {
    float sdd, sod, vox_size, det_size;
    int32_t n_vox, n_det_u, n_det_v, n_proj;
} xr_geometry;

/* Per-condition scan tables the reference keeps in __constant__ / __managed__
 * memory (params.cu:31-36): cond_R = elemR (9 per condition), cond_t = elemT
 * (3 per condition, areaTrans already added), cond_offset = INIT_OFFSET
 * (3 per condition), basis = basisVector (3 per basis vector).  grating is an
 * optional per-condition grating vector (3 per condition; NULL -> z, as in
 * ir.cu:1901).  No fixed caps: n_cond is unbounded, n_basis <= 16 (the reference
 * caps 13 conditions and 7 basis vectors, params.h:26-30). */
typedef struct xr_scan {
    int32_t n_cond, n_basis;
    const float* cond_R;
    const float* cond_t;
    const float* cond_offset;
    const float* basis;
    const float* grating;
} xr_scan;

typedef struct xr_context xr_context;
typedef struct xr_plan xr_plan;

/* ---- context and plans ------------------------------------------------- */

/* Create a context on HIP device `device` (default null stream). */
int xr_context_create(xr_context** out, int device);
void xr_context_destroy(xr_context* ctx);
/* Use `hip_stream` (a hipStream_t, NULL = null stream) for all further work. */
int xr_context_set_stream(xr_context* ctx, void* hip_stream);
int xr_context_synchronize(xr_context* ctx);
/* Copy `bytes` between any two addresses (host or device) on the context's
 * stream and wait for it (hipMemcpyDefault).  A convenience for callers that
 * hold raw device addresses, e.g. the buffers of a batch. */
int xr_memcpy(xr_context* ctx, void* dst, const void* src, size_t bytes);
const char* xr_last_error(const xr_context* ctx);

/* Kernel timing: when enabled every launch is bracketed by HIP events on the
 * context's stream.  kind: 0 forward projection, 1 backprojection, 2 FDK filter,
 * 3 other.  xr_timing_read synchronises and returns the summed milliseconds and
 * the launch count since the last reset. */
int xr_timing_enable(xr_context* ctx, int enable);
int xr_timing_reset(xr_context* ctx);
int xr_timing_read(xr_context* ctx, int kind, double* total_ms, int64_t* count);

/* A plan holds the per-(condition, view) projection tables for one geometry,
 * scan, convention and direction; the L/(factor+1e-5) normalisation table is
 * built on first use (forward projection's correlationProjByLength,
 * ir.cu:1533-1603).  Replaces init_params' constant-memory upload
 * (params.cu:80-95) and the per-thread geometry of rayCasting. */
int xr_plan_create(xr_context* ctx, const xr_geometry* g, const xr_scan* s, int convention, int direction,
                   xr_plan** out);
void xr_plan_destroy(xr_plan* plan);

float xr_default_vox_size(float sdd, float sod, float det_size);

/* Projector tile shape of a plan.  No counterpart in the reference (its
 * kernels launch one thread per voxel column with a fixed blockSize,
 * params.cu:31).  Zero fields select the defaults.  The tile shape changes
 * only the work split, never the arithmetic: every voxel still sums the views
 * of a subset in order and the forward accumulation is exact fixed point, so
 * results are bitwise identical for every setting.
 *   bwd_zpl:  z voxels per lane in the backprojection (1, 2, 4; default by N)
 *   bwd_gx:   column groups along x in a backprojection block of 16 waves
 *             (1, 2, 4, 8, 16; gy = 16/gx; default 4); XTT kinds run 8-wave
 *             blocks, where a gx of 16 is taken as 8 (gy = 8/gx; default 2)
 *   fwd_rows: cap on the detector rows of one forward launch (default: the
 *             most that fit 64 KB of LDS)
 *   fwd_scan: 0 (default): the forward of an untilted view scans, per volume
 *             row, only the columns a per-plan table gives as possibly
 *             touching the detector band; 1: every column of every row
 *   fwd_split: detector columns of one forward work item: 1, 2, 4 split each
 *             16-column band into that many items; 0 (default) splits only
 *             launches with too few items to fill the device (few views) */
typedef struct xr_tiles {
    int32_t bwd_zpl;
    int32_t bwd_gx;
    int32_t fwd_rows;
    int32_t fwd_scan;
    int32_t fwd_split;
} xr_tiles;

int xr_plan_set_tiles(xr_plan* plan, const xr_tiles* tiles);
int xr_plan_get_tiles(const xr_plan* plan, xr_tiles* tiles);

/* Autotuner: times every candidate tile shape of the plan's forward projection
 * (OUT_RATIO) and MLEM backprojection on the given device data (column-layout
 * volume of condition 0, its sinogram) and returns the fastest in `best`
 * (also set on the plan).  vol_cols and sino are not modified.  best_ms
 * (optional) receives the forward + backward milliseconds of the winner.
 * Results are cached in the context per (N, U, V, P, axis). */
int xr_autotune(xr_context* ctx, xr_plan* plan, const float* vol_cols, const float* sino, xr_tiles* best,
                double* best_ms);

/* ---- device-level operators ------------------------------------------- */

/* Reference layout <-> column layout, `channels` volumes of n_vox^3 (device). */
int xr_volume_to_columns(xr_context* ctx, const float* ref_dev, float* col_dev, int n_vox, int channels);
int xr_volume_from_columns(xr_context* ctx, const float* col_dev, float* ref_dev, int n_vox, int channels);

/* Forward projection of the listed views.  Replaces forwardProj /
 * forwardProjXTT (ir.cu:1480-1531, 133-141, 1753-1804) followed by
 * correlationProjByLength (ir.cu:1533-1603) and projRatio / projSubtract
 * (ir.cu:1616-1666).
 *   views: host array of n_views codes cond*n_proj + view (view = sinogram index).
 *   accumulate: optional host array (n_views bytes); 1 -> the splat is added to
 *     the image already in proj (the reference zeroes proj once per epoch,
 *     reconstruct.cu:83), 0 -> the image is overwritten.
 *   vol: column layout; plain: n_cond volumes of N^3 (volume cond*N^3 is used
 *     for condition cond); xtt != 0: n_basis channels of N^3.
 *   proj, sino: n_cond*U*V*P (device).  sino is read for RATIO / RESIDUAL.
 *   loss: optional host out, sum |y - p| over the projected pixels (FP64). */
int xr_project(xr_context* ctx, xr_plan* plan, const float* vol, float* proj, const float* sino,
               const int32_t* views, int n_views, const uint8_t* accumulate, int xtt, int out_mode, double* loss);

/* Data residual of a volume: the normalised forward projection p of the
 * listed views (as xr_project's XR_OUT_NORMALISED), the residual images
 * y - p written to `resid` (the layout of proj in xr_project), and the FP64
 * sums l1 = sum |y - p| (the reference's loss, reconstruct.cu:131) and
 * l2 = sum (y - p)^2 over the listed views (either may be NULL).  No
 * counterpart in the reference beyond the loss: used for the data-residual
 * criterion of the learned warm start. */
int xr_residual(xr_context* ctx, xr_plan* plan, const float* vol, float* resid, const float* sino,
                const int32_t* views, int n_views, int xtt, double* l1, double* l2);

/* xr_project without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as xr_project on XR_CONV_IR plans (other conventions return
 * XR_ERR_INVALID); the view list, accumulate flags, FP64 loss sum
 * and max |voxel| of the call live in a fixed 80.25 KiB device array per GPU
 * (the library's `scratch`), held until the process ends, instead of buffers
 * the library allocates.  Limits: at most 16384 views per call (larger lists
 * return XR_ERR_INVALID); the plan's normalisation table must already exist
 * (any earlier xr_project call on the plan builds it, else XR_ERR_INVALID);
 * no kernel timing.  Every call on one GPU shares that GPU's array, which
 * keeps the last call's contents: calls that overlap on one GPU (other
 * streams, other host threads, other contexts) race. */
int xr_project_noalloc(xr_context* ctx, xr_plan* plan, const float* vol, float* proj, const float* sino,
                       const int32_t* views, int n_views, const uint8_t* accumulate, int xtt, int out_mode,
                       double* loss);

/* Raw backprojection of the listed views into column-layout accumulators:
 * tmp += sum_c c*r, factor += 1 (backwardProj, ir.cu:1605-1614, 1720-1751) or,
 * xtt != 0, per channel tmp_k += v_km^2 sum_c c*r, factor_k += v_km^2
 * (backwardProjXTT, ir.cu:143-152, 1807-1850). */
int xr_backproject(xr_context* ctx, xr_plan* plan, const float* proj, const int32_t* views, int n_views, int xtt,
                   float* tmp, float* factor);

/* Backprojection of the listed views fused with the volume update:
 * MLEM mu = factor==0 ? 0 : mu*tmp/factor (voxelProduct, ir.cu:1637-1654) or ART
 * mu += lambda/subset_size*tmp (voxelPlus, ir.cu:1682-1693); per channel for xtt. */
int xr_backproject_update(xr_context* ctx, xr_plan* plan, const float* proj, const int32_t* views, int n_views,
                          int xtt, int method, float lambda, int subset_size, float* vol);

/* FDK pieces (fdk.cu:13-43, 51-170).  filter: XR_FILTER_*; the plan must use
 * XR_CONV_FDK for the backprojection.  sino/filt: n_cond*U*V*P (device);
 * vol: column layout N^3 of the condition of the listed views. */
int xr_fdk_filter(xr_context* ctx, xr_plan* plan, int filter, const float* sino, float* filt, const int32_t* views,
                  int n_views);
int xr_fdk_backproject(xr_context* ctx, xr_plan* plan, const float* filt, const int32_t* views, int n_views,
                       int overwrite, float* vol);
/* The filter taps the FDK path uses (n_det_u floats, host out): h[|m|] of the
 * symmetric filters, h[m mod U] of the circular Hilbert filter. */
int xr_fdk_taps(const xr_geometry* g, int filter, float* taps_out);

/* Orientation-aware XTT (forwardOrth / backwardOrth, ir.cu:295-423): with an
 * orientation field set, the plan's XTT projections (xr_project / xr_residual
 * with xtt, xr_backproject, xr_backproject_update) rotate the basis per voxel
 * by rodriguesRotation (ir.cu:597-621) about (cos phi, sin phi, 0) with
 * cos theta: coef holds phi then cos theta, each N^3 floats in the column
 * layout (device, read at every launch; the caller keeps it alive).  NULL
 * returns to the plan's fixed basis. */
int xr_plan_set_orientation(xr_plan* plan, const float* coef);

/* Orientation helpers of the orientation-aware XTT loop (orthTwiceReconstruct,
 * reconstruct.cu:956-1239), per voxel on the device, channels nvox floats
 * apart (any voxel order; the mean filter takes the column layout):
 * vol 3 channels (the three-direction coefficients), coef 2 channels (phi,
 * cos theta), est 2 channels (best cosine, iteration), md 3 channels;
 * basis9: the three basis vectors S0, S1, S2 (host, 9 floats).
 * xr_orth_normal: calcNormalVectorThreeDirec (ir.cu:856-974, sign choice rot
 * 0-3) or, with est, calcNormalVectorThreeDirecWithEst (ir.cu:976-1071);
 * updates coef in place, loss = |R e_z x normal| per voxel.
 * xr_orth_update_estimate: updateEstimationByCoef (ir.cu:1073-1118).
 * xr_orth_md: convertNormVector (ir.cu:1448-1480), md = mean(vol) R e_z.
 * xr_orth_mean_filter: meanFiltFiber (ir.cu:1231-1303) over the interior
 * voxels of an n_vox^3 volume; boundary voxels copy the source. */
int xr_orth_normal(xr_context* ctx, const float* basis9, const float* vol, int64_t nvox, float* coef, float* loss,
                   int rot, const float* est);
int xr_orth_update_estimate(xr_context* ctx, const float* basis9, const float* vol, int64_t nvox, float* est,
                            int iter);
int xr_orth_md(xr_context* ctx, const float* vol, const float* coef, int64_t nvox, float* md);
int xr_orth_mean_filter(xr_context* ctx, const float* coef_src, const float* vol, int n_vox, float weight,
                        float* coef_dst);

/* XTT::orthTwiceReconstruct (reconstruct.cu:956-1239): the orientation-aware
 * XTT loop over the three-direction model (n_basis must be 3).  sino[c]:
 * host sinograms; vol[3]: host volumes (reference layout), initial guess in,
 * the kept volume out; md[3]: host main directions out.  iter1 outer
 * iterations of five passes (iter2 epochs, then four of iter2/4), batch
 * subsets, method XR_METHOD_MLEM / XR_METHOD_ART.  seed: the subset-order
 * engine (std::mt19937; the reference seeds it from std::random_device).
 * volume_path / direction_path (either may be NULL): per outer iteration the
 * kept volume and the main directions are written as raw float32 files
 * path + "_sequence<outer>d<i>_NxNxN.raw" (reconstruct.cu:1203-1216).
 * The recorded-estimate buffer starts at zero (the reference leaves it
 * uninitialised). */
int xr_orth_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                        float* const* md, int iter1, int iter2, int batch, int direction, int method, float lambda,
                        uint32_t seed, const char* volume_path, const char* direction_path);

/* Test hook (no counterpart in the reference): after `after` further device
 * allocations, every allocation of the library fails with
 * hipErrorOutOfMemory (the entry point returns XR_ERR_HIP with the message
 * in xr_last_error) until the hook is reset with a negative value.
 * Process wide; for testing error paths from the host. */
void xr_test_fail_allocations(int after);

/* Multi-resolution helpers of the learned warm start (no counterpart in the
 * reference), single-channel column-layout volumes on the device.
 * xr_volume_downsample: coarse (edge n_vox/factor) = mean of each
 * factor^3 block of fine (edge n_vox, a multiple of factor).
 * xr_volume_upsample_add: fine (edge n_coarse*factor) = max(fine +
 * trilinear(coarse_a - coarse_b), floor_value), fine voxel i at coarse
 * coordinate (i + 0.5)/factor - 0.5 clamped to the coarse grid. */
int xr_volume_downsample(xr_context* ctx, const float* fine, int n_vox, int factor, float* coarse);
int xr_volume_upsample_add(xr_context* ctx, const float* coarse_a, const float* coarse_b, int n_coarse, int factor,
                           float* fine, float floor_value);

/* voxelSqrt (ir.cu:1707-1717): v = sqrt(|v|) for n floats (device). */
int xr_volume_sqrt(xr_context* ctx, float* vol, int64_t n);

/* calcEigenVector (direction/src/pca.cu:12-65) for nvox voxels: mu has n_basis
 * channels of chan_stride floats; md and evalue get 3 channels each (device). */
int xr_pca(xr_context* ctx, const float* mu, int n_basis, const float* basis_host, int64_t nvox, int64_t chan_stride,
           float* md, float* evalue);

/* calcEigenVector on host buffers (reference layout or any layout shared by the
 * channels): mu_host has n_basis channels of nvox floats; md_host and
 * evalue_host receive 3 channels of nvox floats each.  Computed on the GPU. */
int xr_pca_host(xr_context* ctx, const float* mu_host, int n_basis, const float* basis_host, int64_t nvox,
                float* md_host, float* evalue_host);

/* Direction post-processing, the consumers of md (direction/src/pca.cu:67-106).
 * md: 3 channels of nvox floats (x, y, z components of the main direction).
 * xr_md_angles (calcAngleFromMD): angle gets 2 channels, phi = atan2(y, x) and
 * theta = atan2(sqrt(x^2 + y^2), z), with the direction flipped when y < 0.
 * xr_md_normalise (directionNormalization): md / |md| where |md| > eps, else 0
 * (the reference uses eps = 0.65).  Device buffers, or host buffers for the
 * _host variants (computed on the GPU). */
int xr_md_angles(xr_context* ctx, const float* md, float* angle, int64_t nvox);
int xr_md_normalise(xr_context* ctx, float* md, int64_t nvox, float eps);
int xr_md_angles_host(xr_context* ctx, const float* md_host, float* angle_host, int64_t nvox);
int xr_md_normalise_host(xr_context* ctx, float* md_host, int64_t nvox, float eps);

/* ---- host-buffer solvers (the reference's host API) ------------------- */

/* Per-epoch loss sums (raw sums of |y-p|; the reference writes them divided by
 * U*V*P to loss.csv, reconstruct.cu:153-155). */
typedef struct xr_stats {
    double* losses; /* caller-provided array of `epochs` doubles, or NULL */
} xr_stats;

/* IR::reconstruct (reconstruct.cu:23-156): OS-EM / OS-ART.  sino[c], vol[c]
 * are host arrays (reference layouts) for each condition; vol is the initial
 * guess on entry and the result on exit. */
int xr_ir_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                      int epochs, int batch, int direction, int method, float lambda, xr_stats* stats);

/* XTT::reconstruct (reconstruct.cu:799-954): vol[k] (n_basis channels) in/out;
 * sino[c] is overwritten with the final ratio / residual images
 * (reconstruct.cu:924-925); md[0..2] receives the principal orientation
 * (calcEigenVector); evalue[0..2] (optional, may be NULL) the normalised
 * eigenvalues. */
int xr_xtt_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                       float* const* md, float* const* evalue, int epochs, int batch, int direction, int method,
                       float lambda, xr_stats* stats);

/* FDK::reconstruct (reconstruct.cu:1651-1724): vol[c] out (zero-initialised,
 * a deviation: the reference leaves its device buffers uninitialised); sino[c]
 * is overwritten with the filtered sinogram, as in the reference. */
int xr_fdk_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                       int direction, int filter);

/* IR::gradReconstruct (reconstruct.cu:158-328), gradient-domain phase
 * contrast.  g: the result's geometry (n_vox = N); sino: ONE U*V*P host
 * sinogram (the reference copies sinogram[0] into every slot, :198-199).  For
 * each condition c and d in {x, y, z}, slot 3c+d reconstructs an (N+1)^3
 * gradient volume (initial value 0.01) by OS-ART (or OS-EM) from the sinogram
 * scaled per view by {cos, sin, 1}[d] of 2 pi n/P (sinogramGradientCoef,
 * ir.cu:585-594) with the XR_CONV_GRAD operator; all slots share the subset
 * order of one default-seeded mt19937_64 (:221, 226).  The x and y slots of
 * condition 0 are then integrated per z slice by xr_poisson_solve into vol
 * (N^3, host, reference layout).  block_size: the reference's blockSize
 * (settings recon_variable.blockSize): its launches cover x, z <
 * min(N+1, ceil(N/blockSize)*blockSize) only (:203-204), and the layers beyond
 * keep 0.01.  grad: NULL, or 3*n_cond host arrays of (N+1)^3 floats that
 * receive the gradient volumes.  stats->losses: 3*n_cond*epochs sums of |y - p|,
 * slot-major.  CCW: the reference's forwardProjGrad writes view images at the
 * signed index (out of its buffer, ir.cu:496-499) and backwardProjGrad turns
 * by |n| (ir.cu:516); here both use the signed angle and image |n|. */
int xr_grad_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, const float* sino, float* vol,
                        float* const* grad, int epochs, int batch, int direction, int method, float lambda,
                        int block_size, xr_stats* stats);

/* poissonSolveLDLT (poisson_cpu.cpp:15-112): per z slice of the n^3 result,
 * the n*n system with identity rows on the slice border and the 5-point
 * Laplacian inside, right side gx(x,y,z) - gx(x+1,y,z) + gy(x,y,z) - gy(x,y+1,z)
 * (FP32, as the reference) on the interior.  gx, gy: (n+1)^3 device arrays,
 * vol: n^3 device array, reference layouts; n >= 3.  Solved exactly by type-I
 * sine transforms in FP64 (the reference factorises the matrix with Eigen's
 * SparseLU in FP32), result rounded to FP32. */
int xr_poisson_solve(xr_context* ctx, int n, const float* gx, const float* gy, float* vol);

/* ---- batched multi-problem solver ---------------------------------------- */

/* Many independent problems of one method (MLEM or ART) solved together:
 * IR problems (one untilted or tilted condition, no basis) and XTT problems
 * (n_cond >= 1 grating conditions and n_basis scattering directions), each
 * with its own geometry, view count and subset count.  Every subset step is
 * one persistent forward launch and one persistent backprojection launch per
 * kind (IR, XTT) over all problems of that kind, and with XR_BATCH_GRAPH the
 * whole epoch loop is captured once as a HIP graph and replayed.  No
 * counterpart in the reference (IR::reconstruct, reconstruct.cu:23-156, and
 * XTT::reconstruct, reconstruct.cu:799-954, solve one problem per call); each
 * problem's result equals its own xr_ir_reconstruct / xr_xtt_reconstruct
 * solve bit for bit (same subset order: one default-seeded mt19937_64 per IR
 * problem, reconstruct.cu:75; a fresh one each epoch for XTT, :858). */
typedef struct xr_batch xr_batch;
#define XR_BATCH_GRAPH 1    /* capture the epoch loop as a HIP graph */
#define XR_BATCH_AUTOTUNE 2 /* tile shapes from xr_autotune per distinct problem shape */

/* geoms[i], scans[i] (scans may be NULL: one untilted condition, no basis;
 * an IR problem has n_cond 1 and n_basis 0, an XTT problem n_basis > 0);
 * batch[i]: subsets of problem i (NULL: 1).  Allocates every device buffer of
 * the batch.  Autotuning (XR_BATCH_AUTOTUNE) applies to the IR problems. */
int xr_batch_create(xr_context* ctx, int n, const xr_geometry* geoms, const xr_scan* scans, const int32_t* batch,
                    int direction, int method, float lambda, int flags, xr_batch** out);
void xr_batch_destroy(xr_batch* b);
/* Device buffers of problem i, owned by the batch: the column-layout volume
 * (N^3, or n_basis channels of N^3 for XTT; initial guess in / iterate out,
 * before any square root) and the measured sinogram (U*V*P per condition). */
int xr_batch_buffers(xr_batch* b, int i, float** vol_cols, float** sino);
/* The plan of problem i (projection tables; use with xr_project to make data). */
xr_plan* xr_batch_plan(xr_batch* b, int i);
/* Tune tile shapes now (per distinct shape) on the current buffers. */
int xr_batch_autotune(xr_batch* b);
/* Run `epochs` epochs on every problem.  losses: optional host array of
 * epochs*n doubles, [epoch][problem] raw sums of |y-p|.  Each problem's
 * subset order continues across runs (runs of a and b epochs visit subsets as
 * one solve of a + b epochs does) until xr_batch_restart. */
int xr_batch_run(xr_batch* b, int epochs, double* losses);
/* Restart every problem's subset order (fresh default-seeded engine), as at
 * the start of a new IR::reconstruct call. */
int xr_batch_restart(xr_batch* b);
/* Milliseconds of the last xr_batch_run on the GPU (HIP events around the
 * epoch loop, excluding the upload of the subset lists). */
int xr_batch_elapsed_ms(const xr_batch* b, double* ms);
/* Host-buffer convenience (reference layouts), IR::reconstruct / XTT::reconstruct
 * for n problems: sino[i], vol[i] host arrays (XTT: the conditions' sinograms
 * and the channels' volumes each one after another); vol[i] in/out.  For XTT
 * problems, as XTT::reconstruct (reconstruct.cu:919-927), the result is
 * sqrt(|v|) and sino[i] is overwritten with the final ratio / residual images;
 * the principal orientation is left to xr_pca. */
int xr_batch_reconstruct(xr_context* ctx, int n, const xr_geometry* geoms, const xr_scan* scans,
                         const int32_t* batch, float* const* sino, float* const* vol, int epochs, int direction,
                         int method, float lambda, int flags, double* losses);

/* ---- learned warm start: 3-level 3-D U-Net -------------------------------- */

/* No counterpart in the reference.  A U-Net of base width C (8..32, multiple
 * of 8): 3x3x3 convolutions + ReLU at widths C, 2C, 4C over three resolution
 * levels (2x2x2 average pooling, nearest x2 upsampling, concatenated skips),
 * BF16 weights and activations on MFMA with FP32 accumulation, and a residual
 * 1x1x1 head: y = x + s*net(x/s), s = mean(x) unless a scale > 0 is given.
 * Volumes are device FP32 in the column layout (any D0 x D1 x D2 with each a
 * multiple of 4; D2 is the fastest axis).  Weights are He-initialised from a
 * Philox stream of `seed`, with a zero head, so a new network is the identity. */
typedef struct xr_unet xr_unet;
int xr_unet_create(xr_context* ctx, int base_channels, uint64_t seed, xr_unet** out);
void xr_unet_destroy(xr_unet* net);
int64_t xr_unet_num_params(const xr_unet* net);
/* Flat FP32 parameters: per layer W[Cout][27][Cin] then b[Cout] (layers 0-9),
 * then the head's w[C] and b. */
int xr_unet_get_params(xr_unet* net, float* host);
int xr_unet_set_params(xr_unet* net, const float* host);
int xr_unet_get_grads(xr_unet* net, float* host);
/* Inference: y = x + s*net(x/s). */
int xr_unet_forward(xr_unet* net, const float* x_dev, float* y_dev, int D0, int D1, int D2, float scale);
/* Training: loss = mean((x/s + net(x/s)) - t/s)^2 (FP64 sum) and its gradient
 * into the network's gradient buffer. */
int xr_unet_loss_grad(xr_unet* net, const float* x_dev, const float* t_dev, int D0, int D1, int D2, float scale,
                      double* loss);
/* One Adam step on the current gradients (bias corrected). */
int xr_unet_adam_step(xr_unet* net, float lr, float beta1, float beta2, float eps);
/* Scale s used by the last forward / loss_grad call. */
float xr_unet_last_scale(const xr_unet* net);
/* Inference fuses the decoder's nearest x2 upsampling into the following
 * convolution's input fetch (default on), so the upsampled half of each skip
 * concatenation is never written; with enable = 0 it is materialised first,
 * as training does.  The output is bitwise the same either way.  No
 * counterpart in the reference. */
int xr_unet_set_fused_upsampling(xr_unet* net, int enable);
/* Raw parameter file: 16-byte header ("XUNT", version 1, C, count) + FP32. */
int xr_unet_save(xr_unet* net, const char* path);
int xr_unet_load(xr_unet* net, const char* path);

/* ---- settings (init_params, params.cu:42-104) -------------------------- */

/* Parsed settings.json entry for one tag, same schema as the reference's file
 * (utility/settings.json).  String fields are NUL-terminated.  The per
 * condition tables are already selected by `index` when `sorting` is true and
 * have areaTrans added (params.cu:74-95). */
typedef struct xr_settings {
    char proj_path[512];
    char vol_path[512];
    char direc_path[512];
    int32_t block_size;
    int32_t n_basis;       /* recon_variable.vector */
    int32_t n_cond;        /* recon_variable.condition */
    float sod, sdd, det_size;
    int32_t n_proj, n_det_u, n_det_v, n_vox;
    int32_t load_index[64];  /* LOAD_INDEX (1-based file indices) */
    float cond_R[9 * 64];
    float cond_t[3 * 64];
    float cond_offset[3 * 64];
    float basis[3 * 16];
} xr_settings;

int xr_load_settings(const char* json_path, const char* tag, xr_settings* out);
/* Geometry/scan views of a loaded settings entry (vox_size from geometry.h:13). */
int xr_settings_geometry(const xr_settings* st, xr_geometry* g, xr_scan* s);

/* ---- synthetic inputs ---------------------------------------------------- */

/* Seeded random ellipsoids (16 doubles each: centre[3] mm, semi-axes[3] mm,
 * rotation[9] row-major body<-world, value), from a SplitMix64 stream of `seed`:
 * centres uniform in a ball of radius centre_frac*fov_radius, semi-axes
 * uniform in [axis_lo, axis_hi]*fov_radius, uniform random rotation, value
 * uniform in [value_lo, value_hi].  Host only. */
int xr_phantom_ellipsoid_params(uint64_t seed, int count, double fov_radius, double centre_frac, double axis_lo,
                                double axis_hi, double value_lo, double value_hi, double* params_out);
/* Rasterise ellipsoids at voxel centres (x-(N+1)/2)*vs into a reference-layout
 * device volume (values of overlapping ellipsoids add). */
int xr_phantom_rasterise(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                         float* vol_ref_dev);

/* Hollow z-axis cylinder (all z) added into a reference-layout device volume:
 * value where r_out - thickness < r <= r_out (mm), voxel centres (x-(N+1)/2)*vs. */
int xr_phantom_cylinder_shell(xr_context* ctx, int n_vox, float vox_size, double r_out, double thickness,
                              double value, float* vol_ref_dev);

/* Seeded fibre bundles (8 doubles each: centre[3] mm, unit direction[3],
 * radius mm, half-length mm): n_axes bundle axes uniform on the sphere
 * (axes_out, 3 doubles each); fibre e follows axis e % n_axes with a direction
 * drawn from a von Mises-Fisher distribution of concentration kappa (Wood's
 * method); centres uniform in a ball of radius centre_frac*fov_radius; radius
 * uniform in [r_lo, r_hi] mm.  SplitMix64 stream of `seed`.  Host only. */
int xr_phantom_fibre_params(uint64_t seed, int count, int n_axes, double kappa, double fov_radius,
                            double centre_frac, double r_lo, double r_hi, double half_length, double* params_out,
                            double* axes_out);
/* Rasterise fibres into n_basis reference-layout channels (device): inside a
 * fibre of direction f, channel k gains |S_k x f|^2 (values add). */
int xr_phantom_fibres(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                      const float* basis_host, int n_basis, float* vol_ref_dev);

/* Scalar fibres (attenuation phantom): `value` added inside every fibre. */
int xr_phantom_fibres_scalar(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                             float value, float* vol_ref_dev);

/* Noise on device arrays of n floats (y may alias the input):
 * Poisson: y = -ln(max(c, 1)/I0), c ~ Poisson(I0 exp(-p));
 * Gaussian visibility: y = x (1 + sigma N(0, 1)).
 * Counter-based (Philox4x32-10 keyed by seed, counter = element index): the
 * result does not depend on launch shape. */
int xr_noise_poisson(xr_context* ctx, float* y_dev, const float* p_dev, int64_t n, double I0, uint64_t seed);
int xr_noise_gaussian(xr_context* ctx, float* y_dev, const float* x_dev, int64_t n, double sigma, uint64_t seed);
/* The driver's sinogram clamp (main.cu:35): values below 0 become 1e-8, for n
 * floats on the device (e.g. after Poisson noise, where -ln(c/I0) < 0 when a
 * count exceeds I0). */
int xr_sino_clamp(xr_context* ctx, float* sino_dev, int64_t n);
/* The Philox4x32-10 block function (host), for known-answer tests. */
void xr_philox4x32(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif /* XR_H */
