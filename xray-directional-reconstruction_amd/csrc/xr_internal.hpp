// Internal declarations shared by the HIP kernels (xr_kernels.hip) and the
// host runtime (xr_runtime.cpp).  Not part of the public C ABI (include/xr.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace xr {

// Detector band width (pixels along u) owned by one forward-projection block.
constexpr int kBandU = 16;
// LDS row stride of a band accumulator: one guard column on each side plus
// one pad column so the stride is odd (conflict-free rows on 32 banks).
constexpr int kBandStride = kBandU + 3;
// Waves per forward-projection block (they share one band accumulator).
constexpr int kFwdWaves = 12;  // two 12-wave blocks per CU (58 KB band accumulator each): 6 waves per SIMD
// The persistent multi-problem forward keeps more state per lane (about 100
// VGPRs: 5 waves per SIMD), so 8-wave blocks fit two per CU where 12-wave ones
// fit one.
constexpr int kFwdBatchWaves = 8;
// The untilted XTT forward keeps K column terms and K coefficient loads per
// lane (about 120 VGPRs: 4 waves per SIMD): 8-wave blocks, two per CU.
constexpr int kFwdXttWaves = 8;
// Forward item order: groups of this many consecutive views, every band of a
// group before the next (see view_grouped).  8 to 48 measure the same at
// 512^3 / 720 views (57-59 ms against 90 ms band-major over all views).
constexpr int kFwdViewGroup = 16;
// 8-byte slot index just past a forward band accumulator of `rows` rows of
// `acc_bytes`-byte words (the per-wave reduction slots follow it).
__host__ __device__ constexpr int fwd_red_slot(int rows, int acc_bytes) {
    return (rows * kBandStride * acc_bytes + 7) / 8;
}
// Max basis vectors (scattering directions) per voxel for the XTT operator.
constexpr int kMaxBasis = 16;

// Per-(condition, view) projection table, computed on the host in FP64 and
// rounded once to FP32.  With centred voxel coordinates x' = x - xc (same for
// y, z) the reference's ray casting (ir.cu:1852-1902, fdk.cu:111-148) reduces to
//   u = hu + nu/den,  v = hv + nv/den,
//   nu = au.(x',y',z',1), nv = av.(x',y',z',1), den = ad.(x',y',z',1).
// d0: source->voxel vector at x'=0 (d = vs*x' + d0, only its direction is used).
// fy: un-rotated voxel y at y'=0 (FDK distance weight, fdk.cu:160).
// G:  grating sensitivity vector (XTT, ir.cu:1901).  axis: 1 when au[2] and
// ad[2] are exactly zero (u constant along z columns).
struct ViewGeom {
    float au[4];
    float av[4];
    float ad[4];
    float d0[3];
    float fy;
    float G[3];
    float axis;
    float pad[4];
};
static_assert(sizeof(ViewGeom) == 96, "ViewGeom layout");

// FP64 pixel-ray geometry for the chord length of correlationProjByLength
// (ir.cu:1533-1597): source, detector axes, box faces.
struct ChordGeom {
    double det[2];  // detector pixel pitch, sdd - sod
    double src[3];
    double eu[3];
    double ev[3];
    double ey[3];
    double lo[3];
    double hi[3];
};

// OUT_RESIDUAL_L2: the residual image of OUT_RESIDUAL, with loss[0] += sum |y - p|
// and loss[1] += sum (y - p)^2 (two FP64 sums).
enum OutMode { OUT_RAW = 0, OUT_NORMALISED = 1, OUT_RATIO = 2, OUT_RESIDUAL = 3, OUT_NORM_TABLE = 4,
               OUT_RESIDUAL_L2 = 5 };
enum FwdKind { FWD_WEIGHT = 0, FWD_PLAIN = 1, FWD_XTT = 2 };
enum BwdKind { BWD_RAW = 0, BWD_MLEM = 1, BWD_ART = 2, BWD_FDK = 3, BWD_XTT_RAW = 4, BWD_XTT_MLEM = 5, BWD_XTT_ART = 6 };

struct Dims {
    int N, U, V, P, K;
    int nz;          // z extent the projectors visit (N; less where the layers above are known zero)
    float xc;        // voxel index of the centred origin ((N+1)/2 IR, (N-1)/2 FDK)
    float hu, hv;    // 0.5f*U, 0.5f*V
    float umax, vmax;  // (float)U - 0.55f, (float)V - 0.55f
    float vs;
    float ratio;     // (vs*vs)/(det*(sod/sdd)*det*(sod/sdd))
    float sod;
    float cfdk;      // 2 pi / P (FP32, fdk.cu:161)
};

struct Basis {
    float S[kMaxBasis][3];
    float S2[kMaxBasis];
};

// Row stride of the untilted XTT weight table: the largest basis count, so a
// kernel compiled for any KM reads whole rows without bounds tests (entries
// k >= K are zero).  A batch's kernels are compiled for its largest K, which
// may exceed a problem's own.
constexpr int kXwStride = kMaxBasis;

struct FwdArgs {
    Dims d;
    Basis b;
    const ViewGeom* vg;      // [n_cond*P]
    const ChordGeom* cg;     // [n_cond*P] (OUT_NORM_TABLE only)
    const int32_t* views;    // launch slot -> cond*P + view
    const uint8_t* accum;    // launch slot -> 1: add the splat to proj's current contents
    const float* vol;        // internal layout; plain: + cond*vol_cond_stride; XTT: K channels of chan_stride
    int64_t vol_cond_stride;
    int64_t chan_stride;
    float* proj;             // [n_cond][P][V][U]
    const float* sino;       // measured data, same layout (RATIO / RESIDUAL)
    const float* norm;       // L/(factor+1e-5), same layout (NORMALISED / RATIO / RESIDUAL)
    const float* g2;         // XTT [n_cond*P][K]: (S_k.G)^2
    const float* xw;         // untilted XTT [n_cond*P][5][kXwStride]: weight constants (see BwdArgs::xw)
    double* loss;            // accumulated sum |y - p| (RATIO / RESIDUAL)
    const float* vmax;       // device scalar: max |voxel value| of vol (1 for the weight pass)
    float wmul;              // bound of a voxel's splat value per unit |voxel value|
    float fsum;              // bound of a pixel's weight sum (1 for the weight pass)
    float* fmax;             // OUT_NORM_TABLE: receives max pixel weight sum (zeroed beforehand), or null
    int code_base;           // proj/sino image of a view = code - code_base (norm uses code)
    int out_mode;
    int n_bands;
    int n_slots;
    int v0, vb;              // detector rows [v0, v0+vb) of the first row chunk (chunk height vb)
    int n_chunks;            // row chunks in one launch: chunk c owns rows v0 + c*vb + [0, vb) (clipped to V)
    int n_sub;               // 1, 2, 4: a work item owns kBandU / n_sub detector columns of its band
    int axis;                // batched launches: 1 -> every view is axis aligned (ViewGeom::axis)
    const int32_t* bands;    // untilted views: [code][band][y] columns xs | xe << 16 to scan, or null (all)
    const float* coef;       // XTT orientation field (forwardOrth): phi, cos theta per voxel, chan_stride apart, or null
    float out_scale = 1.0f;  // factor on each new pixel sum before it joins proj's contents (gradient grid: vs)
    int view_group = 0;      // item order: 0 band-major; G > 0: groups of G views, all bands of a group together
};

struct BwdArgs {
    Dims d;
    Basis b;
    const ViewGeom* vg;
    const int32_t* views;    // cond*P + view, looped in order by every block
    int n_views;
    const float* proj;       // [n_cond][P][V][U] ratio / residual / filtered images
    int code_base;           // image of a view = code - code_base
    float* vol;              // internal layout (K channels of chan_stride for XTT)
    float* tmp;              // BWD_RAW / BWD_XTT_RAW accumulators (internal layout)
    float* factor;
    int64_t chan_stride;
    const float* g2;         // XTT [n_cond*P][K]: (S_k.G)^2
    // untilted XTT [n_cond*P][5][kXwStride], with sig_k = |S_k.G|: sig_k S_k,z,
    // (S_k.G)^2 |S_k|^2, sig_k vs S_k,x, sig_k vs S_k,y, sig_k (d0_x S_k,x + d0_y S_k,y);
    // entries k >= K are zero
    const float* xw;
    float alpha;             // ART step lambda/subsetSize
    int overwrite;           // FDK: 1 -> vol = sum, 0 -> vol += sum
    int zpl, gx, gy;          // tile shape: z voxels per lane (1, 2, 4), column groups along x, y (gx*gy = 16)
    int tiles_x, tiles_y, tiles_z, n_tiles, per_xcd;
    int axis;                 // batched launches: 1 -> every view is axis aligned
    unsigned* vmax_out;       // batched MLEM / ART: atomic max of |updated voxel| (float bits), or null
    const float* coef;        // XTT orientation field (backwardOrth): phi, cos theta per voxel, or null
};

// Tile shape of the plain backprojection (CPW = 2 columns per wave side).
// zpl / gx of 0 select the default for N; fills zpl, gx, gy and the tile
// counts of a.  Returns false for an unsupported shape.
bool backward_shape(BwdArgs& a, int zpl, int gx, bool xtt);
// Default plain-kind tile shape for volume edge N.
int default_zpl(int N);

// Launchers (xr_kernels.hip).  All enqueue on `stream` and return hipSuccess
// or the launch error.
hipError_t launch_forward(const FwdArgs& a, int kind, bool axis, hipStream_t stream);
hipError_t launch_backward(BwdArgs a, int kind, bool axis, hipStream_t stream);
// Persistent multi-problem launches (see k_forward_batch / k_backward_batch).
// items / item_end live in device memory; queue: two zeroed words per launch
// site; grid: resident blocks (batch_grid_*).
// kind FWD_PLAIN or FWD_XTT (kmax: the largest basis count among the items).
hipError_t launch_forward_batch(const FwdArgs* items, const int* item_end, int n_items, unsigned* queue,
                                int ticket_slot, size_t lds_bytes, int grid, int step, int loss_shift, int kind,
                                bool axis_all, hipStream_t stream, int kmax = 0);
// One launch per z depth class: every item of the launch has a.zpl == zpl.
// kind BWD_MLEM / BWD_ART (zpl 1, 2, 4) or BWD_XTT_MLEM / BWD_XTT_ART (zpl 1).
hipError_t launch_backward_batch(const BwdArgs* items, const int* item_end, int n_items, unsigned* queue, int grid,
                                 int step, int kind, int zpl, hipStream_t stream, int kmax = 0);
// Resident blocks per CU for the batched kernels (0 on error).
// axis_all: every item untilted (the untilted-only kernel, kFwdWaves waves per block)
int batch_blocks_per_cu_forward(int kind, size_t lds_bytes, bool axis_all, int kmax = 0);
// Dynamic LDS of a standalone forward launch of vb detector rows.
size_t forward_lds_bytes(int kind, int vb);
// Resident blocks per CU of the standalone forward kernel.
int forward_blocks_per_cu(int kind, bool axis, size_t lds_bytes);
// Sub-bands per detector band (1, 2, 4) for a launch of `items` band items on
// `slots` resident blocks: split while the items would leave the chip under-
// filled (a forced value `tile` > 0 wins).
inline int forward_split(int64_t items, int64_t slots, int tile) {
    if (tile > 0) return tile;
    int s = 1;
    while (s < 4 && items * s * 2 < slots * 3) s *= 2;
    return s;
}
int batch_blocks_per_cu_backward(int kind, int zpl, int kmax = 0);
size_t backward_batch_lds_bytes(int kind);
// out[i] = max |segs[i][0, seg_len[i])| (out zeroed beforehand).
// blocks per segment for k_absmax_seg over n_seg segments of at most max_len floats
int absmax_blocks(int64_t max_len, int n_seg);
hipError_t launch_absmax_seg(const float* const* segs, const int64_t* seg_len, float* out, int n_seg,
                             int blocks_per_seg, hipStream_t stream);
hipError_t launch_fdk_filter(const float* src, float* dst, const float* weight, const float* taps, int U, int V,
                             const int32_t* views, int n_views, int code_base, int circular, hipStream_t stream);
hipError_t launch_fdk_weight(float* weight, int U, int V, float sdd, float det, hipStream_t stream);
hipError_t launch_to_internal(const float* ref, float* internal, int N, int channels, hipStream_t stream);
hipError_t launch_from_internal(const float* internal, float* ref, int N, int channels, hipStream_t stream);
hipError_t launch_sqrt_abs(float* vol, int64_t n, hipStream_t stream);
hipError_t launch_vol_down(const float* fine, float* coarse, int Nc, int f, hipStream_t stream);
hipError_t launch_orth_normal(const float* vol, float* coef, float* loss, int64_t n, int rot, const float* est,
                              const Basis& b, hipStream_t stream);
hipError_t launch_orth_update_est(const float* vol, float* est, int64_t n, int iter, const Basis& b, hipStream_t stream);
hipError_t launch_orth_md(const float* vol, const float* coef, float* md, int64_t n, hipStream_t stream);
hipError_t launch_orth_mean_filter(const float* src, const float* vol, int N, float w, float* dst, hipStream_t stream);
hipError_t launch_vol_up_add(const float* a, const float* b, int Nc, int f, float* fine, float lo, hipStream_t stream);
hipError_t launch_pca(const float* mu, int K, const Basis& b, int64_t nvox, int64_t chan_stride, float* md,
                      float* evalue, hipStream_t stream);
hipError_t launch_fill(float* p, int64_t n, float value, hipStream_t stream);
// gradient-domain phase contrast (xr_grad_kernels.hip)
hipError_t launch_scale_views(const float* src, float* dst, const float* coef, int64_t len_view, int n_views,
                              hipStream_t stream);
hipError_t launch_grad_faces(float* vol, int N1, int ext, float value, hipStream_t stream);
hipError_t launch_poisson_rhs(const float* g0, const float* g1, double* B, int n, hipStream_t stream);
hipError_t launch_poisson_dst(const double* S, const double* e, double* B, double* T, float* vol, int n,
                              hipStream_t stream);
// Direction post-processing on 3 channels of n floats (pca.cu:67-106).
hipError_t launch_md_angles(const float* md, float* angle, int64_t n, hipStream_t stream);
hipError_t launch_md_normalise(float* md, int64_t n, float eps, hipStream_t stream);
// out[0] = max |x| over n floats (out must hold 0 beforehand).
hipError_t launch_absmax(const float* x, int64_t n, float* out, hipStream_t stream);
// Candidate-column table of the forward scan: for every (code, band, volume
// row) the columns [xs, xe) outside which no column can touch the band
// (untilted views; whole rows otherwise), packed xs | xe << 16.
hipError_t launch_band_rows(const ViewGeom* vg, const Dims& d, int n_codes, int n_bands, int32_t* out,
                            hipStream_t stream);
hipError_t launch_clamp_min(float* p, int64_t n, float floor_value, float replacement, hipStream_t stream);
hipError_t launch_ellipsoids(float* vol, int N, float vs, const double* params_dev, int count, hipStream_t stream);
hipError_t launch_cylinder_shell(float* vol, int N, float vs, double r_out, double thick, double value,
                                 hipStream_t stream);
// Fibre phantom: K channels of mu_k = s |S_k x f|^2 inside each fibre (values add).
hipError_t launch_fibres(float* vol, int N, float vs, const double* params_dev, int count, const Basis& b, int K,
                         hipStream_t stream);
// Fibre phantom, scalar: `value` added inside each fibre (values add).
hipError_t launch_fibres_scalar(float* vol, int N, float vs, const double* params_dev, int count, float value,
                                hipStream_t stream);
// y = -ln(max(c, 1)/I0), c ~ Poisson(I0 exp(-p)) (Philox4x32-10 stream of (seed, index)).
hipError_t launch_poisson_noise(float* y, const float* p, int64_t n, double I0, uint64_t seed, hipStream_t stream);
// y = x (1 + sigma N(0,1)) (Box-Muller on Philox4x32-10 of (seed, index)).
hipError_t launch_gaussian_noise(float* y, const float* x, int64_t n, double sigma, uint64_t seed,
                                 hipStream_t stream);

}  // namespace xr
