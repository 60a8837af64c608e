// CPU oracle for the reconstruction hot path — TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the reference's voxel-driven cone-beam
// operators, their MLEM/ART/XTT/FDK orchestration and the per-voxel PCA, so
// that the HIP product path can be checked against it.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the reference results to compare against or the timed CPU baseline.
// The product never links it.
//
// Every function is templated on the working precision T:
//   T = float  : "ref32" — the reference's FP32 arithmetic in its operation
//                order (theta in FP32, cosf/sinf, Matrix3f products in the
//                reference's loop order).  nvcc's FMA contraction and the
//                last ulp of CUDA's cosf/sinf cannot be reproduced here, so
//                ref32 is the reference up to that rounding.
//   T = double : "f64"   — the same operator evaluated in FP64 (theta is the
//                reference's FP32 angle, promoted), used as the exact target.
//
// Reference files followed (paths relative to the reference repository):
//   reconstruct/src/ir.cu:1480-1531   forwardProj / forwardonDevice       (orc_forward_view)
//   reconstruct/src/ir.cu:1533-1603   correlationProjByLength             (orc_chord_length)
//   reconstruct/src/ir.cu:1616-1666   projRatio / projSubtract            (ratio_step)
//   reconstruct/src/ir.cu:1637-1693   voxelProduct / voxelPlus            (update_step)
//   reconstruct/src/ir.cu:1707-1717   voxelSqrt                           (orc_xtt_reconstruct)
//   reconstruct/src/ir.cu:1720-1751   backwardonDevice                    (orc_backward_view)
//   reconstruct/src/ir.cu:1753-1850   forward/backwardXTTonDevice         (orc_forward_view, orc_backward_view)
//   reconstruct/src/ir.cu:295-423     forwardOrth / backwardOrth          (orc_forward_view_orth, orc_backward_view_orth)
//   reconstruct/src/ir.cu:597-621     rodriguesRotation                   (orc_rodrigues)
//   reconstruct/src/ir.cu:856-1118    calcNormalVectorThreeDirec(WithEst), updateEstimationByCoef
//                                                                         (orc_orth_normal, orc_orth_update_estimate)
//   reconstruct/src/ir.cu:1231-1303   meanFiltFiber                       (orc_orth_mean_filter)
//   reconstruct/src/ir.cu:1448-1480   convertNormVector                   (orc_orth_md)
//   reconstruct/src/reconstruct.cu:956-1239   XTT::orthTwiceReconstruct  (orc_orth_reconstruct)
//   reconstruct/src/ir.cu:1852-1902   rayCasting                          (raycast)
//   reconstruct/src/fdk.cu:13-43      calcWeight / projConv               (orc_fdk_*)
//   reconstruct/src/fdk.cu:62-170     FDK backwardonDevice                (orc_fdk_backproject_view)
//   reconstruct/src/reconstruct.cu:23-156     IR::reconstruct             (orc_ir_reconstruct)
//   reconstruct/src/reconstruct.cu:799-954    XTT::reconstruct            (orc_xtt_reconstruct)
//   reconstruct/src/reconstruct.cu:1651-1724  FDK::reconstruct            (orc_fdk_reconstruct)
//   reconstruct/src/reconstruct.cu:1726-1793, fdk.cu:298-375
//                                             FDK::hilbertReconstruct     (orc_fdk_reconstruct, kind 2)
//   reconstruct/src/ir.cu:425-594     forwardProjGrad / backwardProjGrad / sinogramGradientCoef
//                                             (orc_grad_forward_view, orc_grad_backward_view, orc_grad_coef)
//   reconstruct/src/reconstruct.cu:158-300    IR::gradReconstruct up to the Poisson solve (orc_grad_reconstruct);
//                                             the Poisson solve itself (poisson_cpu.cpp:15-112) is
//                                             oracle.poisson_solve in oracle.py
//   reconstruct/include/geometry.h:11-14      voxSize = sod*detSize/sdd
//   reconstruct/include/vec.h:93-109          normalize (eps 1e-8) / norm2
//   direction/src/pca.cu:12-65                calcEigenVector             (orc_pca)
//
// Pinning: see tests/test_oracle_*.py.  Closed forms, invariants (splat/gather
// adjointness, linearity, 90-degree rotation symmetry, MLEM fixed point) and
// textbook routines (numpy eigh / convolve) pin each function.  Parity
// unpinned: the degenerate chord branch of correlationProjByLength when a pixel
// ray meets more than two box faces (the reference overruns a 6-float buffer
// there; this oracle keeps the first two hits), and the eigenvector choice at
// (near-)degenerate eigenvalues in calcEigenVector (Eigen3, unpinned version).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
// Same field order as the reference Geometry (geometry.h:16-24), with the
// detector width and height split (the reference uses one `detect` for both).
typedef struct {
    float sdd, sod, vox_size, det_size;
    int32_t n_vox, n_det_u, n_det_v, n_proj;
} orc_geometry;

// Per-condition tables that the reference keeps in __constant__ memory
// (params.cu:31-36): elemR (9/cond), elemT (3/cond, areaTrans already added),
// INIT_OFFSET (3/cond) and basisVector (3/basis).  `grating` is an optional
// per-condition grating vector (3/cond, NULL -> z axis as in ir.cu:1901).
typedef struct {
    int32_t n_cond, n_basis;
    const float* cond_R;
    const float* cond_t;
    const float* cond_offset;
    const float* basis;
    const float* grating;
} orc_scan;
}

namespace {

constexpr float kPIf = (float)M_PI;

// theta exactly as ir.cu:1860 evaluates it: 2.0f*(float)M_PI*(float)n/(float)P in FP32.
inline float theta_f32(int n, int P) {
    volatile float a = 2.0f * kPIf;
    volatile float b = a * (float)n;
    volatile float c = b / (float)P;
    return c;
}

template <class T> inline T tcos(T a);
template <class T> inline T tsin(T a);
template <> inline float tcos<float>(float a) { return cosf(a); }
template <> inline float tsin<float>(float a) { return sinf(a); }
template <> inline double tcos<double>(double a) { return std::cos(a); }
template <> inline double tsin<double>(double a) { return std::sin(a); }

template <class T> struct V3 {
    T x, y, z;
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> scale(T s, const V3<T>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm2(const V3<T>& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Matrix3X * Vector3X (vec.h:192-200): w[i] accumulated from 0 over j.
template <class T> inline V3<T> mv(const T m[9], const V3<T>& a) {
    T w[3] = {0, 0, 0};
    const T av[3] = {a.x, a.y, a.z};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) w[i] += m[3 * i + j] * av[j];
    return {w[0], w[1], w[2]};
}
// Matrix3X * Matrix3X (vec.h:170-180).
template <class T> inline void mm(const T a[9], const T b[9], T out[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T s = 0;
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            out[3 * i + j] = s;
        }
}
// Vector3X::normalize (vec.h:93-105).
template <class T> inline V3<T> normalized(const V3<T>& a) {
    T value = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    if (value > (T)1e-8f) return {a.x / value, a.y / value, a.z / value};
    return {0, 0, 0};
}

// Per-(cond, view) frame: R = condR * Rz(theta), offset rotated, source.
template <class T> struct Frame {
    T R[9];
    V3<T> off, t, src, eu, ev, G;
};

template <class T>
Frame<T> make_frame(const orc_geometry& g, const orc_scan& s, int cond, int n, int conv) {
    Frame<T> f;
    const T th = (T)theta_f32(n, g.n_proj);
    const T c = tcos<T>(th), sn = tsin<T>(th);
    const T Rz[9] = {c, -sn, 0, sn, c, 0, 0, 0, 1};
    T cR[9];
    for (int i = 0; i < 9; i++) cR[i] = (T)s.cond_R[9 * cond + i];
    mm(cR, Rz, f.R);
    const V3<T> off0 = {(T)s.cond_offset[3 * cond + 0], (T)s.cond_offset[3 * cond + 1], (T)s.cond_offset[3 * cond + 2]};
    f.off = mv(f.R, off0);
    f.t = {(T)s.cond_t[3 * cond + 0], (T)s.cond_t[3 * cond + 1], (T)s.cond_t[3 * cond + 2]};
    // IR: source at R(0,+sod,0) (ir.cu:1874); FDK: R(0,-sod,0) (fdk.cu:125).
    const V3<T> s0 = {0, conv == 0 ? (T)g.sod : -(T)g.sod, 0};
    f.src = mv(f.R, s0);
    f.eu = mv(f.R, V3<T>{1, 0, 0});
    f.ev = mv(f.R, V3<T>{0, 0, 1});
    V3<T> g0 = {0, 0, 1};
    if (s.grating) g0 = {(T)s.grating[3 * cond + 0], (T)s.grating[3 * cond + 1], (T)s.grating[3 * cond + 2]};
    f.G = mv(f.R, g0);
    return f;
}

// Voxel centre relative to the rotation centre: ir.cu:1881-1884 (IR, (x-(N+1)/2)*vs)
// and fdk.cu:132-135 (FDK, (x-(N-1)/2)*vs), both minus the rotated offset and t.
// conv 2 is the gradient grid of forwardProjGrad (ir.cu:459-462): side N+1,
// ((2(x - 0.5) - N + 1)/2)*vs with N the gradient volume's side, and the
// source at R(0,-sod,0) (ir.cu:447) as for FDK.
template <class T> inline T centre_coord(int i, int N, T vs, int conv) {
    if (conv == 0) return ((T)2.0f * (T)i - (T)N - (T)1.0f) * (T)0.5f * vs;
    if (conv == 2) return ((T)2.0f * ((T)i - (T)0.5f) - (T)N + (T)1.0f) * (T)0.5f * vs;
    return ((T)2.0f * (T)i - (T)N + (T)1.0f) * (T)0.5f * vs;
}

template <class T>
inline V3<T> voxel_pos(const orc_geometry& g, const Frame<T>& f, int x, int y, int z, int conv) {
    const T vs = (T)g.vox_size;
    return {centre_coord<T>(x, g.n_vox, vs, conv) - f.off.x - f.t.x,
            centre_coord<T>(y, g.n_vox, vs, conv) - f.off.y - f.t.y,
            centre_coord<T>(z, g.n_vox, vs, conv) - f.off.z - f.t.z};
}

// rayCasting (ir.cu:1852-1902).
template <class T>
inline void raycast(const orc_geometry& g, const Frame<T>& f, int x, int y, int z, int conv, T& u, T& v, V3<T>& B,
                    V3<T>* X_out = nullptr) {
    const V3<T> X = voxel_pos(g, f, x, y, z, conv);
    if (X_out) *X_out = X;
    const V3<T> d = X - f.src;
    const T coeff = -dot(f.src, f.src) / dot(f.src, d);
    const V3<T> p = f.src + scale(coeff, d);
    const T mag = (T)g.sdd / (T)g.sod;
    u = dot(p, f.eu) * mag / (T)g.det_size + (T)0.5f * (T)g.n_det_u;
    v = dot(p, f.ev) * mag / (T)g.det_size + (T)0.5f * (T)g.n_det_v;
    B = normalized(d);
}

// Acceptance window and bilinear weights (ir.cu:1501-1509).
template <class T> struct Splat {
    int64_t iu, iv;
    T c1, c2, c3, c4;  // (iu,iv+1) (iu+1,iv+1) (iu+1,iv) (iu,iv)
};
template <class T> inline bool splat_weights(T u, T v, int U, int V, Splat<T>& s) {
    if (!((T)0.55f < u && u < (T)U - (T)0.55f && (T)0.55f < v && v < (T)V - (T)0.55f)) return false;
    const T ut = u - (T)0.5f, vt = v - (T)0.5f;
    s.iu = (int64_t)std::floor(ut);
    s.iv = (int64_t)std::floor(vt);
    const T fu = ut - (T)s.iu, fv = vt - (T)s.iv;
    s.c1 = ((T)1.0f - fu) * fv;
    s.c2 = fu * fv;
    s.c3 = fu * ((T)1.0f - fv);
    s.c4 = ((T)1.0f - fu) * ((T)1.0f - fv);
    return true;
}

template <class T> inline T proj_ratio(const orc_geometry& g) {
    // ir.cu:1513-1514, left to right.
    const T a = (T)g.vox_size * (T)g.vox_size;
    const T m = (T)g.sod / (T)g.sdd;
    return a / ((T)g.det_size * m * (T)g.det_size * m);
}

// XTT weight v_km = |B x S_k| * |S_k . G| (ir.cu:1786), returned squared.
template <class T> inline T xtt_w2(const V3<T>& B, const V3<T>& S, const V3<T>& G) {
    const T vkm = norm2(cross(B, S)) * std::fabs(dot(S, G));
    return vkm * vkm;
}

template <class T> inline V3<T> basis_vec(const orc_scan& s, int k) {
    return {(T)s.basis[3 * k + 0], (T)s.basis[3 * k + 1], (T)s.basis[3 * k + 2]};
}

// rodriguesRotation (ir.cu:597-621): rotation about the axis (x, y, z) with the
// given cosine and sine, (1 - cos) n n^T + [[cos, -n_z sin, n_y sin],
// [n_z sin, cos, -n_x sin], [-n_y sin, n_x sin, cos]]; identity for a zero axis.
template <class T> void rodrigues(T x, T y, T z, T c, T sn, T R[9]) {
    const T len = std::sqrt(x * x + y * y + z * z);
    if (len < (T)1e-8f) {
        const T I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        std::memcpy(R, I, sizeof(I));
        return;
    }
    const T nx = x / len, ny = y / len, nz = z / len;
    const T r1[9] = {nx * nx, nx * ny, nx * nz, nx * ny, ny * ny, ny * nz, nx * nz, ny * nz, nz * nz};
    const T r2[9] = {c, -nz * sn, ny * sn, nz * sn, c, -nx * sn, -ny * sn, nx * sn, c};
    for (int i = 0; i < 9; i++) R[i] = ((T)1.0f - c) * r1[i] + r2[i];
}

// The per-voxel rotation of forwardOrth / backwardOrth (ir.cu:316-328, 390-400):
// coefficient channel 0 holds phi, channel 1 cos(theta); the axis is
// (cos phi, sin phi, 0) and the angle's sine sqrt(1 - cos^2).
template <class T> void orth_rotation(const float* coef, int64_t idx, int64_t lenV, T R[9]) {
    const T phi = (T)coef[idx], cc = (T)coef[idx + lenV];
    rodrigues<T>(std::cos(phi), std::sin(phi), (T)0, cc, std::sqrt((T)1.0f - cc * cc), R);
}

int n_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

// Raw voxel-driven splat of one view (forwardonDevice / forwardXTTonDevice,
// ir.cu:1490-1531, 1753-1804): proj += c_k*ratio*mu (plain) or c_k*sum_k
// vkm^2*vs*ratio*mu_k (XTT); factor += c_k.  Accumulates into proj/factor (T).
template <class T, class VT>
void forward_raw(const orc_geometry& g, const orc_scan& s, int cond, int n, const VT* vol, int xtt, T* proj,
                 T* factor, const float* coef = nullptr) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const int64_t lenV = (int64_t)N * N * N;
    const Frame<T> f = make_frame<T>(g, s, cond, n, 0);
    const T ratio = proj_ratio<T>(g);
    const int K = xtt ? s.n_basis : 1;
    const int nt = n_threads();
    std::vector<T> pbuf((size_t)nt * U * V, (T)0), fbuf((size_t)nt * U * V, (T)0);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++) {
        T* pp = &pbuf[(size_t)thread_id() * U * V];
        T* ff = &fbuf[(size_t)thread_id() * U * V];
        for (int z = 0; z < N; z++)
            for (int x = 0; x < N; x++) {
                T u, v;
                V3<T> B;
                raycast(g, f, x, y, z, 0, u, v, B);
                Splat<T> w;
                if (!splat_weights(u, v, U, V, w)) continue;
                const int64_t idx = x + (int64_t)N * y + (int64_t)N * N * z;
                T val;
                if (!xtt) {
                    val = ratio * (T)vol[idx];
                } else {
                    // forwardOrth (ir.cu:295-362): the basis rotated per voxel
                    T R[9];
                    if (coef) orth_rotation<T>(coef, idx, lenV, R);
                    val = 0;
                    for (int k = 0; k < K; k++) {
                        const V3<T> S = coef ? mv(R, basis_vec<T>(s, k)) : basis_vec<T>(s, k);
                        const T w2 = xtt_w2(B, S, f.G);
                        val += w2 * (T)g.vox_size * ratio * (T)vol[idx + k * lenV];
                    }
                }
                const int64_t a1 = w.iu + (int64_t)U * (w.iv + 1), a2 = (w.iu + 1) + (int64_t)U * (w.iv + 1);
                const int64_t a3 = (w.iu + 1) + (int64_t)U * w.iv, a4 = w.iu + (int64_t)U * w.iv;
                if (!xtt) {
                    pp[a1] += w.c1 * ratio * (T)vol[idx];
                    pp[a2] += w.c2 * ratio * (T)vol[idx];
                    pp[a3] += w.c3 * ratio * (T)vol[idx];
                    pp[a4] += w.c4 * ratio * (T)vol[idx];
                } else {
                    pp[a1] += w.c1 * val;
                    pp[a2] += w.c2 * val;
                    pp[a3] += w.c3 * val;
                    pp[a4] += w.c4 * val;
                }
                ff[a1] += w.c1;
                ff[a2] += w.c2;
                ff[a3] += w.c3;
                ff[a4] += w.c4;
            }
    }
    for (int t = 0; t < nt; t++)
        for (int64_t i = 0; i < (int64_t)U * V; i++) {
            proj[i] += pbuf[(size_t)t * U * V + i];
            factor[i] += fbuf[(size_t)t * U * V + i];
        }
}

// Chord length of the pixel-centre ray through the reconstruction box
// (correlationProjByLength, ir.cu:1533-1597).
template <class T> T chord_length(const orc_geometry& g, const Frame<T>& f, int u, int v) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const T det = (T)g.det_size;
    const V3<T> pix0 = {((T)u - ((T)U - (T)1.0f) * (T)0.5f) * det, -((T)g.sdd - (T)g.sod),
                        ((T)v - ((T)V - (T)1.0f) * (T)0.5f) * det};
    const V3<T> o2s = f.src;  // Rotate * (0, sod, 0)
    const V3<T> o2p = mv(f.R, pix0);
    const V3<T> s2p = o2s - o2p;
    const T half = ((T)(N + 1) / (T)2.0f) * (T)g.vox_size;
    const T plane[6] = {-half - f.off.x - f.t.x, half - f.off.x - f.t.x, -half - f.off.y - f.t.y,
                        half - f.off.y - f.t.y,  -half - f.off.z - f.t.z, half - f.off.z - f.t.z};
    int cnt = 0;
    T coord[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 0; p < 3; p++)
        for (int sign = 0; sign < 2; sign++) {
            const T tpar = (plane[2 * p + sign] - o2s[p]) / s2p[p];
            const V3<T> pl = scale(tpar, s2p) + o2s;
            const int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
            if (plane[2 * p1] < pl[p1] && pl[p1] < plane[2 * p1 + 1] && plane[2 * p2] < pl[p2] &&
                pl[p2] < plane[2 * p2 + 1]) {
                if (cnt < 2) {
                    coord[3 * cnt + 0] = pl.x;
                    coord[3 * cnt + 1] = pl.y;
                    coord[3 * cnt + 2] = pl.z;
                }
                cnt++;
            }
        }
    return std::sqrt((coord[0] - coord[3]) * (coord[0] - coord[3]) + (coord[1] - coord[4]) * (coord[1] - coord[4]) +
                     (coord[2] - coord[5]) * (coord[2] - coord[5]));
}

// proj *= L / (factor + 1e-5)  (ir.cu:1602).
template <class T> void normalise_view(const orc_geometry& g, const Frame<T>& f, T* proj, const T* factor) {
    const int U = g.n_det_u, V = g.n_det_v;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < V; v++)
        for (int u = 0; u < U; u++) {
            const T L = chord_length(g, f, u, v);
            proj[u + (int64_t)U * v] *= L / (factor[u + (int64_t)U * v] + (T)1e-5f);
        }
}

// Raw gather of one view (backwardonDevice ir.cu:1720-1751, backwardXTTonDevice
// ir.cu:1807-1850): tmp += sum c_k p_k, factor += 1 (plain) or per channel
// tmp_k += vkm^2 sum c p, factor_k += vkm^2 (XTT).  Volume-sized accumulators.
template <class T>
void backward_raw(const orc_geometry& g, const orc_scan& s, int cond, int n, const T* proj, int xtt, T* tmp, T* factor,
                  const float* coef = nullptr) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const int64_t lenV = (int64_t)N * N * N;
    const Frame<T> f = make_frame<T>(g, s, cond, n, 0);
    const int K = xtt ? s.n_basis : 1;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                T u, v;
                V3<T> B;
                raycast(g, f, x, y, z, 0, u, v, B);
                Splat<T> w;
                if (!splat_weights(u, v, U, V, w)) continue;
                const T p1 = proj[w.iu + (int64_t)U * (w.iv + 1)], p2 = proj[(w.iu + 1) + (int64_t)U * (w.iv + 1)];
                const T p3 = proj[(w.iu + 1) + (int64_t)U * w.iv], p4 = proj[w.iu + (int64_t)U * w.iv];
                const int64_t idx = x + (int64_t)N * y + (int64_t)N * N * z;
                if (!xtt) {
                    const T numBack = w.c1 * p1 + w.c2 * p2 + w.c3 * p3 + w.c4 * p4;
                    factor[idx] += (T)1.0f;
                    tmp[idx] += numBack;
                } else {
                    // backwardOrth (ir.cu:365-423): the basis rotated per voxel
                    T R[9];
                    if (coef) orth_rotation<T>(coef, idx, lenV, R);
                    for (int k = 0; k < K; k++) {
                        const V3<T> S = coef ? mv(R, basis_vec<T>(s, k)) : basis_vec<T>(s, k);
                        const T w2 = xtt_w2(B, S, f.G);
                        const T bf = w2 * w.c1 * p1 + w2 * w.c2 * p2 + w2 * w.c3 * p3 + w2 * w.c4 * p4;
                        factor[idx + k * lenV] += w2;
                        tmp[idx + k * lenV] += bf;
                    }
                }
            }
}

// projRatio (ir.cu:1616-1635) / projSubtract (ir.cu:1656-1666) on one view.
template <class T> double ratio_step(int64_t npix, const float* sino, T* proj, int art) {
    double loss = 0;
    for (int64_t i = 0; i < npix; i++) {
        const T y = (T)sino[i];
        loss += (double)std::fabs(y - proj[i]);
        if (art) {
            proj[i] = y - proj[i];
        } else {
            if (proj[i] != (T)0) proj[i] = y / proj[i];
            if (proj[i] > (T)2.0f) proj[i] = (T)2.0f;
        }
    }
    return loss;
}

// voxelProduct (ir.cu:1637-1654) / voxelPlus (ir.cu:1682-1693).
template <class T> void update_step(int64_t n, T* vol, const T* tmp, const T* factor, int art, T alpha) {
    for (int64_t i = 0; i < n; i++) {
        if (art)
            vol[i] = vol[i] + alpha * tmp[i];
        else
            vol[i] = (factor[i] == (T)0) ? (T)0 : vol[i] * tmp[i] / factor[i];
    }
}

// Subset order: std::shuffle with a default-seeded std::mt19937_64, one engine
// per call (IR, reconstruct.cu:75) or re-created per epoch (XTT, :858).
void subset_orders(int batch, int epochs, int per_epoch_engine, int32_t* out) {
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::mt19937_64 rng;
    for (int ep = 0; ep < epochs; ep++) {
        // XTT re-creates the engine each epoch but shuffles the order left by
        // the previous epoch (reconstruct.cu:843-859), so epoch k applies the
        // same permutation k times.
        if (per_epoch_engine) rng = std::mt19937_64();
        std::shuffle(order.begin(), order.end(), rng);
        for (int i = 0; i < batch; i++) out[ep * batch + i] = order[i];
    }
}

// IR::reconstruct (reconstruct.cu:23-156): OS-EM / OS-ART, one volume per cond.
template <class T>
void ir_reconstruct(const orc_geometry& g, const orc_scan& s, const float* sino, const float* vol_in, int epochs,
                    int batch, int dir, int art, float lambda, double* vol_out, double* losses) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v, P = g.n_proj;
    const int64_t lenV = (int64_t)N * N * N, lenP = (int64_t)U * V, lenD = lenP * P;
    const int rotation = dir == 0 ? 1 : -1;
    const int subsetSize = (P + batch - 1) / batch;
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::mt19937_64 rng;
    std::vector<T> vol(lenV), proj(lenD), factorP(lenP), tmp(lenV), factorV(lenV);
    for (int cond = 0; cond < s.n_cond; cond++) {
        for (int64_t i = 0; i < lenV; i++) vol[i] = (T)vol_in[cond * lenV + i];
        const float* y = sino + cond * lenD;
        for (int ep = 0; ep < epochs; ep++) {
            std::shuffle(order.begin(), order.end(), rng);
            double loss = 0;
            std::fill(proj.begin(), proj.end(), (T)0);
            for (int sub : order) {
                for (int so = 0; so < subsetSize; so++) {
                    std::fill(factorP.begin(), factorP.end(), (T)0);
                    const int n = rotation * ((sub + batch * so) % P);
                    const int an = std::abs(n);
                    T* pv = &proj[(int64_t)an * lenP];
                    forward_raw<T, T>(g, s, cond, n, vol.data(), 0, pv, factorP.data());
                    const Frame<T> f = make_frame<T>(g, s, cond, n, 0);
                    normalise_view(g, f, pv, factorP.data());
                    loss += ratio_step(lenP, y + (int64_t)an * lenP, pv, art);
                }
                std::fill(tmp.begin(), tmp.end(), (T)0);
                std::fill(factorV.begin(), factorV.end(), (T)0);
                for (int so = 0; so < subsetSize; so++) {
                    const int n = rotation * ((sub + batch * so) % P);
                    backward_raw<T>(g, s, cond, n, &proj[(int64_t)std::abs(n) * lenP], 0, tmp.data(), factorV.data());
                }
                update_step<T>(lenV, vol.data(), tmp.data(), factorV.data(), art, (T)(lambda / (float)subsetSize));
            }
            if (losses) losses[cond * epochs + ep] = loss;
        }
        for (int64_t i = 0; i < lenV; i++) vol_out[cond * lenV + i] = (double)vol[i];
    }
}

// XTT::reconstruct (reconstruct.cu:799-954) up to the volume square root; the
// final ratio/residual images are returned like the reference's sinogram
// overwrite (reconstruct.cu:924-925).
template <class T>
void xtt_reconstruct(const orc_geometry& g, const orc_scan& s, const float* sino, const float* vol_in, int epochs,
                     int batch, int dir, int art, float lambda, double* vol_out, double* proj_out, double* losses) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v, P = g.n_proj, K = s.n_basis, C = s.n_cond;
    const int64_t lenV = (int64_t)N * N * N, lenP = (int64_t)U * V, lenD = lenP * P;
    const int rotation = dir == 0 ? 1 : -1;
    const int subsetSize = (P + batch - 1) / batch;
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::vector<T> vol(lenV * K), proj(lenD * C), factorP(lenP), tmp(lenV * K), factorV(lenV * K);
    for (int64_t i = 0; i < lenV * K; i++) vol[i] = (T)vol_in[i];
    for (int ep = 0; ep < epochs; ep++) {
        std::mt19937_64 rng;
        std::shuffle(order.begin(), order.end(), rng);
        double loss = 0;
        std::fill(proj.begin(), proj.end(), (T)0);
        for (int sub : order) {
            for (int cond = 0; cond < C; cond++)
                for (int so = 0; so < subsetSize; so++) {
                    std::fill(factorP.begin(), factorP.end(), (T)0);
                    const int n = rotation * ((sub + batch * so) % P);
                    T* pv = &proj[cond * lenD + (int64_t)std::abs(n) * lenP];
                    forward_raw<T, T>(g, s, cond, n, vol.data(), 1, pv, factorP.data());
                    normalise_view(g, make_frame<T>(g, s, cond, n, 0), pv, factorP.data());
                    loss += ratio_step(lenP, sino + cond * lenD + (int64_t)std::abs(n) * lenP, pv, art);
                }
            std::fill(tmp.begin(), tmp.end(), (T)0);
            std::fill(factorV.begin(), factorV.end(), (T)0);
            for (int cond = 0; cond < C; cond++)
                for (int so = 0; so < subsetSize; so++) {
                    const int n = rotation * ((sub + batch * so) % P);
                    backward_raw<T>(g, s, cond, n, &proj[cond * lenD + (int64_t)std::abs(n) * lenP], 1, tmp.data(),
                                    factorV.data());
                }
            update_step<T>(lenV * K, vol.data(), tmp.data(), factorV.data(), art, (T)(lambda / (float)subsetSize));
        }
        if (losses) losses[ep] = loss;
    }
    for (int64_t i = 0; i < lenV * K; i++) {
        // voxelSqrt (ir.cu:1715)
        const T m = vol[i];
        vol_out[i] = (double)((m < (T)0) ? std::sqrt(-m) : std::sqrt(m));
    }
    if (proj_out)
        for (int64_t i = 0; i < lenD * C; i++) proj_out[i] = (double)proj[i];
}

// calcWeight (fdk.cu:13-26).
template <class T> void fdk_weight(const orc_geometry& g, T* w) {
    const int U = g.n_det_u, V = g.n_det_v;
    const T sdd = (T)g.sdd, det = (T)g.det_size;
    for (int v = 0; v < V; v++)
        for (int u = 0; u < U; u++) {
            const T ur = ((T)u + (T)0.5f - (T)U / (T)2.0f) * det;
            const T vr = ((T)v + (T)0.5f - (T)V / (T)2.0f) * det;
            w[u + (int64_t)U * v] = sdd / std::sqrt(sdd * sdd + ur * ur + vr * vr);
        }
}

// Filter taps.  kind 0: the reference's kernel (reconstruct.cu:1694-1699),
// 1/(pi^2 d (1-4m^2)) with d = det*sod/sdd, evaluated as the reference does
// (double product, rounded to float, FP32 reciprocal).  kind 1: Ram-Lak in the
// same scaling convention (sample spacing d and the 1/2 of a full rotation
// folded in): h[0] = 1/(8d), h[odd m] = -1/(2 pi^2 m^2 d), h[even m] = 0.
template <class T> void fdk_taps(const orc_geometry& g, int kind, T* h) {
    const float d = g.det_size * (g.sod / g.sdd);
    for (int m = 0; m < g.n_det_u; m++) {
        if (kind == 0) {
            if (sizeof(T) == sizeof(float))
                h[m] = (T)(1.0f / (float)(M_PI * M_PI * d * (1.0f - 4.0f * (float)(m * m))));
            else
                h[m] = (T)(1.0 / (M_PI * M_PI * (double)d * (1.0 - 4.0 * (double)m * (double)m)));
        } else {
            if (m == 0)
                h[m] = (T)(1.0 / (8.0 * (double)d));
            else if (m % 2 == 1)
                h[m] = (T)(-1.0 / (2.0 * M_PI * M_PI * (double)m * (double)m * (double)d));
            else
                h[m] = (T)0;
        }
    }
}

// cuFFTtoProjection + hilbertFiltering (fdk.cu:298-375) for one view: every
// row x of U samples goes through a forward real-to-complex DFT
// X[k] = sum_j x[j] e^{-2 pi i j k / U} (k = 0 .. U/2), each bin times
// -i / (U/2 + 1), and the unnormalised complex-to-real inverse
// y[n] = Re Y[0] + 2 sum_{0<k<U/2} Re(Y[k] e^{2 pi i k n / U}) + Re Y[U/2] (-1)^n
// (the last term for even U only), which is what cuFFT's C2R computes from a
// half spectrum.  Written as explicit DFT sums (no FFT library); the rows are
// not cosine weighted (hilbertReconstruct leaves calcWeight commented out,
// reconstruct.cu:1763).
template <class T> void hilbert_filter_view(const orc_geometry& g, const float* src, T* dst) {
    const int U = g.n_det_u, V = g.n_det_v, K = U / 2;
    const T c = (T)(-1.0) / (T)(U / 2 + 1);
    std::vector<double> cs((size_t)U), sn((size_t)U);
    for (int m = 0; m < U; m++) {
        cs[m] = std::cos(2.0 * M_PI * (double)m / (double)U);
        sn[m] = std::sin(2.0 * M_PI * (double)m / (double)U);
    }
#pragma omp parallel for schedule(static)
    for (int v = 0; v < V; v++) {
        const float* x = src + (int64_t)U * v;
        std::vector<T> re((size_t)K + 1), im((size_t)K + 1);
        for (int k = 0; k <= K; k++) {
            T a = 0, b = 0;
            for (int j = 0; j < U; j++) {
                const int m = (int)(((int64_t)j * k) % U);
                a += (T)x[j] * (T)cs[m];
                b -= (T)x[j] * (T)sn[m];
            }
            // (a + i b) * (i c) = -c b + i c a
            re[k] = -c * b;
            im[k] = c * a;
        }
        for (int n = 0; n < U; n++) {
            T y = re[0];
            const int top = (U % 2 == 0) ? K - 1 : K;
            for (int k = 1; k <= top; k++) {
                const int m = (int)(((int64_t)k * n) % U);
                y += (T)2 * (re[k] * (T)cs[m] - im[k] * (T)sn[m]);
            }
            if (U % 2 == 0) y += (n % 2 == 0) ? re[K] : -re[K];
            dst[n + (int64_t)U * v] = y;
        }
    }
}

// projConv (fdk.cu:28-43) for one view: dst[u,v] = sum_j h[|u-j|] src[j,v] w[j,v].
template <class T> void fdk_filter_view(const orc_geometry& g, const T* h, const T* w, const float* src, T* dst) {
    const int U = g.n_det_u, V = g.n_det_v;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < V; v++)
        for (int u = 0; u < U; u++) {
            T acc = 0;
            for (int j = 0; j < U; j++) {
                const T hh = (u > j) ? h[u - j] : h[j - u];
                acc += hh * (T)src[j + (int64_t)U * v] * w[j + (int64_t)U * v];
            }
            dst[u + (int64_t)U * v] = acc;
        }
}

// FDK backwardonDevice (fdk.cu:62-170) for one view: vol += Ud^2 * (2 pi / P) * sum c p,
// Ud = sod / (sod + y), y the un-rotated voxel y coordinate (fdk.cu:160).
template <class T, class PT>
void fdk_backproject_view(const orc_geometry& g, const orc_scan& s, int cond, int n, const PT* proj, T* vol) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const Frame<T> f = make_frame<T>(g, s, cond, n, 1);
    const T Cw = (T)2.0f * (T)kPIf / (T)g.n_proj;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                T u, v;
                V3<T> B, X;
                raycast(g, f, x, y, z, 1, u, v, B, &X);
                Splat<T> w;
                if (!splat_weights(u, v, U, V, w)) continue;
                const T Ud = (T)g.sod / ((T)g.sod + X.y);
                const T numBack = w.c1 * (T)proj[w.iu + (int64_t)U * (w.iv + 1)] +
                                  w.c2 * (T)proj[(w.iu + 1) + (int64_t)U * (w.iv + 1)] +
                                  w.c3 * (T)proj[(w.iu + 1) + (int64_t)U * w.iv] +
                                  w.c4 * (T)proj[w.iu + (int64_t)U * w.iv];
                vol[x + (int64_t)N * y + (int64_t)N * N * z] += Ud * Ud * Cw * numBack;
            }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (textbook
// method; the reference calls Eigen's SelfAdjointEigenSolver, pca.cu:33).
// Returns eigenvalues ascending and the matching unit eigenvectors as columns.
void jacobi3(const double A_in[9], double evals[3], double evecs[9]) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = A_in[3 * i + j];
    for (int sweep = 0; sweep < 50; sweep++) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= 1e-300 || off <= 1e-32 * diag) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; k++) {  // A <- A J
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {  // A <- J^T A
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < 3; k++) {  // V <- V J
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - sn * vkq;
                    v[k][q] = sn * vkp + c * vkq;
                }
            }
    }
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int i, int j) { return a[i][i] < a[j][j]; });
    for (int c = 0; c < 3; c++) {
        evals[c] = a[idx[c]][idx[c]];
        for (int r = 0; r < 3; r++) evecs[3 * r + c] = v[r][idx[c]];
    }
}

}  // namespace

// Normal of a voxel's three-direction scattering (ir.cu:912-922, 1042-1052):
// (mu0 S0 - m1 S1) x (mu0 S0 - m2 S2), normalised.
template <class T> V3<T> three_direc_normal(const orc_scan& s, T mu0, T m1, T m2) {
    const V3<T> S0 = basis_vec<T>(s, 0), S1 = basis_vec<T>(s, 1), S2 = basis_vec<T>(s, 2);
    const V3<T> v1 = {mu0 * S0.x - m1 * S1.x, mu0 * S0.y - m1 * S1.y, mu0 * S0.z - m1 * S1.z};
    const V3<T> v2 = {mu0 * S0.x - m2 * S2.x, mu0 * S0.y - m2 * S2.y, mu0 * S0.z - m2 * S2.z};
    return normalized(cross(v1, v2));
}

// calcNormalVectorThreeDirec (ir.cu:856-974; rot 0-3 picks the signs of mu1,
// mu2) and calcNormalVectorThreeDirecWithEst (ir.cu:976-1071; the signs from
// est channel 1: < 0.5, < 1.5, < 2.5, else).  Per voxel: the normal rotated by
// the voxel's current rotation, flipped into z >= 0, normalised; loss = |R e_z
// x normal|; coef <- (atan2 of e_z x normal, e_z . normal).  vol: 3 channels,
// coef: 2 channels, est: 2 channels (or null), lenV voxels each.
template <class T>
void orth_normal(const orc_scan& s, int64_t lenV, const float* vol, float* coef, float* loss, int rot, const float* est) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < lenV; i++) {
        T R[9];
        orth_rotation<T>(coef, i, lenV, R);
        const T mu0 = (T)vol[i], mu1 = (T)vol[i + lenV], mu2 = (T)vol[i + 2 * lenV];
        int r = rot;
        if (est) {
            const T e = (T)est[i + lenV];
            r = e < (T)0.5f ? 0 : (e < (T)1.5f ? 1 : (e < (T)2.5f ? 2 : 3));
        }
        T m1 = mu1, m2 = mu2;
        if (r == 1) {
            m1 = -mu1;
        } else if (r == 2) {
            m1 = -mu1;
            m2 = -mu2;
        } else if (r == 3) {
            m2 = -mu2;
        }
        V3<T> nrm = mv(R, three_direc_normal<T>(s, mu0, m1, m2));
        if (nrm.z < (T)0) nrm = {-nrm.x, -nrm.y, -nrm.z};
        nrm = normalized(nrm);
        const V3<T> ez = {0, 0, 1};
        const V3<T> Rez = mv(R, ez);
        const V3<T> diff = cross(Rez, nrm);
        const V3<T> axis = cross(ez, nrm);
        loss[i] = (float)norm2(diff);
        coef[i] = (float)std::atan2(axis.y, axis.x);
        coef[i + lenV] = (float)dot(ez, nrm);
    }
}

// updateEstimationByCoef (ir.cu:1073-1118): the unrotated normal, flipped into
// z >= 0; where its cosine to e_z exceeds est channel 0, est <- (cos, iter).
template <class T> void orth_update_estimate(const orc_scan& s, int64_t lenV, const float* vol, float* est, int iter) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < lenV; i++) {
        V3<T> nrm = three_direc_normal<T>(s, (T)vol[i], (T)vol[i + lenV], (T)vol[i + 2 * lenV]);
        if (nrm.z < (T)0) nrm = {-nrm.x, -nrm.y, -nrm.z};
        nrm = normalized(nrm);
        const T c = nrm.z;
        if (c > (T)est[i]) {
            est[i] = (float)c;
            est[i + lenV] = (float)iter;
        }
    }
}

// convertNormVector (ir.cu:1448-1480): md_i = mean(mu) (R e_z)_i.
template <class T> void orth_md(int64_t lenV, const float* vol, const float* coef, float* md) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < lenV; i++) {
        const T mu = ((T)vol[i] + (T)vol[i + lenV] + (T)vol[i + 2 * lenV]) / (T)3.0f;
        T R[9];
        orth_rotation<T>(coef, i, lenV, R);
        const V3<T> n = mv(R, V3<T>{0, 0, 1});
        for (int c = 0; c < 3; c++) md[i + c * lenV] = (float)(mu * n[c]);
    }
}

// meanFiltFiber (ir.cu:1231-1303) on the interior voxels (reference layout, x
// fastest): the 27 neighbours' (mu-weighted) orientation vectors, those within
// |cos| > 0.5 of the centre's added with weight `w` (sign-aligned), the sum
// normalised and flipped into z >= 0, stored as (atan2, cos).  The boundary
// voxels of dst are left as they are.
template <class T> void orth_mean_filter(int N, const float* src, const float* vol, float w, float* dst) {
    const int64_t lenV = (int64_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 1; z < N - 1; z++)
        for (int y = 1; y < N - 1; y++)
            for (int x = 1; x < N - 1; x++) {
                V3<T> nb[27];
                int cnt = 0;
                for (int i = -1; i < 2; i++)
                    for (int j = -1; j < 2; j++)
                        for (int k = -1; k < 2; k++) {
                            const int64_t q = (x - i) + (int64_t)N * (y - j) + (int64_t)N * N * (z - k);
                            const T ct = (T)src[q + lenV];
                            const T st = std::sqrt((T)1.0f - ct * ct);
                            const T phi = -(T)kPIf / (T)2.0f + (T)src[q];
                            const T mu = (T)vol[q] + (T)vol[q + lenV] + (T)vol[q + 2 * lenV];
                            nb[cnt++] = {mu * (st * std::cos(phi)), mu * (st * std::sin(phi)), mu * ct};
                        }
                V3<T> c = nb[13];
                const T cut = (T)0.5f;
                for (int i = 0; i < 13; i++) {
                    for (int side = 0; side < 2; side++) {
                        const V3<T>& o = side ? nb[26 - i] : nb[i];
                        const T d = dot(o, nb[13]);
                        if (d > cut)
                            c = {c.x + (T)w * o.x, c.y + (T)w * o.y, c.z + (T)w * o.z};
                        else if (d < -cut)
                            c = {c.x - (T)w * o.x, c.y - (T)w * o.y, c.z - (T)w * o.z};
                    }
                }
                c = normalized(c);
                if (c.z < (T)0) c = {-c.x, -c.y, -c.z};
                const V3<T> ez = {0, 0, 1};
                const V3<T> axis = cross(ez, c);
                const int64_t q = x + (int64_t)N * y + (int64_t)N * N * z;
                dst[q] = (float)std::atan2(axis.y, axis.x);
                dst[q + lenV] = (float)dot(ez, c);
            }
}

// One epoch of the orientation-aware operator over every subset (the subset
// loop of orthTwiceReconstruct, reconstruct.cu:1076-1125): forwardOrth +
// normalisation + ratio per (condition, view), backwardOrth, the update.
template <class T>
void orth_epoch(const orc_geometry& g, const orc_scan& s, const float* sino, const float* coef, int dir, int art,
                float lambda, const std::vector<int>& order, std::vector<T>& vol) {
    const int U = g.n_det_u, V = g.n_det_v, P = g.n_proj, C = s.n_cond, batch = (int)order.size();
    const int64_t lenV = (int64_t)g.n_vox * g.n_vox * g.n_vox, lenP = (int64_t)U * V, lenD = lenP * P;
    const int rotation = dir == 0 ? 1 : -1;
    const int subsetSize = (P + batch - 1) / batch;
    std::vector<T> proj(lenD * C, (T)0), factorP(lenP), tmp(lenV * 3), factorV(lenV * 3);
    for (int sub : order) {
        for (int cond = 0; cond < C; cond++)
            for (int so = 0; so < subsetSize; so++) {
                std::fill(factorP.begin(), factorP.end(), (T)0);
                const int n = rotation * ((sub + batch * so) % P);
                T* pv = &proj[cond * lenD + (int64_t)std::abs(n) * lenP];
                forward_raw<T, T>(g, s, cond, n, vol.data(), 1, pv, factorP.data(), coef);
                normalise_view(g, make_frame<T>(g, s, cond, n, 0), pv, factorP.data());
                ratio_step(lenP, sino + cond * lenD + (int64_t)std::abs(n) * lenP, pv, art);
            }
        std::fill(tmp.begin(), tmp.end(), (T)0);
        std::fill(factorV.begin(), factorV.end(), (T)0);
        for (int cond = 0; cond < C; cond++)
            for (int so = 0; so < subsetSize; so++) {
                const int n = rotation * ((sub + batch * so) % P);
                backward_raw<T>(g, s, cond, n, &proj[cond * lenD + (int64_t)std::abs(n) * lenP], 1, tmp.data(),
                                factorV.data(), coef);
            }
        update_step<T>(lenV * 3, vol.data(), tmp.data(), factorV.data(), art, (T)(lambda / (float)subsetSize));
    }
}

// XTT::orthTwiceReconstruct (reconstruct.cu:956-1239) with a seeded engine and
// a zeroed estimate (see the library's xr_orth_reconstruct for the pass order).
template <class T>
void orth_reconstruct(const orc_geometry& g, const orc_scan& s, const float* sino, const float* vol_in, int iter1,
                      int iter2, int batch, int dir, int art, float lambda, uint32_t seed, double* vol_out,
                      double* md_out) {
    const int64_t lenV = (int64_t)g.n_vox * g.n_vox * g.n_vox;
    std::vector<T> vol(vol_in, vol_in + lenV * 3), kept(vol);
    std::vector<float> store(lenV * 2), coef(lenV * 2), est(lenV * 2, 0.f), loss(lenV), md(lenV * 3), tmpf(lenV * 3);
    for (int64_t i = 0; i < lenV; i++) {
        store[i] = 0.f;
        store[i + lenV] = 1.f;
    }
    coef = store;
    std::mt19937 engine(seed);
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    auto as_float = [&](const std::vector<T>& v) {
        for (int64_t i = 0; i < lenV * 3; i++) tmpf[i] = (float)v[i];
        return tmpf.data();
    };
    for (int outer = 0; outer < iter1; outer++) {
        for (int ep1 = 0; ep1 < 5; ep1++) {
            if (ep1 != 0) {
                vol = kept;
                orth_normal<T>(s, lenV, as_float(vol), coef.data(), loss.data(), ep1 - 1, nullptr);
            }
            for (int64_t i = 0; i < lenV; i++) vol[i] = (T)0.1f;  // fillVolume: channel 0
            const int epochs = (ep1 == 0) ? iter2 : iter2 / 4;
            for (int ep2 = 0; ep2 < epochs; ep2++) {
                std::shuffle(order.begin(), order.end(), engine);
                orth_epoch<T>(g, s, sino, coef.data(), dir, art, lambda, order, vol);
            }
            for (int64_t i = 0; i < lenV * 3; i++) {
                const T m = vol[i];
                vol[i] = (m < (T)0) ? std::sqrt(-m) : std::sqrt(m);  // voxelSqrt
            }
            if (ep1 != 0)
                orth_update_estimate<T>(s, lenV, as_float(vol), est.data(), ep1 - 1);
            else
                kept = vol;
            orth_md<T>(lenV, as_float(kept), store.data(), md.data());
            coef = store;
        }
        vol = kept;
        orth_normal<T>(s, lenV, as_float(vol), coef.data(), loss.data(), 0, est.data());
        store = coef;
        orth_md<T>(lenV, as_float(kept), store.data(), md.data());
    }
    for (int64_t i = 0; i < lenV * 3; i++) {
        vol_out[i] = (double)kept[i];
        md_out[i] = (double)md[i];
    }
}

// ---- gradient-domain phase-contrast reconstruction (IR::gradReconstruct) ----
// g here is the gradient volume's geometry: n_vox = N+1 for an N^3 result.
// ext: the x and z extent the reference's launches cover, min(N+1, ceil(N/bs)*bs)
// for blockSize bs (gridV is sized from the N^3 result, reconstruct.cu:203-204,
// the kernels test against N+1, ir.cu:429).

// forwardProjGrad (ir.cu:425-503): bilinear splat of vs*ratio*voxel, no chord
// normalisation; accumulates into proj (U*V).
template <class T>
void grad_forward_raw(const orc_geometry& g, const orc_scan& s, int cond, int n, const T* vol, int ext, T* proj) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const Frame<T> f = make_frame<T>(g, s, cond, n, 2);
    const T ratio = proj_ratio<T>(g);
    for (int y = 0; y < N; y++)
        for (int z = 0; z < ext; z++)
            for (int x = 0; x < ext; x++) {
                T u, v;
                V3<T> B;
                raycast(g, f, x, y, z, 2, u, v, B);
                Splat<T> w;
                if (!splat_weights(u, v, U, V, w)) continue;
                const int64_t idx = x + (int64_t)N * y + (int64_t)N * N * z;
                const T pr = (T)g.vox_size * ratio * vol[idx];
                proj[w.iu + (int64_t)U * (w.iv + 1)] += w.c1 * pr;
                proj[(w.iu + 1) + (int64_t)U * (w.iv + 1)] += w.c2 * pr;
                proj[(w.iu + 1) + (int64_t)U * w.iv] += w.c3 * pr;
                proj[w.iu + (int64_t)U * w.iv] += w.c4 * pr;
            }
}

// backwardProjGrad (ir.cu:505-583): bilinear gather, +1 per view that lands.
// The angle is that of the signed view n, as in forwardProjGrad.  (For CCW the
// reference's pair is inconsistent: forwardProjGrad writes image n < 0 out of
// its slot, ir.cu:496-499, and backwardProjGrad turns with |n|, ir.cu:516; the
// pair here follows IR::reconstruct: signed angle, image |n|.)
template <class T>
void grad_backward_raw(const orc_geometry& g, const orc_scan& s, int cond, int n, const T* proj, int ext, T* tmp,
                       T* factor) {
    const int N = g.n_vox, U = g.n_det_u, V = g.n_det_v;
    const Frame<T> f = make_frame<T>(g, s, cond, n, 2);
    for (int y = 0; y < N; y++)
        for (int z = 0; z < ext; z++)
            for (int x = 0; x < ext; x++) {
                T u, v;
                V3<T> B;
                raycast(g, f, x, y, z, 2, u, v, B);
                Splat<T> w;
                if (!splat_weights(u, v, U, V, w)) continue;
                const T p1 = proj[w.iu + (int64_t)U * (w.iv + 1)], p2 = proj[(w.iu + 1) + (int64_t)U * (w.iv + 1)];
                const T p3 = proj[(w.iu + 1) + (int64_t)U * w.iv], p4 = proj[w.iu + (int64_t)U * w.iv];
                const int64_t idx = x + (int64_t)N * y + (int64_t)N * N * z;
                factor[idx] += (T)1.0f;
                tmp[idx] += w.c1 * p1 + w.c2 * p2 + w.c3 * p3 + w.c4 * p4;
            }
}

// sinogramGradientCoef (ir.cu:585-594): {cos theta, sin theta, 1}[diff].
template <class T> T grad_coef(int n, int P, int diff) {
    const T th = (T)theta_f32(n, P);
    return diff == 0 ? tcos<T>(th) : diff == 1 ? tsin<T>(th) : (T)1.0f;
}

// IR::gradReconstruct (reconstruct.cu:158-300) up to the Poisson solve.  g:
// the result's geometry (n_vox = N); sino0: one U*V*P sinogram (the reference
// copies sinogram[0] into every slot, :173-174, 198-199).  grads_out: 3*n_cond
// (N+1)^3 doubles, slot 3*cond + diff; losses: epochs sums of |y - p| over all
// slots (the reference's loss1, :232-233, which it never reads back).
template <class T>
void grad_reconstruct(const orc_geometry& g0, const orc_scan& s, const float* sino0, int epochs, int batch, int dir,
                      int art, float lambda, int block_size, double* grads_out, double* losses) {
    orc_geometry g = g0;
    g.n_vox = g0.n_vox + 1;  // geom.voxel + 1 (:161)
    const int N1 = g.n_vox, U = g.n_det_u, V = g.n_det_v, P = g.n_proj, C = s.n_cond;
    const int ext = std::min(N1, (g0.n_vox + block_size - 1) / block_size * block_size);
    const int64_t lenV = (int64_t)N1 * N1 * N1, lenP = (int64_t)U * V, lenD = lenP * P;
    const int rotation = dir == 0 ? 1 : -1;
    const int ss = (P + batch - 1) / batch;
    std::vector<T> vol((size_t)3 * C * lenV, (T)0.01f), sino((size_t)3 * C * lenD), proj((size_t)3 * C * lenD);
    for (int i = 0; i < 3 * C; i++)
        for (int n = 0; n < P; n++) {
            const T c = grad_coef<T>(n, P, i % 3);
            for (int64_t k = 0; k < lenP; k++) sino[i * lenD + n * lenP + k] = c * (T)sino0[n * lenP + k];
        }
    // the reference multiplies in FP32; the slots feed ratio_step as floats
    std::vector<float> sino_f(sino.size());
    for (size_t i = 0; i < sino.size(); i++) sino_f[i] = (float)sino[i];
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::mt19937_64 rng;
    std::vector<T> tmp(lenV), factor(lenV);
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), rng);
        double loss = 0;
        std::fill(proj.begin(), proj.end(), (T)0);
        for (int sub : order) {
            for (int i = 0; i < 3 * C; i++)
                for (int so = 0; so < ss; so++) {
                    const int n = rotation * ((sub + batch * so) % P);
                    T* pv = &proj[i * lenD + (int64_t)std::abs(n) * lenP];
                    grad_forward_raw<T>(g, s, i / 3, n, &vol[i * lenV], ext, pv);
                    loss += ratio_step(lenP, &sino_f[i * lenD + (int64_t)std::abs(n) * lenP], pv, art);
                }
            for (int i = 0; i < 3 * C; i++) {
                std::fill(tmp.begin(), tmp.end(), (T)0);
                std::fill(factor.begin(), factor.end(), (T)0);
                for (int so = 0; so < ss; so++) {
                    const int n = rotation * ((sub + batch * so) % P);
                    grad_backward_raw<T>(g, s, i / 3, n, &proj[i * lenD + (int64_t)std::abs(n) * lenP], ext,
                                         tmp.data(), factor.data());
                }
                // voxelPlus / voxelProduct over the launched x, z range only
                const T alpha = (T)(lambda / (float)ss);
                T* v = &vol[i * lenV];
                for (int z = 0; z < ext; z++)
                    for (int y = 0; y < N1; y++)
                        for (int x = 0; x < ext; x++) {
                            const int64_t idx = x + (int64_t)N1 * y + (int64_t)N1 * N1 * z;
                            if (art)
                                v[idx] = v[idx] + alpha * tmp[idx];
                            else
                                v[idx] = (factor[idx] == (T)0) ? (T)0 : v[idx] * tmp[idx] / factor[idx];
                        }
            }
        }
        if (losses) losses[ep] = loss;
    }
    for (size_t i = 0; i < vol.size(); i++) grads_out[i] = (double)vol[i];
}

extern "C" {

// Reference-preset voxel size (geometry.h:13): sod * detSize / sdd in FP32.
float orc_default_vox_size(float sdd, float sod, float det) { return sod * det / sdd; }

float orc_theta_f32(int n, int P) { return theta_f32(n, P); }

// a2: rayCasting for one voxel.  conv 0 = IR/XTT, 1 = FDK.  out: u, v, B[3], G[3].
void orc_raycast(int prec, const orc_geometry* g, const orc_scan* s, int cond, int x, int y, int z, int n, int conv,
                 double* out8) {
    if (prec == 0) {
        const Frame<float> f = make_frame<float>(*g, *s, cond, n, conv);
        float u, v;
        V3<float> B;
        raycast(*g, f, x, y, z, conv, u, v, B);
        const double o[8] = {u, v, B.x, B.y, B.z, f.G.x, f.G.y, f.G.z};
        std::memcpy(out8, o, sizeof(o));
    } else {
        const Frame<double> f = make_frame<double>(*g, *s, cond, n, conv);
        double u, v;
        V3<double> B;
        raycast(*g, f, x, y, z, conv, u, v, B);
        const double o[8] = {u, v, B.x, B.y, B.z, f.G.x, f.G.y, f.G.z};
        std::memcpy(out8, o, sizeof(o));
    }
}

// a3/a11 (+a4 when normalise): one view n (signed) of cond.  vol: K*N^3 floats in
// the reference layout (x fastest).  out_proj/out_factor: U*V doubles (overwritten).
void orc_forward_view(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n, const float* vol, int xtt,
                      int normalise, double* out_proj, double* out_factor) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v;
    if (prec == 0) {
        std::vector<float> p(lenP, 0.f), f(lenP, 0.f);
        forward_raw<float, float>(*g, *s, cond, n, vol, xtt, p.data(), f.data());
        if (normalise) normalise_view(*g, make_frame<float>(*g, *s, cond, n, 0), p.data(), f.data());
        for (int64_t i = 0; i < lenP; i++) {
            out_proj[i] = p[i];
            if (out_factor) out_factor[i] = f[i];
        }
    } else {
        std::vector<double> p(lenP, 0.0), f(lenP, 0.0);
        forward_raw<double, float>(*g, *s, cond, n, vol, xtt, p.data(), f.data());
        if (normalise) normalise_view(*g, make_frame<double>(*g, *s, cond, n, 0), p.data(), f.data());
        std::memcpy(out_proj, p.data(), lenP * sizeof(double));
        if (out_factor) std::memcpy(out_factor, f.data(), lenP * sizeof(double));
    }
}

// f3 forwardOrth (+a4 when normalise): one view of the orientation-aware XTT
// operator; coef: 2*N^3 floats (phi, cos theta) in the reference layout.
void orc_forward_view_orth(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n, const float* vol,
                           const float* coef, int normalise, double* out_proj, double* out_factor) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v;
    if (prec == 0) {
        std::vector<float> p(lenP, 0.f), f(lenP, 0.f);
        forward_raw<float, float>(*g, *s, cond, n, vol, 1, p.data(), f.data(), coef);
        if (normalise) normalise_view(*g, make_frame<float>(*g, *s, cond, n, 0), p.data(), f.data());
        for (int64_t i = 0; i < lenP; i++) {
            out_proj[i] = p[i];
            if (out_factor) out_factor[i] = f[i];
        }
    } else {
        std::vector<double> p(lenP, 0.0), f(lenP, 0.0);
        forward_raw<double, float>(*g, *s, cond, n, vol, 1, p.data(), f.data(), coef);
        if (normalise) normalise_view(*g, make_frame<double>(*g, *s, cond, n, 0), p.data(), f.data());
        std::memcpy(out_proj, p.data(), lenP * sizeof(double));
        if (out_factor) std::memcpy(out_factor, f.data(), lenP * sizeof(double));
    }
}

// f3 backwardOrth of one view into K*N^3 accumulators (doubles, accumulated).
void orc_backward_view_orth(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n,
                            const float* proj_view, const float* coef, double* tmp, double* factor) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v;
    const int64_t lenV = (int64_t)g->n_vox * g->n_vox * g->n_vox * s->n_basis;
    if (prec == 0) {
        std::vector<float> p(proj_view, proj_view + lenP), t(lenV, 0.f), f(lenV, 0.f);
        backward_raw<float>(*g, *s, cond, n, p.data(), 1, t.data(), f.data(), coef);
        for (int64_t i = 0; i < lenV; i++) {
            tmp[i] += t[i];
            factor[i] += f[i];
        }
    } else {
        std::vector<double> p(proj_view, proj_view + lenP);
        backward_raw<double>(*g, *s, cond, n, p.data(), 1, tmp, factor, coef);
    }
}

void orc_orth_normal(int prec, const orc_scan* s, int64_t lenV, const float* vol, float* coef, float* loss, int rot,
                     const float* est) {
    if (prec == 0)
        orth_normal<float>(*s, lenV, vol, coef, loss, rot, est);
    else
        orth_normal<double>(*s, lenV, vol, coef, loss, rot, est);
}
void orc_orth_update_estimate(int prec, const orc_scan* s, int64_t lenV, const float* vol, float* est, int iter) {
    if (prec == 0)
        orth_update_estimate<float>(*s, lenV, vol, est, iter);
    else
        orth_update_estimate<double>(*s, lenV, vol, est, iter);
}
void orc_orth_md(int prec, int64_t lenV, const float* vol, const float* coef, float* md) {
    if (prec == 0)
        orth_md<float>(lenV, vol, coef, md);
    else
        orth_md<double>(lenV, vol, coef, md);
}
void orc_orth_mean_filter(int prec, int N, const float* src, const float* vol, float w, float* dst) {
    if (prec == 0)
        orth_mean_filter<float>(N, src, vol, w, dst);
    else
        orth_mean_filter<double>(N, src, vol, w, dst);
}

// f3 XTT::orthTwiceReconstruct.  sino: n_cond*U*V*P floats; vol_in: 3*N^3
// floats (reference layout); vol_out, md_out: 3*N^3 doubles.
void orc_orth_reconstruct(int prec, const orc_geometry* g, const orc_scan* s, const float* sino, const float* vol_in,
                          int iter1, int iter2, int batch, int dir, int art, float lambda, uint32_t seed,
                          double* vol_out, double* md_out) {
    if (prec == 0)
        orth_reconstruct<float>(*g, *s, sino, vol_in, iter1, iter2, batch, dir, art, lambda, seed, vol_out, md_out);
    else
        orth_reconstruct<double>(*g, *s, sino, vol_in, iter1, iter2, batch, dir, art, lambda, seed, vol_out, md_out);
}

// rodriguesRotation (ir.cu:597-621) in FP64: R (9 doubles, row major).
void orc_rodrigues(double x, double y, double z, double c, double sn, double* R) { rodrigues<double>(x, y, z, c, sn, R); }

// a4's L for every pixel of view n.
void orc_chord_length(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n, double* out) {
    const int U = g->n_det_u, V = g->n_det_v;
    if (prec == 0) {
        const Frame<float> f = make_frame<float>(*g, *s, cond, n, 0);
        for (int v = 0; v < V; v++)
            for (int u = 0; u < U; u++) out[u + (int64_t)U * v] = chord_length(*g, f, u, v);
    } else {
        const Frame<double> f = make_frame<double>(*g, *s, cond, n, 0);
        for (int v = 0; v < V; v++)
            for (int u = 0; u < U; u++) out[u + (int64_t)U * v] = chord_length(*g, f, u, v);
    }
}

// a7/a12 raw gather of one view into K*N^3 accumulators (reference layout, doubles, accumulated).
void orc_backward_view(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n, const float* proj_view,
                       int xtt, double* tmp, double* factor) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v;
    const int64_t lenV = (int64_t)g->n_vox * g->n_vox * g->n_vox * (xtt ? s->n_basis : 1);
    if (prec == 0) {
        std::vector<float> p(proj_view, proj_view + lenP), t(lenV, 0.f), f(lenV, 0.f);
        backward_raw<float>(*g, *s, cond, n, p.data(), xtt, t.data(), f.data());
        for (int64_t i = 0; i < lenV; i++) {
            tmp[i] += t[i];
            factor[i] += f[i];
        }
    } else {
        std::vector<double> p(proj_view, proj_view + lenP);
        backward_raw<double>(*g, *s, cond, n, p.data(), xtt, tmp, factor);
    }
}

void orc_subset_orders(int batch, int epochs, int per_epoch_engine, int32_t* out) {
    subset_orders(batch, epochs, per_epoch_engine, out);
}

// a5/a6: projRatio / projSubtract on npix pixels.  proj: in = forward
// projection, out = ratio (MLEM) or residual (ART), in the precision `prec`.
// Returns sum |y - p| (the epoch loss term).
double orc_ratio_step(int prec, int64_t npix, const float* sino, double* proj, int art) {
    if (prec == 0) {
        std::vector<float> p(proj, proj + npix);
        const double loss = ratio_step<float>(npix, sino, p.data(), art);
        for (int64_t i = 0; i < npix; i++) proj[i] = p[i];
        return loss;
    }
    return ratio_step<double>(npix, sino, proj, art);
}

// a10: IR::reconstruct.  sino: n_cond*U*V*P floats; vol_in: n_cond*N^3 floats (reference layout).
// dir 0 = CW, 1 = CCW; art 0 = MLEM, 1 = ART.  vol_out: n_cond*N^3 doubles; losses: n_cond*epochs raw sums.
void orc_ir_reconstruct(int prec, const orc_geometry* g, const orc_scan* s, const float* sino, const float* vol_in,
                        int epochs, int batch, int dir, int art, float lambda, double* vol_out, double* losses) {
    if (prec == 0)
        ir_reconstruct<float>(*g, *s, sino, vol_in, epochs, batch, dir, art, lambda, vol_out, losses);
    else
        ir_reconstruct<double>(*g, *s, sino, vol_in, epochs, batch, dir, art, lambda, vol_out, losses);
}

// f4 gradient path.  g: the gradient volume's geometry (n_vox = N+1); vol and
// the accumulators: (N+1)^3 in the reference layout; ext as in grad_forward_raw.
void orc_grad_forward_view(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n, const float* vol,
                           int ext, double* out_proj) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v, lenV = (int64_t)g->n_vox * g->n_vox * g->n_vox;
    if (prec == 0) {
        std::vector<float> p(lenP, 0.f), v(vol, vol + lenV);
        grad_forward_raw<float>(*g, *s, cond, n, v.data(), ext, p.data());
        for (int64_t i = 0; i < lenP; i++) out_proj[i] = p[i];
    } else {
        std::vector<double> v(vol, vol + lenV);
        std::fill(out_proj, out_proj + lenP, 0.0);
        grad_forward_raw<double>(*g, *s, cond, n, v.data(), ext, out_proj);
    }
}

void orc_grad_backward_view(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n,
                            const float* proj_view, int ext, double* tmp, double* factor) {
    const int64_t lenP = (int64_t)g->n_det_u * g->n_det_v, lenV = (int64_t)g->n_vox * g->n_vox * g->n_vox;
    if (prec == 0) {
        std::vector<float> p(proj_view, proj_view + lenP), t(lenV, 0.f), f(lenV, 0.f);
        grad_backward_raw<float>(*g, *s, cond, n, p.data(), ext, t.data(), f.data());
        for (int64_t i = 0; i < lenV; i++) {
            tmp[i] += t[i];
            factor[i] += f[i];
        }
    } else {
        std::vector<double> p(proj_view, proj_view + lenP);
        grad_backward_raw<double>(*g, *s, cond, n, p.data(), ext, tmp, factor);
    }
}

double orc_grad_coef(int prec, int n, int P, int diff) {
    return prec == 0 ? (double)grad_coef<float>(n, P, diff) : grad_coef<double>(n, P, diff);
}

// IR::gradReconstruct up to the Poisson solve.  g: the result's geometry
// (n_vox = N); sino0: U*V*P floats; grads_out: 3*n_cond*(N+1)^3 doubles;
// losses: epochs doubles.  dir 0 = CW, 1 = CCW; art 0 = MLEM, 1 = ART.
void orc_grad_reconstruct(int prec, const orc_geometry* g, const orc_scan* s, const float* sino0, int epochs,
                          int batch, int dir, int art, float lambda, int block_size, double* grads_out,
                          double* losses) {
    if (prec == 0)
        grad_reconstruct<float>(*g, *s, sino0, epochs, batch, dir, art, lambda, block_size, grads_out, losses);
    else
        grad_reconstruct<double>(*g, *s, sino0, epochs, batch, dir, art, lambda, block_size, grads_out, losses);
}

// a14 (+a13): XTT::reconstruct.  sino: n_cond*U*V*P floats; vol_in: K*N^3 floats.
// vol_out: K*N^3 doubles after voxelSqrt; proj_out (optional): n_cond*U*V*P final
// ratio/residual images; losses: epochs raw sums.
void orc_xtt_reconstruct(int prec, const orc_geometry* g, const orc_scan* s, const float* sino, const float* vol_in,
                         int epochs, int batch, int dir, int art, float lambda, double* vol_out, double* proj_out,
                         double* losses) {
    if (prec == 0)
        xtt_reconstruct<float>(*g, *s, sino, vol_in, epochs, batch, dir, art, lambda, vol_out, proj_out, losses);
    else
        xtt_reconstruct<double>(*g, *s, sino, vol_in, epochs, batch, dir, art, lambda, vol_out, proj_out, losses);
}

void orc_fdk_weight(int prec, const orc_geometry* g, double* out) {
    const int64_t n = (int64_t)g->n_det_u * g->n_det_v;
    if (prec == 0) {
        std::vector<float> w(n);
        fdk_weight<float>(*g, w.data());
        for (int64_t i = 0; i < n; i++) out[i] = w[i];
    } else {
        fdk_weight<double>(*g, out);
    }
}

void orc_fdk_taps(int prec, const orc_geometry* g, int kind, double* out) {
    if (prec == 0) {
        std::vector<float> h(g->n_det_u);
        fdk_taps<float>(*g, kind, h.data());
        for (int i = 0; i < g->n_det_u; i++) out[i] = h[i];
    } else {
        fdk_taps<double>(*g, kind, out);
    }
}

// a16-a18 for one view: weighted, filtered detector image (U*V doubles);
// kind 2: the Hilbert filter of hilbertReconstruct (unweighted rows).
void orc_fdk_filter_view(int prec, const orc_geometry* g, int kind, const float* src_view, double* out) {
    const int U = g->n_det_u;
    const int64_t n = (int64_t)U * g->n_det_v;
    if (kind == 2) {
        if (prec == 0) {
            std::vector<float> d(n);
            hilbert_filter_view<float>(*g, src_view, d.data());
            for (int64_t i = 0; i < n; i++) out[i] = d[i];
        } else {
            hilbert_filter_view<double>(*g, src_view, out);
        }
        return;
    }
    if (prec == 0) {
        std::vector<float> h(U), w(n), d(n);
        fdk_taps<float>(*g, kind, h.data());
        fdk_weight<float>(*g, w.data());
        fdk_filter_view<float>(*g, h.data(), w.data(), src_view, d.data());
        for (int64_t i = 0; i < n; i++) out[i] = d[i];
    } else {
        std::vector<double> h(U), w(n);
        fdk_taps<double>(*g, kind, h.data());
        fdk_weight<double>(*g, w.data());
        fdk_filter_view<double>(*g, h.data(), w.data(), src_view, out);
    }
}

// a19 for one view of an already filtered image; accumulates into vol (N^3 doubles).
void orc_fdk_backproject_view(int prec, const orc_geometry* g, const orc_scan* s, int cond, int n,
                              const double* filt_view, double* vol) {
    const int64_t lenV = (int64_t)g->n_vox * g->n_vox * g->n_vox;
    if (prec == 0) {
        std::vector<float> acc(lenV, 0.f);
        std::vector<float> pv(filt_view, filt_view + (int64_t)g->n_det_u * g->n_det_v);
        fdk_backproject_view<float, float>(*g, *s, cond, n, pv.data(), acc.data());
        for (int64_t i = 0; i < lenV; i++) vol[i] += acc[i];
    } else {
        fdk_backproject_view<double, double>(*g, *s, cond, n, filt_view, vol);
    }
}

// a20: FDK::reconstruct with zero-initialised buffers and each condition reading
// its own filtered sinogram.  vol_out: n_cond*N^3 doubles; filt_out (optional):
// n_cond*U*V*P filtered sinograms (the reference returns them through `sinogram`).
void orc_fdk_reconstruct(int prec, const orc_geometry* g, const orc_scan* s, const float* sino, int dir, int kind,
                         double* vol_out, double* filt_out) {
    const int U = g->n_det_u, V = g->n_det_v, P = g->n_proj, N = g->n_vox;
    const int64_t lenP = (int64_t)U * V, lenD = lenP * P, lenV = (int64_t)N * N * N;
    const int rotation = dir == 0 ? 1 : -1;
    std::vector<double> fv(lenP);
    for (int cond = 0; cond < s->n_cond; cond++) {
        std::vector<double> acc(lenV, 0.0);
        std::vector<float> accf(prec == 0 ? lenV : 0, 0.f);
        for (int n = 0; n < P; n++) {
            orc_fdk_filter_view(prec, g, kind, sino + cond * lenD + n * lenP, fv.data());
            if (filt_out) std::memcpy(filt_out + cond * lenD + n * lenP, fv.data(), lenP * sizeof(double));
            if (prec == 0) {
                std::vector<float> pv(fv.begin(), fv.end());
                fdk_backproject_view<float, float>(*g, *s, cond, rotation * n, pv.data(), accf.data());
            } else {
                fdk_backproject_view<double, double>(*g, *s, cond, rotation * n, fv.data(), acc.data());
            }
        }
        for (int64_t i = 0; i < lenV; i++) vol_out[cond * lenV + i] = prec == 0 ? (double)accf[i] : acc[i];
    }
}

// a15: calcEigenVector over nvox voxels.  mu: K blocks of nvox floats; basis: 3K floats.
// md: 3 blocks of nvox; evalue: 3 blocks of nvox (doubles).
void orc_pca(const float* mu, int K, const float* basis, int64_t nvox, double* md, double* evalue) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nvox; i++) {
        double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double mean = 0;
        for (int k = 0; k < K; k++) {
            const double m = mu[k * nvox + i];
            mean += m;
            const double sc[3] = {m * basis[3 * k + 0], m * basis[3 * k + 1], m * basis[3 * k + 2]};
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) C[3 * r + c] += sc[r] * sc[c];
        }
        mean /= K;
        for (int j = 0; j < 9; j++) C[j] /= K;
        double ev[3], vec[9];
        jacobi3(C, ev, vec);
        double e[3] = {vec[0], vec[3], vec[6]};
        const double nrm = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        for (int r = 0; r < 3; r++) e[r] /= nrm;
        const double sign = (e[2] >= 0) ? 1.0 : -1.0;
        for (int r = 0; r < 3; r++) md[r * nvox + i] = sign * mean * e[r];
        const double tot = ev[0] + ev[1] + ev[2];
        for (int r = 0; r < 3; r++) evalue[r * nvox + i] = ev[r] / tot;
    }
}

// Symmetric 3x3 eigen solve (ascending eigenvalues, eigenvectors as columns).
void orc_eig3(const double* A, double* evals, double* evecs) { jacobi3(A, evals, evecs); }

// calcAngleFromMD (direction/src/pca.cu:67-83) in the reference's FP32 host
// arithmetic: md = 3 channels of n floats, angle = 2 channels (phi, theta).
void orc_md_angles(const float* md, float* angle, int64_t n) {
    for (int64_t i = 0; i < n; i++) {
        const float a = md[i], b = md[n + i], c = md[2 * n + i];
        if (b < 0.f) {
            angle[i] = std::atan2(-b, -a);
            angle[n + i] = std::atan2(std::sqrt(b * b + a * a), -c);
        } else {
            angle[i] = std::atan2(b, a);
            angle[n + i] = std::atan2(std::sqrt(b * b + a * a), c);
        }
    }
}

// directionNormalization (direction/src/pca.cu:85-106), eps = 0.65 there.
void orc_md_normalise(float* md, int64_t n, float eps) {
    for (int64_t i = 0; i < n; i++) {
        const float x = md[i], y = md[n + i], z = md[2 * n + i];
        const float nrm = std::sqrt(x * x + y * y + z * z);
        if (nrm > eps) {
            md[i] = x / nrm;
            md[n + i] = y / nrm;
            md[2 * n + i] = z / nrm;
        } else {
            md[i] = md[n + i] = md[2 * n + i] = 0.f;
        }
    }
}

// Ellipsoid phantom rasterised at voxel centres, voxel x at (x-(N+1)/2)*vs (the IR
// convention, ir.cu:1882).  params: 16 doubles per ellipsoid: centre[3] (mm),
// semi-axes[3] (mm), rotation[9] (row-major, body <- world), value.  Values add.
void orc_phantom_ellipsoids(int N, double vs, const double* params, int count, float* vol) {
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const double p[3] = {((double)x - (N + 1) * 0.5) * vs, ((double)y - (N + 1) * 0.5) * vs,
                                     ((double)z - (N + 1) * 0.5) * vs};
                double acc = 0;
                for (int e = 0; e < count; e++) {
                    const double* q = params + 16 * e;
                    const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
                    double r2 = 0;
                    for (int i = 0; i < 3; i++) {
                        const double b = q[6 + 3 * i] * d[0] + q[7 + 3 * i] * d[1] + q[8 + 3 * i] * d[2];
                        r2 += (b / q[3 + i]) * (b / q[3 + i]);
                    }
                    if (r2 <= 1.0) acc += q[15];
                }
                vol[x + (int64_t)N * y + (int64_t)N * N * z] = (float)acc;
            }
}

// Philox4x32-10 block function (Salmon et al., SC 2011), restated.
static void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t a = (uint64_t)0xD2511F53u * c[0], b = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(b >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(a >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)b;
        c[3] = (uint32_t)a;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Uniforms of element idx: pairs of 32-bit words -> 53-bit doubles in (0,1).
struct OrcUniforms {
    uint64_t seed, idx;
    uint32_t draw = 0, w[4] = {0, 0, 0, 0};
    int left = 0;
    OrcUniforms(uint64_t s, uint64_t i) : seed(s), idx(i) {}
    double next() {
        if (!left) {
            w[0] = (uint32_t)idx;
            w[1] = (uint32_t)(idx >> 32);
            w[2] = draw++;
            w[3] = 0;
            philox10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
            left = 2;
        }
        const uint32_t a = left == 2 ? w[0] : w[2], b = left == 2 ? w[1] : w[3];
        left--;
        const uint64_t m = (((uint64_t)a << 21) ^ (uint64_t)b) & ((1ull << 53) - 1);
        return ((double)m + 0.5) / 9007199254740992.0;
    }
};

// Poisson: multiplication method below 10, PTRS (Hoermann 1993) above.
static double orc_poisson(double lam, OrcUniforms& u) {
    if (!(lam > 0)) return 0;
    if (lam < 10) {
        const double L = std::exp(-lam);
        double p = 1, k = 0;
        for (;;) {
            p *= u.next();
            if (p <= L) return k;
            k += 1;
        }
    }
    const double sl = std::sqrt(lam), ll = std::log(lam), b = 0.931 + 2.53 * sl, a = -0.059 + 0.02483 * b;
    const double ia = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = u.next() - 0.5, V = u.next(), us = 0.5 - std::fabs(U);
        const double k = std::floor((2 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(ia) - std::log(a / (us * us) + b) <= -lam + k * ll - std::lgamma(k + 1)) return k;
    }
}

void orc_noise_poisson(float* y, const float* p, int64_t n, double I0, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        OrcUniforms u{seed, (uint64_t)i};
        const double c = orc_poisson(I0 * std::exp(-(double)p[i]), u);
        y[i] = (float)(-std::log(std::max(c, 1.0) / I0));
    }
}

void orc_noise_gaussian(float* y, const float* x, int64_t n, double sigma, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        OrcUniforms u{seed, (uint64_t)i};
        const double u1 = u.next(), u2 = u.next();
        const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
        y[i] = (float)((double)x[i] * (1.0 + sigma * g));
    }
}

// Hollow z cylinder added into vol (reference layout), voxel x at (x-(N+1)/2)*vs.
void orc_phantom_cylinder_shell(int N, double vs, double r_out, double thick, double value, float* vol) {
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const double px = (x - (N + 1) * 0.5) * vs, py = (y - (N + 1) * 0.5) * vs;
                const double r = std::sqrt(px * px + py * py);
                if (r <= r_out && r > r_out - thick) vol[x + (int64_t)N * y + (int64_t)N * N * z] += (float)value;
            }
}

// Fibres (8 doubles each: centre[3], unit dir[3], radius, half-length): channel k
// gains |S_k x f|^2 inside each fibre.  vol: K channels of N^3 (reference layout).
void orc_phantom_fibres(int N, double vs, const double* prm, int count, const float* basis, int K, float* vol) {
    const int64_t N3 = (int64_t)N * N * N;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const double p[3] = {(x - (N + 1) * 0.5) * vs, (y - (N + 1) * 0.5) * vs, (z - (N + 1) * 0.5) * vs};
                std::vector<float> acc(K, 0.f);
                for (int e = 0; e < count; e++) {
                    const double* q = prm + 8 * e;
                    const double d[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
                    const double t = d[0] * q[3] + d[1] * q[4] + d[2] * q[5];
                    if (std::fabs(t) > q[7]) continue;
                    const double r[3] = {d[0] - t * q[3], d[1] - t * q[4], d[2] - t * q[5]};
                    if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] > q[6] * q[6]) continue;
                    for (int k = 0; k < K; k++) {
                        const double S[3] = {basis[3 * k], basis[3 * k + 1], basis[3 * k + 2]};
                        const double c[3] = {S[1] * q[5] - S[2] * q[4], S[2] * q[3] - S[0] * q[5],
                                             S[0] * q[4] - S[1] * q[3]};
                        acc[k] += (float)(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                    }
                }
                const int64_t i = x + (int64_t)N * y + (int64_t)N * N * z;
                for (int k = 0; k < K; k++) vol[i + k * N3] += acc[k];
            }
}

int orc_num_threads(void) { return n_threads(); }

}  // extern "C"
