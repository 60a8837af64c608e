// HIP kernels (gfx950) for the cone-beam reconstruction hot path.
//
// Volume layout in HBM ("column layout"): voxel (x, y, z) of channel k lives at
//   k*chan_stride + z + N*(x + N*y),
// i.e. z-columns are contiguous.  The rotation axis is z, so for an untilted
// scan (condR = I) a z-column projects onto a single detector column pair with
// v advancing ~1 pixel per voxel: 64 lanes walking 64 consecutive z of one
// column read 256 contiguous bytes and hit 64 distinct detector rows.
// Sinograms keep the reference layout u + U*(v + V*view) per condition.
//
// Forward projection (ir.cu:1490-1531, 1753-1804 + 1533-1603 + 1616-1666) is
// voxel-driven, like the reference, but each block OWNS a band of kBandU
// detector columns of one view: it accumulates the bilinear splat of every voxel
// that touches the band in LDS, then writes the band once with the
// L/(factor+1e-5) normalisation and the MLEM ratio / ART residual fused in.
// No global atomics; the sinogram is written exactly once per view.  The LDS
// accumulation is fixed point: each contribution is rounded to a 32-bit integer
// on a per-launch power-of-two scale (from max |voxel value|) and summed with
// 64-bit integer LDS atomics (ds_add_u64).  On gfx950 these run at bank rate
// while FP32 LDS atomics (ds_add_f32) serialise per lane (~20x slower here), and
// integer sums are exact and order independent: the forward projection is
// bitwise deterministic.
//
// Backprojection (ir.cu:1720-1751, 1807-1850, fdk.cu:62-170) is a gather: a
// block owns a tile of voxels, loops over all views of the subset, stages the
// tile's detector footprint in LDS and accumulates in registers; the MLEM/ART
// update (ir.cu:1637-1693) is applied in the same kernel, so each voxel is read
// and written once per subset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "xr_internal.hpp"
#include "xr_random.hpp"

namespace xr {

namespace {

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// Wave-uniform copy of a value loaded through a vector load: the compiler
// then keeps everything derived from it in scalar registers and scalar
// branches (a view's table is the same for every lane).
__device__ __forceinline__ float uniformf(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ int uniformi(int x) { return __builtin_amdgcn_readfirstlane(x); }

struct ColForms {
    float nu, nv, nd;
};

// The projective forms at z' = 0 for column (x', y').  Every kernel evaluates
// u, v through col_forms + voxel_uv so that the splat, the gather and the
// normalisation table agree bit for bit on (u, v) and the acceptance window.
__device__ __forceinline__ ColForms col_forms(const ViewGeom& g, float xp, float yp) {
    ColForms c;
    c.nu = fmaf(g.au[1], yp, fmaf(g.au[0], xp, g.au[3]));
    c.nv = fmaf(g.av[1], yp, fmaf(g.av[0], xp, g.av[3]));
    c.nd = fmaf(g.ad[1], yp, fmaf(g.ad[0], xp, g.ad[3]));
    return c;
}

__device__ __forceinline__ void voxel_uv(const ViewGeom& g, const ColForms& c, float zp, const Dims& d, float& u,
                                         float& v) {
    const float nu = fmaf(g.au[2], zp, c.nu);
    const float nv = fmaf(g.av[2], zp, c.nv);
    const float nd = fmaf(g.ad[2], zp, c.nd);
    const float inv = rcpf(nd);
    u = fmaf(nu, inv, d.hu);
    v = fmaf(nv, inv, d.hv);
}

// Acceptance window 0.55 < u < U-0.55 (ir.cu:1501) and bilinear split with
// floor(u-0.5) (ir.cu:1504-1509).
__device__ __forceinline__ bool accept_u(float u, const Dims& d) { return 0.55f < u && u < d.umax; }
__device__ __forceinline__ bool accept_v(float v, const Dims& d) { return 0.55f < v && v < d.vmax; }
__device__ __forceinline__ void split(float u, int& iu, float& fu) {
    const float t = u - 0.5f;
    const float f = floorf(t);
    iu = (int)f;
    fu = t - f;
}

__device__ __forceinline__ ViewGeom uniform_view(const ViewGeom& v) {
    ViewGeom u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u.au[i] = uniformf(v.au[i]);
        u.av[i] = uniformf(v.av[i]);
        u.ad[i] = uniformf(v.ad[i]);
        u.pad[i] = 0.f;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        u.d0[i] = uniformf(v.d0[i]);
        u.G[i] = uniformf(v.G[i]);
    }
    u.fy = uniformf(v.fy);
    u.axis = uniformf(v.axis);
    return u;
}

typedef __attribute__((address_space(3))) float lds_float;

// Global-memory view of a pointer.  Kernel-argument pointers are known to be
// global, but pointers read from a work-item table in memory are not: without
// this cast their loads are flat accesses (slower, and mixing them with LDS
// accesses in one kernel is what the gfx950 backend handles worst).
// (The host pass only parses device code; it sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <class T>
using gptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}

// Read-only view of a table that no kernel writes (view lists, per-view
// geometry and weight tables): loads through the constant address space at a
// wave-uniform address become scalar loads into SGPRs, instead of vector loads
// plus a readfirstlane (a VALU instruction) per value.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <class T>
using cptr = const T*;
#endif
template <class T>
__device__ __forceinline__ cptr<T> to_const(const T* p) {
    return (cptr<T>)p;
}

__device__ __forceinline__ void lds_add(lds_float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Chord length of the pixel-centre ray through the reconstruction box, FP64,
// following correlationProjByLength (ir.cu:1556-1597): the six faces in order
// x-,x+,y-,y+,z-,z+; a face hit counts when the other two coordinates lie
// strictly inside; L is the distance between the first two hits (one hit: its
// distance from the origin, as the reference's zero-filled buffer gives).
__device__ double chord_length(const ChordGeom& c, int u, int v, const Dims& d) {
    const double cu = ((double)u - ((double)d.U - 1.0) * 0.5) * c.det[0];
    const double cv = ((double)v - ((double)d.V - 1.0) * 0.5) * c.det[0];
    double pix[3], s2p[3];
    for (int i = 0; i < 3; i++) {
        pix[i] = cu * c.eu[i] - c.det[1] * c.ey[i] + cv * c.ev[i];
        s2p[i] = c.src[i] - pix[i];
    }
    int cnt = 0;
    double h0x = 0, h0y = 0, h0z = 0, h1x = 0, h1y = 0, h1z = 0;
#pragma unroll
    for (int p = 0; p < 3; p++) {
#pragma unroll
        for (int sg = 0; sg < 2; sg++) {
            const double plane = sg ? c.hi[p] : c.lo[p];
            const double t = (plane - c.src[p]) / s2p[p];
            double pl[3];
#pragma unroll
            for (int i = 0; i < 3; i++) pl[i] = t * s2p[i] + c.src[i];
            const int p1 = (p + 1) % 3, p2 = (p + 2) % 3;
            if (c.lo[p1] < pl[p1] && pl[p1] < c.hi[p1] && c.lo[p2] < pl[p2] && pl[p2] < c.hi[p2]) {
                if (cnt == 0) {
                    h0x = pl[0];
                    h0y = pl[1];
                    h0z = pl[2];
                } else if (cnt == 1) {
                    h1x = pl[0];
                    h1y = pl[1];
                    h1z = pl[2];
                }
                cnt++;
            }
        }
    }
    const double dx = h0x - h1x, dy = h0y - h1y, dz = h0z - h1z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// Forward projection: one block = (detector band of kBandU columns, one view,
// a chunk of detector rows).  The block's waves (kFwdWaves, or kFwdBatchWaves in
// the persistent multi-problem launch) walk the volume rows y (round
// robin), find the z-columns whose footprint touches the band (ballot over 64 x
// per step), and splat them with lanes = z.  Two columns are in flight per wave
// and all z chunks of a column are loaded before any is used.
// ---------------------------------------------------------------------------
constexpr int kFwdChunks = 8;  // z chunks of 64 loaded per column per step (512 z)
constexpr int kXttChunks = 2;  // untilted XTT: z chunks of 64 (x K channels) loaded per step

typedef __attribute__((address_space(3))) unsigned long long lds_u64;

// Band accumulator word: 64-bit for the weight pass (its own sum bound is what
// it computes), 32-bit otherwise, with the scale chosen from the largest
// per-pixel weight sum so that no pixel sum can reach 2^31 in magnitude.
template <int KIND>
struct FwdAcc {
    typedef unsigned int T;
    typedef int S;
};
template <>
struct FwdAcc<FWD_WEIGHT> {
    typedef unsigned long long T;
    typedef long long S;
};
template <class T>
using lds_ptr = __attribute__((address_space(3))) T*;

struct FwdCol {
    ColForms c;
    float inv, u;
    int64_t base;  // column base index in the column layout
    float dx, dy;  // XTT source->voxel vector (x, y parts)
    int z_beg, z_end;  // z range that can reach this launch's rows (AXIS), else [0, nz)
};

// Value of x in lane `l` (wave-uniform l), into a scalar register.
__device__ __forceinline__ float lane_f(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ int lane_i(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// z range of an untilted column that can reach detector rows [v0, v0+vb):
// v(z') = hv + (nv + av2 z') inv is linear in z'; a voxel reaches those rows
// only if v in [v0 - 0.5, v0 + vb + 0.5).  The range is widened by two voxels
// and every voxel is still tested exactly in the splat.  A chunk of all rows
// keeps [0, nz): the splat's window test rejects voxels that miss the detector.
__device__ __forceinline__ void col_zrange(const ViewGeom& g, const ColForms& c, float inv, const Dims& d, int v0,
                                           int vb, int& z_beg, int& z_end) {
    z_beg = 0;
    z_end = d.nz;
    if (v0 <= 0 && v0 + vb >= d.V) return;
    const float slope = g.av[2] * inv;
    if (fabsf(slope) > 1e-6f) {
        const float za = ((float)v0 - 0.5f - d.hv - c.nv * inv) / slope;
        const float zb = ((float)(v0 + vb) + 0.5f - d.hv - c.nv * inv) / slope;
        const float lo = fminf(za, zb) + d.xc - 2.0f, hi = fmaxf(za, zb) + d.xc + 2.0f;
        z_beg = (int)fmaxf(0.0f, floorf(lo));
        z_end = (int)fminf((float)d.nz, ceilf(hi) + 1.0f);
        if (!(lo == lo) || !(hi == hi)) {
            z_beg = 0;
            z_end = d.nz;
        }
    }
}

// The same z range from t(z') = v - 0.5 = B z' + A, z' = z - xc: rows
// [v0, v0+vb) need t in [v0 - 1, v0 + vb); widened by two voxels (the
// rounding of A, B differs from the forms above by far less), tested exactly
// in the splat.
__device__ __forceinline__ void col_zrange_t(float B, float A, const Dims& d, int v0, int vb, int& z_beg,
                                             int& z_end) {
    z_beg = 0;
    z_end = d.nz;
    if (fabsf(B) > 1e-6f) {
        const float za = ((float)(v0 - 1) - A) / B;
        const float zb = ((float)(v0 + vb) - A) / B;
        const float lo = fminf(za, zb) + d.xc - 2.0f, hi = fmaxf(za, zb) + d.xc + 2.0f;
        z_beg = (int)fmaxf(0.0f, floorf(lo));
        z_end = (int)fminf((float)d.nz, ceilf(hi) + 1.0f);
        if (!(lo == lo) || !(hi == hi)) {
            z_beg = 0;
            z_end = d.nz;
        }
    }
}

template <bool AXIS, int KIND>
__device__ __forceinline__ FwdCol fwd_col(const FwdArgs& a, const ViewGeom& g, int xcol, int y, int v0, int vb) {
    const Dims& d = a.d;
    FwdCol f;
    const float xp = (float)xcol - d.xc, yp = (float)y - d.xc;
    f.c = col_forms(g, xp, yp);
    f.inv = AXIS ? rcpf(f.c.nd) : 0.f;
    f.u = AXIS ? fmaf(f.c.nu, f.inv, d.hu) : 0.f;
    f.base = ((int64_t)y * d.N + xcol) * d.N;
    f.dx = (KIND == FWD_XTT) ? fmaf(d.vs, xp, g.d0[0]) : 0.f;
    f.dy = (KIND == FWD_XTT) ? fmaf(d.vs, yp, g.d0[1]) : 0.f;
    f.z_beg = 0;
    f.z_end = d.nz;
    if (AXIS) col_zrange(g, f.c, f.inv, d, v0, vb, f.z_beg, f.z_end);
    f.z_beg = uniformi(f.z_beg);
    f.z_end = uniformi(f.z_end);
    return f;
}

// Splat one voxel (z of column f, value val) into the band: fixed point with
// scale 2^S, 64-bit integer LDS atomics.
template <bool AXIS, class AccT, class AccS>
__device__ __forceinline__ void fwd_splat(const FwdArgs& a, const ViewGeom& g, const FwdCol& f, int z, float val,
                                          float scale, lds_ptr<AccT> acc, int u0, int v0, int vb) {
    const Dims& d = a.d;
    const float zp = (float)z - d.xc;
    float u, v;
    if (AXIS) {
        u = f.u;
        v = fmaf(fmaf(g.av[2], zp, f.c.nv), f.inv, d.hv);
    } else {
        voxel_uv(g, f.c, zp, d, u, v);
    }
    int iu, iv;
    float fu, fv;
    split(u, iu, fu);
    split(v, iv, fv);
    bool ok = z < f.z_end && accept_u(u, d) && accept_v(v, d) && iv >= v0 - 1 && iv < v0 + vb;
    if (!AXIS) ok = ok && iu >= u0 - 1 && iu < u0 + kBandU;
    if (ok) {
        const float vs = val * scale;
        const float c1 = (1.0f - fu) * fv, c2 = fu * fv;
        const float c3 = fu * (1.0f - fv), c4 = (1.0f - fu) * (1.0f - fv);
        lds_ptr<AccT> p = acc + (iv - v0 + 1) * kBandStride + (iu - u0 + 1);
        const int q4 = __float2int_rn(c4 * vs), q3 = __float2int_rn(c3 * vs);
        const int q1 = __float2int_rn(c1 * vs), q2 = __float2int_rn(c2 * vs);
        __hip_atomic_fetch_add(p, (AccT)(AccS)q4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(p + 1, (AccT)(AccS)q3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(p + kBandStride, (AccT)(AccS)q1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(p + kBandStride + 1, (AccT)(AccS)q2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// floor(x + 0.5) in one instruction (round half up; x is far inside int range).
__device__ __forceinline__ int round_half_up(float x) {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// floor(x) as an integer in one instruction (x is far inside int range).
__device__ __forceinline__ int floor_int(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Untilted splat of one 64-z chunk of one column.  u, hence (iu, fu), is
// constant along the column, so the weights factor into (1-fu, fu) x
// (1-fv, fv) with wu0 = (1-fu)*s, wu1 = fu*s carrying the fixed-point scale,
// and v is affine in z: t = v - 0.5 = B*z_lane + A (B, A per column).  The
// acceptance window 0.55 < v < V-0.55 and the chunk's row range reduce to
// lo < t < hi (lo >= 0.05, so an accepted t is positive and fract(t) =
// t - floor(t) exactly).  Branch free: a rejected voxel adds zero at `spare`,
// a lane-dependent row of the band (one shared address would serialise the
// wave's atomics).  row0 = the LDS byte address of band row floor(t) = 0 of
// the column, so an accepted voxel's first row is row0 + floor(t) * row pitch.
template <class AccT, class AccS>
__device__ __forceinline__ void splat_untilted(float t, float val, unsigned row0, lds_ptr<AccT> spare, float wu0,
                                               float wu1, float lo, float hi) {
    const bool ok = t > lo && t < hi;
    const float fv = __builtin_amdgcn_fractf(t);
    const float w = ok ? val : 0.0f;
    const float g1 = fv * w, g0 = w - g1;
    const unsigned at = __umul24((unsigned)floor_int(t), (unsigned)(kBandStride * sizeof(AccT))) + row0;
    lds_ptr<AccT> p = ok ? (lds_ptr<AccT>)(uintptr_t)at : spare;
    const int q4 = round_half_up(wu0 * g0), q3 = round_half_up(wu1 * g0);
    const int q1 = round_half_up(wu0 * g1), q2 = round_half_up(wu1 * g1);
    __hip_atomic_fetch_add(p, (AccT)(AccS)q4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + 1, (AccT)(AccS)q3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + kBandStride, (AccT)(AccS)q1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + kBandStride + 1, (AccT)(AccS)q2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// splat_untilted for a voxel known to lie in the window (0.05 < t, so fract(t)
// = t - floor(t) exactly): the same terms at the same addresses, without the
// test and the selects.
template <class AccT, class AccS>
__device__ __forceinline__ void splat_inside(float t, float val, unsigned row0, float wu0, float wu1) {
    const float fv = __builtin_amdgcn_fractf(t);
    const float g1 = fv * val, g0 = val - g1;
    const unsigned at = __umul24((unsigned)floor_int(t), (unsigned)(kBandStride * sizeof(AccT))) + row0;
    lds_ptr<AccT> p = (lds_ptr<AccT>)(uintptr_t)at;
    const int q4 = round_half_up(wu0 * g0), q3 = round_half_up(wu1 * g0);
    const int q1 = round_half_up(wu0 * g1), q2 = round_half_up(wu1 * g1);
    __hip_atomic_fetch_add(p, (AccT)(AccS)q4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + 1, (AccT)(AccS)q3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + kBandStride, (AccT)(AccS)q1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(p + kBandStride + 1, (AccT)(AccS)q2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// XTT voxel value: sum_k |B x S_k|^2 (S_k.G)^2 mu_k * vs * ratio (ir.cu:1780-1794),
// with |B x S|^2 = |S|^2 - (d.S)^2/|d|^2 for the source->voxel vector d.
__device__ __forceinline__ float xtt_value(const FwdArgs& a, const FwdCol& f, const ViewGeom& g, int z,
                                           const float* g2, const float* mu) {
    const Dims& d = a.d;
    const float zp = (float)z - d.xc;
    const float dz = fmaf(d.vs, zp, g.d0[2]);
    const float d2 = fmaf(dz, dz, fmaf(f.dy, f.dy, f.dx * f.dx));
    const float id2 = rcpf(d2);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kMaxBasis; k++) {
        if (k < d.K) {
            const float ds = fmaf(dz, a.b.S[k][2], fmaf(f.dy, a.b.S[k][1], f.dx * a.b.S[k][0]));
            const float cr2 = fmaxf(fmaf(-ds * ds, id2, a.b.S2[k]), 0.f);
            s = fmaf(cr2 * g2[k], mu[k], s);
        }
    }
    return s * (d.vs * d.ratio);
}

// The per-voxel basis rotation of forwardOrth / backwardOrth (ir.cu:316-328,
// 390-400): rodriguesRotation (ir.cu:597-621) about the axis (cos phi, sin phi,
// 0) with cosine cc and sine sqrt(1 - cc^2), (1 - cc) n n^T + [[cc, -n_z s,
// n_y s], [n_z s, cc, -n_x s], [-n_y s, n_x s, cc]] with n_z = 0.
__device__ __forceinline__ void orth_rotation(float phi, float cc, float R[9]) {
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float sn = sqrtf(1.0f - cc * cc);
    const float len = sqrtf(cp * cp + sp * sp);
    if (len < 1e-8f) {
        const float I[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
#pragma unroll
        for (int i = 0; i < 9; i++) R[i] = I[i];
        return;
    }
    const float nx = cp / len, ny = sp / len, om = 1.0f - cc;
    R[0] = om * nx * nx + cc;
    R[1] = om * nx * ny;
    R[2] = ny * sn;
    R[3] = om * nx * ny;
    R[4] = om * ny * ny + cc;
    R[5] = -nx * sn;
    R[6] = -ny * sn;
    R[7] = nx * sn;
    R[8] = cc;
}

// Orientation-aware XTT weight of channel k at a voxel: S' = R S_k, w =
// max(|S'|^2 - (d.S')^2 / |d|^2, 0) (S'.G)^2, i.e. |B x S'|^2 (S'.G)^2
// (ir.cu:352-359, 410-417).
__device__ __forceinline__ float orth_weight(const float R[9], const Basis& b, int k, float dx, float dy, float dz,
                                             float id2, const ViewGeom& g) {
    const float s0 = fmaf(R[0], b.S[k][0], fmaf(R[1], b.S[k][1], R[2] * b.S[k][2]));
    const float s1 = fmaf(R[3], b.S[k][0], fmaf(R[4], b.S[k][1], R[5] * b.S[k][2]));
    const float s2 = fmaf(R[6], b.S[k][0], fmaf(R[7], b.S[k][1], R[8] * b.S[k][2]));
    const float ds = fmaf(dz, s2, fmaf(dy, s1, dx * s0));
    const float ss = fmaf(s2, s2, fmaf(s1, s1, s0 * s0));
    const float gs = fmaf(g.G[2], s2, fmaf(g.G[1], s1, g.G[0] * s0));
    return fmaxf(fmaf(-ds * ds, id2, ss), 0.f) * (gs * gs);
}

// forwardOrth voxel value: sum_k |B x R S_k|^2 (R S_k . G)^2 mu_k * vs * ratio.
__device__ __forceinline__ float orth_value(const FwdArgs& a, const FwdCol& f, const ViewGeom& g, int z,
                                            const float* mu, float phi, float cc) {
    const Dims& d = a.d;
    const float zp = (float)z - d.xc;
    const float dz = fmaf(d.vs, zp, g.d0[2]);
    const float id2 = rcpf(fmaf(dz, dz, fmaf(f.dy, f.dy, f.dx * f.dx)));
    float R[9];
    orth_rotation(phi, cc, R);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kMaxBasis; k++)
        if (k < d.K) s = fmaf(orth_weight(R, a.b, k, f.dx, f.dy, dz, id2, g), mu[k], s);
    return s * (d.vs * d.ratio);
}

// Untilted view, volume row y' = yp: u(x') = hu + (au0 x' + cu) / (ad0 x' + cd)
// is linear-fractional in x', hence monotone while the denominator is
// positive, so the columns whose u falls in [ua, ub] form an interval.  The
// u of a column is evaluated exactly as the scan's candidate test does; its
// rounding error (a few ulp, far below kRowMargin) can break monotonicity by
// that much, so every decision below keeps kRowMargin clear of the band:
//  - u at both ends of the row below ua (or both above ub): no column can be
//    a candidate, the row is empty;
//  - otherwise the closed-form solution, widened by two columns, is a guess
//    (it degenerates where u hardly varies along the row), accepted only if
//    u lies beyond the band on the proper side at the first column left of it
//    and at the first column right of it;
//  - anything else (denominator not positive, non-finite u, a rejected guess)
//    scans the whole row.
// The scan still tests every column of [xs, xe) exactly, so the result does
// not depend on the guess.
constexpr float kRowMargin = 0.01f;

__device__ __forceinline__ float row_u(const ViewGeom& g, const Dims& d, int x, float yp) {
    const ColForms c = col_forms(g, (float)x - d.xc, yp);
    return fmaf(c.nu, rcpf(c.nd), d.hu);
}

__device__ __forceinline__ void band_columns(const ViewGeom& g, const Dims& d, float yp, float ua, float ub, int& xs,
                                             int& xe) {
    xs = 0;
    xe = d.N;
    const float cu = fmaf(g.au[1], yp, g.au[3]), cd = fmaf(g.ad[1], yp, g.ad[3]);
    const float x0 = -d.xc, x1 = (float)(d.N - 1) - d.xc;
    if (!(fmaf(g.ad[0], x0, cd) > 0.f && fmaf(g.ad[0], x1, cd) > 0.f)) return;
    const float u_first = row_u(g, d, 0, yp), u_last = row_u(g, d, d.N - 1, yp);
    if (!(fabsf(u_first) < 1e30f && fabsf(u_last) < 1e30f)) return;
    const float lo_u = ua - kRowMargin, hi_u = ub + kRowMargin;
    if (fmaxf(u_first, u_last) < lo_u || fminf(u_first, u_last) > hi_u) {
        xe = 0;  // the whole row misses the band
        return;
    }
    if (!(u_first != u_last)) return;
    const bool up = u_last > u_first;
    // x' with u(x') = u: (u - hu)(ad0 x' + cd) = au0 x' + cu
    const float ka = ua - d.hu, kb = ub - d.hu;
    const float xa = (cu - ka * cd) / (ka * g.ad[0] - g.au[0]);
    const float xb = (cu - kb * cd) / (kb * g.ad[0] - g.au[0]);
    const float lo = fminf(xa, xb) + d.xc - 2.0f, hi = fmaxf(xa, xb) + d.xc + 2.0f;
    if (!(lo == lo) || !(hi == hi)) return;
    const int s = (int)fminf(fmaxf(floorf(lo), 0.0f), (float)d.N);
    const int e = (int)fminf(fmaxf(ceilf(hi) + 1.0f, (float)s), (float)d.N);
    if (s > 0) {
        const float u = row_u(g, d, s - 1, yp);
        if (!(up ? u < lo_u : u > hi_u)) return;
    }
    if (e < d.N) {
        const float u = row_u(g, d, e, yp);
        if (!(up ? u > hi_u : u < lo_u)) return;
    }
    xs = s;
    xe = e;
}

// One thread per (code, band, volume row): the scan bounds of forward_item.
// A candidate of band b has iu in [16b - 1, 16b + 16), i.e. u in
// [16b - 0.5, 16b + 15.5), inside [ua, ub] below.
__global__ void k_band_rows(const ViewGeom* __restrict__ vg, const Dims d, int n_codes, int n_bands,
                            int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n_codes * n_bands * d.N) return;
    const int y = (int)(i % d.N), band = (int)((i / d.N) % n_bands), code = (int)(i / ((int64_t)d.N * n_bands));
    const ViewGeom g = vg[code];
    int xs = 0, xe = d.N;
    if (g.axis != 0.f) {
        const float u0 = (float)(band * kBandU);
        band_columns(g, d, (float)y - d.xc, u0 - 0.5f, u0 + (float)kBandU + 0.5f, xs, xe);
    }
    out[i] = xs | (xe << 16);
}

// One (band, view) work item o = band * n_slots + slot of a forward launch.
// step selects the view list views + step*n_slots (a solver's subset step);
// loss receives the item's share of sum |y - p|.
template <int KIND, bool AXIS, int WAVES, int KM = kMaxBasis>
__device__ __forceinline__ void forward_item(const FwdArgs& a, int o, int step, double* loss,
                                             unsigned long long* acc_raw, int v0, int vb) {
    typedef typename FwdAcc<KIND>::T AccT;
    typedef typename FwdAcc<KIND>::S AccS;
    const lds_ptr<AccT> acc = (lds_ptr<AccT>)acc_raw;
    const Dims& d = a.d;
    // Item o = (sub-band j, slot).  Sub-band j = (band, sub) owns the kBandU/n_sub
    // detector columns us + [0, W) of the band's accumulator; a split launch
    // hands out the central (heaviest) sub-bands first.
    static_assert(kBandU == 16, "sub-band shifts assume 16-column bands");
    const int nsub = a.n_sub > 1 ? a.n_sub : 1;
    const int lw = nsub == 4 ? 2 : (nsub == 2 ? 3 : 4);  // log2 of W
    int j = o / a.n_slots;
    const int slot = o % a.n_slots;
    if (nsub > 1) {
        const int mid = (a.n_bands * nsub) >> 1;
        j = (j & 1) ? mid - ((j + 1) >> 1) : mid + (j >> 1);
    }
    const int band = j >> (4 - lw), sub = j & (nsub - 1);
    const int W = 1 << lw, cs = sub << lw;  // the item's columns cs + [0, W) of the band
    const auto views = to_const(a.views + (int64_t)step * a.n_slots);
    const auto accum = to_global(a.accum ? a.accum + (int64_t)step * a.n_slots : nullptr);
    const auto proj = to_global(a.proj);
    const int code = views[slot];
    const int cond = code / d.P;
    const ViewGeom g = to_const(a.vg)[code];
    const int u0 = band * kBandU;
    const int rows = vb + 2;
    for (int i = threadIdx.x; i < rows * kBandStride; i += 64 * WAVES) acc[i] = 0ull;
    // Fixed-point scale.  Weight pass (64-bit words): a power of two 2^S with
    // the largest term < 2^30 * 2^-S.  Other passes (32-bit words): the bound
    // of a pixel sum (wmul * max|voxel| * fsum, each factor an upper bound
    // with its own margin) maps to 2^31 (1 - 2^-10), so every sum of rounded
    // terms stays inside the signed 32-bit range while using all but the
    // last fraction of a bit of it.
    const float bound = fmaxf(a.wmul * (*to_global(a.vmax)) * a.fsum, 1e-30f);
    float scale, unscale;
    if (KIND == FWD_WEIGHT) {
        int e = 0;
        (void)frexpf(bound, &e);
        const int S = min(max(30 - e, -100), 100);
        scale = ldexpf(1.0f, S);
        unscale = ldexpf(1.0f, -S);
    } else {
        scale = fminf(fmaxf(2145386496.0f / bound, 0x1p-100f), 0x1p100f);
        unscale = a.out_scale / scale;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = uniformi(threadIdx.x >> 6);
    const auto vol = to_global(a.vol + (KIND == FWD_PLAIN ? (int64_t)cond * a.vol_cond_stride : 0));
    float g2[kMaxBasis];
    // untilted XTT: the view's row of the per-plan weight table (BwdArgs::xw)
    constexpr int xws = kXwStride;
    const auto xrow = to_const(a.xw ? a.xw + (int64_t)code * 5 * xws : nullptr);
    if (KIND == FWD_XTT && !AXIS) {
#pragma unroll
        for (int k = 0; k < kMaxBasis; k++) g2[k] = (k < d.K) ? to_const(a.g2)[(int64_t)code * d.K + k] : 0.f;
    }
    // value scale folded into the untilted splat weights: ratio (ir.cu:1513-1518),
    // XTT vs * ratio (ir.cu:1791)
    const float scalev = (KIND == FWD_PLAIN) ? d.ratio : (KIND == FWD_XTT ? d.vs * d.ratio : 1.f);
    const float zlo = -d.xc, zhi = (float)(d.N - 1) - d.xc;
    const float band_lo = (float)(u0 + cs) - 0.5f, band_hi = (float)(u0 + cs + W) + 0.5f;
    // untilted path: lane offsets of the z chunks, and the window of t = v - 0.5:
    // 0.55 < v < V - 0.55 (ir.cu:1501) and v0 - 1 <= floor(t) < v0 + vb
    float lanef[kFwdChunks];
#pragma unroll
    for (int c = 0; c < kFwdChunks; c++) lanef[c] = (float)(c * 64 + lane);
    const float t_lo = (v0 >= 2) ? __int_as_float(__float_as_int((float)(v0 - 1)) - 1) : 0.05f;
    const float t_hi = fminf(d.vmax - 0.5f, (float)(v0 + vb));
    const int spare_row = min(lane, vb);  // rows spare_row, spare_row + 1 lie in the band
    const bool rows_all = v0 <= 0 && v0 + vb >= d.V;  // this launch covers every detector row

    const auto bands = to_global(a.bands ? a.bands + ((int64_t)code * a.n_bands + band) * d.N : nullptr);
    for (int y = wave; y < d.N; y += WAVES) {
        const float yp = (float)y - d.xc;
        int xs = 0, xe = d.N;  // columns of this row to scan (the test below is exact)
        if (AXIS && a.bands) {
            const int v = uniformi(bands[y]);
            xs = v & 0xffff;
            xe = v >> 16;
        }
        for (int x0 = xs & ~63; x0 < xe; x0 += 64) {
            const int x = x0 + lane;
            bool cand = false;
            // untilted: this lane's column terms, kept for the candidate loop (which
            // reads a candidate's terms from its lane): t = v - 0.5 = B (z - xc) + A,
            // the two u weights with the fixed-point scale, the band column, the z range
            float sB = 0.f, sA = 0.f, sw0 = 0.f, sw1 = 0.f;
            int scp = 0;
            // untilted XTT: |d|^2 - dz^2 of the column's source->voxel vector d and
            // sig_k (d_x S_k,x + d_y S_k,y), so that sig_k (d.S_k) = dz sig_k S_k,z + xck_k
            float xcxy = 0.f, xck[KM];
            if (KIND == FWD_XTT && AXIS) {
                const float xp = (float)x - d.xc;
                const float dx = fmaf(d.vs, xp, g.d0[0]), dy = fmaf(d.vs, yp, g.d0[1]);
                xcxy = fmaf(dy, dy, dx * dx);
#pragma unroll
                for (int k = 0; k < KM; k++)
                    xck[k] = fmaf(xp, xrow[2 * xws + k], fmaf(yp, xrow[3 * xws + k], xrow[4 * xws + k]));
            }
            if (x < d.N) {
                const ColForms c = col_forms(g, (float)x - d.xc, yp);
                if (AXIS) {
                    const float inv = rcpf(c.nd);
                    const float u = fmaf(c.nu, inv, d.hu);
                    if (accept_u(u, d)) {
                        int iu;
                        float fu;
                        split(u, iu, fu);
                        cand = (unsigned)(iu - (u0 + cs - 1)) <= (unsigned)W;
                        sB = g.av[2] * inv;
                        sA = fmaf(c.nv, inv, d.hv - 0.5f);
                        sw0 = (1.0f - fu) * (scale * scalev);
                        sw1 = fu * (scale * scalev);
                        scp = iu - u0 + 1;
                    }
                } else {
                    // u is linear-fractional (monotone) along the column while den > 0.
                    const float nd0 = fmaf(g.ad[2], zlo, c.nd), nd1 = fmaf(g.ad[2], zhi, c.nd);
                    if (!(nd0 > 0.f && nd1 > 0.f)) {
                        cand = true;
                    } else {
                        const float ua = fmaf(fmaf(g.au[2], zlo, c.nu), rcpf(nd0), d.hu);
                        const float ub = fmaf(fmaf(g.au[2], zhi, c.nu), rcpf(nd1), d.hu);
                        const float lo = fminf(ua, ub) - 0.01f, hi = fmaxf(ua, ub) + 0.01f;
                        cand = (hi > band_lo) && (lo < band_hi) && (hi > 0.55f) && (lo < d.umax);
                        if (ua != ua || ub != ub) cand = true;
                    }
                }
            }
            unsigned long long m = __ballot(cand);
            if (KIND == FWD_XTT && AXIS) {
                // Untilted XTT: a candidate's column terms come from its lane; per voxel
                // the value sum_k mu_k (S_k.G)^2 |B x S_k|^2 (ir.cu:1780-1794) is
                // sum_k mu_k (g2s_k - q_k^2 / |d|^2), q_k = sig_k (d.S_k) affine in z;
                // vs * ratio rides in the splat weights.
                // EXACT: d.K == KM (the k loops need no bound).  k is the outer loop of a
                // chunk pair, so only one k's constants are live at a time.
                auto candidates = [&](auto exact_c) {
                    constexpr bool EXACT = decltype(exact_c)::value;
                    while (m) {
                        const int b = __ffsll(m) - 1;
                        m &= m - 1;
                        const float Bc = lane_f(sB, b), Ac = lane_f(sA, b);
                        const float w0 = lane_f(sw0, b), w1 = lane_f(sw1, b);
                        const float cxy = lane_f(xcxy, b);
                        const lds_ptr<AccT> cp = acc + lane_i(scp, b);
                        int zb = 0, ze = d.nz;
                        if (!rows_all) col_zrange_t(Bc, Ac, d, v0, vb, zb, ze);
                        const int64_t cbase = ((int64_t)y * d.N + x0 + b) * d.N;
                        const unsigned row0 =
                            (unsigned)(uintptr_t)cp + (unsigned)((1 - v0) * kBandStride * (int)sizeof(AccT));
                        const lds_ptr<AccT> spare = cp + spare_row * kBandStride;
                        const bool whole = (d.nz & 63) == 0 && zb == 0 && ze == d.nz;
                        // every voxel of the column inside the window (t monotone in z)
                        const float ta = fmaf(Bc, -d.xc, Ac), tb = fmaf(Bc, (float)(d.nz - 1) - d.xc, Ac);
                        const bool inside = whole && fminf(ta, tb) > t_lo && fmaxf(ta, tb) < t_hi;
                        const auto col = vol + cbase + lane;
                        for (int zg = zb & ~63; zg < ze; zg += 64 * kXttChunks) {
                            float mu[kXttChunks][KM];
                            if (EXACT && whole && zg + 64 * kXttChunks <= d.nz) {
                                // every lane in range: unmasked loads, one address per channel
                                // walked from the previous one (the chunks are immediate offsets)
#pragma unroll
                                for (int k = 0; k < KM; k++)
#pragma unroll
                                    for (int c = 0; c < kXttChunks; c++)
                                        mu[c][k] = col[zg + c * 64 + (int64_t)k * a.chan_stride];
                            } else {
#pragma unroll
                                for (int c = 0; c < kXttChunks; c++) {
                                    const int z = zg + c * 64 + lane;
                                    const bool in = whole ? (zg + c * 64 < d.nz) : (z >= zb && z < ze);
#pragma unroll
                                    for (int k = 0; k < KM; k++)
                                        mu[c][k] = (in && (EXACT || k < d.K))
                                                       ? col[zg + c * 64 + (int64_t)k * a.chan_stride]
                                                       : 0.f;
                                }
                            }
                            const float zoff = (float)zg - d.xc;
                            float dz[kXttChunks], id2[kXttChunks], val[kXttChunks];
#pragma unroll
                            for (int c = 0; c < kXttChunks; c++) {
                                dz[c] = fmaf(d.vs, lanef[c] + zoff, g.d0[2]);
                                id2[c] = rcpf(fmaf(dz[c], dz[c], cxy));
                                val[c] = 0.f;
                            }
#pragma unroll
                            for (int k = 0; k < KM; k++) {
                                if (EXACT || k < d.K) {
                                    const float ck = lane_f(xck[k], b), sS2 = xrow[k], g2s = xrow[xws + k];
#pragma unroll
                                    for (int c = 0; c < kXttChunks; c++) {
                                        const float q = fmaf(dz[c], sS2, ck);
                                        val[c] = fmaf(fmaf(-q * q, id2[c], g2s), mu[c][k], val[c]);
                                    }
                                }
                            }
#pragma unroll
                            for (int c = 0; c < kXttChunks; c++) {
                                if (zg + c * 64 >= ze) break;
                                const float t = fmaf(Bc, lanef[c] + zoff, Ac);
                                if (inside)
                                    splat_inside<AccT, AccS>(t, val[c], row0, w0, w1);
                                else
                                    splat_untilted<AccT, AccS>(t, val[c], row0, spare, w0, w1, t_lo, t_hi);
                            }
                        }
                    }
                };
                if (d.K == KM)
                    candidates(std::true_type());
                else
                    candidates(std::false_type());
            } else if (KIND == FWD_XTT) {
                while (m) {
                    const int bit = __ffsll(m) - 1;
                    m &= m - 1;
                    const FwdCol f = fwd_col<AXIS, KIND>(a, g, x0 + bit, y, v0, vb);
                    for (int z0 = f.z_beg; z0 < f.z_end; z0 += 64) {
                        const int z = z0 + lane;
                        float mu[kMaxBasis];
#pragma unroll
                        for (int k = 0; k < kMaxBasis; k++)
                            mu[k] = (k < d.K && z < f.z_end) ? vol[f.base + z + k * a.chan_stride] : 0.f;
                        float val;
                        if (a.coef) {
                            const bool in = z < f.z_end;
                            const float phi = in ? a.coef[f.base + z] : 0.f;
                            const float cc = in ? a.coef[f.base + z + a.chan_stride] : 1.f;
                            val = orth_value(a, f, g, z, mu, phi, cc);
                        } else {
                            val = xtt_value(a, f, g, z, g2, mu);
                        }
                        fwd_splat<AXIS, AccT, AccS>(a, g, f, z, val, scale, acc, u0, v0, vb);
                    }
                }
            } else {
                while (m) {
                    const int b0 = __ffsll(m) - 1;
                    m &= m - 1;
                    const bool two = m != 0;
                    const int b1 = two ? __ffsll(m) - 1 : b0;
                    if (two) m &= m - 1;
                    if constexpr (AXIS) {
                        // the candidates' terms, computed by their lanes in the scan
                        const float B0 = lane_f(sB, b0), B1 = lane_f(sB, b1);
                        const float A0 = lane_f(sA, b0), A1 = lane_f(sA, b1);
                        const float w00 = lane_f(sw0, b0), w01 = lane_f(sw1, b0);
                        const float w10 = lane_f(sw0, b1), w11 = lane_f(sw1, b1);
                        const lds_ptr<AccT> cp0 = acc + lane_i(scp, b0), cp1 = acc + lane_i(scp, b1);
                        int zb0 = 0, ze0 = d.nz, zb1 = 0, ze1 = d.nz;
                        if (!rows_all) {
                            col_zrange_t(B0, A0, d, v0, vb, zb0, ze0);
                            col_zrange_t(B1, A1, d, v0, vb, zb1, ze1);
                        }
                        const int zb = min(zb0, zb1), ze = max(ze0, ze1);
                        const int64_t base0 = ((int64_t)y * d.N + x0 + b0) * d.N;
                        const int64_t base1 = ((int64_t)y * d.N + x0 + b1) * d.N;
                        // LDS byte address of band row floor(t) = 0 of each column, and each
                        // lane's spare row for rejected voxels
                        const unsigned row00 =
                            (unsigned)(uintptr_t)cp0 + (unsigned)((1 - v0) * kBandStride * (int)sizeof(AccT));
                        const unsigned row01 =
                            (unsigned)(uintptr_t)cp1 + (unsigned)((1 - v0) * kBandStride * (int)sizeof(AccT));
                        const lds_ptr<AccT> spare0 = cp0 + spare_row * kBandStride;
                        const lds_ptr<AccT> spare1 = cp1 + spare_row * kBandStride;
                        // t = v - 0.5 = B*(z - xc) + (nv*inv + hv - 0.5); z - xc is exact in
                        // FP32 (half-integers below 2^23), so t depends on z alone and not on
                        // where the chunk group starts.
                        // Whole columns (both z ranges are [0, N), N a multiple of 64): every
                        // loaded z is in range, so the loads need no masks and share one
                        // address per column.  Wave-uniform.
                        const bool whole = KIND != FWD_WEIGHT && (d.nz & 63) == 0 && zb0 == 0 && ze0 == d.nz &&
                                           zb1 == 0 && ze1 == d.nz;
                        const auto col0 = vol + base0 + lane, col1 = vol + base1 + lane;
                        // Inside: every voxel of a whole column lies in the window (t at the
                        // first and last z bound every lane's t: z - xc is exact and fmaf is
                        // monotone in it), so the splat needs no window test.
                        auto inside_col = [&](float B, float A) {
                            const float ta = fmaf(B, -d.xc, A), tb = fmaf(B, (float)(d.nz - 1) - d.xc, A);
                            return fminf(ta, tb) > t_lo && fmaxf(ta, tb) < t_hi;
                        };
                        const bool inside = whole && inside_col(B0, A0) && (!two || inside_col(B1, A1));
                        // one or two columns, whole or clipped, inside the window or not:
                        // template constants, so that no per-chunk test of any remains
                        auto columns = [&](auto two_c, auto whole_c, auto inside_c) {
                            constexpr bool TWO = decltype(two_c)::value;
                            constexpr bool WHOLE = decltype(whole_c)::value;
                            constexpr bool INSIDE = decltype(inside_c)::value;
                            // z chunks start at multiples of 64 so that a lane always holds the
                            // same z residue: the splat's lane pairing (hence its rounding) does
                            // not depend on where the row chunk clips the column.
                            for (int zg = zb & ~63; zg < ze; zg += 64 * kFwdChunks) {
                                float m0[kFwdChunks], m1[kFwdChunks];
                                if (WHOLE && zg + 64 * kFwdChunks <= d.nz) {
                                    // all chunks in range: plain loads at immediate offsets
#pragma unroll
                                    for (int c = 0; c < kFwdChunks; c++) {
                                        m0[c] = col0[zg + c * 64];
                                        if (TWO) m1[c] = col1[zg + c * 64];
                                    }
                                } else {
#pragma unroll
                                for (int c = 0; c < kFwdChunks; c++) {
                                    const int z = zg + c * 64 + lane;
                                    if (WHOLE) {
                                        const bool in = zg + c * 64 < d.nz;
                                        m0[c] = in ? col0[zg + c * 64] : 0.f;
                                        if (TWO) m1[c] = in ? col1[zg + c * 64] : 0.f;
                                    } else if (KIND == FWD_WEIGHT) {
                                        // zero outside the column's z range: the splat has no z test
                                        m0[c] = (z < ze0 && z >= zb0) ? 1.f : 0.f;
                                        if (TWO) m1[c] = (z < ze1 && z >= zb1) ? 1.f : 0.f;
                                    } else {
                                        m0[c] = (z < ze0 && z >= zb0) ? vol[base0 + z] : 0.f;
                                        if (TWO) m1[c] = (z < ze1 && z >= zb1) ? vol[base1 + z] : 0.f;
                                    }
                                }
                                }
                                const float zoff = (float)zg - d.xc;
#pragma unroll
                                for (int c = 0; c < kFwdChunks; c++) {
                                    if (zg + c * 64 >= ze) break;
                                    const float zp = lanef[c] + zoff;
                                    if (INSIDE) {
                                        splat_inside<AccT, AccS>(fmaf(B0, zp, A0), m0[c], row00, w00, w01);
                                        if (TWO) splat_inside<AccT, AccS>(fmaf(B1, zp, A1), m1[c], row01, w10, w11);
                                    } else {
                                        splat_untilted<AccT, AccS>(fmaf(B0, zp, A0), m0[c], row00, spare0, w00, w01,
                                                                   t_lo, t_hi);
                                        if (TWO)
                                            splat_untilted<AccT, AccS>(fmaf(B1, zp, A1), m1[c], row01, spare1, w10,
                                                                       w11, t_lo, t_hi);
                                    }
                                }
                            }
                        };
                        using T_ = std::true_type;
                        using F_ = std::false_type;
                        if (two) {
                            if (inside)
                                columns(T_(), T_(), T_());
                            else if (whole)
                                columns(T_(), T_(), F_());
                            else
                                columns(T_(), F_(), F_());
                        } else {
                            if (inside)
                                columns(F_(), T_(), T_());
                            else if (whole)
                                columns(F_(), T_(), F_());
                            else
                                columns(F_(), F_(), F_());
                        }
                        continue;
                    }
                    const FwdCol f0 = fwd_col<AXIS, KIND>(a, g, x0 + b0, y, v0, vb);
                    const FwdCol f1 = fwd_col<AXIS, KIND>(a, g, x0 + b1, y, v0, vb);
                    const int zb = min(f0.z_beg, f1.z_beg), ze = max(f0.z_end, f1.z_end);
                    for (int zg = zb & ~63; zg < ze; zg += 64 * kFwdChunks) {
                        float m0[kFwdChunks], m1[kFwdChunks];
#pragma unroll
                        for (int c = 0; c < kFwdChunks; c++) {
                            const int z = zg + c * 64 + lane;
                            if (KIND == FWD_WEIGHT) {
                                m0[c] = (z < f0.z_end && z >= f0.z_beg) ? 1.f : 0.f;
                                m1[c] = (two && z < f1.z_end && z >= f1.z_beg) ? 1.f : 0.f;
                            } else {
                                m0[c] = (z < f0.z_end && z >= f0.z_beg) ? vol[f0.base + z] : 0.f;
                                m1[c] = (two && z < f1.z_end && z >= f1.z_beg) ? vol[f1.base + z] : 0.f;
                            }
                        }
#pragma unroll
                        for (int c = 0; c < kFwdChunks; c++) {
                            if (zg + c * 64 >= ze) break;
                            const int z = zg + c * 64 + lane;
                            if (z >= f0.z_beg) fwd_splat<AXIS, AccT, AccS>(a, g, f0, z, m0[c] * scalev, scale, acc, u0, v0, vb);
                        }
                        if (two) {
#pragma unroll
                            for (int c = 0; c < kFwdChunks; c++) {
                                if (zg + c * 64 >= ze) break;
                                const int z = zg + c * 64 + lane;
                                if (z >= f1.z_beg) fwd_splat<AXIS, AccT, AccS>(a, g, f1, z, m1[c] * scalev, scale, acc, u0, v0, vb);
                            }
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    double lsum = 0.0, l2sum = 0.0;  // sum |y - p|, sum (y - p)^2 (OUT_RESIDUAL_L2)
    float fmax_loc = 0.0f;  // weight pass: largest per-pixel weight sum
    const int64_t img = (int64_t)(code - a.code_base) * d.V * d.U;
    const int64_t nimg = (int64_t)code * d.V * d.U;
    const bool acc_old = (KIND != FWD_WEIGHT) && accum && accum[slot];
    for (int i = threadIdx.x; i < (vb << lw); i += 64 * WAVES) {
        const int r = i >> lw, cc = cs + (i & (W - 1));
        const int uu = u0 + cc, vv = v0 + r;
        if (uu >= d.U || vv >= d.V) continue;
        const float s = (float)(AccS)acc[(r + 1) * kBandStride + cc + 1] * unscale;
        const int64_t pix = img + (int64_t)vv * d.U + uu;
        const int64_t npix = nimg + (int64_t)vv * d.U + uu;
        if (a.out_mode == OUT_NORM_TABLE) {
            const float L = (float)chord_length(to_global(a.cg)[code], uu, vv, d);
            proj[pix] = L / (s + 1e-5f);
            fmax_loc = fmaxf(fmax_loc, s);
            continue;
        }
        float p = (acc_old ? proj[pix] : 0.f) + s;
        if (a.out_mode == OUT_RAW) {
            proj[pix] = p;
            continue;
        }
        p = p * to_global(a.norm)[npix];
        if (a.out_mode == OUT_NORMALISED) {
            proj[pix] = p;
        } else {
            const float yv = to_global(a.sino)[pix];
            lsum += (double)fabsf(yv - p);
            if (a.out_mode == OUT_RESIDUAL || a.out_mode == OUT_RESIDUAL_L2) {
                const float rv = yv - p;
                proj[pix] = rv;
                l2sum = fma((double)rv, (double)rv, l2sum);
            } else {
                float q = p;
                if (q != 0.0f) q = yv / q;
                if (q > 2.0f) q = 2.0f;
                proj[pix] = q;
            }
        }
    }
    if (a.out_mode == OUT_NORM_TABLE && a.fmax) {
        for (int off = 32; off > 0; off >>= 1) fmax_loc = fmaxf(fmax_loc, __shfl_xor(fmax_loc, off));
        if (lane == 0) atomicMax((unsigned int*)a.fmax, __float_as_uint(fmax_loc));
    }
    if (a.out_mode == OUT_RATIO || a.out_mode == OUT_RESIDUAL || a.out_mode == OUT_RESIDUAL_L2) {
        // Cross-wave sums through LDS past the band accumulator (8-byte aligned
        // slots: WAVES for the L1 sum, then WAVES for the squared sum).
        double* red = reinterpret_cast<double*>(acc_raw + fwd_red_slot(rows, sizeof(AccT)));
        const bool l2 = a.out_mode == OUT_RESIDUAL_L2;
        lsum = wave_sum(lsum);
        if (l2) l2sum = wave_sum(l2sum);
        if (lane == 0) {
            red[wave] = lsum;
            if (l2) red[WAVES + wave] = l2sum;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0, t2 = 0.0;
            for (int w = 0; w < WAVES; w++) t += red[w];
            atomicAdd(loss, t);
            if (l2) {
                for (int w = 0; w < WAVES; w++) t2 += red[WAVES + w];
                atomicAdd(loss + 1, t2);
            }
        }
    }
}

// Item order of a forward launch.  Items (band j, slot) are numbered
// j*n_slots + slot (band-major).  With view_group G > 0 the launch walks groups
// of G consecutive slots, every band of a group before the next group, and G
// slots of one band next to each other: blocks running at the same time then
// read the slabs of one band at G neighbouring views, which overlap, so a
// column fetched by one is found in L2 by the others (at 512^3 / 720 views the
// forward reads 2.6x less from HBM than band-major over all views).  The order
// changes no result: every item writes its own pixels.
__device__ __forceinline__ int view_grouped(const FwdArgs& a, int oi) {
    if (a.view_group <= 0 || a.n_sub > 1) return oi;
    const int nb = a.n_bands, G = a.view_group;
    const int full = (a.n_slots / G) * G * nb;
    int jj, sl;
    if (oi < full) {
        const int grp = oi / (nb * G), r = oi - grp * nb * G;
        jj = r / G;
        sl = grp * G + (r - jj * G);
    } else {
        const int r = oi - full, gl = a.n_slots - (a.n_slots / G) * G;
        jj = r / gl;
        sl = (a.n_slots / G) * G + (r - jj * gl);
    }
    return jj * a.n_slots + sl;
}

template <int KIND, bool AXIS, int KM = kMaxBasis, int WAVES = kFwdWaves>
__global__ __launch_bounds__(64 * WAVES) void k_forward(const FwdArgs a) {
    extern __shared__ unsigned long long acc_raw[];
    const int per_chunk = a.n_bands * max(a.n_sub, 1) * a.n_slots;
    const int nblk = per_chunk * a.n_chunks;
    const int per_xcd = (nblk + 7) / 8;
    // XCD-chunked order: each XCD walks a contiguous run of (row chunk, band,
    // view) triples in the view-grouped order of view_grouped, so concurrent
    // blocks on one L2 share a few bands at adjacent views.  A split launch
    // (few views) lists its heaviest items first; there blocks take items in
    // launch order, so that consecutive (round-robin dispatched) blocks spread
    // them over the XCDs.
    const int o = a.n_sub > 1 ? (int)blockIdx.x : (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (o >= nblk) return;
    const int chunk = o / per_chunk;
    const int v0 = a.v0 + chunk * a.vb;
    const int vb = min(a.vb, a.d.V - v0);
    forward_item<KIND, AXIS, WAVES, KM>(a, view_grouped(a, o - chunk * per_chunk), 0, a.loss, acc_raw, v0, vb);
}

// Work queue of a persistent launch: q[0] hands out tickets 0, 1, 2, ...; a
// block stops at its first ticket >= total and counts itself in q[1]; the last
// block to stop resets both words, so the next launch (or graph replay) starts
// from zero without a host-side clear.  Every block stops after one failed
// ticket: there is no wait on other blocks.
__device__ __forceinline__ unsigned take_ticket(unsigned* q, unsigned total) {
    const unsigned t = __hip_atomic_fetch_add(q, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t >= total) {
        const unsigned done = __hip_atomic_fetch_add(q + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done == gridDim.x - 1) {
            __hip_atomic_store(q, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(q + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    return t;
}

// Item holding ticket t: the first i with item_end[i] > t (binary search).
__device__ __forceinline__ int find_item(const int* __restrict__ item_end, int n_items, int t) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (item_end[mid] > t)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Persistent forward launch over many problems: items[i] is one (problem,
// detector-row chunk) launch description with item_end[i] - item_end[i-1]
// (band, view) blocks.  A fixed grid of resident blocks pulls tickets in order,
// so the long items listed first are started first and the short ones fill
// the tail.  ticket_slot: index of a free 8-byte LDS word past every item's
// band accumulator.
template <int KIND, int AXES, int WAVES, int KM = kMaxBasis>
__device__ __forceinline__ void forward_batch_loop(const FwdArgs* __restrict__ items,
                                                   const int* __restrict__ item_end, int n_items, unsigned* queue,
                                                   int ticket_slot, int step, int loss_shift) {
    extern __shared__ unsigned long long acc_raw[];
    lds_u64* const tk = (lds_u64*)acc_raw + ticket_slot;
    const int total = item_end[n_items - 1];
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) *tk = take_ticket(queue, (unsigned)total);
        __syncthreads();
        // uniform across the block: keeps the item's arguments in scalar registers
        const int t = __builtin_amdgcn_readfirstlane((int)*tk);
        if (t >= total) break;
        const int i = find_item(item_end, n_items, t);
        const FwdArgs& a = items[i];
        const int o = view_grouped(a, t - (i ? item_end[i - 1] : 0));
        double* loss = a.loss ? a.loss + loss_shift : nullptr;
        if (AXES == 1 || a.axis)
            forward_item<KIND, true, WAVES, KM>(a, o, step, loss, acc_raw, a.v0, a.vb);
        else
            forward_item<KIND, false, WAVES, KM>(a, o, step, loss, acc_raw, a.v0, a.vb);
    }
}

// Persistent forward of XTT problems (forwardProjXTT, ir.cu:1753-1804): the
// same loop with the directional splat, KM >= every item's basis count.
template <int KM>
__global__ __launch_bounds__(64 * kFwdXttWaves) void k_forward_batch_xtt(const FwdArgs* __restrict__ items,
                                                                       const int* __restrict__ item_end,
                                                                       int n_items, unsigned* queue, int ticket_slot,
                                                                       int step, int loss_shift) {
    forward_batch_loop<FWD_XTT, 0, kFwdXttWaves, KM>(items, item_end, n_items, queue, ticket_slot, step, loss_shift);
}

template <int KIND>
__global__ __launch_bounds__(64 * kFwdBatchWaves) void k_forward_batch(const FwdArgs* __restrict__ items,
                                                                  const int* __restrict__ item_end, int n_items,
                                                                  unsigned* queue, int ticket_slot, int step,
                                                                  int loss_shift) {
    forward_batch_loop<KIND, 0, kFwdBatchWaves>(items, item_end, n_items, queue, ticket_slot, step, loss_shift);
}

// The same loop when every item is untilted: without the tilted path the
// kernel needs fewer registers, so 12-wave blocks fit two per CU.
template <int KIND>
__global__ __launch_bounds__(64 * kFwdWaves) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_forward_batch_axis(const FwdArgs* __restrict__ items,
                                                                      const int* __restrict__ item_end, int n_items,
                                                                      unsigned* queue, int ticket_slot, int step,
                                                                      int loss_shift) {
    forward_batch_loop<KIND, 1, kFwdWaves>(items, item_end, n_items, queue, ticket_slot, step, loss_shift);
}

// ---------------------------------------------------------------------------
// Backprojection gather with fused update.  Block = BwdBlock waves laid out
// as a gx x gy grid of CPW x CPW column groups; every lane owns ZPL voxels
// (z = z0 + lane + 64 q) in each of its CPW^2 columns, so the per-column
// geometry of a view is computed once per ZPL voxels.  The block's detector
// footprint for each view is staged in LDS: the untilted paths fetch view s+2
// into registers while view s is accumulated (four buffers, a barrier every
// second view), the general path view s+1 (two buffers, a barrier per view).
// The update (MLEM / ART / FDK / raw) is applied at the end: each voxel is
// read and written once per subset.
// ---------------------------------------------------------------------------
template <int KIND>
struct IsXtt {
    static constexpr bool value = (KIND == BWD_XTT_RAW || KIND == BWD_XTT_MLEM || KIND == BWD_XTT_ART);
};

// Block shape per kind.  The plain kinds (about 110 VGPRs, 4 waves per SIMD)
// run one 16-wave block per CU; XTT runs 8-wave blocks, two per CU, so one
// block accumulates while the other waits at its per-view barrier (its
// footprints are small: 8 columns x 64 z).  Measured: XTT 150 against 167 ms
// per backprojection at 256^3 x 13 coefficients, 4 x 360 views; the plain
// kinds in 8-wave blocks (half the tile, 6144-float staging) 70 against 67 ms
// at 512^3 / 720 views.
// The untilted paths stage each view two views ahead in four buffers, so
// that a block meets a barrier every second view; the general (tilted) path
// double buffers in the first two.
template <int KIND>
struct BwdBlock {
    static constexpr int waves = IsXtt<KIND>::value ? 8 : 16;
    static constexpr int threads = 64 * waves;
    static constexpr int nbuf = 4;                                      // staging buffers
    static constexpr int stage_buf = IsXtt<KIND>::value ? 4000 : 8000;  // floats per staging buffer
    static constexpr int stage_per = stage_buf / threads;              // staged floats per thread
};

struct Foot {
    int iu_lo, iv_lo, W, H, Ws, R;  // R: footprint rows staged per pass of the block's threads
    bool staged, empty;
};

template <int KIND>
__device__ __forceinline__ Foot foot_from_extent(float umn, float umx, float vmn, float vmx, int bad, const Dims& d);

// Footprint of the block box on view g: the projective map's extremes over a box
// are at its corners while den > 0 there (all lanes compute it identically).
template <int KIND>
__device__ __forceinline__ Foot footprint(const ViewGeom& g, const Dims& d, float bx0, float bx1, float by0, float by1,
                                          float bz0, float bz1, int lane) {
    const int ci = lane & 7;
    float cu, cv, cd;
    {
        const ColForms c = col_forms(g, (ci & 1) ? bx1 : bx0, (ci & 2) ? by1 : by0);
        const float czp = (ci & 4) ? bz1 : bz0;
        cd = fmaf(g.ad[2], czp, c.nd);
        voxel_uv(g, c, czp, d, cu, cv);
    }
    float umn = cu, umx = cu, vmn = cv, vmx = cv;
    int bad = (!(cd > 0.f) || !(fabsf(cu) < 1e7f) || !(fabsf(cv) < 1e7f)) ? 1 : 0;
    for (int off = 1; off < 8; off <<= 1) {
        umn = fminf(umn, __shfl_xor(umn, off));
        umx = fmaxf(umx, __shfl_xor(umx, off));
        vmn = fminf(vmn, __shfl_xor(vmn, off));
        vmx = fmaxf(vmx, __shfl_xor(vmx, off));
        bad |= __shfl_xor(bad, off);
    }
    return foot_from_extent<KIND>(umn, umx, vmn, vmx, bad, d);
}

// Footprint from the extremes of the projected corners (bad: a corner behind
// the source or far off the detector, so the extremes bound nothing).
template <int KIND>
__device__ __forceinline__ Foot foot_from_extent(float umn, float umx, float vmn, float vmx, int bad, const Dims& d) {
    Foot f{0, 0, 0, 0, 0, 0, false, false};
    if (bad) return f;  // direct (unstaged) reads
    f.iu_lo = max((int)floorf(umn - 0.5f) - 1, 0);
    const int iu_hi = min((int)floorf(umx - 0.5f) + 2, d.U - 1);
    f.iv_lo = max((int)floorf(vmn - 0.5f) - 1, 0);
    const int iv_hi = min((int)floorf(vmx - 0.5f) + 2, d.V - 1);
    f.W = iu_hi - f.iu_lo + 1;
    f.H = iv_hi - f.iv_lo + 1;
    if (f.W <= 1 || f.H <= 1) {
        f.empty = true;  // entirely off the detector: nothing is accepted
        return f;
    }
    f.Ws = f.W | 1;
    f.R = BwdBlock<KIND>::threads / f.W;
    // the last word of the second buffer is reserved (work-queue ticket of the
    // persistent launch), so a staged footprint holds at most stage_buf-1 floats
    f.staged = (f.Ws * f.H <= BwdBlock<KIND>::stage_buf - 1) && (f.H <= BwdBlock<KIND>::stage_per * f.R);
    return f;
}

// Word 3 of a raw buffer descriptor on gfx9-family GPUs (32-bit data, no
// swizzle, stride 0): the range check is on the byte offset alone.
constexpr int kBufferWord3 = 0x00020000;

// Staging of one view's footprint (W x H texels at (iu_lo, iv_lo) of the
// view's U-wide image) into an LDS buffer of row stride Ws, in passes of
// R = threads / W rows: thread t owns column c = t % W of rows r0 + k R,
// r0 = t / W (one division per view).  Only the passes the footprint needs
// are issued and stored (uniform tests); only the last needs a per-lane row
// test.  The LDS contents are those of a plain row-by-row copy.
// BUF: a pass is one raw buffer load at a per-lane byte offset whose range is
// the view's image, so a row below the image's last reads 0 (it lies outside
// the footprint and is never stored): no address select, no predication, one
// add per pass.  Otherwise (kernels whose scalar registers are already
// scarce: the buffer descriptor takes four) a global load per pass, lanes
// outside the footprint reading its first texel.
template <int NP, bool BUF>
struct FootStage {
    float pre[NP];
    int r0, lds0;  // lds0 = r0 Ws + c; r0 >= R: the thread stages nothing

    __device__ __forceinline__ void fetch(const float* img, int img_bytes, const Foot& f, int U) {
        r0 = (int)((float)threadIdx.x * rcpf((float)f.W));  // within 1 of t / W: corrected below
        int c = (int)threadIdx.x - r0 * f.W;
        if (c < 0) {
            r0--;
            c += f.W;
        } else if (c >= f.W) {
            r0++;
            c -= f.W;
        }
        lds0 = r0 * f.Ws + c;
        if constexpr (BUF) {
            // the image (one view's, the same for the whole block) in scalar
            // registers, as a buffer descriptor must be
            const uint64_t ip = (uint64_t)(uintptr_t)img;
            const uint64_t up = ((uint64_t)(unsigned)uniformi((int)(ip >> 32)) << 32) | (unsigned)uniformi((int)ip);
            const auto rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)up, (short)0, uniformi(img_bytes),
                                                                kBufferWord3);
            unsigned voff = ((unsigned)(f.iv_lo + r0) * (unsigned)U + (unsigned)(f.iu_lo + c)) * 4u;
            // opaque to the optimiser, so that pass k adds the uniform k * step
            // (one add) instead of recomputing (row + k R) * U (a quarter-rate multiply)
            asm volatile("" : "+v"(voff));
            const unsigned step = (unsigned)f.R * (unsigned)U * 4u;
#pragma unroll
            for (int k = 0; k < NP; k++)
                if (k * f.R < f.H)
                    pre[k] =
                        __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff + k * step, 0, 0));
        } else {
            const auto g = to_global(img);
            const bool act = r0 < f.R;
            const auto src = g + (int64_t)(f.iv_lo + r0) * U + f.iu_lo + c;
            const auto safe = g + (int64_t)f.iv_lo * U + f.iu_lo;  // inside the image
            const int64_t step = (int64_t)f.R * U;
#pragma unroll
            for (int k = 0; k < NP; k++)
                if (k * f.R < f.H) pre[k] = *((act && r0 + k * f.R < f.H) ? src + k * step : safe);
        }
    }
    __device__ __forceinline__ void store(lds_float* buf, const Foot& f) const {
        if (r0 >= f.R) return;
        const int rs = f.R * f.Ws;
#pragma unroll
        for (int k = 0; k < NP; k++) {
            if ((k + 1) * f.R <= f.H)
                buf[lds0 + k * rs] = pre[k];
            else if (k * f.R < f.H && r0 + k * f.R < f.H)
                buf[lds0 + k * rs] = pre[k];
        }
    }
};

// Fused update of a backprojection tile: every voxel of the lane's columns is
// read and written once (raw accumulate, MLEM ratio, ART step or FDK sum).
// Invariant relied on by the solvers: the MLEM / ART tiles of a launch cover
// and write every voxel of the volume, so the maximum of the values written
// here is the volume's max |voxel| exactly, and the next forward takes its
// fixed-point scale from it.  A tile skipped for any reason (say, outside the
// field of view) would leave that bound too small.  (Layers z >= nz are not
// visited; a plan sets nz < N only where those layers are held at zero.)
template <int KIND, int CPW, int ZPL, int KA>
__device__ __forceinline__ float backward_update(const BwdArgs& a, const float (&tmp)[CPW * CPW][ZPL][KA],
                                                 const float (&fac)[CPW * CPW][ZPL][KA], const int (&nwhole)[CPW * CPW],
                                                 int xb, int yb, int z0, int lane) {
    // returns max |updated voxel| of this lane (MLEM / ART kinds)
    constexpr bool XTT = IsXtt<KIND>::value;
    const Dims& d = a.d;
    const auto vol = to_global(a.vol);
    float vmx = 0.f;
#pragma unroll
    for (int j = 0; j < CPW; j++) {
#pragma unroll
        for (int i = 0; i < CPW; i++) {
            const int x = xb + i, y = yb + j;
            if (x >= d.N || y >= d.N) continue;
            const int cidx = j * CPW + i;
#pragma unroll
            for (int q = 0; q < ZPL; q++) {
                const int z = z0 + lane + 64 * q;
                if (z >= d.N) continue;
                const int64_t idx = ((int64_t)y * d.N + x) * d.N + z;
#pragma unroll
                for (int k = 0; k < KA; k++) {
                    if (XTT && k >= d.K) continue;
                    const int64_t vi = idx + (int64_t)k * a.chan_stride;
                    const float t = tmp[cidx][q][k], fc = fac[cidx][q][k] + (float)nwhole[cidx];
                    if (KIND == BWD_RAW || KIND == BWD_XTT_RAW) {
                        to_global(a.tmp)[vi] += t;
                        to_global(a.factor)[vi] += fc;
                    } else if (KIND == BWD_MLEM || KIND == BWD_XTT_MLEM) {
                        const float mu = vol[vi];
                        const float nv = (fc == 0.0f) ? 0.0f : mu * t / fc;
                        vol[vi] = nv;
                        vmx = fmaxf(vmx, fabsf(nv));
                    } else if (KIND == BWD_ART || KIND == BWD_XTT_ART) {
                        const float nv = vol[vi] + a.alpha * t;
                        vol[vi] = nv;
                        vmx = fmaxf(vmx, fabsf(nv));
                    } else {  // BWD_FDK
                        vol[vi] = a.overwrite ? t : vol[vi] + t;
                    }
                }
            }
        }
    }
    return vmx;
}

// ---------------------------------------------------------------------------
// Untilted backprojection of the plain kinds (raw, MLEM, ART, FDK).  The
// per-view quantities of a wave are the same in all its lanes: the block
// footprint (8 projected corners) and, for each of the wave's 4 columns, the
// column's u split, the affine t(z) = v - 0.5 and the acceptance window.  They
// are computed lane-parallel, 16 views at a time, into LDS: the footprints of
// a batch by wave 0 (lane = view x corner pair, published by the per-view
// barrier two views before the batch starts), the column terms by every wave
// for its own columns (lane = view x column).  The view loop reads them back
// with broadcast LDS loads.  Arithmetic and accumulation order are those of
// the general path, so results are bitwise the same.
// ---------------------------------------------------------------------------
constexpr int kBwdBatch = 16;    // views per batch of per-view terms
constexpr int kFootWords = 8;    // iu_lo, iv_lo, W, H, Ws, R, staged | empty << 1, -
constexpr int kTermWords = 8;    // Bc, Ac, fu, hi (+inf: whole z extent accepted), wcol, lu, iu, base
constexpr int kTermCols = 4;     // columns per wave (CPW = 2)
constexpr int kBwdFootArea = 2 * kBwdBatch * kFootWords;          // words, block-wide, two batches
constexpr int kBwdTermArea = kBwdBatch * kTermCols * kTermWords;  // words per wave
template <int KIND>
constexpr int bwd_extra_words() {
    return kBwdFootArea + BwdBlock<KIND>::waves * kBwdTermArea;
}
static_assert(64 % kTermCols == 0 && 64 / kTermCols >= kBwdBatch, "one lane per (view, column) of a batch");

typedef __attribute__((address_space(3))) int lds_int;

template <int KIND, int ZPL>
__device__ __forceinline__ float backward_item_untilted(const BwdArgs& a, int o, int step, float* stage) {
    constexpr int CPW = 2, NC = CPW * CPW;
    static_assert(NC == kTermCols, "term layout");
    const Dims& d = a.d;
    const auto views = to_global(a.views + (int64_t)step * a.n_views);
    const auto vg = to_global(a.vg);
    const int gx = a.gx;
    const int bcx = CPW * gx, bcy = CPW * a.gy;  // block columns along x, y
    const int tx = o % a.tiles_x, ty = (o / a.tiles_x) % a.tiles_y, tz = o / (a.tiles_x * a.tiles_y);
    // the wave index as a wave-uniform value: the column tests and counts below stay scalar
    const int lane = threadIdx.x & 63, wave = uniformi(threadIdx.x >> 6);
    const int xb = tx * bcx + (wave % gx) * CPW, yb = ty * bcy + (wave / gx) * CPW;
    const int z0 = tz * 64 * ZPL;
    const float bx0 = (float)(tx * bcx) - d.xc, bx1 = (float)min(tx * bcx + bcx - 1, d.N - 1) - d.xc;
    const float by0 = (float)(ty * bcy) - d.xc, by1 = (float)min(ty * bcy + bcy - 1, d.N - 1) - d.xc;
    const float bz0 = (float)z0 - d.xc, bz1 = (float)min(z0 + 64 * ZPL - 1, d.N - 1) - d.xc;
    const int nv = a.n_views;
    if (nv == 0) return 0.f;

    float tmp[NC][ZPL][1], fac[NC][ZPL][1];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < ZPL; q++) {
            tmp[c][q][0] = 0.f;
            fac[c][q][0] = 0.f;
        }
    int nwhole[NC];  // views accepting a column's whole z extent (added to the counts at the end)
#pragma unroll
    for (int c = 0; c < NC; c++) nwhole[c] = 0;
    float zpq[ZPL];
#pragma unroll
    for (int q = 0; q < ZPL; q++) zpq[q] = (float)(z0 + lane + 64 * q) - d.xc;

    lds_float* const lds = (lds_float*)stage;
    lds_int* const foot = (lds_int*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf);
    lds_float* const term =
        (lds_float*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf + kBwdFootArea + wave * kBwdTermArea);

    // footprints of views b0 .. b0+15 into buffer buf: lane = 4 x view + corner
    // pair (x, y corner; both z corners), extremes over the 4 lanes of a view
    auto foot_batch = [&](int b0, int buf) {
        const int j = lane >> 2, c = lane & 3;
        const ViewGeom g = vg[views[min(b0 + j, nv - 1)]];
        const ColForms cf = col_forms(g, (c & 1) ? bx1 : bx0, (c & 2) ? by1 : by0);
        float umn = 0.f, umx = 0.f, vmn = 0.f, vmx = 0.f;
        int bad = 0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float czp = k ? bz1 : bz0;
            const float cd = fmaf(g.ad[2], czp, cf.nd);
            float cu, cv;
            voxel_uv(g, cf, czp, d, cu, cv);
            umn = k ? fminf(umn, cu) : cu;
            umx = k ? fmaxf(umx, cu) : cu;
            vmn = k ? fminf(vmn, cv) : cv;
            vmx = k ? fmaxf(vmx, cv) : cv;
            bad |= (!(cd > 0.f) || !(fabsf(cu) < 1e7f) || !(fabsf(cv) < 1e7f)) ? 1 : 0;
        }
        for (int off = 1; off < 4; off <<= 1) {
            umn = fminf(umn, __shfl_xor(umn, off));
            umx = fmaxf(umx, __shfl_xor(umx, off));
            vmn = fminf(vmn, __shfl_xor(vmn, off));
            vmx = fmaxf(vmx, __shfl_xor(vmx, off));
            bad |= __shfl_xor(bad, off);
        }
        const Foot f = foot_from_extent<KIND>(umn, umx, vmn, vmx, bad, d);
        if (c == 0) {
            lds_int* p = foot + (buf * kBwdBatch + j) * kFootWords;
            p[0] = f.iu_lo;
            p[1] = f.iv_lo;
            p[2] = f.W;
            p[3] = f.H;
            p[4] = f.Ws;
            p[5] = f.R;
            p[6] = (f.staged ? 1 : 0) | (f.empty ? 2 : 0);
        }
    };
    auto load_foot = [&](int buf, int j) {
        const lds_int* p = foot + (buf * kBwdBatch + j) * kFootWords;
        Foot f;
        f.iu_lo = uniformi(p[0]);
        f.iv_lo = uniformi(p[1]);
        f.W = uniformi(p[2]);
        f.H = uniformi(p[3]);
        f.Ws = uniformi(p[4]);
        f.R = uniformi(p[5]);
        const int fl = uniformi(p[6]);
        f.staged = (fl & 1) != 0;
        f.empty = (fl & 2) != 0;
        return f;
    };
    // column terms of views b0 .. b0+15 for this wave's columns: lane = 4 x view + column
    auto term_batch = [&](int b0, int buf) {
        const int j = lane >> 2, cidx = lane & 3;
        const ViewGeom g = vg[views[min(b0 + j, nv - 1)]];
        const lds_int* fp = foot + (buf * kBwdBatch + j) * kFootWords;
        const int iu_lo = fp[0], iv_lo = fp[1], W = fp[2], Ws = fp[4], fl = fp[6];
        const float xp = (float)(xb + (cidx & 1)) - d.xc, yp = (float)(yb + (cidx >> 1)) - d.xc;
        const ColForms c = col_forms(g, xp, yp);
        const float inv_c = rcpf(c.nd);
        const float u_c = fmaf(c.nu, inv_c, d.hu);
        int iu;
        float fu;
        split(u_c, iu, fu);
        const float Bc = g.av[2] * inv_c;
        const float Ac = fmaf(c.nv, inv_c, d.hv - 0.5f);
        const float hi = accept_u(u_c, d) ? d.vmax - 0.5f : -1.0f;
        const int lu = min(max(iu - iu_lo, 0), W - 2);
        float wcol = 0.f;
        if (KIND == BWD_FDK) {
            // vol += Ud^2 * (2 pi/P) * sum c p, Ud = sod/(sod + y_unrotated) (fdk.cu:160-169)
            const float Xy = fmaf(d.vs, yp, g.fy);
            const float Ud = d.sod / (d.sod + Xy);
            wcol = Ud * Ud * d.cfdk;
        }
        // whole extent accepted: t at the tile's first and last z bound every
        // lane's t (fmaf is monotone in z)
        const float ta = fmaf(Bc, bz0, Ac);
        const float tb = fmaf(Bc, (float)(z0 + 64 * ZPL - 1) - d.xc, Ac);
        const bool whole = (fl & 1) && fminf(ta, tb) > 0.05f && fmaxf(ta, tb) < hi;
        lds_float* tp = term + (j * kTermCols + cidx) * kTermWords;
        lds_int* ti = (lds_int*)tp;
        tp[0] = Bc;
        tp[1] = Ac;
        tp[2] = fu;
        tp[3] = whole ? __int_as_float(0x7f800000) : hi;
        tp[4] = wcol;
        ti[5] = lu;
        ti[6] = iu;
        ti[7] = (lu - iv_lo * Ws) * 4;  // LDS byte offset of texel (iu, 0) in a staging buffer
    };
    // a wave's own LDS writes, then its lanes read each other's words
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // staging of one view's footprint: prefetch into registers, store later
    FootStage<BwdBlock<KIND>::stage_per, true> st;
    const int img_bytes = d.V * d.U * 4;
    auto fetch = [&](const Foot& f, int code) {
        st.fetch(a.proj + (int64_t)(code - a.code_base) * d.V * d.U, img_bytes, f, d.U);
    };

    // View v is staged in buffer v % 4 two views ahead (fetched into registers
    // during view v-2, stored after it) and the block meets a barrier after
    // every odd view.  A store into buffer v % 4 then follows, across at least
    // one barrier, the last reads of view v-4 there, and precedes the reads of
    // view v across one; footprint batches are written four views ahead.
    static_assert(BwdBlock<KIND>::nbuf == 4, "four staging buffers");
    constexpr int SB = BwdBlock<KIND>::stage_buf;
    if (wave == 0) foot_batch(0, 0);
    __syncthreads();
    term_batch(0, 0);
    int code = views[0], code1 = 0;
    Foot f = load_foot(0, 0), f1{0, 0, 0, 0, 0, 0, false, false};
    if (f.staged) {
        fetch(f, code);
        st.store(lds, f);
    }
    if (nv > 1) {
        code1 = views[1];
        f1 = load_foot(0, 1);
        if (f1.staged) {
            fetch(f1, code1);
            st.store(lds + SB, f1);
        }
    }
    __syncthreads();

    for (int s = 0; s < nv; s++) {
        const int j = s & (kBwdBatch - 1);
        if (j == 0 && s > 0) term_batch(s, (s / kBwdBatch) & 1);
        wave_sync();
        // the footprints of the batch after this one: read from view s + 2 on
        if (wave == 0 && j == kBwdBatch - 4 && s + 4 < nv) foot_batch(s + 4, ((s + 4) / kBwdBatch) & 1);
        lds_float* cur = lds + (s & 3) * SB;
        lds_float* nxt = lds + ((s + 2) & 3) * SB;
        int ncode = 0;
        Foot nf{0, 0, 0, 0, 0, 0, false, false};
        if (s + 2 < nv) {
            ncode = views[s + 2];
            nf = load_foot(((s + 2) / kBwdBatch) & 1, (s + 2) & (kBwdBatch - 1));
            if (nf.staged) fetch(nf, ncode);
        }
        if (!f.empty) {
            const auto img = to_global(a.proj + (int64_t)(code - a.code_base) * d.V * d.U);
#pragma unroll
            for (int cidx = 0; cidx < NC; cidx++) {
                const int x = xb + (cidx & 1), y = yb + (cidx >> 1);
                if (x >= d.N || y >= d.N) continue;
                const lds_float* tp = term + (j * kTermCols + cidx) * kTermWords;
                const lds_int* ti = (const lds_int*)tp;
                const float Bc = tp[0], Ac = tp[1], fu = tp[2], hi = tp[3], wcol = tp[4];
                const int lu = ti[5], iu = ti[6], boff = ti[7];
                const float a0 = 1.0f - fu;
                // the column's samples; both branches leave them in nbq and the
                // accumulation after them is shared, so the accumulators are
                // updated in place (no copies where the branches meet)
                float nbq[ZPL];
                if (uniformi(__float_as_int(hi)) == 0x7f800000) {
                    // all reads of the column first (LDS latency), then the sums
                    const unsigned ws4 = (unsigned)f.Ws * 4u;
                    const unsigned base = (unsigned)(uintptr_t)cur + (unsigned)boff;
                    float fv[ZPL], r[ZPL][4];
#pragma unroll
                    for (int q = 0; q < ZPL; q++) {
                        const float t = fmaf(Bc, zpq[q], Ac);
                        fv[q] = __builtin_amdgcn_fractf(t);
                        const unsigned at = __umul24((unsigned)floor_int(t), ws4) + base;
                        const lds_float* p0 = (const lds_float*)(uintptr_t)at;
                        const lds_float* p1 = (const lds_float*)(uintptr_t)(at + ws4);
                        r[q][0] = p0[0];
                        r[q][1] = p0[1];
                        r[q][2] = p1[0];
                        r[q][3] = p1[1];
                    }
#pragma unroll
                    for (int q = 0; q < ZPL; q++) {
                        const float top = fmaf(fu, r[q][1], a0 * r[q][0]);
                        const float bot = fmaf(fu, r[q][3], a0 * r[q][2]);
                        nbq[q] = fmaf(fv[q], bot - top, top);
                    }
                    if (KIND != BWD_FDK) nwhole[cidx]++;
                } else {
                    // Window test per voxel: 0.55 < v < V - 0.55 becomes 0.05 < t < hi
                    // (hi < 0 rejects the column).  Branch free: a rejected voxel reads
                    // an in-bounds texel and adds zero.  Lanes with z >= N accumulate
                    // but are never stored.
#pragma unroll
                    for (int q = 0; q < ZPL; q++) {
                        const float t = fmaf(Bc, zpq[q], Ac);
                        const bool ok = t > 0.05f && t < hi;
                        const float fl = floorf(t);
                        const float fv = t - fl;
                        const int iv = (int)fl;
                        float r00, r01, r10, r11;
                        if (f.staged) {
                            const int qi = (ok ? iv - f.iv_lo : 0) * f.Ws + lu;
                            r00 = cur[qi];
                            r01 = cur[qi + 1];
                            r10 = cur[qi + f.Ws];
                            r11 = cur[qi + f.Ws + 1];
                        } else {
                            // direct reads: a distinct (nontemporal) load form next to the LDS reads
                            const int64_t qi = (int64_t)min(max(iv, 0), d.V - 2) * d.U + min(max(iu, 0), d.U - 2);
                            r00 = __builtin_nontemporal_load(img + qi);
                            r01 = __builtin_nontemporal_load(img + qi + 1);
                            r10 = __builtin_nontemporal_load(img + qi + d.U);
                            r11 = __builtin_nontemporal_load(img + qi + d.U + 1);
                        }
                        const float top = fmaf(fu, r01, a0 * r00);
                        const float bot = fmaf(fu, r11, a0 * r10);
                        nbq[q] = ok ? fmaf(fv, bot - top, top) : 0.0f;
                        if (KIND != BWD_FDK) fac[cidx][q][0] += ok ? 1.0f : 0.0f;
                    }
                }
#pragma unroll
                for (int q = 0; q < ZPL; q++) {
                    if (KIND == BWD_FDK)
                        tmp[cidx][q][0] = fmaf(wcol, nbq[q], tmp[cidx][q][0]);
                    else
                        tmp[cidx][q][0] += nbq[q];
                }
            }
        }
        if (s + 2 < nv && nf.staged) st.store(nxt, nf);
        code = code1;
        f = f1;
        code1 = ncode;
        f1 = nf;
        if (s & 1) __syncthreads();
    }
    return backward_update<KIND, CPW, ZPL, 1>(a, tmp, fac, nwhole, xb, yb, z0, lane);
}

// ---------------------------------------------------------------------------
// Untilted XTT backprojection.  A wave owns one column and ZPL voxels per lane
// in each of its K channels.  Per view and voxel the work is the bilinear
// gather, dz and 1/|d|^2, then per channel w_k = g2s_k - q_k^2/|d|^2 with
// q_k = dz sig_k S_k,z + ca_k (ir.cu:1832-1840) and two accumulations.
// Everything else is the same for all lanes of a wave and is computed
// lane-parallel 16 views at a time, as in the plain untilted path: the block
// footprints by wave 0 (lane = view x corner pair), the column terms and the
// K column constants ca_k = sig_k (d_x S_k,x + d_y S_k,y) by every wave for
// its own column (lane = view x channel quarter), all into LDS; the view loop
// reads them back with broadcast LDS loads.  The view's sig_k S_k,z and g2s_k
// are scalar loads of the plan's weight table.  Footprints are staged in
// LDS two views ahead in four buffers, with as many fetch passes as the
// footprint needs (at most kXttStagePer; a taller one is read from global
// memory directly).
// ---------------------------------------------------------------------------
constexpr int kXttStagePer = 4;

template <int KIND, int ZPL, int KM>
__device__ __forceinline__ float backward_item_untilted_xtt(const BwdArgs& a, int o, int step, float* stage) {
    constexpr int KP = (KM + 3) & ~3;  // channel constants per view, padded for 16-byte LDS loads
    constexpr int TW = 8 + KP;         // term words per view: Bc, Ac, fu, hi, cxy, lu, iu, boff, ca[KP]
    static_assert(kBwdBatch * TW <= kBwdTermArea, "term layout");
    const Dims& d = a.d;
    constexpr int xws = kXwStride;  // >= KM: whole rows, zero padded
    const auto views = to_global(a.views + (int64_t)step * a.n_views);
    const auto vg = to_global(a.vg);
    const auto vgc = to_const(a.vg);
    const int gx = a.gx;
    const int tx = o % a.tiles_x, ty = (o / a.tiles_x) % a.tiles_y, tz = o / (a.tiles_x * a.tiles_y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int xb = tx * gx + (wave % gx), yb = ty * a.gy + (wave / gx);
    const int z0 = tz * 64 * ZPL;
    const float bx0 = (float)(tx * gx) - d.xc, bx1 = (float)min(tx * gx + gx - 1, d.N - 1) - d.xc;
    const float by0 = (float)(ty * a.gy) - d.xc, by1 = (float)min(ty * a.gy + a.gy - 1, d.N - 1) - d.xc;
    const float bz0 = (float)z0 - d.xc, bz1 = (float)min(z0 + 64 * ZPL - 1, d.N - 1) - d.xc;
    const float xp = (float)xb - d.xc, yp = (float)yb - d.xc;
    const bool col_in = xb < d.N && yb < d.N;
    const int nv = a.n_views;
    if (nv == 0) return 0.f;

    float tmp[1][ZPL][KM], fac[1][ZPL][KM];
#pragma unroll
    for (int q = 0; q < ZPL; q++)
#pragma unroll
        for (int k = 0; k < KM; k++) {
            tmp[0][q][k] = 0.f;
            fac[0][q][k] = 0.f;
        }
    const int nwhole[1] = {0};
    float zpq[ZPL];
#pragma unroll
    for (int q = 0; q < ZPL; q++) zpq[q] = (float)(z0 + lane + 64 * q) - d.xc;

    lds_float* const lds = (lds_float*)stage;
    lds_int* const foot = (lds_int*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf);
    lds_float* const term =
        (lds_float*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf + kBwdFootArea + wave * kBwdTermArea);
    // staged only when the footprint fits the fetch passes of this path
    auto staged_here = [](int fl, int H, int R) { return (fl & 1) != 0 && H <= kXttStagePer * R; };

    auto foot_batch = [&](int b0, int buf) {
        const int j = lane >> 2, c = lane & 3;
        const ViewGeom g = vg[views[min(b0 + j, nv - 1)]];
        const ColForms cf = col_forms(g, (c & 1) ? bx1 : bx0, (c & 2) ? by1 : by0);
        float umn = 0.f, umx = 0.f, vmn = 0.f, vmx = 0.f;
        int bad = 0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float czp = k ? bz1 : bz0;
            const float cd = fmaf(g.ad[2], czp, cf.nd);
            float cu, cv;
            voxel_uv(g, cf, czp, d, cu, cv);
            umn = k ? fminf(umn, cu) : cu;
            umx = k ? fmaxf(umx, cu) : cu;
            vmn = k ? fminf(vmn, cv) : cv;
            vmx = k ? fmaxf(vmx, cv) : cv;
            bad |= (!(cd > 0.f) || !(fabsf(cu) < 1e7f) || !(fabsf(cv) < 1e7f)) ? 1 : 0;
        }
        for (int off = 1; off < 4; off <<= 1) {
            umn = fminf(umn, __shfl_xor(umn, off));
            umx = fmaxf(umx, __shfl_xor(umx, off));
            vmn = fminf(vmn, __shfl_xor(vmn, off));
            vmx = fmaxf(vmx, __shfl_xor(vmx, off));
            bad |= __shfl_xor(bad, off);
        }
        const Foot f = foot_from_extent<KIND>(umn, umx, vmn, vmx, bad, d);
        if (c == 0) {
            lds_int* p = foot + (buf * kBwdBatch + j) * kFootWords;
            p[0] = f.iu_lo;
            p[1] = f.iv_lo;
            p[2] = f.W;
            p[3] = f.H;
            p[4] = f.Ws;
            p[5] = f.R;
            p[6] = (f.staged ? 1 : 0) | (f.empty ? 2 : 0);
        }
    };
    auto load_foot = [&](int buf, int j) {
        const lds_int* p = foot + (buf * kBwdBatch + j) * kFootWords;
        Foot f;
        f.iu_lo = uniformi(p[0]);
        f.iv_lo = uniformi(p[1]);
        f.W = uniformi(p[2]);
        f.H = uniformi(p[3]);
        f.Ws = uniformi(p[4]);
        f.R = uniformi(p[5]);
        const int fl = uniformi(p[6]);
        f.staged = staged_here(fl, f.H, f.R);
        f.empty = (fl & 2) != 0;
        return f;
    };
    // column terms of views b0 .. b0+15: lane = 4 x view + channel quarter
    auto term_batch = [&](int b0, int buf) {
        const int j = lane >> 2, p = lane & 3;
        const int code = views[min(b0 + j, nv - 1)];
        const ViewGeom g = vg[code];
        const lds_int* fp = foot + (buf * kBwdBatch + j) * kFootWords;
        const int iu_lo = fp[0], iv_lo = fp[1], W = fp[2], H = fp[3], Ws = fp[4], R = fp[5], fl = fp[6];
        const ColForms c = col_forms(g, xp, yp);
        const float inv_c = rcpf(c.nd);
        const float u_c = fmaf(c.nu, inv_c, d.hu);
        int iu;
        float fu;
        split(u_c, iu, fu);
        const float Bc = g.av[2] * inv_c;
        const float Ac = fmaf(c.nv, inv_c, d.hv - 0.5f);
        const float hi = accept_u(u_c, d) ? d.vmax - 0.5f : -1.0f;
        const int lu = min(max(iu - iu_lo, 0), W - 2);
        const float dx = fmaf(d.vs, xp, g.d0[0]);
        const float dy = fmaf(d.vs, yp, g.d0[1]);
        const float ta = fmaf(Bc, bz0, Ac);
        const float tb = fmaf(Bc, (float)(z0 + 64 * ZPL - 1) - d.xc, Ac);
        const bool whole = staged_here(fl, H, R) && fminf(ta, tb) > 0.05f && fmaxf(ta, tb) < hi;
        lds_float* tp = term + j * TW;
        lds_int* ti = (lds_int*)tp;
        if (p == 0) {
            tp[0] = Bc;
            tp[1] = Ac;
            tp[2] = fu;
            tp[3] = whole ? __int_as_float(0x7f800000) : hi;
            tp[4] = fmaf(dy, dy, dx * dx);
            ti[5] = lu;
            ti[6] = iu;
            ti[7] = (lu - iv_lo * Ws) * 4;  // LDS byte offset of texel (iu, 0) in a staging buffer
        }
        const auto row = to_global(a.xw + (int64_t)code * 5 * xws);
#pragma unroll
        for (int m = 0; m < KP / 4; m++) {
            const int k = p + 4 * m;
            tp[8 + k] = fmaf(xp, row[2 * xws + k], fmaf(yp, row[3 * xws + k], row[4 * xws + k]));  // 0 for k >= K
        }
    };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    float pre[kXttStagePer];
    int pidx[kXttStagePer];
    auto fetch = [&](const Foot& f, int code) {
        const auto img = to_global(a.proj + (int64_t)(code - a.code_base) * d.V * d.U);
        int r0 = (int)((float)threadIdx.x * rcpf((float)f.W));  // within 1 of t / W: corrected below
        int c = (int)threadIdx.x - r0 * f.W;
        if (c < 0) {
            r0--;
            c += f.W;
        } else if (c >= f.W) {
            r0++;
            c -= f.W;
        }
        const bool act = r0 < f.R;
        const auto src = img + (int64_t)(f.iv_lo + r0) * d.U + f.iu_lo + c;
        const int64_t rstep = (int64_t)f.R * d.U;
#pragma unroll
        for (int k = 0; k < kXttStagePer; k++) {
            pidx[k] = -1;
            if (k * f.R < f.H) {  // uniform: only the passes the footprint needs
                const int r = r0 + k * f.R;
                const bool in = act && r < f.H;
                pidx[k] = in ? r * f.Ws + c : -1;
                pre[k] = in ? src[k * rstep] : 0.f;
            }
        }
    };
    auto store = [&](const Foot& f, lds_float* buf) {
#pragma unroll
        for (int k = 0; k < kXttStagePer; k++)
            if (k * f.R < f.H && pidx[k] >= 0) buf[pidx[k]] = pre[k];
    };

    // the staging schedule of the plain untilted path: view v in buffer v % 4,
    // fetched during view v-2, a barrier after every odd view, footprint
    // batches written four views ahead
    static_assert(BwdBlock<KIND>::nbuf == 4, "four staging buffers");
    constexpr int SB = BwdBlock<KIND>::stage_buf;
    if (wave == 0) foot_batch(0, 0);
    __syncthreads();
    term_batch(0, 0);
    int code = views[0], code1 = 0;
    Foot f = load_foot(0, 0), f1{0, 0, 0, 0, 0, 0, false, false};
    if (f.staged) {
        fetch(f, code);
        store(f, lds);
    }
    if (nv > 1) {
        code1 = views[1];
        f1 = load_foot(0, 1);
        if (f1.staged) {
            fetch(f1, code1);
            store(f1, lds + SB);
        }
    }
    __syncthreads();

    for (int s = 0; s < nv; s++) {
        const int j = s & (kBwdBatch - 1);
        if (j == 0 && s > 0) term_batch(s, (s / kBwdBatch) & 1);
        wave_sync();
        if (wave == 0 && j == kBwdBatch - 4 && s + 4 < nv) foot_batch(s + 4, ((s + 4) / kBwdBatch) & 1);
        lds_float* cur = lds + (s & 3) * SB;
        lds_float* nxt = lds + ((s + 2) & 3) * SB;
        int ncode = 0;
        Foot nf{0, 0, 0, 0, 0, 0, false, false};
        if (s + 2 < nv) {
            ncode = views[s + 2];
            nf = load_foot(((s + 2) / kBwdBatch) & 1, (s + 2) & (kBwdBatch - 1));
            if (nf.staged) fetch(nf, ncode);
        }
        if (!f.empty && col_in) {
            const lds_float* tp = term + j * TW;
            const lds_int* ti = (const lds_int*)tp;
            const float Bc = tp[0], Ac = tp[1], fu = tp[2], hi = tp[3], cxy = tp[4];
            const int lu = ti[5], iu = ti[6], boff = ti[7];
            const float a0 = 1.0f - fu;
            const float d0z = vgc[code].d0[2];
            const auto row = to_const(a.xw + (int64_t)code * 5 * xws);
            float ca[KM], sz[KM], g2s[KM];
#pragma unroll
            for (int k = 0; k < KM; k++) {
                ca[k] = tp[8 + k];
                sz[k] = row[k];
                g2s[k] = row[xws + k];
            }
            // channels k >= K have zero constants (table padding): they add zero and
            // are never stored
            auto accumulate = [&](int q, float nb) {
                const float dz = fmaf(d.vs, zpq[q], d0z);
                const float id2 = rcpf(fmaf(dz, dz, cxy));
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    const float qk = fmaf(dz, sz[k], ca[k]);
                    const float w = fmaf(-qk * qk, id2, g2s[k]);
                    tmp[0][q][k] = fmaf(w, nb, tmp[0][q][k]);
                    fac[0][q][k] += w;
                }
            };
            if (uniformi(__float_as_int(hi)) == 0x7f800000) {
                // the whole z extent accepted: all reads first (LDS latency), then the sums
                const unsigned ws4 = (unsigned)f.Ws * 4u;
                const unsigned base = (unsigned)(uintptr_t)cur + (unsigned)boff;
                float fv[ZPL], r[ZPL][4];
#pragma unroll
                for (int q = 0; q < ZPL; q++) {
                    const float t = fmaf(Bc, zpq[q], Ac);
                    fv[q] = __builtin_amdgcn_fractf(t);
                    const unsigned at = __umul24((unsigned)floor_int(t), ws4) + base;
                    const lds_float* p0 = (const lds_float*)(uintptr_t)at;
                    const lds_float* p1 = (const lds_float*)(uintptr_t)(at + ws4);
                    r[q][0] = p0[0];
                    r[q][1] = p0[1];
                    r[q][2] = p1[0];
                    r[q][3] = p1[1];
                }
#pragma unroll
                for (int q = 0; q < ZPL; q++) {
                    const float top = fmaf(fu, r[q][1], a0 * r[q][0]);
                    const float bot = fmaf(fu, r[q][3], a0 * r[q][2]);
                    accumulate(q, fmaf(fv[q], bot - top, top));
                }
            } else {
                // window test per voxel, 0.05 < t < hi (hi < 0 rejects the column); a
                // rejected voxel reads an in-bounds texel and skips the sums.  Lanes
                // with z >= N accumulate but are never stored.
                const auto img = to_global(a.proj + (int64_t)(code - a.code_base) * d.V * d.U);
#pragma unroll
                for (int q = 0; q < ZPL; q++) {
                    const float t = fmaf(Bc, zpq[q], Ac);
                    const bool ok = t > 0.05f && t < hi;
                    const float fl = floorf(t);
                    const float fv = t - fl;
                    const int iv = (int)fl;
                    float r00, r01, r10, r11;
                    if (f.staged) {
                        const int qi = (ok ? iv - f.iv_lo : 0) * f.Ws + lu;
                        r00 = cur[qi];
                        r01 = cur[qi + 1];
                        r10 = cur[qi + f.Ws];
                        r11 = cur[qi + f.Ws + 1];
                    } else {
                        const int64_t qi = (int64_t)min(max(iv, 0), d.V - 2) * d.U + min(max(iu, 0), d.U - 2);
                        r00 = __builtin_nontemporal_load(img + qi);
                        r01 = __builtin_nontemporal_load(img + qi + 1);
                        r10 = __builtin_nontemporal_load(img + qi + d.U);
                        r11 = __builtin_nontemporal_load(img + qi + d.U + 1);
                    }
                    const float top = fmaf(fu, r01, a0 * r00);
                    const float bot = fmaf(fu, r11, a0 * r10);
                    if (ok) accumulate(q, fmaf(fv, bot - top, top));
                }
            }
        }
        if (s + 2 < nv && nf.staged) store(nf, nxt);
        code = code1;
        f = f1;
        code1 = ncode;
        f1 = nf;
        if (s & 1) __syncthreads();
    }
    return backward_update<KIND, 1, ZPL, KM>(a, tmp, fac, nwhole, xb, yb, z0, lane);
}

// One voxel tile o of a backprojection launch; step selects the view list
// views + step*n_views (a solver's subset step).  stage: nbuf x stage_buf floats, then the footprint and term areas.
// Returns the lane's max |updated voxel| (MLEM / ART kinds).
template <int KIND, bool AXIS, int CPW, int ZPL, int KM>
__device__ __forceinline__ float backward_item(const BwdArgs& a, int o, int step, float* stage) {
    constexpr bool XTT = IsXtt<KIND>::value;
    constexpr int NC = CPW * CPW;
    constexpr int KA = XTT ? KM : 1;
    if constexpr (AXIS && !XTT && CPW == 2) {
        return backward_item_untilted<KIND, ZPL>(a, o, step, stage);
    }
    if constexpr (AXIS && XTT && CPW == 1) {
        return backward_item_untilted_xtt<KIND, ZPL, KM>(a, o, step, stage);
    }
    const Dims& d = a.d;
    const auto views = to_const(a.views + (int64_t)step * a.n_views);
    const auto vg = to_const(a.vg);
    const auto vol = to_global(a.vol);
    const int gx = a.gx;
    const int bcx = CPW * gx, bcy = CPW * a.gy;  // block columns along x, y
    const int tx = o % a.tiles_x, ty = (o / a.tiles_x) % a.tiles_y, tz = o / (a.tiles_x * a.tiles_y);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int xb = tx * bcx + (wave % gx) * CPW, yb = ty * bcy + (wave / gx) * CPW;
    const int z0 = tz * 64 * ZPL;
    const float bx0 = (float)(tx * bcx) - d.xc, bx1 = (float)min(tx * bcx + bcx - 1, d.N - 1) - d.xc;
    const float by0 = (float)(ty * bcy) - d.xc, by1 = (float)min(ty * bcy + bcy - 1, d.N - 1) - d.xc;
    const float bz0 = (float)z0 - d.xc, bz1 = (float)min(z0 + 64 * ZPL - 1, d.N - 1) - d.xc;

    float tmp[NC][ZPL][KA], fac[NC][ZPL][KA];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < ZPL; q++)
#pragma unroll
            for (int k = 0; k < KA; k++) {
                tmp[c][q][k] = 0.f;
                fac[c][q][k] = 0.f;
            }
    // untilted: views that accept the whole z extent of a column (counted once
    // per column, added to the per-voxel counts at the end; all integers, so
    // the sums are exact in any order)
    int nwhole[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) nwhole[c] = 0;

    // staging of one view's footprint: prefetch into registers, store later
    FootStage<BwdBlock<KIND>::stage_per, !XTT> st;
    const int img_bytes = d.V * d.U * 4;
    auto fetch = [&](const Foot& f, int code) {
        st.fetch(a.proj + (int64_t)(code - a.code_base) * d.V * d.U, img_bytes, f, d.U);
    };

    if (a.n_views == 0) return 0.f;
    // view-independent per-lane z terms
    float zpq[ZPL];
#pragma unroll
    for (int q = 0; q < ZPL; q++) zpq[q] = (float)(z0 + lane + 64 * q) - d.xc;
    int code = views[0];
    ViewGeom g = vg[code];
    Foot f = footprint<KIND>(g, d, bx0, bx1, by0, by1, bz0, bz1, lane);
    lds_float* const lds = (lds_float*)stage;
    if (f.staged) {
        fetch(f, code);
        st.store(lds, f);
    }
    __syncthreads();

    for (int s = 0; s < a.n_views; s++) {
        lds_float* cur = lds + (s & 1) * BwdBlock<KIND>::stage_buf;
        lds_float* nxt = lds + ((s + 1) & 1) * BwdBlock<KIND>::stage_buf;
        // prefetch the next view's footprint
        int ncode = 0;
        ViewGeom ng;
        Foot nf{0, 0, 0, 0, 0, 0, false, false};
        if (s + 1 < a.n_views) {
            ncode = views[s + 1];
            ng = vg[ncode];
            nf = footprint<KIND>(ng, d, bx0, bx1, by0, by1, bz0, bz1, lane);
            if (nf.staged) fetch(nf, ncode);
        }
        if (!f.empty) {
            const auto img = to_global(a.proj + (int64_t)(code - a.code_base) * d.V * d.U);
            float g2[KA];
            if (XTT && !AXIS) {
#pragma unroll
                for (int k = 0; k < KA; k++) g2[k] = (k < d.K) ? to_const(a.g2)[(int64_t)code * d.K + k] : 0.f;
            }
#pragma unroll
            for (int j = 0; j < CPW; j++) {
#pragma unroll
                for (int i = 0; i < CPW; i++) {
                    const int x = xb + i, y = yb + j;
                    if (x >= d.N || y >= d.N) continue;
                    const int cidx = j * CPW + i;
                    const float xp = (float)x - d.xc, yp = (float)y - d.xc;
                    const ColForms c = col_forms(g, xp, yp);
                    float inv_c = 0.f, u_c = 0.f;
                    if (AXIS) {
                        inv_c = rcpf(c.nd);
                        u_c = fmaf(c.nu, inv_c, d.hu);
                    }
                    float wcol = 0.f;
                    if (KIND == BWD_FDK) {
                        // vol += Ud^2 * (2 pi/P) * sum c p, Ud = sod/(sod + y_unrotated) (fdk.cu:160-169)
                        const float Xy = fmaf(d.vs, yp, g.fy);
                        const float Ud = d.sod / (d.sod + Xy);
                        wcol = Ud * Ud * d.cfdk;
                    }
                    float dx = 0.f, dy = 0.f;
                    if (XTT) {
                        dx = fmaf(d.vs, xp, g.d0[0]);
                        dy = fmaf(d.vs, yp, g.d0[1]);
                    }
                    if (AXIS && !XTT) {
                        // Untilted: u, hence (iu, fu), is constant along the column, the
                        // bilinear weights factor into (1-fu, fu) x (1-fv, fv), and
                        // t = v - 0.5 is affine in z.  The acceptance window 0.55 < u, v <
                        // U, V - 0.55 becomes 0.05 < t < hi (hi < 0 rejects the column).
                        // Branch free: a rejected voxel reads an in-bounds texel and adds
                        // zero.  Lanes with z >= N accumulate but are never stored.
                        int iu;
                        float fu;
                        split(u_c, iu, fu);
                        const float a0 = 1.0f - fu;
                        const float Bc = g.av[2] * inv_c;
                        const float Ac = fmaf(c.nv, inv_c, d.hv - 0.5f);
                        const float hi = accept_u(u_c, d) ? d.vmax - 0.5f : -1.0f;
                        const int lu = min(max(iu - f.iu_lo, 0), f.W - 2);
                        if (f.staged) {
                            // Whole extent accepted: t at the tile's first and last z bound
                            // every lane's t (fmaf is monotone in z), so one wave-uniform
                            // test replaces the per-voxel window test and the select.
                            const float ta = fmaf(Bc, (float)z0 - d.xc, Ac);
                            const float tb = fmaf(Bc, (float)(z0 + 64 * ZPL - 1) - d.xc, Ac);
                            if (uniformi(fminf(ta, tb) > 0.05f && fmaxf(ta, tb) < hi)) {
                                // LDS byte address of texel (iu, iv): iv*Ws*4 + base
                                const unsigned ws4 = (unsigned)f.Ws * 4u;
                                const unsigned base =
                                    (unsigned)(uintptr_t)cur + (unsigned)((lu - f.iv_lo * f.Ws) * 4);
                                // all reads of the column first (LDS latency), then the sums
                                float fv[ZPL], r[ZPL][4];
#pragma unroll
                                for (int q = 0; q < ZPL; q++) {
                                    const float t = fmaf(Bc, zpq[q], Ac);
                                    fv[q] = __builtin_amdgcn_fractf(t);
                                    const unsigned at = __umul24((unsigned)floor_int(t), ws4) + base;
                                    const lds_float* p0 = (const lds_float*)(uintptr_t)at;
                                    const lds_float* p1 = (const lds_float*)(uintptr_t)(at + ws4);
                                    r[q][0] = p0[0];
                                    r[q][1] = p0[1];
                                    r[q][2] = p1[0];
                                    r[q][3] = p1[1];
                                }
#pragma unroll
                                for (int q = 0; q < ZPL; q++) {
                                    const float top = fmaf(fu, r[q][1], a0 * r[q][0]);
                                    const float bot = fmaf(fu, r[q][3], a0 * r[q][2]);
                                    const float nb = fmaf(fv[q], bot - top, top);
                                    if (KIND == BWD_FDK)
                                        tmp[cidx][q][0] = fmaf(wcol, nb, tmp[cidx][q][0]);
                                    else
                                        tmp[cidx][q][0] += nb;
                                }
                                if (KIND != BWD_FDK) nwhole[cidx]++;
                                continue;
                            }
                        }
#pragma unroll
                        for (int q = 0; q < ZPL; q++) {
                            const float t = fmaf(Bc, zpq[q], Ac);
                            const bool ok = t > 0.05f && t < hi;
                            const float fl = floorf(t);
                            const float fv = t - fl;
                            const int iv = (int)fl;
                            float r00, r01, r10, r11;
                            if (f.staged) {
                                const int qi = (ok ? iv - f.iv_lo : 0) * f.Ws + lu;
                                r00 = cur[qi];
                                r01 = cur[qi + 1];
                                r10 = cur[qi + f.Ws];
                                r11 = cur[qi + f.Ws + 1];
                            } else {
                                // direct reads (see the general path below)
                                const int64_t qi = (int64_t)min(max(iv, 0), d.V - 2) * d.U + min(max(iu, 0), d.U - 2);
                                r00 = __builtin_nontemporal_load(img + qi);
                                r01 = __builtin_nontemporal_load(img + qi + 1);
                                r10 = __builtin_nontemporal_load(img + qi + d.U);
                                r11 = __builtin_nontemporal_load(img + qi + d.U + 1);
                            }
                            const float top = fmaf(fu, r01, a0 * r00);
                            const float bot = fmaf(fu, r11, a0 * r10);
                            const float nb = ok ? fmaf(fv, bot - top, top) : 0.0f;
                            if (KIND == BWD_FDK) {
                                tmp[cidx][q][0] = fmaf(wcol, nb, tmp[cidx][q][0]);
                            } else {
                                tmp[cidx][q][0] += nb;
                                fac[cidx][q][0] += ok ? 1.0f : 0.0f;
                            }
                        }
                        continue;
                    }
#pragma unroll
                    for (int q = 0; q < ZPL; q++) {
                        const int z = z0 + lane + 64 * q;
                        const float zp = (float)z - d.xc;
                        float u, v;
                        if (AXIS) {
                            u = u_c;
                            v = fmaf(fmaf(g.av[2], zp, c.nv), inv_c, d.hv);
                        } else {
                            voxel_uv(g, c, zp, d, u, v);
                        }
                        if (!(z < d.N && accept_u(u, d) && accept_v(v, d))) continue;
                        int iu, iv;
                        float fu, fv;
                        split(u, iu, fu);
                        split(v, iv, fv);
                        float r00, r01, r10, r11;  // (iu,iv) (iu+1,iv) (iu,iv+1) (iu+1,iv+1)
                        if (f.staged) {
                            const int lu = min(max(iu - f.iu_lo, 0), f.W - 2);
                            const int lv = min(max(iv - f.iv_lo, 0), f.H - 2);
                            const int qi = lv * f.Ws + lu;
                            r00 = cur[qi];
                            r01 = cur[qi + 1];
                            r10 = cur[qi + f.Ws];
                            r11 = cur[qi + f.Ws + 1];
                        } else {
                            // Footprint too large for LDS: read the image directly.  These
                            // loads must not be merged with the LDS reads above into one
                            // flat access (the gfx950 backend rejects that form), hence the
                            // distinct nontemporal load.
                            const int64_t qi = (int64_t)iv * d.U + iu;
                            r00 = __builtin_nontemporal_load(img + qi);
                            r01 = __builtin_nontemporal_load(img + qi + 1);
                            r10 = __builtin_nontemporal_load(img + qi + d.U);
                            r11 = __builtin_nontemporal_load(img + qi + d.U + 1);
                        }
                        const float c1 = (1.0f - fu) * fv, c2 = fu * fv;
                        const float c3 = fu * (1.0f - fv), c4 = (1.0f - fu) * (1.0f - fv);
                        const float nb = c1 * r10 + c2 * r11 + c3 * r01 + c4 * r00;
                        if (KIND == BWD_FDK) {
                            tmp[cidx][q][0] = fmaf(wcol, nb, tmp[cidx][q][0]);
                        } else if (!XTT) {
                            tmp[cidx][q][0] += nb;
                            fac[cidx][q][0] += 1.0f;
                        } else {
                            const float dz = fmaf(d.vs, zp, g.d0[2]);
                            const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                            const float id2 = rcpf(d2);
                            if (a.coef) {
                                // backwardOrth: the basis rotated per voxel
                                const int64_t vi = ((int64_t)y * d.N + x) * d.N + z;
                                float R[9];
                                orth_rotation(a.coef[vi], a.coef[vi + a.chan_stride], R);
#pragma unroll
                                for (int k = 0; k < KA; k++) {
                                    if (k < d.K) {
                                        const float w = orth_weight(R, a.b, k, dx, dy, dz, id2, g);
                                        tmp[cidx][q][k] = fmaf(w, nb, tmp[cidx][q][k]);
                                        fac[cidx][q][k] += w;
                                    }
                                }
                                continue;
                            }
#pragma unroll
                            for (int k = 0; k < KA; k++) {
                                if (k < d.K) {
                                    const float ds = fmaf(dz, a.b.S[k][2], fmaf(dy, a.b.S[k][1], dx * a.b.S[k][0]));
                                    const float w = fmaxf(fmaf(-ds * ds, id2, a.b.S2[k]), 0.f) * g2[k];
                                    tmp[cidx][q][k] = fmaf(w, nb, tmp[cidx][q][k]);
                                    fac[cidx][q][k] += w;
                                }
                            }
                        }
                    }
                }
            }
        }
        if (s + 1 < a.n_views) {
            if (nf.staged) st.store(nxt, nf);
            code = ncode;
            g = ng;
            f = nf;
        }
        __syncthreads();
    }
    return backward_update<KIND, CPW, ZPL, KA>(a, tmp, fac, nwhole, xb, yb, z0, lane);
}

template <int KIND, bool AXIS, int CPW, int ZPL, int KM>
// at least 4 waves per SIMD (at most 128 VGPRs): two 8-wave blocks, or one
// 16-wave block, per CU
__global__ __launch_bounds__(BwdBlock<KIND>::threads) __attribute__((amdgpu_waves_per_eu(4))) void k_backward(
    const BwdArgs a) {
    extern __shared__ float stage[];  // nbuf x stage_buf staging, footprint and term areas
    const int o = (blockIdx.x % 8) * a.per_xcd + blockIdx.x / 8;
    if (o >= a.n_tiles) return;
    float m = backward_item<KIND, AXIS, CPW, ZPL, KM>(a, o, 0, stage);
    if constexpr (KIND == BWD_MLEM || KIND == BWD_ART) {
        if (a.vmax_out) {
            // the tile's max |updated voxel|: one atomic per block
            lds_float* const wmax =
                (lds_float*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf + bwd_extra_words<KIND>());
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
            __syncthreads();
            if (threadIdx.x == 0) {
                float t = 0.f;
                for (int w = 0; w < BwdBlock<KIND>::waves; w++) t = fmaxf(t, wmax[w]);
                __hip_atomic_fetch_max(a.vmax_out, __float_as_uint(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Persistent backprojection over many problems (plain MLEM / ART kinds): the
// same work queue as k_forward_batch; items[i] is one problem with
// item_end[i] - item_end[i-1] tiles, and its own tile shape (zpl, gx).
template <int KIND, int ZPL, int CPW = 2, int KM = 1>
__global__ __launch_bounds__(BwdBlock<KIND>::threads) __attribute__((amdgpu_waves_per_eu(4))) void k_backward_batch(
    const BwdArgs* __restrict__ items,
                                                                const int* __restrict__ item_end, int n_items,
                                                                unsigned* queue, int step) {
    extern __shared__ float stage[];  // nbuf x stage_buf; the last word of buffer 1 holds the ticket
    typedef __attribute__((address_space(3))) unsigned lds_uint;
    lds_uint* const tk = (lds_uint*)stage + (2 * BwdBlock<KIND>::stage_buf - 1);
    // past the staging and term areas: the last tile's per-wave max |voxel| and its item
    lds_float* const wmax =
        (lds_float*)(stage + BwdBlock<KIND>::nbuf * BwdBlock<KIND>::stage_buf + bwd_extra_words<KIND>());
    lds_int* const witem = (lds_int*)(wmax + BwdBlock<KIND>::waves);
    const int total = item_end[n_items - 1];
    if (threadIdx.x == 0) *witem = -1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) {
            // the previous tile's max |updated voxel| -> its problem's next-step bound
            const int pi = *witem;
            unsigned* const dst = pi >= 0 ? items[pi].vmax_out : nullptr;
            if (dst) {
                float m = 0.f;
                for (int w = 0; w < BwdBlock<KIND>::waves; w++) m = fmaxf(m, wmax[w]);
                __hip_atomic_fetch_max(dst, __float_as_uint(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            *witem = -1;  // flushed: only a tile that sets it again sends a max
            *tk = take_ticket(queue, (unsigned)total);
        }
        __syncthreads();
        // uniform across the block: keeps the item's arguments in scalar registers
        const int t = __builtin_amdgcn_readfirstlane((int)*tk);
        if (t >= total) break;
        const int i = find_item(item_end, n_items, t);
        const BwdArgs& a = items[i];
        const int o = t - (i ? item_end[i - 1] : 0);
        float m = a.axis ? backward_item<KIND, true, CPW, ZPL, KM>(a, o, step, stage)
                         : backward_item<KIND, false, CPW, ZPL, KM>(a, o, step, stage);
        if (a.vmax_out) {
            for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
            if (threadIdx.x == 0) *witem = i;
        }
    }
}

// max |x| over blocks b, b + nb, ... of 256 threads: 16-byte loads, four in
// flight per thread; scalar loads for an unaligned base and for the tail.
__device__ __forceinline__ float absmax4(const float4 a) {
    return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
}

__device__ __forceinline__ float absmax_range(const float* __restrict__ x, int64_t n, int64_t b, int64_t nb) {
    float m = 0.f;
    const int64_t n4 = (reinterpret_cast<uintptr_t>(x) & 15) == 0 ? n >> 2 : 0;
    const float4* __restrict__ x4 = reinterpret_cast<const float4*>(x);
    const int64_t stride = nb * 256;
    int64_t i = b * 256 + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 a0 = x4[i], a1 = x4[i + stride], a2 = x4[i + 2 * stride], a3 = x4[i + 3 * stride];
        m = fmaxf(m, fmaxf(fmaxf(absmax4(a0), absmax4(a1)), fmaxf(absmax4(a2), absmax4(a3))));
    }
    for (; i < n4; i += stride) m = fmaxf(m, absmax4(x4[i]));
    for (int64_t j = 4 * n4 + b * 256 + threadIdx.x; j < n; j += stride) m = fmaxf(m, fabsf(x[j]));
    return m;
}

// Maximum over a block of 256 threads, valid in thread 0 (one atomic per
// block: same-address atomics from thousands of waves serialise).
__device__ __forceinline__ float block_max256(float m) {
    __shared__ float part[4];
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    return fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
}

// Per-segment max |x|: segment i is segs[i][0, seg_len[i]), its maximum goes
// to out[i] (zeroed beforehand; non-negative floats order like their bit
// patterns); grid (blocks per segment, segments).
__global__ __launch_bounds__(256) void k_absmax_seg(const float* const* __restrict__ segs,
                                                    const int64_t* __restrict__ seg_len, float* out) {
    const float m = block_max256(absmax_range(segs[blockIdx.y], seg_len[blockIdx.y], blockIdx.x, gridDim.x));
    if (threadIdx.x == 0) atomicMax((unsigned int*)(out + blockIdx.y), __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// FDK row filter (projConv, fdk.cu:28-43): dst[u] = sum_j h(u - j) w[j] src[j]
// for every row (view, v) of the launch, as one FP32 matrix product on the
// matrix cores: C[row][u] = sum_j A[row][j] T[j][u], A the weighted rows and
// T[j][u] = h(u - j) taken from a tap window in LDS (the Toeplitz operand is
// never stored).  h(m) = taps[|m|] (a linear convolution) or taps[m mod U]
// (circular), zero for |m| >= U.  Block tile 64 rows x 128 columns, four
// waves of 32 x 64, K steps of 32 samples.  v_mfma_f32_16x16x4f32 operands:
// A from lane (row lane&15, k lane>>4), B from lane (k lane>>4, column
// lane&15); the result lane holds column lane&15, rows 4*(lane>>4) + r.
// ---------------------------------------------------------------------------
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int kFiltBM = 64, kFiltBN = 128, kFiltBK = 32, kFiltAStride = kFiltBK + 1;
constexpr int kFiltWin = kFiltBN + kFiltBK - 1;  // taps one K step of one block touches

__global__ __launch_bounds__(256) void k_fdk_filter(const float* __restrict__ src, float* __restrict__ dst,
                                                   const float* __restrict__ weight, const float* __restrict__ taps,
                                                   int U, int V, const int32_t* __restrict__ views, int n_rows,
                                                   int row_base, int code_base, int circular) {
    __shared__ float As[kFiltBM * kFiltAStride];
    __shared__ float Hs[kFiltWin + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = row_base + blockIdx.y * kFiltBM, col0 = blockIdx.x * kFiltBN;
    // this thread loads row lr, samples lj .. lj+7 of each K step
    const int lr = tid >> 2, lj = (tid & 3) * 8;
    int64_t roff = -1;
    int wv = 0;
    if (row0 + lr < n_rows) {
        const int slot = (row0 + lr) / V;
        wv = row0 + lr - slot * V;
        roff = ((int64_t)(views[slot] - code_base) * V + wv) * U;
    }
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;  // the wave's 32 x 64 sub-tile
    const int li = lane & 15, lk = lane >> 4;
    f32x4_t acc[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int j0 = 0; j0 < U; j0 += kFiltBK) {
        __syncthreads();  // the previous K step is done with As and Hs
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int j = j0 + lj + q;
            float x = 0.f;
            if (roff >= 0 && j < U) {
                x = src[roff + j];
                if (weight) x *= weight[(int64_t)wv * U + j];
            }
            As[lr * kFiltAStride + lj + q] = x;
        }
        // Hs[t] = h(m), m = col0 - j0 - (BK - 1) + t
        if (tid < kFiltWin) {
            const int m = col0 - j0 - (kFiltBK - 1) + tid, am = abs(m);
            Hs[tid] = (am < U) ? taps[circular ? ((m % U) + U) % U : am] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kFiltBK; kk += 4) {
            float a[2], b[4];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) a[mt] = As[(wm + mt * 16 + li) * kFiltAStride + kk + lk];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) b[nt] = Hs[wn + nt * 16 + li - kk - lk + (kFiltBK - 1)];
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 4; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = row0 + wm + mt * 16 + 4 * lk + r;
            if (row >= n_rows) continue;
            const int slot = row / V, v = row - slot * V;
            const int64_t off = ((int64_t)(views[slot] - code_base) * V + v) * U;
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                const int col = col0 + wn + nt * 16 + li;
                if (col < U) dst[off + col] = acc[mt][nt][r];
            }
        }
}

// calcWeight (fdk.cu:13-26).
__global__ void k_fdk_weight(float* w, int U, int V, float sdd, float det) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x, v = blockIdx.y;
    if (u >= U) return;
    const float ur = ((float)u + 0.5f - (float)U / 2.0f) * det;
    const float vr = ((float)v + 0.5f - (float)V / 2.0f) * det;
    w[(int64_t)v * U + u] = sdd / sqrtf(sdd * sdd + ur * ur + vr * vr);
}

// Reference layout (x + N*y + N^2*z) <-> column layout (z + N*(x + N*y)):
// per y it is a transpose of the (z, x) plane; 32x32 tiles through LDS.
template <bool TO_INTERNAL>
__global__ __launch_bounds__(256) void k_transpose(const float* __restrict__ in, float* __restrict__ out, int N) {
    __shared__ float t[32][33];
    const int64_t N2 = (int64_t)N * N, N3 = N2 * N;
    const int y = blockIdx.z % N, ch = blockIdx.z / N;
    const float* src = in + ch * N3;
    float* dst = out + ch * N3;
    const int a0 = blockIdx.x * 32, b0 = blockIdx.y * 32;  // TO_INTERNAL: a = x, b = z
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int a = a0 + threadIdx.x, b = b0 + r;
        if (a < N && b < N) {
            // TO_INTERNAL: read ref[x=a, y, z=b]; else read internal[z=a, x=b, y]
            t[r][threadIdx.x] = TO_INTERNAL ? src[a + (int64_t)N * y + N2 * b] : src[a + (int64_t)N * b + N2 * y];
        }
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int b = b0 + threadIdx.x, a = a0 + r;
        if (a < N && b < N) {
            if (TO_INTERNAL)
                dst[b + (int64_t)N * a + N2 * y] = t[threadIdx.x][r];  // internal[z=b, x=a, y]
            else
                dst[b + (int64_t)N * y + N2 * a] = t[threadIdx.x][r];  // ref[x=b, y, z=a]
        }
    }
}

// ---- orientation helpers of the orientation-aware XTT loop ------------------
// Per voxel, channels nvox apart.  Normal of the three-direction scattering
// (ir.cu:912-922): (mu0 S0 - m1 S1) x (mu0 S0 - m2 S2), normalised (zero if
// its length is at most 1e-8, vec.h:94-104).
__device__ __forceinline__ void three_direc_normal(const Basis& b, float mu0, float m1, float m2, float n[3]) {
    float v1[3], v2[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        v1[i] = mu0 * b.S[0][i] - m1 * b.S[1][i];
        v2[i] = mu0 * b.S[0][i] - m2 * b.S[2][i];
    }
    n[0] = v1[1] * v2[2] - v1[2] * v2[1];
    n[1] = v1[2] * v2[0] - v1[0] * v2[2];
    n[2] = v1[0] * v2[1] - v1[1] * v2[0];
    const float len = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
#pragma unroll
    for (int i = 0; i < 3; i++) n[i] = len > 1e-8f ? n[i] / len : 0.f;
}

// calcNormalVectorThreeDirec (ir.cu:856-974; rot picks the signs of mu1, mu2)
// and calcNormalVectorThreeDirecWithEst (ir.cu:976-1071; the signs from est
// channel 1).  The normal is rotated by the voxel's current rotation, flipped
// into z >= 0 and normalised; loss = |R e_z x n|; coef <- (atan2(n_x, -n_y),
// n_z), the rotation that carries e_z onto n.
__global__ __launch_bounds__(256) void k_orth_normal(const float* __restrict__ vol, float* __restrict__ coef,
                                                     float* __restrict__ loss, int64_t n, int rot,
                                                     const float* __restrict__ est, Basis b) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float R[9];
    orth_rotation(coef[i], coef[i + n], R);
    const float mu0 = vol[i], mu1 = vol[i + n], mu2 = vol[i + 2 * n];
    int r = rot;
    if (est) {
        const float e = est[i + n];
        r = e < 0.5f ? 0 : (e < 1.5f ? 1 : (e < 2.5f ? 2 : 3));
    }
    const float m1 = (r == 1 || r == 2) ? -mu1 : mu1;
    const float m2 = (r == 2 || r == 3) ? -mu2 : mu2;
    float n0[3], nr[3];
    three_direc_normal(b, mu0, m1, m2, n0);
#pragma unroll
    for (int k = 0; k < 3; k++) nr[k] = R[3 * k] * n0[0] + R[3 * k + 1] * n0[1] + R[3 * k + 2] * n0[2];
    if (nr[2] < 0.f) {
        nr[0] = -nr[0];
        nr[1] = -nr[1];
        nr[2] = -nr[2];
    }
    const float len = sqrtf(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) nr[k] = len > 1e-8f ? nr[k] / len : 0.f;
    // (R e_z) x n
    const float dx = R[5] * nr[2] - R[8] * nr[1], dy = R[8] * nr[0] - R[2] * nr[2], dz = R[2] * nr[1] - R[5] * nr[0];
    loss[i] = sqrtf(dx * dx + dy * dy + dz * dz);
    coef[i] = atan2f(nr[0], -nr[1]);  // e_z x n = (-n_y, n_x, 0)
    coef[i + n] = nr[2];
}

// updateEstimationByCoef (ir.cu:1073-1118).
__global__ __launch_bounds__(256) void k_orth_update_est(const float* __restrict__ vol, float* __restrict__ est,
                                                         int64_t n, int iter, Basis b) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float nr[3];
    three_direc_normal(b, vol[i], vol[i + n], vol[i + 2 * n], nr);
    if (nr[2] < 0.f) {
        nr[0] = -nr[0];
        nr[1] = -nr[1];
        nr[2] = -nr[2];
    }
    const float len = sqrtf(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    const float c = len > 1e-8f ? nr[2] / len : 0.f;
    if (c > est[i]) {
        est[i] = c;
        est[i + n] = (float)iter;
    }
}

// convertNormVector (ir.cu:1448-1480): md = mean(mu) R e_z.
__global__ __launch_bounds__(256) void k_orth_md(const float* __restrict__ vol, const float* __restrict__ coef,
                                                 float* __restrict__ md, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float mu = (vol[i] + vol[i + n] + vol[i + 2 * n]) / 3.0f;
    float R[9];
    orth_rotation(coef[i], coef[i + n], R);
    md[i] = mu * R[2];
    md[i + n] = mu * R[5];
    md[i + 2 * n] = mu * R[8];
}

// meanFiltFiber (ir.cu:1231-1303) on a column-layout volume of edge N: the
// 27 neighbours' mu-weighted orientation vectors (sin t cos(phi - pi/2),
// sin t sin(phi - pi/2), cos t), those with |dot| > 0.5 against the centre
// added with weight w (sign aligned), normalised, flipped into z >= 0, stored
// as (atan2, cos).  Boundary voxels copy the source.
__global__ __launch_bounds__(256) void k_orth_mean_filter(const float* __restrict__ src, const float* __restrict__ vol,
                                                          int N, float w, float* __restrict__ dst) {
    const int64_t n = (int64_t)N * N * N;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % N), x = (int)((i / N) % N), y = (int)(i / ((int64_t)N * N));
    if (x < 1 || y < 1 || z < 1 || x > N - 2 || y > N - 2 || z > N - 2) {
        dst[i] = src[i];
        dst[i + n] = src[i + n];
        return;
    }
    auto vec = [&](int di, int dj, int dk, float v[3]) {
        const int64_t q = ((int64_t)(y - dj) * N + (x - di)) * N + (z - dk);
        const float ct = src[q + n];
        const float st = sqrtf(1.0f - ct * ct);
        float sp, cp;
        sincosf(-(float)M_PI / 2.0f + src[q], &sp, &cp);
        const float mu = vol[q] + vol[q + n] + vol[q + 2 * n];
        v[0] = mu * (st * cp);
        v[1] = mu * (st * sp);
        v[2] = mu * ct;
    };
    float c[3], m[3];
    vec(0, 0, 0, m);
    c[0] = m[0];
    c[1] = m[1];
    c[2] = m[2];
    // the reference's order: neighbour cnt and 26 - cnt for cnt = 0 .. 12
    for (int cnt = 0; cnt < 13; cnt++) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int q = side ? 26 - cnt : cnt;
            float o[3];
            vec(q / 9 - 1, (q / 3) % 3 - 1, q % 3 - 1, o);
            const float d = o[0] * m[0] + o[1] * m[1] + o[2] * m[2];
            const float sgn = d > 0.5f ? 1.f : (d < -0.5f ? -1.f : 0.f);
            if (sgn != 0.f) {
                c[0] += sgn * w * o[0];
                c[1] += sgn * w * o[1];
                c[2] += sgn * w * o[2];
            }
        }
    }
    const float len = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] = len > 1e-8f ? c[k] / len : 0.f;
    if (c[2] < 0.f) {
        c[0] = -c[0];
        c[1] = -c[1];
        c[2] = -c[2];
    }
    dst[i] = atan2f(c[0], -c[1]);
    dst[i + n] = c[2];
}

// Mean over f x f x f blocks of a column-layout volume of edge N = Nc * f
// (one thread per coarse voxel, z fastest: a wave reads 64 f consecutive
// floats of each fine column it visits).
__global__ __launch_bounds__(256) void k_vol_down(const float* __restrict__ fine, float* __restrict__ coarse, int Nc,
                                                  int f) {
    const int64_t n = (int64_t)Nc * Nc * Nc;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int zc = (int)(i % Nc), xc = (int)((i / Nc) % Nc), yc = (int)(i / ((int64_t)Nc * Nc));
    const int64_t N = (int64_t)Nc * f;
    float s = 0.f;
    for (int dy = 0; dy < f; dy++)
        for (int dx = 0; dx < f; dx++) {
            const float* col = fine + ((int64_t)(yc * f + dy) * N + (xc * f + dx)) * N + (int64_t)zc * f;
            for (int dz = 0; dz < f; dz++) s += col[dz];
        }
    coarse[i] = s / (float)(f * f * f);
}

// fine = max(fine + up(a - b), lo): the difference of two coarse column-layout
// volumes of edge Nc, interpolated trilinearly onto the fine grid of edge
// Nc * f (fine index i sits at coarse coordinate (i + 0.5) / f - 0.5, clamped
// to the coarse grid), added to the fine volume and floored at lo.
__global__ __launch_bounds__(256) void k_vol_up_add(const float* __restrict__ a, const float* __restrict__ b, int Nc,
                                                    int f, float* __restrict__ fine, float lo) {
    const int64_t N = (int64_t)Nc * f;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N * N * N) return;
    const int z = (int)(i % N), x = (int)((i / N) % N), y = (int)(i / (N * N));
    auto axis = [&](int j, int& j0, int& j1, float& w) {
        const float c = fmaxf(((float)j + 0.5f) / (float)f - 0.5f, 0.f);
        j0 = min((int)c, Nc - 1);
        j1 = min(j0 + 1, Nc - 1);
        w = c - (float)j0;
        if (j0 == Nc - 1) w = 0.f;
    };
    int z0, z1, x0, x1, y0, y1;
    float wz, wx, wy;
    axis(z, z0, z1, wz);
    axis(x, x0, x1, wx);
    axis(y, y0, y1, wy);
    auto at = [&](int yy, int xx, int zz) {
        const int64_t k = ((int64_t)yy * Nc + xx) * Nc + zz;
        return a[k] - b[k];
    };
    auto lerp = [](float u, float v, float w) { return fmaf(w, v - u, u); };
    const float c00 = lerp(at(y0, x0, z0), at(y0, x0, z1), wz), c01 = lerp(at(y0, x1, z0), at(y0, x1, z1), wz);
    const float c10 = lerp(at(y1, x0, z0), at(y1, x0, z1), wz), c11 = lerp(at(y1, x1, z0), at(y1, x1, z1), wz);
    const float d = lerp(lerp(c00, c01, wx), lerp(c10, c11, wx), wy);
    fine[i] = fmaxf(fine[i] + d, lo);
}

__global__ void k_sqrt_abs(float* v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = v[i];
    v[i] = (m < 0.0f) ? sqrtf(-m) : sqrtf(m);  // voxelSqrt (ir.cu:1715)
}

// calcAngleFromMD (pca.cu:67-83): azimuth phi and polar angle theta of the
// main direction, the direction flipped into y >= 0 first.  The squares and the
// sums are rounded separately (no contraction into FMAs), as the reference's
// host code evaluates them.
__global__ __launch_bounds__(256) void k_md_angles(const float* __restrict__ md, float* __restrict__ ang,
                                                   int64_t n) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = md[i], b = md[n + i], c = md[2 * n + i];
    const float r = sqrtf(b * b + a * a);
    if (b < 0.f) {
        ang[i] = atan2f(-b, -a);
        ang[n + i] = atan2f(r, -c);
    } else {
        ang[i] = atan2f(b, a);
        ang[n + i] = atan2f(r, c);
    }
}

// directionNormalization (pca.cu:85-106): unit main direction where its norm
// exceeds eps (0.65 in the reference), zero elsewhere.
__global__ __launch_bounds__(256) void k_md_normalise(float* __restrict__ md, int64_t n, float eps) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = md[i], y = md[n + i], z = md[2 * n + i];
    const float nrm = sqrtf(x * x + y * y + z * z);
    const bool keep = nrm > eps;
    md[i] = keep ? x / nrm : 0.f;
    md[n + i] = keep ? y / nrm : 0.f;
    md[2 * n + i] = keep ? z / nrm : 0.f;
}

// max |x| over n floats into out[0] (zeroed beforehand).
__global__ __launch_bounds__(256) void k_absmax(const float* __restrict__ x, int64_t n, float* out) {
    const float m = block_max256(absmax_range(x, n, blockIdx.x, gridDim.x));
    if (threadIdx.x == 0) atomicMax((unsigned int*)out, __float_as_uint(m));
}

__global__ void k_fill(float* p, int64_t n, float value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// value < floor -> replacement (main.cu:35 clamps the sinogram: <0 -> 1e-8).
__global__ void k_clamp_min(float* p, int64_t n, float floor_value, float replacement) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && p[i] < floor_value) p[i] = replacement;
}

// Per-voxel principal orientation (calcEigenVector, pca.cu:12-65):
// C = (1/K) sum_k (mu_k S_k)(mu_k S_k)^T, eigenvector of the smallest
// eigenvalue, sign so that z >= 0, md = sign * mean(mu) * e_min, eigenvalues
// normalised to sum 1.  Closed-form (trigonometric) eigenvalues in FP64 and the
// eigenvector from cross products of the rows of C - lambda_min I.
__global__ void k_pca(const float* __restrict__ mu, int K, const Basis b, int64_t nvox, int64_t cs,
                      float* __restrict__ md, float* __restrict__ ev) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox) return;
    double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0, mean = 0;
    for (int k = 0; k < K; k++) {
        const double m = mu[i + k * cs];
        mean += m;
        const double s0 = m * b.S[k][0], s1 = m * b.S[k][1], s2 = m * b.S[k][2];
        a00 += s0 * s0;
        a01 += s0 * s1;
        a02 += s0 * s2;
        a11 += s1 * s1;
        a12 += s1 * s2;
        a22 += s2 * s2;
    }
    const double invK = 1.0 / K;
    mean *= invK;
    a00 *= invK;
    a01 *= invK;
    a02 *= invK;
    a11 *= invK;
    a12 *= invK;
    a22 *= invK;
    double l0, l1, l2;  // ascending
    double e0 = 1, e1 = 0, e2 = 0;
    const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
    const double q = (a00 + a11 + a22) / 3.0;
    const double p2 = (a00 - q) * (a00 - q) + (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + 2.0 * p1;
    const double scale = fmax(fabs(a00), fmax(fabs(a11), fabs(a22))) + sqrt(p1);
    if (p1 <= 1e-30 * scale * scale) {
        // diagonal: sort the diagonal ascending, eigenvectors are the axes
        double dv[3] = {a00, a11, a22};
        int id[3] = {0, 1, 2};
        for (int s = 0; s < 2; s++)
            for (int t = 0; t < 2 - s; t++)
                if (dv[id[t + 1]] < dv[id[t]]) {
                    const int tt = id[t];
                    id[t] = id[t + 1];
                    id[t + 1] = tt;
                }
        l0 = dv[id[0]];
        l1 = dv[id[1]];
        l2 = dv[id[2]];
        e0 = id[0] == 0;
        e1 = id[0] == 1;
        e2 = id[0] == 2;
    } else {
        const double p = sqrt(p2 / 6.0);
        const double ip = 1.0 / p;
        const double b00 = (a00 - q) * ip, b11 = (a11 - q) * ip, b22 = (a22 - q) * ip;
        const double b01 = a01 * ip, b02 = a02 * ip, b12 = a12 * ip;
        double r = 0.5 * (b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02));
        r = fmin(1.0, fmax(-1.0, r));
        const double phi = acos(r) / 3.0;
        l2 = q + 2.0 * p * cos(phi);
        l0 = q + 2.0 * p * cos(phi + 2.0943951023931954923);
        l1 = 3.0 * q - l0 - l2;
        // rows of C - l0 I; the eigenvector is orthogonal to all of them
        const double r0[3] = {a00 - l0, a01, a02}, r1[3] = {a01, a11 - l0, a12}, r2[3] = {a02, a12, a22 - l0};
        const double c01[3] = {r0[1] * r1[2] - r0[2] * r1[1], r0[2] * r1[0] - r0[0] * r1[2], r0[0] * r1[1] - r0[1] * r1[0]};
        const double c02[3] = {r0[1] * r2[2] - r0[2] * r2[1], r0[2] * r2[0] - r0[0] * r2[2], r0[0] * r2[1] - r0[1] * r2[0]};
        const double c12[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
        const double n01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
        const double n02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
        const double n12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
        const double* best = c01;
        double nb = n01;
        if (n02 > nb) {
            best = c02;
            nb = n02;
        }
        if (n12 > nb) {
            best = c12;
            nb = n12;
        }
        if (nb > 1e-30 * scale * scale * scale * scale) {
            const double in = 1.0 / sqrt(nb);
            e0 = best[0] * in;
            e1 = best[1] * in;
            e2 = best[2] * in;
        } else {
            // lambda_min is (at least) double: any unit vector orthogonal to the
            // largest row of C - l0 I spans the eigenspace.
            const double n0 = r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2];
            const double n1 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
            const double n2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
            const double* rr = r0;
            if (n1 > n0 && n1 >= n2) rr = r1;
            if (n2 > n0 && n2 > n1) rr = r2;
            // cross with the axis least aligned with rr
            const double ax = fabs(rr[0]), ay = fabs(rr[1]), az = fabs(rr[2]);
            double t[3] = {0, 0, 0};
            if (ax <= ay && ax <= az)
                t[0] = 1;
            else if (ay <= az)
                t[1] = 1;
            else
                t[2] = 1;
            double w0 = rr[1] * t[2] - rr[2] * t[1], w1 = rr[2] * t[0] - rr[0] * t[2], w2 = rr[0] * t[1] - rr[1] * t[0];
            const double wn = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
            if (wn > 0) {
                e0 = w0 / wn;
                e1 = w1 / wn;
                e2 = w2 / wn;
            }
        }
    }
    const double sg = (e2 >= 0) ? 1.0 : -1.0;
    md[i] = (float)(sg * mean * e0);
    md[i + cs] = (float)(sg * mean * e1);
    md[i + 2 * cs] = (float)(sg * mean * e2);
    const float f0 = (float)l0, f1 = (float)l1, f2 = (float)l2;
    const float tot = f0 + f1 + f2;
    ev[i] = f0 / tot;
    ev[i + cs] = f1 / tot;
    ev[i + 2 * cs] = f2 / tot;
}

// Ellipsoid phantom in the reference layout, voxel x at (x-(N+1)/2)*vs.
// params: 16 doubles per ellipsoid (centre[3], semi-axes[3], rotation[9], value).
__global__ void k_ellipsoids(float* vol, int N, double vs, const double* __restrict__ prm, int count) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= N) return;
    const double p0 = ((double)x - (N + 1) * 0.5) * vs, p1 = ((double)y - (N + 1) * 0.5) * vs,
                 p2 = ((double)z - (N + 1) * 0.5) * vs;
    double acc = 0;
    for (int e = 0; e < count; e++) {
        const double* q = prm + 16 * e;
        const double d0 = p0 - q[0], d1 = p1 - q[1], d2 = p2 - q[2];
        double r2 = 0;
        for (int i = 0; i < 3; i++) {
            const double bb = q[6 + 3 * i] * d0 + q[7 + 3 * i] * d1 + q[8 + 3 * i] * d2;
            r2 += (bb / q[3 + i]) * (bb / q[3 + i]);
        }
        if (r2 <= 1.0) acc += q[15];
    }
    vol[x + (int64_t)N * y + (int64_t)N * N * z] = (float)acc;
}

// Hollow cylinder about the z axis (all z): value where r_out - thick < r <= r_out,
// voxel centres (x-(N+1)/2)*vs as for the ellipsoids.  Adds into vol (reference layout).
__global__ void k_cylinder_shell(float* vol, int N, double vs, double r_out, double thick, double value) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= N) return;
    const double px = ((double)x - (N + 1) * 0.5) * vs, py = ((double)y - (N + 1) * 0.5) * vs;
    const double r = sqrt(px * px + py * py);
    if (r <= r_out && r > r_out - thick) vol[x + (int64_t)N * y + (int64_t)N * N * z] += (float)value;
}

// Fibres: params 8 doubles each (centre[3] mm, unit direction f[3], radius mm,
// half-length mm).  Inside a fibre every channel k gains s_k = |S_k x f|^2
// (scattering strength 1 per unit direction-weight; values add).  Reference
// layout, K channels of N^3.
__global__ void k_fibres(float* vol, int N, double vs, const double* __restrict__ prm, int count, const Basis b,
                         int K) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= N) return;
    const double p0 = ((double)x - (N + 1) * 0.5) * vs, p1 = ((double)y - (N + 1) * 0.5) * vs,
                 p2 = ((double)z - (N + 1) * 0.5) * vs;
    float acc[kMaxBasis];
    for (int k = 0; k < kMaxBasis; k++) acc[k] = 0.f;
    for (int e = 0; e < count; e++) {
        const double* q = prm + 8 * e;
        const double d0 = p0 - q[0], d1 = p1 - q[1], d2 = p2 - q[2];
        const double t = d0 * q[3] + d1 * q[4] + d2 * q[5];
        if (fabs(t) > q[7]) continue;
        const double r0 = d0 - t * q[3], r1 = d1 - t * q[4], r2 = d2 - t * q[5];
        if (r0 * r0 + r1 * r1 + r2 * r2 > q[6] * q[6]) continue;
        for (int k = 0; k < K; k++) {
            const double c0 = b.S[k][1] * q[5] - b.S[k][2] * q[4];
            const double c1 = b.S[k][2] * q[3] - b.S[k][0] * q[5];
            const double c2 = b.S[k][0] * q[4] - b.S[k][1] * q[3];
            acc[k] += (float)(c0 * c0 + c1 * c1 + c2 * c2);
        }
    }
    const int64_t i = x + (int64_t)N * y + (int64_t)N * N * z, N3 = (int64_t)N * N * N;
    for (int k = 0; k < K; k++) vol[i + k * N3] += acc[k];
}

// Scalar fibres: `value` added inside every fibre (attenuation phantom).
__global__ void k_fibres_scalar(float* vol, int N, double vs, const double* __restrict__ prm, int count,
                                float value) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= N) return;
    const double p0 = ((double)x - (N + 1) * 0.5) * vs, p1 = ((double)y - (N + 1) * 0.5) * vs,
                 p2 = ((double)z - (N + 1) * 0.5) * vs;
    float acc = 0.f;
    for (int e = 0; e < count; e++) {
        const double* q = prm + 8 * e;
        const double d0 = p0 - q[0], d1 = p1 - q[1], d2 = p2 - q[2];
        const double t = d0 * q[3] + d1 * q[4] + d2 * q[5];
        if (fabs(t) > q[7]) continue;
        const double r0 = d0 - t * q[3], r1 = d1 - t * q[4], r2 = d2 - t * q[5];
        if (r0 * r0 + r1 * r1 + r2 * r2 > q[6] * q[6]) continue;
        acc += value;
    }
    vol[x + (int64_t)N * y + (int64_t)N * N * z] += acc;
}

__global__ void k_poisson_noise(float* y, const float* p, int64_t n, double I0, uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    UniformStream us(seed, (uint64_t)i);
    const double c = poisson_sample(I0 * exp(-(double)p[i]), us);
    y[i] = (float)(-log(fmax(c, 1.0) / I0));
}

__global__ void k_gaussian_noise(float* y, const float* x, int64_t n, double sigma, uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    UniformStream us(seed, (uint64_t)i);
    const double u1 = us.next(), u2 = us.next();
    const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);  // Box-Muller
    y[i] = (float)((double)x[i] * (1.0 + sigma * g));
}

inline unsigned grid1(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_forward(const FwdArgs& a, int kind, bool axis, hipStream_t stream) {
    const int nblk = a.n_bands * std::max(a.n_sub, 1) * a.n_slots * a.n_chunks;
    if (nblk == 0) return hipSuccess;
    const unsigned grid = (unsigned)(8 * ((nblk + 7) / 8));
    const size_t lds = forward_lds_bytes(kind, a.vb);
#define XR_FWD(KD, A) hipLaunchKernelGGL((k_forward<KD, A>), dim3(grid), dim3(64 * kFwdWaves), lds, stream, a)
    if (kind == FWD_WEIGHT) {
        if (axis) XR_FWD(FWD_WEIGHT, true); else XR_FWD(FWD_WEIGHT, false);
    } else if (kind == FWD_PLAIN) {
        if (axis) XR_FWD(FWD_PLAIN, true); else XR_FWD(FWD_PLAIN, false);
    } else {
        // basis count rounded up to a compiled size (registers of the untilted path)
#define XR_FWD_K(KM)                                                                                     \
    hipLaunchKernelGGL((k_forward<FWD_XTT, true, KM, kFwdXttWaves>), dim3(grid), dim3(64 * kFwdXttWaves), lds, \
                       stream, a)
        if (!axis)
            XR_FWD(FWD_XTT, false);
        else if (a.d.K <= 4)
            XR_FWD_K(4);
        else if (a.d.K <= 8)
            XR_FWD_K(8);
        else if (a.d.K <= 13)
            XR_FWD_K(13);
        else
            XR_FWD_K(16);
#undef XR_FWD_K
    }
#undef XR_FWD
    return hipGetLastError();
}

int default_zpl(int N) { return N > 128 ? 4 : (N > 64 ? 2 : 1); }

// staging buffers, per-view footprint and term areas, per-wave max slots (and
// the persistent launch's item slot)
template <int KD>
constexpr size_t backward_lds_bytes(bool batch) {
    return (BwdBlock<KD>::nbuf * BwdBlock<KD>::stage_buf + bwd_extra_words<KD>() + BwdBlock<KD>::waves +
            (batch ? 1 : 0)) *
           sizeof(float);
}

inline bool bwd_xtt(int kind) { return kind == BWD_XTT_RAW || kind == BWD_XTT_MLEM || kind == BWD_XTT_ART; }

int backward_threads(int kind) {
    return bwd_xtt(kind) ? BwdBlock<BWD_XTT_MLEM>::threads : BwdBlock<BWD_MLEM>::threads;
}

size_t backward_batch_lds_bytes(int kind) {
    return bwd_xtt(kind) ? backward_lds_bytes<BWD_XTT_MLEM>(true) : backward_lds_bytes<BWD_MLEM>(true);
}

bool backward_shape(BwdArgs& a, int zpl, int gx, bool xtt) {
    const int N = a.d.N;
    // Plain / FDK: 2x2 columns x zpl z per lane, 16 waves as gx x gy column
    // groups (a (2gx) x (2gy)-column tile, 64*zpl z).  XTT (2K accumulators per
    // voxel): one column x zpl (1 or 2) z per lane, 8 waves.
    const int CPW = xtt ? 1 : 2;
    if (xtt && zpl == 0) zpl = 1;
    if (xtt && zpl > 2) return false;
    if (zpl == 0) zpl = default_zpl(N);
    // default column layout: 8 x 8 columns (plain); 2 x 4 for XTT (measured
    // 144.6 against 146.2 ms for 4 x 2 at 256^3 x 13, 4 x 360 views)
    if (gx == 0) gx = xtt ? 2 : 4;
    if (!(zpl == 1 || zpl == 2 || zpl == 4)) return false;
    if (!(gx == 1 || gx == 2 || gx == 4 || gx == 8 || gx == 16)) return false;
    const int waves = xtt ? BwdBlock<BWD_XTT_MLEM>::waves : BwdBlock<BWD_MLEM>::waves;
    gx = std::min(gx, waves);  // XTT blocks have 8 waves: gx 16 is taken as 8
    a.zpl = zpl;
    a.gx = gx;
    a.gy = waves / gx;
    a.tiles_x = (N + CPW * a.gx - 1) / (CPW * a.gx);
    a.tiles_y = (N + CPW * a.gy - 1) / (CPW * a.gy);
    a.tiles_z = (a.d.nz + 64 * zpl - 1) / (64 * zpl);  // z >= nz: never visited
    a.n_tiles = a.tiles_x * a.tiles_y * a.tiles_z;
    a.per_xcd = (a.n_tiles + 7) / 8;
    return true;
}

hipError_t launch_backward(BwdArgs a, int kind, bool axis, hipStream_t stream) {
    const bool xtt = (kind == BWD_XTT_RAW || kind == BWD_XTT_MLEM || kind == BWD_XTT_ART);
    if (!backward_shape(a, a.zpl, a.gx, xtt)) return hipErrorInvalidValue;
    const int zpl = a.zpl;
    const unsigned grid = (unsigned)(8 * a.per_xcd);
#define XR_BWD(KD, A, C, Z, KM)                                                                          \
    hipLaunchKernelGGL((k_backward<KD, A, C, Z, KM>), dim3(grid), dim3(BwdBlock<KD>::threads),          \
                       backward_lds_bytes<KD>(false), stream, a)
#define XR_BWD_AX(KD, C, Z, KM)         \
    do {                                \
        if (axis)                       \
            XR_BWD(KD, true, C, Z, KM); \
        else                            \
            XR_BWD(KD, false, C, Z, KM); \
    } while (0)
#define XR_BWD_PLAIN(KD)             \
    do {                             \
        if (zpl == 4)                \
            XR_BWD_AX(KD, 2, 4, 1);  \
        else if (zpl == 2)           \
            XR_BWD_AX(KD, 2, 2, 1);  \
        else                         \
            XR_BWD_AX(KD, 2, 1, 1);  \
    } while (0)
#define XR_BWD_XTT_Z(KD, Z)              \
    do {                                 \
        if (a.d.K <= 4)                  \
            XR_BWD_AX(KD, 1, Z, 4);      \
        else if (a.d.K <= 8)             \
            XR_BWD_AX(KD, 1, Z, 8);      \
        else if (a.d.K <= 13)            \
            XR_BWD_AX(KD, 1, Z, 13);     \
        else                             \
            XR_BWD_AX(KD, 1, Z, 16);     \
    } while (0)
#define XR_BWD_XTT(KD)             \
    do {                           \
        if (zpl == 2)              \
            XR_BWD_XTT_Z(KD, 2);   \
        else                       \
            XR_BWD_XTT_Z(KD, 1);   \
    } while (0)
    switch (kind) {
        case BWD_RAW: XR_BWD_PLAIN(BWD_RAW); break;
        case BWD_MLEM: XR_BWD_PLAIN(BWD_MLEM); break;
        case BWD_ART: XR_BWD_PLAIN(BWD_ART); break;
        case BWD_FDK: XR_BWD_PLAIN(BWD_FDK); break;
        case BWD_XTT_RAW: XR_BWD_XTT(BWD_XTT_RAW); break;
        case BWD_XTT_MLEM: XR_BWD_XTT(BWD_XTT_MLEM); break;
        case BWD_XTT_ART: XR_BWD_XTT(BWD_XTT_ART); break;
        default: return hipErrorInvalidValue;
    }
#undef XR_BWD_XTT
#undef XR_BWD_XTT_Z
#undef XR_BWD_PLAIN
#undef XR_BWD_AX
#undef XR_BWD
    return hipGetLastError();
}

// Basis count rounded up to a compiled size of the XTT kernels.
inline int xtt_km(int kmax) { return kmax <= 4 ? 4 : (kmax <= 8 ? 8 : (kmax <= 13 ? 13 : 16)); }

const void* forward_batch_xtt_kernel(int kmax) {
    switch (xtt_km(kmax)) {
        case 4: return (const void*)k_forward_batch_xtt<4>;
        case 8: return (const void*)k_forward_batch_xtt<8>;
        case 13: return (const void*)k_forward_batch_xtt<13>;
        default: return (const void*)k_forward_batch_xtt<16>;
    }
}

hipError_t launch_forward_batch(const FwdArgs* items, const int* item_end, int n_items, unsigned* queue,
                                int ticket_slot, size_t lds_bytes, int grid, int step, int loss_shift, int kind,
                                bool axis_all, hipStream_t stream, int kmax) {
    if (n_items <= 0 || grid <= 0) return hipSuccess;
    if (kind == FWD_XTT) {
        void* args[] = {(void*)&items, (void*)&item_end, (void*)&n_items, (void*)&queue, (void*)&ticket_slot,
                        (void*)&step, (void*)&loss_shift};
        return hipLaunchKernel(forward_batch_xtt_kernel(kmax), dim3(grid), dim3(64 * kFwdXttWaves), args, lds_bytes,
                               stream);
    }
    if (kind != FWD_PLAIN) return hipErrorInvalidValue;
    if (axis_all)
        hipLaunchKernelGGL((k_forward_batch_axis<FWD_PLAIN>), dim3(grid), dim3(64 * kFwdWaves), lds_bytes, stream,
                           items, item_end, n_items, queue, ticket_slot, step, loss_shift);
    else
        hipLaunchKernelGGL((k_forward_batch<FWD_PLAIN>), dim3(grid), dim3(64 * kFwdBatchWaves), lds_bytes, stream, items,
                           item_end, n_items, queue, ticket_slot, step, loss_shift);
    return hipGetLastError();
}

// Batched backprojection kernel of (kind, zpl) (XTT kinds: zpl 1, one column
// per wave, kmax basis vectors at most), or nullptr.
const void* backward_batch_kernel(int kind, int zpl, int kmax) {
#define XR_BB(KD)                                                        \
    if (kind == KD) {                                                    \
        if (zpl == 4) return (const void*)k_backward_batch<KD, 4>;       \
        if (zpl == 2) return (const void*)k_backward_batch<KD, 2>;       \
        if (zpl == 1) return (const void*)k_backward_batch<KD, 1>;       \
    }
    XR_BB(BWD_MLEM)
    XR_BB(BWD_ART)
#undef XR_BB
#define XR_BBX(KD)                                                                \
    if (kind == KD && zpl == 1) {                                                 \
        switch (xtt_km(kmax)) {                                                   \
            case 4: return (const void*)k_backward_batch<KD, 1, 1, 4>;            \
            case 8: return (const void*)k_backward_batch<KD, 1, 1, 8>;            \
            case 13: return (const void*)k_backward_batch<KD, 1, 1, 13>;          \
            default: return (const void*)k_backward_batch<KD, 1, 1, 16>;          \
        }                                                                         \
    }
    XR_BBX(BWD_XTT_MLEM)
    XR_BBX(BWD_XTT_ART)
#undef XR_BBX
    return nullptr;
}

hipError_t launch_backward_batch(const BwdArgs* items, const int* item_end, int n_items, unsigned* queue, int grid,
                                 int step, int kind, int zpl, hipStream_t stream, int kmax) {
    if (n_items <= 0 || grid <= 0) return hipSuccess;
    const void* k = backward_batch_kernel(kind, zpl, kmax);
    if (!k) return hipErrorInvalidValue;
    void* args[] = {(void*)&items, (void*)&item_end, (void*)&n_items, (void*)&queue, (void*)&step};
    return hipLaunchKernel(k, dim3(grid), dim3(backward_threads(kind)), args, backward_batch_lds_bytes(kind), stream);
}

int batch_blocks_per_cu_forward(int kind, size_t lds_bytes, bool axis_all, int kmax) {
    int n = 0;
    if (kind == FWD_XTT) {
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, forward_batch_xtt_kernel(kmax),
                                                                          64 * kFwdXttWaves, lds_bytes);
        return e == hipSuccess ? n : 0;
    }
    if (kind != FWD_PLAIN) return 0;
    const hipError_t e =
        axis_all ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)k_forward_batch_axis<FWD_PLAIN>,
                                                                64 * kFwdWaves, lds_bytes)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)k_forward_batch<FWD_PLAIN>,
                                                                64 * kFwdBatchWaves, lds_bytes);
    return e == hipSuccess ? n : 0;
}

size_t forward_lds_bytes(int kind, int vb) {
    const int acc_bytes = (kind == FWD_WEIGHT) ? 8 : 4;
    return ((size_t)fwd_red_slot(vb + 2, acc_bytes) + 2 * kFwdWaves) * sizeof(unsigned long long);
}

int forward_blocks_per_cu(int kind, bool axis, size_t lds_bytes) {
    int n = 0;
    const void* k = nullptr;
    if (kind == FWD_WEIGHT) k = axis ? (const void*)k_forward<FWD_WEIGHT, true> : (const void*)k_forward<FWD_WEIGHT, false>;
    else if (kind == FWD_PLAIN) k = axis ? (const void*)k_forward<FWD_PLAIN, true> : (const void*)k_forward<FWD_PLAIN, false>;
    else k = axis ? (const void*)k_forward<FWD_XTT, true, 13, kFwdXttWaves> : (const void*)k_forward<FWD_XTT, false>;
    const int threads = (kind == FWD_XTT && axis) ? 64 * kFwdXttWaves : 64 * kFwdWaves;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, lds_bytes);
    return e == hipSuccess ? n : 0;
}

int batch_blocks_per_cu_backward(int kind, int zpl, int kmax) {
    int n = 0;
    const void* k = backward_batch_kernel(kind, zpl, kmax);
    if (!k) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, backward_threads(kind), backward_batch_lds_bytes(kind)) !=
        hipSuccess)
        return 0;
    return n;
}

int absmax_blocks(int64_t max_len, int n_seg) {
    // at least 64 floats per thread in the largest segment; about 2048 blocks in all
    const int64_t want = (max_len + 256 * 64 - 1) / (256 * 64);
    return (int)std::max<int64_t>(1, std::min<int64_t>(want, std::max(1, 2048 / std::max(n_seg, 1))));
}

hipError_t launch_absmax_seg(const float* const* segs, const int64_t* seg_len, float* out, int n_seg,
                             int blocks_per_seg, hipStream_t stream) {
    if (n_seg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_absmax_seg, dim3((unsigned)blocks_per_seg, (unsigned)n_seg), dim3(256), 0, stream, segs,
                       seg_len, out);
    return hipGetLastError();
}

hipError_t launch_fdk_filter(const float* src, float* dst, const float* weight, const float* taps, int U, int V,
                             const int32_t* views, int n_views, int code_base, int circular, hipStream_t stream) {
    const int64_t n_rows = (int64_t)n_views * V;
    if (n_rows > INT32_MAX - kFiltBM) return hipErrorInvalidValue;
    // row tiles in launches of at most 65535 tiles (the grid's y limit)
    const int64_t per_launch = (int64_t)65535 * kFiltBM;
    for (int64_t base = 0; base < n_rows; base += per_launch) {
        const int64_t rows = std::min(per_launch, n_rows - base);
        const dim3 grid((U + kFiltBN - 1) / kFiltBN, (unsigned)((rows + kFiltBM - 1) / kFiltBM));
        hipLaunchKernelGGL(k_fdk_filter, grid, dim3(256), 0, stream, src, dst, weight, taps, U, V, views, (int)n_rows,
                           (int)base, code_base, circular);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_fdk_weight(float* weight, int U, int V, float sdd, float det, hipStream_t stream) {
    hipLaunchKernelGGL(k_fdk_weight, dim3(grid1(U, 256), V), dim3(256), 0, stream, weight, U, V, sdd, det);
    return hipGetLastError();
}

hipError_t launch_to_internal(const float* ref, float* internal, int N, int channels, hipStream_t stream) {
    hipLaunchKernelGGL((k_transpose<true>), dim3(grid1(N, 32), grid1(N, 32), (unsigned)N * channels), dim3(32, 8), 0,
                       stream, ref, internal, N);
    return hipGetLastError();
}

hipError_t launch_from_internal(const float* internal, float* ref, int N, int channels, hipStream_t stream) {
    hipLaunchKernelGGL((k_transpose<false>), dim3(grid1(N, 32), grid1(N, 32), (unsigned)N * channels), dim3(32, 8), 0,
                       stream, internal, ref, N);
    return hipGetLastError();
}

hipError_t launch_orth_normal(const float* vol, float* coef, float* loss, int64_t n, int rot, const float* est,
                              const Basis& b, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_orth_normal, dim3(grid1(n, 256)), dim3(256), 0, stream, vol, coef, loss, n, rot, est, b);
    return hipGetLastError();
}

hipError_t launch_orth_update_est(const float* vol, float* est, int64_t n, int iter, const Basis& b,
                                  hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_orth_update_est, dim3(grid1(n, 256)), dim3(256), 0, stream, vol, est, n, iter, b);
    return hipGetLastError();
}

hipError_t launch_orth_md(const float* vol, const float* coef, float* md, int64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_orth_md, dim3(grid1(n, 256)), dim3(256), 0, stream, vol, coef, md, n);
    return hipGetLastError();
}

hipError_t launch_orth_mean_filter(const float* src, const float* vol, int N, float w, float* dst, hipStream_t stream) {
    const int64_t n = (int64_t)N * N * N;
    hipLaunchKernelGGL(k_orth_mean_filter, dim3(grid1(n, 256)), dim3(256), 0, stream, src, vol, N, w, dst);
    return hipGetLastError();
}

hipError_t launch_vol_down(const float* fine, float* coarse, int Nc, int f, hipStream_t stream) {
    const int64_t n = (int64_t)Nc * Nc * Nc;
    hipLaunchKernelGGL(k_vol_down, dim3(grid1(n, 256)), dim3(256), 0, stream, fine, coarse, Nc, f);
    return hipGetLastError();
}

hipError_t launch_vol_up_add(const float* a, const float* b, int Nc, int f, float* fine, float lo,
                             hipStream_t stream) {
    const int64_t N = (int64_t)Nc * f;
    hipLaunchKernelGGL(k_vol_up_add, dim3(grid1(N * N * N, 256)), dim3(256), 0, stream, a, b, Nc, f, fine, lo);
    return hipGetLastError();
}

hipError_t launch_sqrt_abs(float* vol, int64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sqrt_abs, dim3(grid1(n, 256)), dim3(256), 0, stream, vol, n);
    return hipGetLastError();
}

hipError_t launch_md_angles(const float* md, float* angle, int64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_md_angles, dim3(grid1(n, 256)), dim3(256), 0, stream, md, angle, n);
    return hipGetLastError();
}

hipError_t launch_md_normalise(float* md, int64_t n, float eps, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_md_normalise, dim3(grid1(n, 256)), dim3(256), 0, stream, md, n, eps);
    return hipGetLastError();
}

hipError_t launch_pca(const float* mu, int K, const Basis& b, int64_t nvox, int64_t chan_stride, float* md,
                      float* evalue, hipStream_t stream) {
    if (nvox == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pca, dim3(grid1(nvox, 256)), dim3(256), 0, stream, mu, K, b, nvox, chan_stride, md, evalue);
    return hipGetLastError();
}

hipError_t launch_band_rows(const ViewGeom* vg, const Dims& d, int n_codes, int n_bands, int32_t* out,
                            hipStream_t stream) {
    const int64_t n = (int64_t)n_codes * n_bands * d.N;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_band_rows, dim3(grid1(n, 256)), dim3(256), 0, stream, vg, d, n_codes, n_bands, out);
    return hipGetLastError();
}

hipError_t launch_absmax(const float* x, int64_t n, float* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)absmax_blocks(n, 1)), dim3(256), 0, stream, x, n, out);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, int64_t n, float value, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(grid1(n, 256)), dim3(256), 0, stream, p, n, value);
    return hipGetLastError();
}

hipError_t launch_clamp_min(float* p, int64_t n, float floor_value, float replacement, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clamp_min, dim3(grid1(n, 256)), dim3(256), 0, stream, p, n, floor_value, replacement);
    return hipGetLastError();
}

hipError_t launch_ellipsoids(float* vol, int N, float vs, const double* params_dev, int count, hipStream_t stream) {
    hipLaunchKernelGGL(k_ellipsoids, dim3(grid1(N, 64), N, N), dim3(64), 0, stream, vol, N, (double)vs, params_dev,
                       count);
    return hipGetLastError();
}

hipError_t launch_cylinder_shell(float* vol, int N, float vs, double r_out, double thick, double value,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(k_cylinder_shell, dim3(grid1(N, 64), N, N), dim3(64), 0, stream, vol, N, (double)vs, r_out,
                       thick, value);
    return hipGetLastError();
}

hipError_t launch_fibres(float* vol, int N, float vs, const double* params_dev, int count, const Basis& b, int K,
                         hipStream_t stream) {
    hipLaunchKernelGGL(k_fibres, dim3(grid1(N, 64), N, N), dim3(64), 0, stream, vol, N, (double)vs, params_dev, count,
                       b, K);
    return hipGetLastError();
}

hipError_t launch_fibres_scalar(float* vol, int N, float vs, const double* params_dev, int count, float value,
                                hipStream_t stream) {
    hipLaunchKernelGGL(k_fibres_scalar, dim3(grid1(N, 64), N, N), dim3(64), 0, stream, vol, N, (double)vs,
                       params_dev, count, value);
    return hipGetLastError();
}

hipError_t launch_poisson_noise(float* y, const float* p, int64_t n, double I0, uint64_t seed, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_poisson_noise, dim3(grid1(n, 256)), dim3(256), 0, stream, y, p, n, I0, seed);
    return hipGetLastError();
}

hipError_t launch_gaussian_noise(float* y, const float* x, int64_t n, double sigma, uint64_t seed,
                                 hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gaussian_noise, dim3(grid1(n, 256)), dim3(256), 0, stream, y, x, n, sigma, seed);
    return hipGetLastError();
}

}  // namespace xr
