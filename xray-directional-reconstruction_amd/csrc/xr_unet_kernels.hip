// HIP kernels (gfx950) of the learned warm start: a 3-level 3-D U-Net on
// channels-last BF16 activations with FP32 accumulation.  No counterpart in
// the reference (it has no learned component).
//
// Activations: act[v * ld + off + c], v = voxel index of the column layout
// (z fastest: v = z + N*(x + N*y)), c < C, BF16.  A tensor can be a channel
// slice (off, ld) of a wider buffer, which is how skip connections are
// concatenated without copies.
//
// 3x3x3 convolution, stride 1, zero padding 1, as an implicit GEMM on MFMA
// (v_mfma_f32_16x16x32_bf16): M = voxels, N = output channels, K = 27*Cin with
// k = tap*Cin + ci, tap = (d0+1)*9 + (d1+1)*3 + (d2+1).  Cin is a multiple of
// 8 so the 8 consecutive k of one MFMA lane lie in one tap: one 16-byte load
// of 8 channels of one neighbour voxel.  Weights are [Cout16][Kpad] BF16 (k
// fastest, zero padded).  The data gradient is the same kernel on the output
// gradient with mirrored taps and transposed channels (weights [Cin16][K'pad],
// k' = (26 - tap)*Cout + co); the weight gradient contracts over voxels with
// both operands transposed through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "xr_unet_internal.hpp"

namespace xr {
namespace nn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) bf16x8 lds_bf16x8;
typedef __attribute__((address_space(3))) __bf16 lds_bf16;

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
    return z;
}

// Tap t -> offsets (d0, d1, d2) in {-1, 0, 1}.
__device__ __forceinline__ void tap_offsets(int t, int& d0, int& d1, int& d2) {
    d0 = t / 9 - 1;
    d1 = (t / 3) % 3 - 1;
    d2 = t % 3 - 1;
}

// ---------------------------------------------------------------------------
// Forward / data-gradient convolution.  Block: 4 waves, 128 consecutive
// voxels (32 per wave as two 16-row MFMA tiles) x all output channels (NT
// tiles of 16).
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void k_conv3(const ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t V = (int64_t)a.D0 * a.D1 * a.D2;
    const int64_t vbase = (int64_t)blockIdx.x * 128 + wave * 32;
    if (vbase >= V) return;
    const int row = lane & 15, kg = lane >> 4;
    int p0[2], p1[2], p2[2];
    bool vin[2];
#pragma unroll
    for (int ms = 0; ms < 2; ms++) {
        const int64_t v = vbase + ms * 16 + row;
        vin[ms] = v < V;
        const int64_t vv = vin[ms] ? v : 0;
        p2[ms] = (int)(vv % a.D2);
        p1[ms] = (int)((vv / a.D2) % a.D1);
        p0[ms] = (int)(vv / ((int64_t)a.D2 * a.D1));
    }
    f32x4 acc[2][NT];
#pragma unroll
    for (int ms = 0; ms < 2; ms++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[ms][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = 27 * a.Cin;
    const __bf16* __restrict__ in = a.in + a.in_off;
    for (int kc = 0; kc < a.Kpad; kc += 32) {
        const int k0 = kc + 8 * kg;
        const int tap = k0 / a.Cin;
        const int ci = k0 - tap * a.Cin;
        int d0, d1, d2;
        tap_offsets(tap, d0, d1, d2);
        bf16x8 A[2];
#pragma unroll
        for (int ms = 0; ms < 2; ms++) {
            const int n0 = p0[ms] + d0, n1 = p1[ms] + d1, n2 = p2[ms] + d2;
            const bool ok = vin[ms] && k0 < K && (unsigned)n0 < (unsigned)a.D0 && (unsigned)n1 < (unsigned)a.D1 &&
                            (unsigned)n2 < (unsigned)a.D2;
            const int64_t nv = ((int64_t)n0 * a.D1 + n1) * a.D2 + n2;
            A[ms] = ok ? *(const bf16x8*)(in + nv * a.in_ld + ci) : zero8();
        }
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const bf16x8 B = *(const bf16x8*)(a.w + (int64_t)(nt * 16 + row) * a.Kpad + k0);
#pragma unroll
            for (int ms = 0; ms < 2; ms++) acc[ms][nt] = mfma(A[ms], B, acc[ms][nt]);
        }
    }
    // D layout of the 16x16 MFMA: lane holds column (lane & 15), rows 4*kg + i.
#pragma unroll
    for (int ms = 0; ms < 2; ms++) {
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const int co = nt * 16 + row;
            if (co >= a.Cout) continue;
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t v = vbase + ms * 16 + 4 * kg + i;
                if (v >= V) continue;
                float val = acc[ms][nt][i] + b;
                if (a.relu) val = fmaxf(val, 0.f);
                if (a.mask && !((float)a.mask[v * a.mask_ld + a.mask_off + co] > 0.f)) val = 0.f;
                __bf16* o = a.out + v * a.out_ld + a.out_off + co;
                if (a.accumulate) val += (float)*o;
                *o = (__bf16)val;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward convolution, LDS-tiled.  Block: 4 waves, an output tile of
// 2 (D0) x 4 (D1) x 32 (D2) voxels; wave w owns D0 row w/2, D1 columns
// 2(w%2) and 2(w%2)+1, and both 16-voxel halves of D2: four 16-row MFMA
// tiles.  K runs over chunks of 8 input channels; for each chunk the input
// halo (4 x 6 x 34 voxels x 8 channels, 13 KB) is staged in LDS once and read
// by all 27 taps, instead of one global load per tap.  K within a chunk:
// k = tap*8 + c, 28 taps with the last one zero (7 MFMA steps of 32: lane
// group kg takes tap 4s + kg).
// ---------------------------------------------------------------------------
constexpr int kCT0 = 2, kCT1 = 4, kCT2 = 32;
constexpr int kCH0 = kCT0 + 2, kCH1 = kCT1 + 2, kCH2 = kCT2 + 2;
constexpr int kCHalo = kCH0 * kCH1 * kCH2;
constexpr int kTapsPad = 28;

// CPS: input-channel chunks staged per barrier.  Wide inputs (the decoder's
// concatenations, 96-byte voxel records) stage two chunks at once, so that
// both 16-byte pieces of a record are fetched together while its cache line
// is resident, instead of the line being fetched again for the next chunk.
// UP: the first a.up_chunks chunks come from the half-resolution tensor a.up
// at the halo voxel's parent (nearest x2 upsampling), the rest from a.in, so
// the decoder's concatenation is never written; the halo in LDS, and hence
// the result, is the same as with the concatenation materialised.
template <int NT, int CPS, bool UP>
__global__ __launch_bounds__(256) void k_conv3_tiled(const ConvArgs a) {
    __shared__ bf16x8 halo[CPS * kCHalo];
    const lds_bf16x8* const hl = (const lds_bf16x8*)halo;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int nb2 = a.D2 / kCT2, nb1 = a.D1 / kCT1;
    const int b2 = blockIdx.x % nb2, b1 = (blockIdx.x / nb2) % nb1, b0 = blockIdx.x / (nb2 * nb1);
    const int o0 = b0 * kCT0, o1 = b1 * kCT1, o2 = b2 * kCT2;
    // halo coordinates (+1) of this lane's output voxels: tile m = 2*(x half) + (z half)
    const int h0 = (wave >> 1) + 1;
    int h1[4], h2[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        h1[m] = 2 * (wave & 1) + (m >> 1) + 1;
        h2[m] = 16 * (m & 1) + row + 1;
    }
    f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[m][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* __restrict__ in = a.in + a.in_off;
    const int nchunk = a.Cin / 8;
    // halo voxels of this thread: i = threadIdx.x + 256 j; their global offsets
    // (-1: outside the volume, zero padding) are the same for every chunk
    constexpr int kPer = (kCHalo + 255) / 256;
    int64_t hoff[kPer], uoff[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = threadIdx.x + 256 * j;
        const int q2 = i % kCH2, q1 = (i / kCH2) % kCH1, q0 = i / (kCH2 * kCH1);
        const int g0 = o0 + q0 - 1, g1 = o1 + q1 - 1, g2 = o2 + q2 - 1;
        const bool ok = i < kCHalo && (unsigned)g0 < (unsigned)a.D0 && (unsigned)g1 < (unsigned)a.D1 &&
                        (unsigned)g2 < (unsigned)a.D2;
        hoff[j] = ok ? (((int64_t)g0 * a.D1 + g1) * a.D2 + g2) * a.in_ld : -1;
        if (UP)
            uoff[j] = ok ? (((int64_t)(g0 >> 1) * (a.D1 >> 1) + (g1 >> 1)) * (a.D2 >> 1) + (g2 >> 1)) * a.up_ld : -1;
    }
    const __bf16* __restrict__ up = UP ? a.up + a.up_off : nullptr;
    // stage st holds chunks st*CPS .. st*CPS + CPS-1 (chunks past the last are zero)
    bf16x8 pre[CPS][kPer];
    auto fetch = [&](int st) {
#pragma unroll
        for (int c = 0; c < CPS; c++) {
            const int cc = st * CPS + c;
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const __bf16* src = (UP && cc < a.up_chunks) ? up + uoff[j] + cc * 8
                                                             : in + hoff[j] + (cc - (UP ? a.up_chunks : 0)) * 8;
                pre[c][j] = (cc < nchunk && hoff[j] >= 0) ? *(const bf16x8*)src : zero8();
            }
        }
    };
    auto store = [&] {
#pragma unroll
        for (int c = 0; c < CPS; c++)
#pragma unroll
            for (int j = 0; j < kPer; j++)
                if (threadIdx.x + 256 * j < kCHalo) halo[c * kCHalo + threadIdx.x + 256 * j] = pre[c][j];
    };
    const int nstage = (nchunk + CPS - 1) / CPS;
    fetch(0);
    for (int st = 0; st < nstage; st++) {
        __syncthreads();  // the previous stage's reads are done
        store();
        __syncthreads();
        if (st + 1 < nstage) fetch(st + 1);  // in flight during this stage's MFMAs
#pragma unroll 1
        for (int c = 0; c < CPS; c++) {
            const int cc = st * CPS + c;
            if (cc >= nchunk) break;
            const __bf16* __restrict__ wc = a.wt + cc * (kTapsPad * 8);
            const lds_bf16x8* const hc = hl + c * kCHalo;
#pragma unroll
            for (int s = 0; s < kTapsPad / 4; s++) {
                const int tap = 4 * s + kg;
                int d0, d1, d2;
                tap_offsets(min(tap, 26), d0, d1, d2);  // tap 27 has zero weights
                bf16x8 A[4];
#pragma unroll
                for (int m = 0; m < 4; m++) A[m] = hc[((h0 + d0) * kCH1 + h1[m] + d1) * kCH2 + h2[m] + d2];
#pragma unroll
                for (int nt = 0; nt < NT; nt++) {
                    const bf16x8 B = *(const bf16x8*)(wc + (int64_t)(nt * 16 + row) * a.Ktpad + tap * 8);
#pragma unroll
                    for (int m = 0; m < 4; m++) acc[m][nt] = mfma(A[m], B, acc[m][nt]);
                }
            }
        }
    }
    // D layout of the 16x16 MFMA: lane holds column (lane & 15), rows 4*kg + i.
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int64_t vrow = ((int64_t)(o0 + h0 - 1) * a.D1 + (o1 + h1[m] - 1)) * a.D2 + o2 + 16 * (m & 1);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const int co = nt * 16 + row;
            if (co >= a.Cout) continue;
            const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int64_t v = vrow + 4 * kg + i;
                float val = acc[m][nt][i] + b;
                if (a.relu) val = fmaxf(val, 0.f);
                if (a.mask && !((float)a.mask[v * a.mask_ld + a.mask_off + co] > 0.f)) val = 0.f;
                __bf16* o = a.out + v * a.out_ld + a.out_off + co;
                if (a.accumulate) val += (float)*o;
                *o = (__bf16)val;
            }
        }
    }
}

// Master FP32 weights [Cout][27][Cin] -> tiled-order BF16 [Cout16][Ktpad].
__global__ void k_pack_tiled(const float* __restrict__ W, int Cin, int Cout, __bf16* wt, int Ktpad, int cout16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)cout16 * Ktpad) return;
    const int co = (int)(i / Ktpad), k = (int)(i % Ktpad);
    const int cc = k / (kTapsPad * 8), tap = (k / 8) % kTapsPad, c = k % 8;
    const int ci = cc * 8 + c;
    wt[i] = (__bf16)((co < Cout && tap < 27 && ci < Cin) ? W[((int64_t)co * 27 + tap) * Cin + ci] : 0.f);
}

// ---------------------------------------------------------------------------
// Weight gradient: dw[co][k] += sum_v gout[v][co] * im2col(in)[v][k].
// Block: 4 waves; all output channels (MT tiles of 16) x 64 k columns (16 per
// wave) over a slab of voxels, 32 voxels per step.  Both operands are staged
// transposed in LDS ([channel][voxel]) so that each MFMA lane reads 8
// consecutive voxels with one 16-byte LDS load.  Partial sums are added to dw
// with FP32 atomics (the summation order of the slabs is not fixed).
// ---------------------------------------------------------------------------
constexpr int kWgVox = 32;          // voxels per step
constexpr int kWgRow = kWgVox + 8;  // LDS row stride (bf16), padded

template <int MT>
__global__ __launch_bounds__(256) void k_conv3_wgrad(const WgradArgs a) {
    extern __shared__ __bf16 wg_smem[];
    lds_bf16* const gT = (lds_bf16*)wg_smem;                    // [MT*16][kWgRow]
    lds_bf16* const iT = (lds_bf16*)wg_smem + MT * 16 * kWgRow;  // [64][kWgRow]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int kbase = blockIdx.x * 64;
    const int64_t V = (int64_t)a.D0 * a.D1 * a.D2;
    const int64_t v_beg = (int64_t)blockIdx.y * a.vox_per_block;
    const int64_t v_end = v_beg + a.vox_per_block < V ? v_beg + a.vox_per_block : V;
    const int K = 27 * a.Cin;
    const int cg_out = a.Cout / 8;  // 8-channel groups of gout
    // rows of gT beyond Cout stay zero
    for (int i = tid; i < MT * 16 * kWgRow; i += 256) gT[i] = (__bf16)0.0f;
    f32x4 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this thread's im2col column group (fixed over the slab): 32 voxels x 8 groups
    const int iv = tid >> 3, ig = tid & 7;
    const int k0 = kbase + 8 * ig;
    const bool kin = k0 < K;
    const int tap = kin ? k0 / a.Cin : 0;
    const int ci = kin ? k0 - tap * a.Cin : 0;
    int d0, d1, d2;
    tap_offsets(tap, d0, d1, d2);
    for (int64_t v0 = v_beg; v0 < v_end; v0 += kWgVox) {
        __syncthreads();
        // stage gout[v0 .. v0+31][0 .. Cout) transposed
        for (int i = tid; i < kWgVox * cg_out; i += 256) {
            const int vv = i / cg_out, g = i - vv * cg_out;
            const int64_t v = v0 + vv;
            const bf16x8 x = (v < v_end) ? *(const bf16x8*)(a.gout + v * a.g_ld + a.g_off + 8 * g) : zero8();
#pragma unroll
            for (int j = 0; j < 8; j++) gT[(8 * g + j) * kWgRow + vv] = x[j];
        }
        // stage im2col[v0 .. v0+31][kbase .. kbase+63] transposed
        {
            const int64_t v = v0 + iv;
            bf16x8 x = zero8();
            if (kin && v < v_end) {
                const int p2 = (int)(v % a.D2), p1 = (int)((v / a.D2) % a.D1), p0 = (int)(v / ((int64_t)a.D2 * a.D1));
                const int n0 = p0 + d0, n1 = p1 + d1, n2 = p2 + d2;
                if ((unsigned)n0 < (unsigned)a.D0 && (unsigned)n1 < (unsigned)a.D1 && (unsigned)n2 < (unsigned)a.D2) {
                    const int64_t nv = ((int64_t)n0 * a.D1 + n1) * a.D2 + n2;
                    x = *(const bf16x8*)(a.in + nv * a.in_ld + a.in_off + ci);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) iT[(8 * ig + j) * kWgRow + iv] = x[j];
        }
        __syncthreads();
        const bf16x8 B = *(const lds_bf16x8*)(iT + (16 * wave + row) * kWgRow + 8 * kg);
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
            const bf16x8 A = *(const lds_bf16x8*)(gT + (mt * 16 + row) * kWgRow + 8 * kg);
            acc[mt] = mfma(A, B, acc[mt]);
        }
    }
    const int k = kbase + 16 * wave + row;
    if (k >= K) return;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = mt * 16 + 4 * kg + i;
            if (co < a.Cout) atomicAdd(a.dw + (int64_t)co * a.ldw + k, acc[mt][i]);
        }
    }
}

// db[c] += sum_v g[v][c]; block = C threads (one per channel) over a voxel slab.
__global__ void k_bias_grad(const __bf16* __restrict__ g, int ld, int off, int C, int64_t V, int64_t per_block,
                            float* db) {
    const int c = threadIdx.x;
    if (c >= C) return;
    const int64_t v0 = (int64_t)blockIdx.x * per_block;
    const int64_t v1 = v0 + per_block < V ? v0 + per_block : V;
    float s = 0.f;
    for (int64_t v = v0; v < v1; v++) s += (float)g[v * ld + off + c];
    atomicAdd(db + c, s);
}

// 2x2x2 average pooling: out[p][c] = mean of the 8 children (dims halve).
// Index arithmetic of the elementwise kernels below: 32-bit (I = unsigned)
// when the launch's element count fits, 64-bit otherwise (64-bit division is
// several times the cost and these kernels are otherwise pure data movement).
template <class I>
__global__ void k_pool(const __bf16* __restrict__ in, int in_ld, int in_off, int C, int D0, int D1, int D2,
                       __bf16* __restrict__ out) {
    const int H0 = D0 / 2, H1 = D1 / 2, H2 = D2 / 2;
    const int G = C / 8;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    const I n = (I)H0 * H1 * H2 * G;
    if (i >= n) return;
    const int g = (int)(i % G);
    const int64_t p = i / G;
    const int q2 = (int)((I)p % H2), q1 = (int)(((I)p / H2) % H1), q0 = (int)((I)p / ((I)H2 * H1));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 8; c++) {
        const int64_t v = ((int64_t)(2 * q0 + (c >> 2)) * D1 + (2 * q1 + ((c >> 1) & 1))) * D2 + (2 * q2 + (c & 1));
        const bf16x8 x = *(const bf16x8*)(in + v * in_ld + in_off + 8 * g);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += (float)x[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = (__bf16)(0.125f * s[j]);
    *(bf16x8*)(out + p * C + 8 * g) = o;
}

// Pooling backward into a (masked) slice: gin[child][c] (+)= gp[p][c]/8 * (mask[child][c] > 0).
template <class I>
__global__ void k_pool_bwd(const __bf16* __restrict__ gp, int C, int D0, int D1, int D2, __bf16* gin, int g_ld,
                           int g_off, const __bf16* __restrict__ mask, int m_ld, int m_off, int accumulate) {
    const int G = C / 8;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    const I n = (I)D0 * D1 * D2 * G;
    if (i >= n) return;
    const int g = (int)(i % G);
    const int64_t v = i / G;
    const int p2 = (int)((I)v % D2), p1 = (int)(((I)v / D2) % D1), p0 = (int)((I)v / ((I)D2 * D1));
    const int64_t p = ((int64_t)(p0 / 2) * (D1 / 2) + p1 / 2) * (D2 / 2) + p2 / 2;
    const bf16x8 x = *(const bf16x8*)(gp + p * C + 8 * g);
    const bf16x8 m = *(const bf16x8*)(mask + v * m_ld + m_off + 8 * g);
    bf16x8* o = (bf16x8*)(gin + v * g_ld + g_off + 8 * g);
    bf16x8 old = accumulate ? *o : zero8();
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float val = ((float)m[j] > 0.f) ? 0.125f * (float)x[j] : 0.f;
        if (accumulate) val += (float)old[j];
        r[j] = (__bf16)val;
    }
    *o = r;
}

// Nearest-neighbour upsampling x2 into a channel slice: out[child][off + c] = low[parent][c].
template <class I>
__global__ void k_upsample(const __bf16* __restrict__ low, int C, int D0, int D1, int D2, __bf16* out, int o_ld,
                           int o_off) {
    const int G = C / 8;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    const I n = (I)D0 * D1 * D2 * G;
    if (i >= n) return;
    const int g = (int)(i % G);
    const int64_t v = i / G;
    const int p2 = (int)((I)v % D2), p1 = (int)(((I)v / D2) % D1), p0 = (int)((I)v / ((I)D2 * D1));
    const int64_t p = ((int64_t)(p0 / 2) * (D1 / 2) + p1 / 2) * (D2 / 2) + p2 / 2;
    *(bf16x8*)(out + v * o_ld + o_off + 8 * g) = *(const bf16x8*)(low + p * C + 8 * g);
}

// Upsampling backward: glow[p][c] = sum of the 8 children's g[child][off + c].
template <class I>
__global__ void k_upsample_bwd(const __bf16* __restrict__ g, int g_ld, int g_off, int C, int D0, int D1, int D2,
                               __bf16* __restrict__ glow) {
    const int H0 = D0 / 2, H1 = D1 / 2, H2 = D2 / 2;
    const int G = C / 8;
    const I i = (I)blockIdx.x * blockDim.x + threadIdx.x;
    const I n = (I)H0 * H1 * H2 * G;
    if (i >= n) return;
    const int gg = (int)(i % G);
    const int64_t p = i / G;
    const int q2 = (int)((I)p % H2), q1 = (int)(((I)p / H2) % H1), q0 = (int)((I)p / ((I)H2 * H1));
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 8; c++) {
        const int64_t v = ((int64_t)(2 * q0 + (c >> 2)) * D1 + (2 * q1 + ((c >> 1) & 1))) * D2 + (2 * q2 + (c & 1));
        const bf16x8 x = *(const bf16x8*)(g + v * g_ld + g_off + 8 * gg);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += (float)x[j];
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = (__bf16)s[j];
    *(bf16x8*)(glow + p * C + 8 * gg) = o;
}

// Network input: x8[v] = (bf16(x[v] * inv_scale), 0, ..., 0).
__global__ void k_input8(const float* __restrict__ x, int64_t V, float inv_scale, __bf16* __restrict__ x8) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    bf16x8 o = zero8();
    o[0] = (__bf16)(x[v] * inv_scale);
    *(bf16x8*)(x8 + v * 8) = o;
}

// Output head (1x1x1 conv, C -> 1) with the residual: y = x + scale*(w.e[v] + b),
// w[C] = b.  Loss mode (t != null): r = (x/scale + w.e + b) - t/scale,
// loss += r^2, g[v] = 2 r / V (normalised units).
template <int C>
__global__ __launch_bounds__(256) void k_head(const __bf16* __restrict__ e, int64_t V, const float* __restrict__ w,
                                              const float* __restrict__ x, float scale, float* __restrict__ y,
                                              const float* __restrict__ t, float* g, double* loss) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (v < V) {
        float f = w[C];
#pragma unroll
        for (int c = 0; c < C; c += 8) {
            const bf16x8 a = *(const bf16x8*)(e + v * C + c);
#pragma unroll
            for (int j = 0; j < 8; j++) f = fmaf(w[c + j], (float)a[j], f);
        }
        const float inv = 1.0f / scale;
        if (y) y[v] = x[v] + scale * f;
        if (t) {
            const float r = (x[v] * inv + f) - t[v] * inv;
            l = (double)r * (double)r;
            g[v] = 2.0f * r / (float)V;
        }
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
        if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
    }
}

// Head backward: ge[v][c] = g[v] w[c] (e[v][c] > 0); dw[c] += sum_v g[v] e[v][c];
// dw[C] (bias) += sum_v g[v].  Wave sums, then one atomic per value per block.
template <int C>
__global__ __launch_bounds__(256) void k_head_bwd(const __bf16* __restrict__ e, int64_t V, const float* __restrict__ w,
                                                  const float* __restrict__ g, __bf16* __restrict__ ge, float* dw) {
    __shared__ float red[C + 1][4];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float part[C + 1];
#pragma unroll
    for (int c = 0; c <= C; c++) part[c] = 0.f;
    if (v < V) {
        const float gv = g[v];
#pragma unroll
        for (int c = 0; c < C; c += 8) {
            const bf16x8 a = *(const bf16x8*)(e + v * C + c);
            bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                o[j] = (__bf16)(((float)a[j] > 0.f) ? gv * w[c + j] : 0.f);
                part[c + j] = gv * (float)a[j];
            }
            *(bf16x8*)(ge + v * C + c) = o;
        }
        part[C] = gv;
    }
#pragma unroll
    for (int c = 0; c <= C; c++) {
        float s = part[c];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[c][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x <= C) {
        const int c = threadIdx.x;
        atomicAdd(dw + c, red[c][0] + red[c][1] + red[c][2] + red[c][3]);
    }
}

// Master FP32 weights [Cout][27][Cin] -> forward BF16 [Cout16][Kpad] and
// data-gradient BF16 [Cin16][K'pad] (k' = (26 - tap)*Cout + co).
__global__ void k_pack(const float* __restrict__ W, int Cin, int Cout, __bf16* wf, int Kpad, int cout16,
                       __bf16* wd, int Kdpad, int cin16) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nf = (int64_t)cout16 * Kpad;
    const int64_t nd = (int64_t)cin16 * Kdpad;
    const int K = 27 * Cin, Kd = 27 * Cout;
    if (i < nf) {
        const int co = (int)(i / Kpad), k = (int)(i % Kpad);
        wf[i] = (__bf16)((co < Cout && k < K) ? W[(int64_t)co * K + k] : 0.f);
    } else if (wd && i < nf + nd) {
        const int64_t j = i - nf;
        const int ci = (int)(j / Kdpad), kd = (int)(j % Kdpad);
        float val = 0.f;
        if (ci < Cin && kd < Kd) {
            const int tp = kd / Cout, co = kd % Cout;
            val = W[(int64_t)co * K + (26 - tp) * Cin + ci];
        }
        wd[j] = (__bf16)val;
    }
}

// Adam (bias-corrected step sizes precomputed on the host).
__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float b1, float b2, float lr_t,
                       float eps_t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
}

// sum x over n floats (FP64 accumulation).
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ x, int64_t n, double* out) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s += x[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

bool conv3_tiled_ok(const ConvArgs& a) {
    const int nt = (a.Cout + 15) / 16;
    return a.wt && a.D0 % kCT0 == 0 && a.D1 % kCT1 == 0 && a.D2 % kCT2 == 0 &&
           a.Ktpad == a.Cin / 8 * kTapsPad * 8 && nt <= 6;
}

hipError_t launch_conv3(const ConvArgs& a, hipStream_t s) {
    const int64_t V = (int64_t)a.D0 * a.D1 * a.D2;
    if (V == 0) return hipSuccess;
    if (a.Cin % 8 != 0 || a.in_ld % 8 != 0 || a.in_off % 8 != 0 || a.Kpad % 32 != 0 || a.Kpad < 27 * a.Cin)
        return hipErrorInvalidValue;
    const int nt = (a.Cout + 15) / 16;
    const bool up = a.up_chunks > 0;
    if (up && (!conv3_tiled_ok(a) || !a.up || a.up_chunks > a.Cin / 8 || a.up_ld % 8 != 0 || a.up_off % 8 != 0 ||
               a.up_off + 8 * a.up_chunks > a.up_ld))
        return hipErrorInvalidValue;
    if (conv3_tiled_ok(a)) {
        const dim3 tgrid((unsigned)(V / (kCT0 * kCT1 * kCT2)));
        // several chunks per stage for inputs of 32 or more channels
        const int cps = (a.Cin >= 32 && nt <= 3) ? (nt <= 2 ? 3 : 2) : 1;
        switch (nt) {
#define XR_CONVT_UP(N, C)                                                                  \
    if (up) hipLaunchKernelGGL((k_conv3_tiled<N, C, true>), tgrid, dim3(256), 0, s, a); \
    else hipLaunchKernelGGL((k_conv3_tiled<N, C, false>), tgrid, dim3(256), 0, s, a);
#define XR_CONVT(N)                 \
    case N:                         \
        if (cps == 3) {             \
            XR_CONVT_UP(N, 3)       \
        } else if (cps == 2) {      \
            XR_CONVT_UP(N, 2)       \
        } else {                    \
            XR_CONVT_UP(N, 1)       \
        }                           \
        break;
            XR_CONVT(1)
            XR_CONVT(2)
            XR_CONVT(3)
            XR_CONVT(4)
            XR_CONVT(5)
            XR_CONVT(6)
#undef XR_CONVT
#undef XR_CONVT_UP
        }
        return hipGetLastError();
    }
    const dim3 grid(blocks(V, 128));
    switch (nt) {
#define XR_CONV(N) \
    case N: hipLaunchKernelGGL(k_conv3<N>, grid, dim3(256), 0, s, a); break;
        XR_CONV(1)
        XR_CONV(2)
        XR_CONV(3)
        XR_CONV(4)
        XR_CONV(5)
        XR_CONV(6)
        XR_CONV(7)
        XR_CONV(8)
        XR_CONV(9)
#undef XR_CONV
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_conv3_wgrad(WgradArgs a, hipStream_t s) {
    const int64_t V = (int64_t)a.D0 * a.D1 * a.D2;
    if (V == 0) return hipSuccess;
    if (a.Cin % 8 != 0 || a.Cout % 8 != 0 || a.in_ld % 8 != 0 || a.in_off % 8 != 0 || a.g_ld % 8 != 0 ||
        a.g_off % 8 != 0 || a.ldw < 27 * a.Cin)
        return hipErrorInvalidValue;
    const int mt = (a.Cout + 15) / 16;
    const int kblocks = (27 * a.Cin + 63) / 64;
    // slabs: about 2048 blocks in all, whole 32-voxel steps
    int64_t slabs = std::max<int64_t>(1, std::min<int64_t>((V + kWgVox - 1) / kWgVox, 2048 / kblocks + 1));
    int64_t per = (V + slabs - 1) / slabs;
    per = ((per + kWgVox - 1) / kWgVox) * kWgVox;
    slabs = (V + per - 1) / per;
    a.vox_per_block = per;
    const dim3 grid((unsigned)kblocks, (unsigned)slabs);
    const size_t lds = (size_t)(mt * 16 + 64) * kWgRow * sizeof(__bf16);
    switch (mt) {
#define XR_WG(M) \
    case M: hipLaunchKernelGGL(k_conv3_wgrad<M>, grid, dim3(256), lds, s, a); break;
        XR_WG(1)
        XR_WG(2)
        XR_WG(3)
        XR_WG(4)
        XR_WG(5)
        XR_WG(6)
        XR_WG(7)
        XR_WG(8)
        XR_WG(9)
#undef XR_WG
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bias_grad(const __bf16* g, int ld, int off, int C, int64_t V, float* db, hipStream_t s) {
    if (V == 0) return hipSuccess;
    if (C > 1024) return hipErrorInvalidValue;
    const int64_t nb = std::min<int64_t>(1024, (V + 255) / 256);
    const int64_t per = (V + nb - 1) / nb;
    hipLaunchKernelGGL(k_bias_grad, dim3(blocks(V, (int)per)), dim3(((C + 63) / 64) * 64), 0, s, g, ld, off, C, V,
                       per, db);
    return hipGetLastError();
}

hipError_t launch_pool(const __bf16* in, int in_ld, int in_off, int C, int D0, int D1, int D2, __bf16* out,
                       hipStream_t s) {
    if (C % 8 || in_ld % 8 || in_off % 8 || (D0 | D1 | D2) & 1) return hipErrorInvalidValue;
    const int64_t n = (int64_t)(D0 / 2) * (D1 / 2) * (D2 / 2) * (C / 8);
    if (n == 0) return hipSuccess;
    if (n < (int64_t)1 << 31)
        hipLaunchKernelGGL(k_pool<unsigned>, dim3(blocks(n, 256)), dim3(256), 0, s, in, in_ld, in_off, C, D0, D1, D2,
                           out);
    else
        hipLaunchKernelGGL(k_pool<int64_t>, dim3(blocks(n, 256)), dim3(256), 0, s, in, in_ld, in_off, C, D0, D1, D2,
                           out);
    return hipGetLastError();
}

hipError_t launch_pool_bwd(const __bf16* gp, int C, int D0, int D1, int D2, __bf16* gin, int g_ld, int g_off,
                           const __bf16* mask, int m_ld, int m_off, int accumulate, hipStream_t s) {
    if (C % 8 || g_ld % 8 || g_off % 8 || m_ld % 8 || m_off % 8 || (D0 | D1 | D2) & 1) return hipErrorInvalidValue;
    const int64_t n = (int64_t)D0 * D1 * D2 * (C / 8);
    if (n == 0) return hipSuccess;
    if (n < (int64_t)1 << 31)
        hipLaunchKernelGGL(k_pool_bwd<unsigned>, dim3(blocks(n, 256)), dim3(256), 0, s, gp, C, D0, D1, D2, gin, g_ld,
                           g_off, mask, m_ld, m_off, accumulate);
    else
        hipLaunchKernelGGL(k_pool_bwd<int64_t>, dim3(blocks(n, 256)), dim3(256), 0, s, gp, C, D0, D1, D2, gin, g_ld,
                           g_off, mask, m_ld, m_off, accumulate);
    return hipGetLastError();
}

hipError_t launch_upsample(const __bf16* low, int C, int D0, int D1, int D2, __bf16* out, int o_ld, int o_off,
                           hipStream_t s) {
    if (C % 8 || o_ld % 8 || o_off % 8 || (D0 | D1 | D2) & 1) return hipErrorInvalidValue;
    const int64_t n = (int64_t)D0 * D1 * D2 * (C / 8);
    if (n == 0) return hipSuccess;
    if (n < (int64_t)1 << 31)
        hipLaunchKernelGGL(k_upsample<unsigned>, dim3(blocks(n, 256)), dim3(256), 0, s, low, C, D0, D1, D2, out, o_ld,
                           o_off);
    else
        hipLaunchKernelGGL(k_upsample<int64_t>, dim3(blocks(n, 256)), dim3(256), 0, s, low, C, D0, D1, D2, out, o_ld,
                           o_off);
    return hipGetLastError();
}

hipError_t launch_upsample_bwd(const __bf16* g, int g_ld, int g_off, int C, int D0, int D1, int D2, __bf16* glow,
                               hipStream_t s) {
    if (C % 8 || g_ld % 8 || g_off % 8 || (D0 | D1 | D2) & 1) return hipErrorInvalidValue;
    const int64_t n = (int64_t)(D0 / 2) * (D1 / 2) * (D2 / 2) * (C / 8);
    if (n == 0) return hipSuccess;
    if (n < (int64_t)1 << 31)
        hipLaunchKernelGGL(k_upsample_bwd<unsigned>, dim3(blocks(n, 256)), dim3(256), 0, s, g, g_ld, g_off, C, D0, D1,
                           D2, glow);
    else
        hipLaunchKernelGGL(k_upsample_bwd<int64_t>, dim3(blocks(n, 256)), dim3(256), 0, s, g, g_ld, g_off, C, D0, D1,
                           D2, glow);
    return hipGetLastError();
}

hipError_t launch_input8(const float* x, int64_t V, float inv_scale, __bf16* x8, hipStream_t s) {
    if (V == 0) return hipSuccess;
    hipLaunchKernelGGL(k_input8, dim3(blocks(V, 256)), dim3(256), 0, s, x, V, inv_scale, x8);
    return hipGetLastError();
}

hipError_t launch_head(const __bf16* e, int C, int64_t V, const float* w, const float* x, float scale, float* y,
                       const float* t, float* g, double* loss, hipStream_t s) {
    if (V == 0) return hipSuccess;
    switch (C) {
#define XR_HEAD(N) \
    case N: hipLaunchKernelGGL(k_head<N>, dim3(blocks(V, 256)), dim3(256), 0, s, e, V, w, x, scale, y, t, g, loss); break;
        XR_HEAD(8)
        XR_HEAD(16)
        XR_HEAD(24)
        XR_HEAD(32)
#undef XR_HEAD
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_head_bwd(const __bf16* e, int C, int64_t V, const float* w, const float* g, __bf16* ge,
                           float* dw, hipStream_t s) {
    if (V == 0) return hipSuccess;
    switch (C) {
#define XR_HEADB(N) \
    case N: hipLaunchKernelGGL(k_head_bwd<N>, dim3(blocks(V, 256)), dim3(256), 0, s, e, V, w, g, ge, dw); break;
        XR_HEADB(8)
        XR_HEADB(16)
        XR_HEADB(24)
        XR_HEADB(32)
#undef XR_HEADB
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_pack(const float* W, int Cin, int Cout, __bf16* wf, int Kpad, __bf16* wd, int Kdpad,
                       hipStream_t s) {
    const int cout16 = (Cout + 15) / 16 * 16, cin16 = (Cin + 15) / 16 * 16;
    const int64_t n = (int64_t)cout16 * Kpad + (wd ? (int64_t)cin16 * Kdpad : 0);
    hipLaunchKernelGGL(k_pack, dim3(blocks(n, 256)), dim3(256), 0, s, W, Cin, Cout, wf, Kpad, cout16, wd, Kdpad,
                       cin16);
    return hipGetLastError();
}

hipError_t launch_pack_tiled(const float* W, int Cin, int Cout, __bf16* wt, int Ktpad, hipStream_t s) {
    const int cout16 = (Cout + 15) / 16 * 16;
    const int64_t n = (int64_t)cout16 * Ktpad;
    hipLaunchKernelGGL(k_pack_tiled, dim3(blocks(n, 256)), dim3(256), 0, s, W, Cin, Cout, wt, Ktpad, cout16);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t n, float b1, float b2, float lr_t,
                       float eps_t, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, s, p, g, m, v, n, b1, b2, lr_t, eps_t);
    return hipGetLastError();
}

hipError_t launch_sum(const float* x, int64_t n, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int64_t nb = std::min<int64_t>(2048, (n + 255) / 256);
    hipLaunchKernelGGL(k_sum, dim3((unsigned)nb), dim3(256), 0, s, x, n, out);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace xr
