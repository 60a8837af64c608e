// Kernels of the gradient-domain phase-contrast reconstruction
// (IR::gradReconstruct, reconstruct.cu:158-328) that the projectors do not
// cover: the per-view sinogram factor, the faces of the gradient volume that
// the reference's launches leave alone, and the Poisson integration of the x
// and y gradients (poissonSolveLDLT, poisson_cpu.cpp:15-112).
//
// The Poisson system of one z slice is the 5-point Laplacian with identity
// rows on the slice border and zero right side there, i.e. the Dirichlet
// problem on the (n-2)^2 interior.  The type-I discrete sine transform
// S[j][k] = sin(pi (j+1)(k+1) / (M+1)), M = n-2, diagonalises it exactly:
//   F = (2/(M+1))^2 S ((S B S) ./ L) S,  L[k][l] = e[k] + e[l],
//   e[k] = 2 - 2 cos(pi (k+1) / (M+1)),
// so a slice costs four M x M matrix products.  They run in FP64 on the
// matrix cores (batched GEMM, 64x64 tiles), every slice of the volume in one
// launch per product; the right sides are formed in FP32 as the reference
// forms them, and the result is rounded to FP32 once.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "xr_internal.hpp"

namespace xr {

namespace {

// dst[view][pix] = coef[view] * src[view][pix] (sinogramGradientCoef, ir.cu:585-594).
__global__ void k_scale_views(const float* __restrict__ src, float* __restrict__ dst,
                              const float* __restrict__ coef, int64_t len_view, int64_t total) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t view = i / len_view;
        dst[i] = coef[view] * src[i];
    }
}

// Voxels with x >= ext or z >= ext of an N1^3 volume in the column layout
// (z + N1*(x + N1*y)) set to `value`: the layers the reference's launches do
// not reach (reconstruct.cu:203-204 size the grid from N, the kernels run to
// N1 = N+1).
__global__ void k_grad_faces(float* __restrict__ vol, int N1, int ext, float value) {
    const int layers = N1 - ext;
    const int64_t plane = (int64_t)N1 * N1;
    const int64_t per_face = (int64_t)layers * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * per_face;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int face = (int)(i / per_face);
        const int64_t r = i - face * per_face;
        const int layer = (int)(r / plane);
        const int64_t q = r - layer * plane;
        const int a = (int)(q / N1), b = (int)(q - (int64_t)a * N1);  // a: y, b: the other free axis
        int64_t idx;
        if (face == 0)
            idx = b + (int64_t)N1 * ((ext + layer) + (int64_t)N1 * a);  // x = ext + layer, z = b
        else
            idx = (ext + layer) + (int64_t)N1 * (b + (int64_t)N1 * a);  // z = ext + layer, x = b
        vol[idx] = value;
    }
}

// Right sides of one z slice each (poisson_cpu.cpp:98-105), FP32 arithmetic in
// the reference's order, stored as FP64 interior matrices B[z][y-1][x-1].
// g0, g1: (n+1)^3 in the reference layout (x fastest).
__global__ void k_poisson_rhs(const float* __restrict__ g0, const float* __restrict__ g1, double* __restrict__ B,
                              int n) {
    const int M = n - 2;
    const int64_t n1 = n + 1;
    const int64_t total = (int64_t)n * M * M;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(i / ((int64_t)M * M));
        const int64_t r = i - (int64_t)z * M * M;
        const int y = (int)(r / M) + 1, x = (int)(r - (int64_t)(y - 1) * M) + 1;
        const int64_t o = x + n1 * (y + n1 * z);
        const float b = ((g0[o] - g0[o + 1]) + g1[o]) - g1[o + n1];
        B[i] = (double)b;
    }
}

constexpr int kTile = 64, kTk = 16, kGemmThreads = 256;
typedef double f64x4_t __attribute__((ext_vector_type(4)));

enum GemmOut { GEMM_STORE = 0, GEMM_SPECTRAL = 1, GEMM_VOLUME = 2 };

// C_b = A_b * B_b for M x M row-major FP64 matrices, b = blockIdx.z; a stride
// of 0 shares one matrix over the batch.  OUT: GEMM_STORE writes C;
// GEMM_SPECTRAL writes C[i][j] * scale / (e[i] + e[j]); GEMM_VOLUME writes
// (float)C[y][x] to vol[(x+1) + n*((y+1) + n*b)], n = M+2.
// On the matrix cores: a 64 x 64 block tile, four waves of 32 x 32, each a
// 2 x 2 set of v_mfma_f64_16x16x4f64 tiles; A from lane (row lane&15, k
// lane>>4), B from lane (k lane>>4, column lane&15); result register r of a
// lane holds column lane&15, row 4*r + (lane>>4) (unlike the FP32 and BF16
// 16x16 forms, whose rows are 4*(lane>>4) + r).  K steps of 16 staged
// through LDS.
template <int OUT>
__global__ void __launch_bounds__(kGemmThreads) k_gemm64(const double* __restrict__ A, int64_t sa,
                                                         const double* __restrict__ Bm, int64_t sb,
                                                         double* __restrict__ C, float* __restrict__ vol,
                                                         const double* __restrict__ e, double scale, int M) {
    __shared__ double As[kTile][kTk + 1];  // As[row][k]
    __shared__ double Bs[kTk][kTile + 1];  // Bs[k][col]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = blockIdx.y * kTile, col0 = blockIdx.x * kTile;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t b = blockIdx.z;
    const double* Ab = A + b * sa;
    const double* Bb = Bm + b * sb;
    f64x4_t acc[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++) acc[mt][nt] = f64x4_t{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < M; k0 += kTk) {
        __syncthreads();  // the previous K step is done with As and Bs
#pragma unroll
        for (int r = 0; r < 4; r++) {
            // A: 64 rows x 16 k, k fastest in memory
            const int m = (tid >> 4) + 16 * r, kk = tid & 15;
            const int gr = row0 + m, gk = k0 + kk;
            As[m][kk] = (gr < M && gk < M) ? Ab[(int64_t)gr * M + gk] : 0.0;
            // B: 16 k x 64 columns, columns fastest
            const int nn = tid & 63, kb = (tid >> 6) + 4 * r;
            const int gc = col0 + nn, gkb = k0 + kb;
            Bs[kb][nn] = (gkb < M && gc < M) ? Bb[(int64_t)gkb * M + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTk; kk += 4) {
            double a[2], bv[2];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) a[mt] = As[wm + mt * 16 + li][kk + lk];
#pragma unroll
            for (int nt = 0; nt < 2; nt++) bv[nt] = Bs[kk + lk][wn + nt * 16 + li];
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mt], bv[nt], acc[mt][nt], 0, 0, 0);
        }
    }
    const int n = M + 2;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = row0 + wm + mt * 16 + 4 * r + lk;
            if (row >= M) continue;
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int c = col0 + wn + nt * 16 + li;
                if (c >= M) continue;
                const double v = acc[mt][nt][r];
                if (OUT == GEMM_STORE) {
                    C[b * (int64_t)M * M + (int64_t)row * M + c] = v;
                } else if (OUT == GEMM_SPECTRAL) {
                    C[b * (int64_t)M * M + (int64_t)row * M + c] = v * scale / (e[row] + e[c]);
                } else {
                    vol[(c + 1) + (int64_t)n * ((row + 1) + (int64_t)n * b)] = (float)v;
                }
            }
        }
}

inline int grid_1d(int64_t total, int threads) {
    const int64_t g = (total + threads - 1) / threads;
    return (int)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

}  // namespace

hipError_t launch_scale_views(const float* src, float* dst, const float* coef, int64_t len_view, int n_views,
                              hipStream_t stream) {
    const int64_t total = len_view * n_views;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_views, dim3(grid_1d(total, 256)), dim3(256), 0, stream, src, dst, coef, len_view,
                       total);
    return hipGetLastError();
}

hipError_t launch_grad_faces(float* vol, int N1, int ext, float value, hipStream_t stream) {
    if (ext >= N1) return hipSuccess;
    const int64_t total = 2 * (int64_t)(N1 - ext) * N1 * N1;
    hipLaunchKernelGGL(k_grad_faces, dim3(grid_1d(total, 256)), dim3(256), 0, stream, vol, N1, ext, value);
    return hipGetLastError();
}

hipError_t launch_poisson_rhs(const float* g0, const float* g1, double* B, int n, hipStream_t stream) {
    const int64_t total = (int64_t)n * (n - 2) * (n - 2);
    hipLaunchKernelGGL(k_poisson_rhs, dim3(grid_1d(total, 256)), dim3(256), 0, stream, g0, g1, B, n);
    return hipGetLastError();
}

// out = (2/(M+1))^2 S ((S B S) ./ L) S per slice, written into the interior of
// the n^3 volume `vol` (its border is left to the caller).  B and T: n slices
// of M x M doubles each; B is overwritten.
hipError_t launch_poisson_dst(const double* S, const double* e, double* B, double* T, float* vol, int n,
                              hipStream_t stream) {
    const int M = n - 2;
    const int64_t MM = (int64_t)M * M;
    const double h = (double)(M + 1);
    const double scale = (2.0 / h) * (2.0 / h);
    const dim3 grid((M + kTile - 1) / kTile, (M + kTile - 1) / kTile, n), block(kGemmThreads);
    // T = S B (over y), B = (T S) ./ L (over x), T = S B, vol = T S
    hipLaunchKernelGGL(k_gemm64<GEMM_STORE>, grid, block, 0, stream, S, (int64_t)0, B, MM, T, nullptr, nullptr, 0.0,
                       M);
    hipLaunchKernelGGL(k_gemm64<GEMM_SPECTRAL>, grid, block, 0, stream, T, MM, S, (int64_t)0, B, nullptr, e, scale,
                       M);
    hipLaunchKernelGGL(k_gemm64<GEMM_STORE>, grid, block, 0, stream, S, (int64_t)0, B, MM, T, nullptr, nullptr, 0.0,
                       M);
    hipLaunchKernelGGL(k_gemm64<GEMM_VOLUME>, grid, block, 0, stream, T, MM, S, (int64_t)0, nullptr, vol, nullptr,
                       0.0, M);
    return hipGetLastError();
}

}  // namespace xr
