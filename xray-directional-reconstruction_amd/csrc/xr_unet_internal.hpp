// Internal declarations of the U-Net kernels (xr_unet_kernels.hip) used by the
// host side (xr_unet.cpp).  Not part of the public C ABI (include/xr.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace xr {
namespace nn {

// 3x3x3 convolution (forward, or data gradient with packed mirrored weights).
// out[v][out_off + co] = relu?(sum_k w[co][k] im2col(in)[v][k] + bias[co]),
// optionally masked by (mask[v][mask_off + co] > 0) and added to out.
struct ConvArgs {
    const __bf16* in;
    int in_ld, in_off, Cin;
    const __bf16* w;  // [round16(Cout)][Kpad], k = tap*Cin + ci
    int Kpad;
    const float* bias;  // [Cout] or null
    __bf16* out;
    int out_ld, out_off, Cout;
    const __bf16* mask;  // or null
    int mask_ld, mask_off;
    int D0, D1, D2;  // spatial dims, D2 fastest
    int relu, accumulate;
    // optional: the same weights as [round16(Cout)][Ktpad] with k = chunk*224 +
    // tap*8 + c (chunk of 8 input channels, 28 taps with tap 27 zero), for the
    // LDS-tiled kernel (used when the dims are multiples of its tile)
    const __bf16* wt;
    int Ktpad;
    // optional (LDS-tiled kernel only): a concatenation [upsample(up) | in]
    // read in place, with the nearest x2 upsampling fused into the halo fetch.
    // Input channels ci < 8*up_chunks are read from the half-resolution tensor
    // up[(v0/2, v1/2, v2/2) * up_ld + up_off + ci] (dims D0/2 x D1/2 x D2/2),
    // the others from in[v * in_ld + in_off + ci - 8*up_chunks].
    const __bf16* up;
    int up_ld, up_off, up_chunks;
};

// True when launch_conv3 runs a on the LDS-tiled kernel (the only one that
// takes a fused upsampling source).
bool conv3_tiled_ok(const ConvArgs& a);

// dw[co * ldw + k] += sum_v gout[v][g_off + co] im2col(in)[v][k]  (k < 27*Cin).
struct WgradArgs {
    const __bf16* in;
    int in_ld, in_off, Cin;
    const __bf16* gout;
    int g_ld, g_off, Cout;
    float* dw;
    int ldw;
    int D0, D1, D2;
    int64_t vox_per_block;  // set by the launcher
};

hipError_t launch_conv3(const ConvArgs& a, hipStream_t s);
hipError_t launch_conv3_wgrad(WgradArgs a, hipStream_t s);
hipError_t launch_bias_grad(const __bf16* g, int ld, int off, int C, int64_t V, float* db, hipStream_t s);
hipError_t launch_pool(const __bf16* in, int in_ld, int in_off, int C, int D0, int D1, int D2, __bf16* out,
                       hipStream_t s);
hipError_t launch_pool_bwd(const __bf16* gp, int C, int D0, int D1, int D2, __bf16* gin, int g_ld, int g_off,
                           const __bf16* mask, int m_ld, int m_off, int accumulate, hipStream_t s);
hipError_t launch_upsample(const __bf16* low, int C, int D0, int D1, int D2, __bf16* out, int o_ld, int o_off,
                           hipStream_t s);
hipError_t launch_upsample_bwd(const __bf16* g, int g_ld, int g_off, int C, int D0, int D1, int D2, __bf16* glow,
                               hipStream_t s);
hipError_t launch_input8(const float* x, int64_t V, float inv_scale, __bf16* x8, hipStream_t s);
hipError_t launch_head(const __bf16* e, int C, int64_t V, const float* w, const float* x, float scale, float* y,
                       const float* t, float* g, double* loss, hipStream_t s);
hipError_t launch_head_bwd(const __bf16* e, int C, int64_t V, const float* w, const float* g, __bf16* ge,
                           float* dw, hipStream_t s);
hipError_t launch_pack(const float* W, int Cin, int Cout, __bf16* wf, int Kpad, __bf16* wd, int Kdpad,
                       hipStream_t s);
// Tiled-order forward weights (see ConvArgs::wt); Ktpad = (Cin / 8) * 224.
hipError_t launch_pack_tiled(const float* W, int Cin, int Cout, __bf16* wt, int Ktpad, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t n, float b1, float b2, float lr_t,
                       float eps_t, hipStream_t s);
hipError_t launch_sum(const float* x, int64_t n, double* out, hipStream_t s);

}  // namespace nn
}  // namespace xr
