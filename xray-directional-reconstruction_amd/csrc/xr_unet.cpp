// Learned warm start: host side of the 3-level 3-D U-Net (C ABI in
// include/xr.h).  No counterpart in the reference.
//
// Network (base width C, 3x3x3 convolutions + ReLU, 2x2x2 average pooling,
// nearest x2 upsampling, channel concatenation of the skips):
//   level 1:  x8 (8: x/s and 7 zeros) -L0-> a1 (C) -L1-> skip1 (C)
//   level 2:  pool(skip1) -L2-> b1 (2C) -L3-> skip2 (2C)
//   level 3:  pool(skip2) -L4-> c1 (4C) -L5-> c2 (4C)
//   level 2:  [up(c2), skip2] (6C) -L6-> d1 (2C) -L7-> d2 (2C)
//   level 1:  [up(d2), skip1] (3C) -L8-> e1 (C) -L9-> e2 (C)
//   head:     y = x + s * (w . e2 + b)          (1x1x1, residual, FP32)
// s is the input's mean (or a caller-given scale), so the network sees
// volumes of unit mean.  The head starts at zero: an untrained network is the
// identity.  Parameters are FP32 masters (flat: per layer W[Cout][27][Cin],
// b[Cout]; head w[C], b), packed to BF16 for the kernels after every update.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/xr.h"
#include "xr_host.hpp"
#include "xr_random.hpp"
#include "xr_unet_internal.hpp"

using namespace xr;
using namespace xr::host;
using namespace xr::nn;

namespace {

constexpr int kLayers = 10;

struct Layer {
    int Cin, Cout;
    int64_t w_off, b_off;  // offsets in the flat parameter array
    int Kpad, Kdpad;       // packed BF16 row lengths (forward, data gradient)
    __bf16* wf = nullptr;  // [round16(Cout)][Kpad]
    __bf16* wd = nullptr;  // [round16(Cin)][Kdpad]
    int Ktpad;             // tiled-order forward row length: (Cin / 8) * 28 taps * 8
    __bf16* wt = nullptr;  // [round16(Cout)][Ktpad]
};

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

struct Bf16Buf {
    __bf16* p = nullptr;
    size_t n = 0;
    ~Bf16Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t count) {
        if (count <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = device_malloc((void**)&p, count * sizeof(__bf16));
        if (e == hipSuccess) n = count;
        return e;
    }
};

}  // namespace

struct xr_unet {
    xr_context* ctx = nullptr;
    int C = 24;
    Layer L[kLayers];
    int64_t n_params = 0, head_off = 0;
    float* d_p = nullptr;  // FP32 parameters
    float* d_g = nullptr;  // gradients
    float* d_m = nullptr;  // Adam moments
    float* d_v = nullptr;
    int64_t step = 0;
    double* d_red = nullptr;  // [0] sum (scale), [1] loss
    float last_scale = 1.f;
    bool fuse_upsampling = true;  // inference only (training reads the concatenations)
    // activations (forward) and gradients (backward), sized for the last dims
    Bf16Buf x8, a1, cat1, p1, b1, cat2, p2, c1, c2, d1, d2, e1, e2;
    Bf16Buf g_e2, g_e1, g_cat1, g_d2, g_d1, g_cat2, g_c2, g_c1, g_p2, g_b1, g_p1, g_a1;
    float* d_ghead = nullptr;
    size_t ghead_n = 0;

    ~xr_unet() {
        if (ctx) (void)hipStreamSynchronize(ctx->stream);
        for (auto& l : L) {
            if (l.wf) (void)hipFree(l.wf);
            if (l.wd) (void)hipFree(l.wd);
            if (l.wt) (void)hipFree(l.wt);
        }
        for (float* p : {d_p, d_g, d_m, d_v, d_ghead})
            if (p) (void)hipFree(p);
        if (d_red) (void)hipFree(d_red);
    }
};

namespace {

int pack_all(xr_unet* n) {
    xr_context* ctx = n->ctx;
    for (int i = 0; i < kLayers; i++) {
        Layer& l = n->L[i];
        const hipError_t e = launch_pack(n->d_p + l.w_off, l.Cin, l.Cout, l.wf, l.Kpad, i == 0 ? nullptr : l.wd,
                                         l.Kdpad, ctx->stream);
        if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "weight packing failed: %s", hipGetErrorString(e));
        const hipError_t e2 = launch_pack_tiled(n->d_p + l.w_off, l.Cin, l.Cout, l.wt, l.Ktpad, ctx->stream);
        if (e2 != hipSuccess) return fail(ctx, XR_ERR_HIP, "weight packing failed: %s", hipGetErrorString(e2));
    }
    return XR_OK;
}

int check_dims(xr_context* ctx, int D0, int D1, int D2) {
    if (D0 <= 0 || D1 <= 0 || D2 <= 0 || D0 % 4 || D1 % 4 || D2 % 4)
        return fail(ctx, XR_ERR_INVALID, "U-Net dims must be positive multiples of 4 (got %d x %d x %d)", D0, D1, D2);
    return XR_OK;
}

int ensure_buffers(xr_unet* n, int64_t V1, bool training) {
    xr_context* ctx = n->ctx;
    const int C = n->C;
    const int64_t V2 = V1 / 8, V3 = V1 / 64;
    XR_HIP(ctx, n->x8.ensure(V1 * 8));
    XR_HIP(ctx, n->a1.ensure(V1 * C));
    XR_HIP(ctx, n->cat1.ensure(V1 * 3 * C));
    XR_HIP(ctx, n->p1.ensure(V2 * C));
    XR_HIP(ctx, n->b1.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->cat2.ensure(V2 * 6 * C));
    XR_HIP(ctx, n->p2.ensure(V3 * 2 * C));
    XR_HIP(ctx, n->c1.ensure(V3 * 4 * C));
    XR_HIP(ctx, n->c2.ensure(V3 * 4 * C));
    XR_HIP(ctx, n->d1.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->d2.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->e1.ensure(V1 * C));
    XR_HIP(ctx, n->e2.ensure(V1 * C));
    if (!training) return XR_OK;
    XR_HIP(ctx, n->g_e2.ensure(V1 * C));
    XR_HIP(ctx, n->g_e1.ensure(V1 * C));
    XR_HIP(ctx, n->g_cat1.ensure(V1 * 3 * C));
    XR_HIP(ctx, n->g_d2.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->g_d1.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->g_cat2.ensure(V2 * 6 * C));
    XR_HIP(ctx, n->g_c2.ensure(V3 * 4 * C));
    XR_HIP(ctx, n->g_c1.ensure(V3 * 4 * C));
    XR_HIP(ctx, n->g_p2.ensure(V3 * 2 * C));
    XR_HIP(ctx, n->g_b1.ensure(V2 * 2 * C));
    XR_HIP(ctx, n->g_p1.ensure(V2 * C));
    XR_HIP(ctx, n->g_a1.ensure(V1 * C));
    if (n->ghead_n < (size_t)V1) {
        if (n->d_ghead) (void)hipFree(n->d_ghead);
        n->d_ghead = nullptr;
        n->ghead_n = 0;
        XR_HIP(ctx, device_malloc((void**)&n->d_ghead, sizeof(float) * V1));
        n->ghead_n = V1;
    }
    return XR_OK;
}

#define XR_K(ctx, expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return fail((ctx), XR_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Forward convolution of layer i: in (ld, off) at dims -> out (ld, off), ReLU.
ConvArgs conv_args(xr_unet* n, int i, const __bf16* in, int in_ld, int in_off, __bf16* out, int out_ld,
                   int out_off, int D0, int D1, int D2) {
    const Layer& l = n->L[i];
    ConvArgs a{};
    a.in = in;
    a.in_ld = in_ld;
    a.in_off = in_off;
    a.Cin = l.Cin;
    a.w = l.wf;
    a.Kpad = l.Kpad;
    a.wt = l.wt;
    a.Ktpad = l.Ktpad;
    a.bias = n->d_p + l.b_off;
    a.out = out;
    a.out_ld = out_ld;
    a.out_off = out_off;
    a.Cout = l.Cout;
    a.D0 = D0;
    a.D1 = D1;
    a.D2 = D2;
    a.relu = 1;
    return a;
}

int conv_fwd(xr_unet* n, int i, const __bf16* in, int in_ld, int in_off, __bf16* out, int out_ld, int out_off,
             int D0, int D1, int D2) {
    XR_K(n->ctx, launch_conv3(conv_args(n, i, in, in_ld, in_off, out, out_ld, out_off, D0, D1, D2), n->ctx->stream));
    return XR_OK;
}

// Decoder convolution i over the concatenation [upsample(low) | skip] (low:
// half-resolution, up_C channels).  fused: the concatenation is read in place,
// the nearest x2 upsampling in the convolution's halo fetch (the caller has
// checked that the tiled kernel takes the dims); otherwise the upsampled
// channels are first written into cat, whose channels [up_C, ...) hold the
// skip (training: the weight gradient reads cat).  The output is the same.
int conv_up(xr_unet* n, int i, bool fused, const __bf16* low, int up_C, const __bf16* skip, int skip_ld,
            __bf16* cat, int cat_ld, __bf16* out, int out_ld, int D0, int D1, int D2) {
    ConvArgs a;
    if (fused) {
        a = conv_args(n, i, skip, skip_ld, 0, out, out_ld, 0, D0, D1, D2);
        a.up = low;
        a.up_ld = up_C;
        a.up_off = 0;
        a.up_chunks = up_C / 8;
    } else {
        XR_K(n->ctx, launch_upsample(low, up_C, D0, D1, D2, cat, cat_ld, 0, n->ctx->stream));
        a = conv_args(n, i, cat, cat_ld, 0, out, out_ld, 0, D0, D1, D2);
    }
    XR_K(n->ctx, launch_conv3(a, n->ctx->stream));
    return XR_OK;
}

// Backward of layer i given its output gradient gout (already masked by the
// ReLU): weight and bias gradients, and (gin != null) the input gradient
// masked by (mask > 0) when mask is given.
int conv_bwd(xr_unet* n, int i, const __bf16* in, int in_ld, int in_off, const __bf16* gout, int g_ld, int g_off,
             __bf16* gin, int gin_ld, const __bf16* mask, int m_ld, int m_off, int D0, int D1, int D2) {
    xr_context* ctx = n->ctx;
    const Layer& l = n->L[i];
    WgradArgs w{};
    w.in = in;
    w.in_ld = in_ld;
    w.in_off = in_off;
    w.Cin = l.Cin;
    w.gout = gout;
    w.g_ld = g_ld;
    w.g_off = g_off;
    w.Cout = l.Cout;
    w.dw = n->d_g + l.w_off;
    w.ldw = 27 * l.Cin;
    w.D0 = D0;
    w.D1 = D1;
    w.D2 = D2;
    XR_K(ctx, launch_conv3_wgrad(w, ctx->stream));
    const int64_t V = (int64_t)D0 * D1 * D2;
    XR_K(ctx, launch_bias_grad(gout, g_ld, g_off, l.Cout, V, n->d_g + l.b_off, ctx->stream));
    if (!gin) return XR_OK;
    ConvArgs a{};
    a.in = gout;
    a.in_ld = g_ld;
    a.in_off = g_off;
    a.Cin = l.Cout;
    a.w = l.wd;
    a.Kpad = l.Kdpad;
    a.out = gin;
    a.out_ld = gin_ld;
    a.out_off = 0;
    a.Cout = l.Cin;
    a.mask = mask;
    a.mask_ld = m_ld;
    a.mask_off = m_off;
    a.D0 = D0;
    a.D1 = D1;
    a.D2 = D2;
    XR_K(ctx, launch_conv3(a, ctx->stream));
    return XR_OK;
}

// Forward pass; with t != null also the loss (FP64 into d_red[1]) and the
// head gradient.
int forward(xr_unet* n, const float* x, float* y, const float* t, int D0, int D1, int D2, float scale) {
    xr_context* ctx = n->ctx;
    const int C = n->C;
    const int64_t V1 = (int64_t)D0 * D1 * D2;
    const int E0 = D0 / 2, E1 = D1 / 2, E2 = D2 / 2, F0 = D0 / 4, F1 = D1 / 4, F2 = D2 / 4;
    if (!(scale > 0.f)) {
        XR_K(ctx, launch_fill(reinterpret_cast<float*>(n->d_red), 4, 0.0f, ctx->stream));
        XR_K(ctx, launch_sum(x, V1, n->d_red, ctx->stream));
        double sum = 0.0;
        XR_HIP(ctx, hipMemcpyAsync(&sum, n->d_red, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        scale = (float)(sum / (double)V1);
        if (!(scale > 0.f)) scale = 1.f;
    }
    n->last_scale = scale;
    XR_K(ctx, launch_input8(x, V1, 1.0f / scale, n->x8.p, ctx->stream));
    // Inference reads each decoder concatenation in place (fused): the skips
    // go to compact buffers that are free until the decoder's last
    // convolution of their level (e2 at level 0, d2 at level 1); training
    // writes them into the concatenations cat1 / cat2, which it keeps.
    const bool fuse = !t && n->fuse_upsampling;
    const bool fuse0 = fuse && conv3_tiled_ok(conv_args(n, 8, nullptr, 3 * C, 0, nullptr, C, 0, D0, D1, D2));
    const bool fuse1 = fuse && conv3_tiled_ok(conv_args(n, 6, nullptr, 6 * C, 0, nullptr, 2 * C, 0, E0, E1, E2));
    __bf16* const skip0 = fuse0 ? n->e2.p : n->cat1.p;
    const int s0_ld = fuse0 ? C : 3 * C, s0_off = fuse0 ? 0 : 2 * C;
    __bf16* const skip1 = fuse1 ? n->d2.p : n->cat2.p;
    const int s1_ld = fuse1 ? 2 * C : 6 * C, s1_off = fuse1 ? 0 : 4 * C;
    XR_TRY(conv_fwd(n, 0, n->x8.p, 8, 0, n->a1.p, C, 0, D0, D1, D2));
    XR_TRY(conv_fwd(n, 1, n->a1.p, C, 0, skip0, s0_ld, s0_off, D0, D1, D2));
    XR_K(ctx, launch_pool(skip0, s0_ld, s0_off, C, D0, D1, D2, n->p1.p, ctx->stream));
    XR_TRY(conv_fwd(n, 2, n->p1.p, C, 0, n->b1.p, 2 * C, 0, E0, E1, E2));
    XR_TRY(conv_fwd(n, 3, n->b1.p, 2 * C, 0, skip1, s1_ld, s1_off, E0, E1, E2));
    XR_K(ctx, launch_pool(skip1, s1_ld, s1_off, 2 * C, E0, E1, E2, n->p2.p, ctx->stream));
    XR_TRY(conv_fwd(n, 4, n->p2.p, 2 * C, 0, n->c1.p, 4 * C, 0, F0, F1, F2));
    XR_TRY(conv_fwd(n, 5, n->c1.p, 4 * C, 0, n->c2.p, 4 * C, 0, F0, F1, F2));
    XR_TRY(conv_up(n, 6, fuse1, n->c2.p, 4 * C, skip1, s1_ld, n->cat2.p, 6 * C, n->d1.p, 2 * C, E0, E1, E2));
    XR_TRY(conv_fwd(n, 7, n->d1.p, 2 * C, 0, n->d2.p, 2 * C, 0, E0, E1, E2));
    XR_TRY(conv_up(n, 8, fuse0, n->d2.p, 2 * C, skip0, s0_ld, n->cat1.p, 3 * C, n->e1.p, C, D0, D1, D2));
    XR_TRY(conv_fwd(n, 9, n->e1.p, C, 0, n->e2.p, C, 0, D0, D1, D2));
    if (t) XR_K(ctx, launch_fill(reinterpret_cast<float*>(n->d_red + 1), 2, 0.0f, ctx->stream));
    XR_K(ctx, launch_head(n->e2.p, C, V1, n->d_p + n->head_off, x, scale, y, t, t ? n->d_ghead : nullptr,
                          t ? n->d_red + 1 : nullptr, ctx->stream));
    return XR_OK;
}

int backward(xr_unet* n, int D0, int D1, int D2) {
    xr_context* ctx = n->ctx;
    const int C = n->C;
    const int64_t V1 = (int64_t)D0 * D1 * D2;
    const int E0 = D0 / 2, E1 = D1 / 2, E2 = D2 / 2, F0 = D0 / 4, F1 = D1 / 4, F2 = D2 / 4;
    XR_K(ctx, launch_fill(n->d_g, n->n_params, 0.0f, ctx->stream));
    XR_K(ctx, launch_head_bwd(n->e2.p, C, V1, n->d_p + n->head_off, n->d_ghead, n->g_e2.p, n->d_g + n->head_off,
                              ctx->stream));
    // level 1 (up path)
    XR_TRY(conv_bwd(n, 9, n->e1.p, C, 0, n->g_e2.p, C, 0, n->g_e1.p, C, n->e1.p, C, 0, D0, D1, D2));
    XR_TRY(conv_bwd(n, 8, n->cat1.p, 3 * C, 0, n->g_e1.p, C, 0, n->g_cat1.p, 3 * C, n->cat1.p, 3 * C, 0, D0, D1,
                    D2));
    XR_K(ctx, launch_upsample_bwd(n->g_cat1.p, 3 * C, 0, 2 * C, D0, D1, D2, n->g_d2.p, ctx->stream));
    // level 2 (up path)
    XR_TRY(conv_bwd(n, 7, n->d1.p, 2 * C, 0, n->g_d2.p, 2 * C, 0, n->g_d1.p, 2 * C, n->d1.p, 2 * C, 0, E0, E1, E2));
    XR_TRY(conv_bwd(n, 6, n->cat2.p, 6 * C, 0, n->g_d1.p, 2 * C, 0, n->g_cat2.p, 6 * C, n->cat2.p, 6 * C, 0, E0, E1,
                    E2));
    XR_K(ctx, launch_upsample_bwd(n->g_cat2.p, 6 * C, 0, 4 * C, E0, E1, E2, n->g_c2.p, ctx->stream));
    // level 3
    XR_TRY(conv_bwd(n, 5, n->c1.p, 4 * C, 0, n->g_c2.p, 4 * C, 0, n->g_c1.p, 4 * C, n->c1.p, 4 * C, 0, F0, F1, F2));
    XR_TRY(conv_bwd(n, 4, n->p2.p, 2 * C, 0, n->g_c1.p, 4 * C, 0, n->g_p2.p, 2 * C, nullptr, 0, 0, F0, F1, F2));
    XR_K(ctx, launch_pool_bwd(n->g_p2.p, 2 * C, E0, E1, E2, n->g_cat2.p, 6 * C, 4 * C, n->cat2.p, 6 * C, 4 * C, 1,
                              ctx->stream));
    // level 2 (down path): skip2's gradient is the slice [4C, 6C) of g_cat2
    XR_TRY(conv_bwd(n, 3, n->b1.p, 2 * C, 0, n->g_cat2.p, 6 * C, 4 * C, n->g_b1.p, 2 * C, n->b1.p, 2 * C, 0, E0, E1,
                    E2));
    XR_TRY(conv_bwd(n, 2, n->p1.p, C, 0, n->g_b1.p, 2 * C, 0, n->g_p1.p, C, nullptr, 0, 0, E0, E1, E2));
    XR_K(ctx, launch_pool_bwd(n->g_p1.p, C, D0, D1, D2, n->g_cat1.p, 3 * C, 2 * C, n->cat1.p, 3 * C, 2 * C, 1,
                              ctx->stream));
    // level 1 (down path): skip1's gradient is the slice [2C, 3C) of g_cat1
    XR_TRY(conv_bwd(n, 1, n->a1.p, C, 0, n->g_cat1.p, 3 * C, 2 * C, n->g_a1.p, C, n->a1.p, C, 0, D0, D1, D2));
    XR_TRY(conv_bwd(n, 0, n->x8.p, 8, 0, n->g_a1.p, C, 0, nullptr, 0, nullptr, 0, 0, D0, D1, D2));
    return XR_OK;
}

}  // namespace

extern "C" {

int xr_unet_create(xr_context* ctx, int base_channels, uint64_t seed, xr_unet** out) {
    if (!ctx || !out) return XR_ERR_INVALID;
    *out = nullptr;
    const int C = base_channels;
    if (C <= 0 || C % 8 || C > 32) return fail(ctx, XR_ERR_INVALID, "base channels must be 8, 16, 24 or 32");
    std::unique_ptr<xr_unet> n(new (std::nothrow) xr_unet());
    if (!n) return fail(ctx, XR_ERR_ALLOC, "U-Net allocation failed");
    n->ctx = ctx;
    n->C = C;
    const int io[kLayers][2] = {{8, C},         {C, C},         {C, 2 * C},     {2 * C, 2 * C}, {2 * C, 4 * C},
                                {4 * C, 4 * C}, {6 * C, 2 * C}, {2 * C, 2 * C}, {3 * C, C},     {C, C}};
    int64_t off = 0;
    for (int i = 0; i < kLayers; i++) {
        Layer& l = n->L[i];
        l.Cin = io[i][0];
        l.Cout = io[i][1];
        l.w_off = off;
        off += (int64_t)l.Cout * 27 * l.Cin;
        l.b_off = off;
        off += l.Cout;
        l.Kpad = round_up(27 * l.Cin, 32);
        l.Kdpad = round_up(27 * l.Cout, 32);
        l.Ktpad = l.Cin / 8 * 28 * 8;
    }
    n->head_off = off;
    off += C + 1;
    n->n_params = off;
    for (int i = 0; i < kLayers; i++) {
        Layer& l = n->L[i];
        XR_HIP(ctx, device_malloc((void**)&l.wf, sizeof(__bf16) * (size_t)round_up(l.Cout, 16) * l.Kpad));
        XR_HIP(ctx, device_malloc((void**)&l.wd, sizeof(__bf16) * (size_t)round_up(l.Cin, 16) * l.Kdpad));
        XR_HIP(ctx, device_malloc((void**)&l.wt, sizeof(__bf16) * (size_t)round_up(l.Cout, 16) * l.Ktpad));
    }
    for (float** p : {&n->d_p, &n->d_g, &n->d_m, &n->d_v}) XR_HIP(ctx, device_malloc((void**)p, sizeof(float) * off));
    XR_HIP(ctx, device_malloc((void**)&n->d_red, sizeof(double) * 2));
    // He-normal weights from a Philox stream of `seed` (Box-Muller), zero
    // biases; layer 0 only for its real input channel; zero head.
    std::vector<float> h(off, 0.0f);
    for (int i = 0; i < kLayers; i++) {
        const Layer& l = n->L[i];
        const int fan_in = 27 * (i == 0 ? 1 : l.Cin);
        const double sd = std::sqrt(2.0 / fan_in);
        for (int64_t j = 0; j < (int64_t)l.Cout * 27 * l.Cin; j++) {
            const int ci = (int)(j % l.Cin);
            if (i == 0 && ci != 0) continue;
            uint32_t r[4];
            philox4x32(seed, (uint64_t)(l.w_off + j), 0, r);
            const double u1 = ((double)r[0] + 1.0) / 4294967296.0, u2 = (double)r[1] / 4294967296.0;
            h[l.w_off + j] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
        }
    }
    XR_HIP(ctx, hipMemcpy(n->d_p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice));
    XR_HIP(ctx, hipMemset(n->d_g, 0, sizeof(float) * off));
    XR_HIP(ctx, hipMemset(n->d_m, 0, sizeof(float) * off));
    XR_HIP(ctx, hipMemset(n->d_v, 0, sizeof(float) * off));
    XR_TRY(pack_all(n.get()));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = n.release();
    return XR_OK;
}

void xr_unet_destroy(xr_unet* n) { delete n; }

int64_t xr_unet_num_params(const xr_unet* n) { return n ? n->n_params : 0; }

int xr_unet_get_params(xr_unet* n, float* host) {
    if (!n || !host) return XR_ERR_INVALID;
    XR_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    XR_HIP(n->ctx, hipMemcpy(host, n->d_p, sizeof(float) * n->n_params, hipMemcpyDeviceToHost));
    return XR_OK;
}

int xr_unet_set_params(xr_unet* n, const float* host) {
    if (!n || !host) return XR_ERR_INVALID;
    XR_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    XR_HIP(n->ctx, hipMemcpy(n->d_p, host, sizeof(float) * n->n_params, hipMemcpyHostToDevice));
    XR_TRY(pack_all(n));
    XR_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return XR_OK;
}

int xr_unet_get_grads(xr_unet* n, float* host) {
    if (!n || !host) return XR_ERR_INVALID;
    XR_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    XR_HIP(n->ctx, hipMemcpy(host, n->d_g, sizeof(float) * n->n_params, hipMemcpyDeviceToHost));
    return XR_OK;
}

int xr_unet_forward(xr_unet* n, const float* x, float* y, int D0, int D1, int D2, float scale) {
    if (!n || !x || !y) return XR_ERR_INVALID;
    XR_TRY(check_dims(n->ctx, D0, D1, D2));
    XR_TRY(ensure_buffers(n, (int64_t)D0 * D1 * D2, false));
    return forward(n, x, y, nullptr, D0, D1, D2, scale);
}

int xr_unet_loss_grad(xr_unet* n, const float* x, const float* t, int D0, int D1, int D2, float scale,
                      double* loss) {
    if (!n || !x || !t) return XR_ERR_INVALID;
    xr_context* ctx = n->ctx;
    XR_TRY(check_dims(ctx, D0, D1, D2));
    XR_TRY(ensure_buffers(n, (int64_t)D0 * D1 * D2, true));
    XR_TRY(forward(n, x, nullptr, t, D0, D1, D2, scale));
    XR_TRY(backward(n, D0, D1, D2));
    if (loss) {
        double l = 0.0;
        XR_HIP(ctx, hipMemcpyAsync(&l, n->d_red + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *loss = l / (double)((int64_t)D0 * D1 * D2);
    }
    return XR_OK;
}

int xr_unet_adam_step(xr_unet* n, float lr, float beta1, float beta2, float eps) {
    if (!n) return XR_ERR_INVALID;
    xr_context* ctx = n->ctx;
    n->step++;
    const double c1 = 1.0 - std::pow((double)beta1, (double)n->step);
    const double c2 = 1.0 - std::pow((double)beta2, (double)n->step);
    // p -= lr/c1 * m / (sqrt(v/c2) + eps) = lr*sqrt(c2)/c1 * m / (sqrt(v) + eps*sqrt(c2))
    const float lr_t = (float)(lr * std::sqrt(c2) / c1), eps_t = (float)(eps * std::sqrt(c2));
    XR_K(ctx, launch_adam(n->d_p, n->d_g, n->d_m, n->d_v, n->n_params, beta1, beta2, lr_t, eps_t, ctx->stream));
    return pack_all(n);
}

float xr_unet_last_scale(const xr_unet* n) { return n ? n->last_scale : 0.f; }

int xr_unet_set_fused_upsampling(xr_unet* n, int enable) {
    if (!n) return XR_ERR_INVALID;
    n->fuse_upsampling = enable != 0;
    return XR_OK;
}

int xr_unet_save(xr_unet* n, const char* path) {
    if (!n || !path) return XR_ERR_INVALID;
    std::vector<float> h(n->n_params);
    XR_TRY(xr_unet_get_params(n, h.data()));
    FILE* f = fopen(path, "wb");
    if (!f) return fail(n->ctx, XR_ERR_IO, "cannot write %s", path);
    const int32_t hdr[4] = {0x544e5558 /* "XUNT" */, 1, n->C, (int32_t)n->n_params};
    const bool ok = fwrite(hdr, sizeof(hdr), 1, f) == 1 && fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
    fclose(f);
    if (!ok) return fail(n->ctx, XR_ERR_IO, "short write to %s", path);
    return XR_OK;
}

int xr_unet_load(xr_unet* n, const char* path) {
    if (!n || !path) return XR_ERR_INVALID;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(n->ctx, XR_ERR_IO, "cannot read %s", path);
    int32_t hdr[4] = {0, 0, 0, 0};
    std::vector<float> h(n->n_params);
    bool ok = fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == 0x544e5558 && hdr[1] == 1 && hdr[2] == n->C &&
              hdr[3] == (int32_t)n->n_params;
    ok = ok && fread(h.data(), sizeof(float), h.size(), f) == h.size();
    fclose(f);
    if (!ok) return fail(n->ctx, XR_ERR_IO, "%s is not a U-Net file of this width", path);
    return xr_unet_set_params(n, h.data());
}

}  // extern "C"
