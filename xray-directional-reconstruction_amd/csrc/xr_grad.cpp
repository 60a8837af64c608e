// Gradient-domain phase-contrast reconstruction (IR::gradReconstruct,
// reconstruct.cu:158-328) and its Poisson integration (poissonSolveLDLT,
// poisson_cpu.cpp:15-112) behind the C ABI.
//
// The gradient slots reuse the projector kernels through an XR_CONV_GRAD plan
// of side N+1: the same voxel-driven splat / gather, with the source on the
// -y side, the half-voxel-shifted grid and a constant vox_size in place of the
// chord normalisation.  Slots run one after another; each is an OS-ART (or
// OS-EM) solve like xr_ir_reconstruct's, with the subset order of a fresh
// default-seeded mt19937_64, which is the order the reference's single engine
// gives every slot (it shuffles once per epoch for all of them).
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "xr_host.hpp"

using namespace xr;
using namespace xr::host;

namespace {

// subset s holds views (s + batch*j) % P, j < ceil(P/batch) (reconstruct.cu:240-242)
void subset_views(int sub, int batch, int P, std::vector<int>& out) {
    const int ss = (P + batch - 1) / batch;
    out.resize(ss);
    for (int j = 0; j < ss; j++) out[j] = (sub + batch * j) % P;
}

}  // namespace

int xr_poisson_solve(xr_context* ctx, int n, const float* gx, const float* gy, float* vol) {
    if (!ctx || !gx || !gy || !vol) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (n < 3 || n > 4096) return fail(ctx, XR_ERR_INVALID, "poisson solve needs 3 <= n <= 4096 (got %d)", n);
    const int M = n - 2;
    const double h = (double)(M + 1);
    // S[j][k] = sin(pi (j+1)(k+1) / h), with the product reduced mod 2h so the
    // argument stays in [0, 2 pi)
    std::vector<double> S((size_t)M * M), e(M);
    for (int j = 0; j < M; j++) {
        for (int k = 0; k < M; k++) {
            const int64_t m = ((int64_t)(j + 1) * (k + 1)) % (2 * (int64_t)(M + 1));
            S[(size_t)j * M + k] = std::sin(M_PI * (double)m / h);
        }
        e[j] = 2.0 - 2.0 * std::cos(M_PI * (double)(j + 1) / h);
    }
    const int64_t slab = (int64_t)n * M * M;
    DevBuf<double> dS, de, dB, dT;
    XR_HIP(ctx, dS.alloc(S.size()));
    XR_HIP(ctx, de.alloc(e.size()));
    XR_HIP(ctx, dB.alloc(slab));
    XR_HIP(ctx, dT.alloc(slab));
    XR_HIP(ctx, hipMemcpyAsync(dS.p, S.data(), S.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    XR_HIP(ctx, hipMemcpyAsync(de.p, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // the slice borders are the identity rows with zero right side: zero
    XR_TRY(timed(ctx, 3, [&] { return launch_fill(vol, (int64_t)n * n * n, 0.0f, ctx->stream); }));
    XR_TRY(timed(ctx, 3, [&] { return launch_poisson_rhs(gx, gy, dB.p, n, ctx->stream); }));
    XR_TRY(timed(ctx, 3, [&] { return launch_poisson_dst(dS.p, de.p, dB.p, dT.p, vol, n, ctx->stream); }));
    // the workspace is freed on return: wait for the kernels that read it
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_grad_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, const float* sino, float* vol,
                        float* const* grad, int epochs, int batch, int direction, int method, float lambda,
                        int block_size, xr_stats* stats) {
    if (!ctx) return XR_ERR_INVALID;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (!sino || !vol) return fail(ctx, XR_ERR_INVALID, "sinogram / volume missing");
    if (g->n_vox < 3) return fail(ctx, XR_ERR_INVALID, "gradient reconstruction needs N >= 3");
    if (epochs < 0 || batch <= 0 || batch > g->n_proj) return fail(ctx, XR_ERR_INVALID, "bad epochs / batch");
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    if (block_size <= 0) return fail(ctx, XR_ERR_INVALID, "block_size must be positive");
    const int N = g->n_vox, N1 = N + 1, P = g->n_proj, C = s->n_cond;
    xr_geometry g1 = *g;
    g1.n_vox = N1;  // geom.voxel + 1 (reconstruct.cu:161)
    xr_plan* plan = nullptr;
    XR_TRY(xr_plan_create(ctx, &g1, s, XR_CONV_GRAD, direction, &plan));
    std::unique_ptr<xr_plan, void (*)(xr_plan*)> guard(plan, xr_plan_destroy);
    const int ext = (int)std::min<int64_t>(N1, ((int64_t)N + block_size - 1) / block_size * block_size);
    // the layers z >= ext are held at zero while projecting: no projector pass
    // visits them, the weight pass of the plan's tables included (a 513-voxel
    // column is then 8 chunks of 64, not 9)
    plan->dims.nz = ext;
    XR_TRY(ensure_norm(plan));
    const int64_t V1 = (int64_t)N1 * N1 * N1, lenP = (int64_t)g->n_det_u * g->n_det_v, lenD = lenP * P;
    const int ss = (P + batch - 1) / batch;
    const size_t per_epoch = (size_t)batch * ss;
    DevBuf<float> dref, dvol, dsino0, dsino, dproj, dvmax, dcoef, dgx, dgy, dout;
    DevBuf<int32_t> dcodes;
    DevBuf<uint8_t> dflags;
    XR_HIP(ctx, dref.alloc(V1));
    XR_HIP(ctx, dvol.alloc(V1));
    XR_HIP(ctx, dsino0.alloc(lenD));
    XR_HIP(ctx, dsino.alloc(lenD));
    XR_HIP(ctx, dproj.alloc(lenD));
    XR_HIP(ctx, dvmax.alloc(1));
    XR_HIP(ctx, dcoef.alloc(P));
    XR_HIP(ctx, dgx.alloc(V1));
    XR_HIP(ctx, dgy.alloc(V1));
    XR_HIP(ctx, dout.alloc((int64_t)N * N * N));
    XR_HIP(ctx, dcodes.alloc(per_epoch * std::max(epochs, 1)));
    XR_HIP(ctx, dflags.alloc(per_epoch * std::max(epochs, 1)));
    XR_HIP(ctx, hipMemcpyAsync(dsino0.p, sino, sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
    const float alpha = (method == XR_METHOD_ART) ? lambda / (float)ss : 0.f;
    const int kind = (method == XR_METHOD_MLEM) ? BWD_MLEM : BWD_ART;
    const int out_mode = (method == XR_METHOD_MLEM) ? OUT_RATIO : OUT_RESIDUAL;
    std::vector<int32_t> codes(per_epoch * std::max(epochs, 1));
    std::vector<uint8_t> flags(codes.size());
    std::vector<float> coef(P);
    std::vector<int> sv;
    for (int c = 0; c < C; c++) {
        // the subset order every slot sees; views repeated within an epoch
        // accumulate onto their image (proj is zeroed once per epoch, :231)
        std::vector<int> order(batch);
        for (int i = 0; i < batch; i++) order[i] = i;
        std::mt19937_64 rng;
        for (int ep = 0; ep < epochs; ep++) {
            std::shuffle(order.begin(), order.end(), rng);
            std::vector<char> visited(P, 0);
            for (int pos = 0; pos < batch; pos++) {
                subset_views(order[pos], batch, P, sv);
                for (int j = 0; j < ss; j++) {
                    const size_t i = (size_t)ep * per_epoch + (size_t)pos * ss + j;
                    codes[i] = c * P + sv[j];
                    flags[i] = visited[sv[j]];
                    visited[sv[j]] = 1;
                }
            }
        }
        if (epochs > 0) {
            XR_HIP(ctx, hipMemcpyAsync(dcodes.p, codes.data(), codes.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                       ctx->stream));
            XR_HIP(ctx, hipMemcpyAsync(dflags.p, flags.data(), flags.size(), hipMemcpyHostToDevice, ctx->stream));
        }
        for (int d = 0; d < 3; d++) {
            const int slot = 3 * c + d;
            // sinogramGradientCoef: {cos, sin, 1}[d] of the FP32 angle of view n >= 0
            for (int n = 0; n < P; n++) {
                const double th = (double)theta_f32(n, P);
                coef[n] = d == 0 ? (float)std::cos(th) : d == 1 ? (float)std::sin(th) : 1.0f;
            }
            XR_HIP(ctx, hipMemcpyAsync(dcoef.p, coef.data(), sizeof(float) * P, hipMemcpyHostToDevice, ctx->stream));
            XR_TRY(timed(ctx, 3, [&] { return launch_scale_views(dsino0.p, dsino.p, dcoef.p, lenP, P, ctx->stream); }));
            // 0.01 everywhere (:168); the layers outside the launches are held
            // at zero while projecting and get their 0.01 back at the end
            XR_HIP(ctx, launch_fill(dvol.p, V1, 0.01f, ctx->stream));
            XR_HIP(ctx, launch_grad_faces(dvol.p, N1, ext, 0.0f, ctx->stream));
            bool vmax_known = false;
            for (int ep = 0; ep < epochs; ep++) {
                XR_HIP(ctx, hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
                for (int pos = 0; pos < batch; pos++) {
                    const size_t off = (size_t)ep * per_epoch + (size_t)pos * ss;
                    XR_TRY(forward_launch(plan, FWD_PLAIN, dvol.p, 0, dproj.p, dsino.p, c * P, dcodes.p + off,
                                          dflags.p + off, ss, out_mode, ctx->d_loss, vmax_known ? dvmax.p : nullptr));
                    XR_HIP(ctx, launch_fill(dvmax.p, 1, 0.0f, ctx->stream));
                    XR_TRY(backward_launch(plan, kind, dproj.p, c * P, dcodes.p + off, ss, dvol.p, nullptr, nullptr,
                                           alpha, 0, reinterpret_cast<unsigned*>(dvmax.p)));
                    // the update touched the skipped layers too: back to zero (the
                    // max left in dvmax may count them, a bound that stays safe)
                    XR_HIP(ctx, launch_grad_faces(dvol.p, N1, ext, 0.0f, ctx->stream));
                    vmax_known = true;
                }
                if (stats && stats->losses)
                    XR_HIP(ctx, hipMemcpyAsync(&stats->losses[(size_t)slot * epochs + ep], ctx->d_loss,
                                               sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            }
            XR_HIP(ctx, launch_grad_faces(dvol.p, N1, ext, 0.01f, ctx->stream));
            float* dst = (c == 0 && d == 0) ? dgx.p : (c == 0 && d == 1) ? dgy.p : dref.p;
            XR_TRY(xr_volume_from_columns(ctx, dvol.p, dst, N1, 1));
            if (grad && grad[slot])
                XR_HIP(ctx, hipMemcpyAsync(grad[slot], dst, sizeof(float) * V1, hipMemcpyDeviceToHost, ctx->stream));
        }
    }
    XR_TRY(xr_poisson_solve(ctx, N, dgx.p, dgy.p, dout.p));
    XR_HIP(ctx, hipMemcpyAsync(vol, dout.p, sizeof(float) * N * N * N, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}
