// Host-side internals shared by the C ABI runtime (xr_runtime.cpp) and the
// batched solver / autotuner (xr_batch.cpp): context and plan objects, error
// helpers and the forward / backward launch helpers.  Not part of the public
// C ABI (include/xr.h).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "../../include/xr.h"
#include "xr_internal.hpp"

struct xr_context {
    int device = 0;
    int n_cu = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    std::string err;
    int32_t* d_codes = nullptr;  // per-call view lists
    size_t codes_cap = 0;
    uint8_t* d_flags = nullptr;  // per-call accumulate flags
    size_t flags_cap = 0;
    double* d_loss = nullptr;
    float* d_scal = nullptr;  // [0] = 1.0f, [1] = max |voxel value| of the last forward launch
    bool timing = false;
    struct Ev {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;
    std::vector<hipEvent_t> pool;
    // autotuner results per (N, U, V, P, axis)
    std::vector<std::pair<std::array<int, 5>, xr_tiles>> tune_cache;
};

struct xr_plan {
    xr_context* ctx = nullptr;
    xr_geometry g{};
    int n_cond = 0, n_basis = 0, conv = 0, dir = 0;
    std::vector<float> cond_R, cond_t, cond_offset, basis, grating;
    std::vector<char> axis_cond;  // 1: every view of the condition has au[2] = ad[2] = 0
    xr::ViewGeom* d_vg = nullptr;
    xr::ChordGeom* d_cg = nullptr;
    float* d_g2 = nullptr;    // [n_cond*P][K]: (S_k.G)^2
    float* d_xw = nullptr;    // [n_cond*P][5][kXwStride]: XTT weight constants of each view (BwdArgs::xw)
    float* d_norm = nullptr;  // [n_cond][P][V][U]: L/(factor+1e-5)
    int32_t* d_bands = nullptr;  // [n_cond*P][n_bands][N]: candidate columns xs | xe << 16 of each volume row
    xr::Dims dims{};
    xr::Basis bas{};
    float xtt_wsum = 0.f;  // sum_k |S_k|^2 max_view (S_k.G)^2: XTT splat value per unit |mu|
    float xtt_wsum_orth = 0.f;  // sum_k |S_k|^4 max_view |G|^2: the same bound under any rotation of the basis
    const float* d_orient = nullptr;  // caller's orientation field (xr_plan_set_orientation), or null
    xr_tiles tiles{};      // projector tile shape (zeros: defaults)
    float max_factor = 0.f;     // largest per-pixel splat weight sum (set with d_norm)
    float* d_fmax = nullptr;    // device scalar receiving it during the weight pass
};

namespace xr {
namespace host {

int fail(xr_context* ctx, int code, const char* fmt, ...);
// 2.0f * (float)M_PI * (float)n / (float)P in FP32, step by step (ir.cu:1860)
float theta_f32(int n, int P);

#define XR_HIP(ctx, call)                                                                                      \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) return fail((ctx), XR_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define XR_TRY(expr)          \
    do {                      \
        int r_ = (expr);      \
        if (r_ != XR_OK) return r_; \
    } while (0)

// Every device allocation of the library goes through device_malloc, so a
// test can make allocations fail from the host (xr_test_fail_allocations)
// and check that each entry point reports the error and leaves the library
// usable.
hipError_t device_malloc(void** p, size_t bytes);

template <class T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        if (count == 0) return hipSuccess;
        return device_malloc((void**)&p, count * sizeof(T));
    }
};

int check_geometry(xr_context* ctx, const xr_geometry* g);
int check_scan(xr_context* ctx, const xr_scan* s);
hipEvent_t get_event(xr_context* ctx);

// Launch wrapper: brackets the launch with events when timing is enabled.
template <class F>
int timed(xr_context* ctx, int kind, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->timing) {
        a = get_event(ctx);
        b = get_event(ctx);
        if (a && b) XR_HIP(ctx, hipEventRecord(a, ctx->stream));
    }
    const hipError_t e = launch();
    if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    if (ctx->timing && a && b) {
        XR_HIP(ctx, hipEventRecord(b, ctx->stream));
        ctx->events.push_back({kind, a, b});
    }
    return XR_OK;
}

// vol_cond_stride: plain volumes of several conditions side by side (N^3), or 0
// for a single volume; code_base: proj/sino hold the images of codes >= code_base.
// vmax_known: device max |voxel| of vol already measured (the solver's
// previous backprojection), or null to measure it here.
// no_alloc: the call makes no allocation of any kind (the plan's tables must
// already exist; max |voxel| goes to vmax_buf, cleared by a fill kernel; no
// timing events).
int forward_launch(xr_plan* p, int kind, const float* vol, int64_t vol_cond_stride, float* proj, const float* sino,
                   int code_base, const int32_t* d_codes, const uint8_t* d_flags, int n_slots, int out_mode,
                   double* d_loss, const float* vmax_known = nullptr, float* vmax_buf = nullptr,
                   bool no_alloc = false);
// vmax_out (MLEM / ART kinds): receives the atomic max of |updated voxel| as
// float bits; the caller clears it first.
int backward_launch(xr_plan* p, int kind, const float* proj, int code_base, const int32_t* d_codes, int n_views,
                    float* vol, float* tmp, float* factor, float alpha, int overwrite, unsigned* vmax_out = nullptr);
// Per-pixel L/(factor + 1e-5) table of the plan (computed once).
int ensure_norm(xr_plan* p);
// Detector rows per forward row chunk for the plan's tile setting and the
// accumulator width of `kind` (FWD_WEIGHT: 64-bit words, else 32-bit).
int forward_rows(const xr_plan* p, int kind);

}  // namespace host
}  // namespace xr
