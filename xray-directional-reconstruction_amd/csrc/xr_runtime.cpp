// Host runtime behind the C ABI (include/xr.h): contexts, plans (per-view
// projection tables in FP64), device-level operators and the IR / XTT / FDK
// solvers that mirror reconstruct/src/reconstruct.cu.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <atomic>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/xr.h"
#include "xr_host.hpp"
#include "xr_internal.hpp"
#include "xr_random.hpp"

using namespace xr;

// Scratch of xr_project_noalloc: the view list, the accumulate flags, the
// FP64 loss sum and max |voxel| of one call live in this array instead of
// buffers the library allocates.  Each device holds its own copy (found with
// hipGetSymbolAddress at context creation); all calls on one device share it.
namespace {
constexpr int kNoAllocMaxViews = 16384;
constexpr size_t kScratchLoss = 0;                                    // double
constexpr size_t kScratchVmax = 8;                                    // float
constexpr size_t kScratchCodes = 256;                                 // int32 x kNoAllocMaxViews
constexpr size_t kScratchFlags = kScratchCodes + 4 * kNoAllocMaxViews;  // uint8 x kNoAllocMaxViews
constexpr size_t kScratchBytes = kScratchFlags + kNoAllocMaxViews;     // 80.25 KiB
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

// Device address of this device's scratch; the device must be current.
hipError_t find_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<unsigned char*>(ptr);
    return hipSuccess;
}

unsigned char* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}
}  // namespace

namespace xr {
namespace host {

std::atomic<int> g_alloc_budget{-1};  // allocations left before forced failures, or -1 (off)

hipError_t device_malloc(void** p, size_t bytes) {
    int left = g_alloc_budget.load();
    while (left >= 0) {
        if (left == 0) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        if (g_alloc_budget.compare_exchange_weak(left, left - 1)) break;
    }
    return hipMalloc(p, bytes);
}


int fail(xr_context* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

float theta_f32(int n, int P) {
    // ir.cu:1860: 2.0f * (float)M_PI * (float)n / (float)P, each step in FP32.
    volatile float a = 2.0f * (float)M_PI;
    volatile float b = a * (float)n;
    volatile float c = b / (float)P;
    return c;
}

void mat3mul(const double* a, const double* b, double* o) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}
void mat3vec(const double* m, const double* v, double* o) {
    for (int i = 0; i < 3; i++) o[i] = m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
}
double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Per-view projective tables (see ViewGeom).  Derivation from rayCasting
// (ir.cu:1852-1902): p = s + coeff*(X - s), coeff = -(s.s)/(s.(X - s)) and
// s.eu = s.ev = 0 give u = U/2 + (sdd/det) (X.eu) / (sod - shat.X), with
// X = vs*x' - (R*offset + t) and shat = +-R*y (IR +, FDK -).  Numerator and
// denominator are both divided by sod so that den ~ 1.
void build_view(const xr_plan& p, int cond, int view, ViewGeom& vg, ChordGeom& cg) {
    const xr_geometry& g = p.g;
    const int n = (p.dir == XR_CW ? 1 : -1) * view;
    const double th = (double)theta_f32(n, g.n_proj);
    const double c = std::cos(th), s = std::sin(th);
    const double Rz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
    double cR[9], R[9];
    for (int i = 0; i < 9; i++) cR[i] = (double)p.cond_R[9 * cond + i];
    mat3mul(cR, Rz, R);
    const double off0[3] = {p.cond_offset[3 * cond + 0], p.cond_offset[3 * cond + 1], p.cond_offset[3 * cond + 2]};
    double off[3];
    mat3vec(R, off0, off);
    const double t[3] = {p.cond_t[3 * cond + 0], p.cond_t[3 * cond + 1], p.cond_t[3 * cond + 2]};
    const double eu[3] = {R[0], R[3], R[6]}, ey[3] = {R[1], R[4], R[7]}, ev[3] = {R[2], R[5], R[8]};
    const double sgn = (p.conv == XR_CONV_IR) ? 1.0 : -1.0;
    const double sh[3] = {sgn * ey[0], sgn * ey[1], sgn * ey[2]};
    const double Xc[3] = {-(off[0] + t[0]), -(off[1] + t[1]), -(off[2] + t[2])};
    const double sod = g.sod, vs = g.vox_size;
    const double k = (double)g.sdd / (double)g.det_size / sod;
    for (int i = 0; i < 3; i++) {
        vg.au[i] = (float)(k * vs * eu[i]);
        vg.av[i] = (float)(k * vs * ev[i]);
        vg.ad[i] = (float)(-vs / sod * sh[i]);
    }
    vg.au[3] = (float)(k * dot3(eu, Xc));
    vg.av[3] = (float)(k * dot3(ev, Xc));
    vg.ad[3] = (float)(1.0 - dot3(sh, Xc) / sod);
    for (int i = 0; i < 3; i++) vg.d0[i] = (float)(Xc[i] - sod * sh[i]);
    vg.fy = (float)Xc[1];
    double g0[3] = {0, 0, 1};
    if (!p.grating.empty())
        for (int i = 0; i < 3; i++) g0[i] = p.grating[3 * cond + i];
    double G[3];
    mat3vec(R, g0, G);
    for (int i = 0; i < 3; i++) vg.G[i] = (float)G[i];
    vg.axis = (vg.au[2] == 0.0f && vg.ad[2] == 0.0f) ? 1.0f : 0.0f;
    for (int i = 0; i < 4; i++) vg.pad[i] = 0.0f;
    // chord geometry of the IR source side (ir.cu:1556-1574)
    cg.det[0] = g.det_size;
    cg.det[1] = (double)g.sdd - (double)g.sod;
    const double half = ((double)(g.n_vox + 1) / 2.0) * vs;
    for (int i = 0; i < 3; i++) {
        cg.src[i] = sod * ey[i];
        cg.eu[i] = eu[i];
        cg.ev[i] = ev[i];
        cg.ey[i] = ey[i];
        cg.lo[i] = -half - off[i] - t[i];
        cg.hi[i] = half - off[i] - t[i];
    }
}

int check_geometry(xr_context* ctx, const xr_geometry* g) {
    if (!g) return fail(ctx, XR_ERR_INVALID, "geometry is NULL");
    if (g->n_vox <= 0 || g->n_det_u <= 1 || g->n_det_v <= 1 || g->n_proj <= 0)
        return fail(ctx, XR_ERR_INVALID, "invalid sizes N=%d U=%d V=%d P=%d", g->n_vox, g->n_det_u, g->n_det_v,
                    g->n_proj);
    if (g->n_vox > 4096 || g->n_det_u > 16384 || g->n_det_v > 16384)
        return fail(ctx, XR_ERR_INVALID, "sizes beyond the supported range");
    if (!(g->sod > 0) || !(g->sdd > 0) || !(g->det_size > 0) || !(g->vox_size > 0))
        return fail(ctx, XR_ERR_INVALID, "distances and pitches must be positive");
    return XR_OK;
}

int check_scan(xr_context* ctx, const xr_scan* s) {
    if (!s) return fail(ctx, XR_ERR_INVALID, "scan is NULL");
    if (s->n_cond <= 0) return fail(ctx, XR_ERR_INVALID, "n_cond must be positive");
    if (!s->cond_R || !s->cond_t || !s->cond_offset) return fail(ctx, XR_ERR_INVALID, "scan tables missing");
    if (s->n_basis < 0 || s->n_basis > kMaxBasis)
        return fail(ctx, XR_ERR_INVALID, "n_basis must be in [0, %d]", kMaxBasis);
    if (s->n_basis > 0 && !s->basis) return fail(ctx, XR_ERR_INVALID, "basis vectors missing");
    return XR_OK;
}

int ensure_codes(xr_context* ctx, size_t n) {
    if (n <= ctx->codes_cap) return XR_OK;
    if (ctx->d_codes) (void)hipFree(ctx->d_codes);
    ctx->d_codes = nullptr;
    ctx->codes_cap = 0;  // stays empty if the allocation fails
    XR_HIP(ctx, device_malloc((void**)&ctx->d_codes, n * sizeof(int32_t)));
    ctx->codes_cap = n;
    return XR_OK;
}
int ensure_flags(xr_context* ctx, size_t n) {
    if (n <= ctx->flags_cap) return XR_OK;
    if (ctx->d_flags) (void)hipFree(ctx->d_flags);
    ctx->d_flags = nullptr;
    ctx->flags_cap = 0;
    XR_HIP(ctx, device_malloc((void**)&ctx->d_flags, n));
    ctx->flags_cap = n;
    return XR_OK;
}

hipEvent_t get_event(xr_context* ctx) {
    if (!ctx->pool.empty()) {
        hipEvent_t e = ctx->pool.back();
        ctx->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}


// vol_cond_stride: plain volumes of several conditions side by side (N^3), or 0
// for a single volume; code_base: proj/sino hold the images of codes >= code_base.
int forward_launch(xr_plan* p, int kind, const float* vol, int64_t vol_cond_stride, float* proj, const float* sino,
                   int code_base, const int32_t* d_codes, const uint8_t* d_flags, int n_slots, int out_mode,
                   double* d_loss, const float* vmax_known, float* vmax_buf, bool no_alloc) {
    xr_context* ctx = p->ctx;
    FwdArgs a{};
    a.d = p->dims;
    a.b = p->bas;
    a.vg = p->d_vg;
    a.cg = p->d_cg;
    a.views = d_codes;
    a.accum = d_flags;
    a.vol = vol;
    // forwardProjGrad splats vs*ratio*voxel (ir.cu:489-491), and a repeated view
    // adds that to the image it already holds (no normalisation afterwards)
    if (p->conv == XR_CONV_GRAD) a.out_scale = p->g.vox_size;
    a.view_group = kFwdViewGroup;
    const int64_t N3 = (int64_t)p->g.n_vox * p->g.n_vox * p->g.n_vox;
    a.vol_cond_stride = vol_cond_stride;
    a.chan_stride = N3;
    a.proj = proj;
    a.sino = sino;
    a.norm = p->d_norm;
    a.bands = p->tiles.fwd_scan ? nullptr : p->d_bands;
    a.code_base = code_base;
    a.g2 = p->d_g2;
    a.xw = p->d_xw;
    a.loss = d_loss;
    a.out_mode = out_mode;
    a.n_bands = (p->g.n_det_u + kBandU - 1) / kBandU;
    a.n_slots = n_slots;
    // Fixed-point scale of the LDS accumulation: bound of a voxel's splat value
    // (wmul * vmax) times the bound of a pixel's weight sum (fsum).
    if (kind == FWD_WEIGHT) {
        a.vmax = ctx->d_scal;  // 1.0
        a.wmul = 1.0f;
        a.fsum = 1.0f;
        a.fmax = (out_mode == OUT_NORM_TABLE) ? p->d_fmax : nullptr;
    } else {
        if (no_alloc) {
            if (!p->d_norm)
                return fail(ctx, XR_ERR_INVALID,
                            "the plan's normalisation table does not exist yet (a first xr_project call builds it)");
        } else {
            XR_TRY(ensure_norm(p));
        }
        a.fsum = p->max_factor * 1.01f + 1.0f;
        if (vmax_known) {
            a.vmax = vmax_known;
        } else {
            const int64_t nvals = (kind == FWD_XTT) ? N3 * p->n_basis : (vol_cond_stride ? N3 * p->n_cond : N3);
            float* vb = vmax_buf ? vmax_buf : ctx->d_scal + 1;
            if (no_alloc) {
                XR_HIP(ctx, launch_fill(vb, 1, 0.0f, ctx->stream));
                XR_HIP(ctx, launch_absmax(vol, nvals, vb, ctx->stream));
            } else {
                XR_HIP(ctx, hipMemsetAsync(vb, 0, sizeof(float), ctx->stream));
                XR_TRY(timed(ctx, 3, [&] { return launch_absmax(vol, nvals, vb, ctx->stream); }));
            }
            a.vmax = vb;
        }
        const float wsum = p->d_orient ? p->xtt_wsum_orth : p->xtt_wsum;
        a.wmul = (kind == FWD_XTT) ? p->dims.vs * p->dims.ratio * wsum * 1.0001f : p->dims.ratio * 1.0001f;
    }
    // All conditions in one launch share the axis specialisation only if all are axis-aligned.
    bool axis = true;
    for (int c = 0; c < p->n_cond; c++) axis = axis && p->axis_cond[c];
    if (kind == FWD_XTT && p->d_orient) {
        a.coef = p->d_orient;  // forwardOrth: per-voxel rotation, general path
        axis = false;
    }
    // Detector rows in equal chunks, all in one launch (untilted scans walk
    // only the z range that reaches a chunk's rows).
    const int V = p->g.n_det_v;
    const int rows = forward_rows(p, kind);
    a.v0 = 0;
    a.vb = rows;
    a.n_chunks = (V + rows - 1) / rows;
    // Few views (an ordered subset) leave most CUs idle: split the bands.
    const int64_t items = (int64_t)a.n_bands * a.n_slots * a.n_chunks;
    a.n_sub = p->tiles.fwd_split;
    if (a.n_sub == 0 && items < 4096) {
        if (ctx->n_cu == 0)
            XR_HIP(ctx, hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
        const int64_t slots = (int64_t)forward_blocks_per_cu(kind, axis, forward_lds_bytes(kind, rows)) * ctx->n_cu;
        a.n_sub = forward_split(items, slots, 0);
    }
    a.n_sub = std::max(a.n_sub, 1);
    if (no_alloc) {
        XR_HIP(ctx, launch_forward(a, kind, axis, ctx->stream));
        return XR_OK;
    }
    return timed(ctx, 0, [&] { return launch_forward(a, kind, axis, ctx->stream); });
}

int forward_rows(const xr_plan* p, int kind) {
    // Detector rows per chunk: the band accumulator (kBandStride words per
    // row, two guard rows, the reduction slots) stays within 64 KB of LDS;
    // taller detectors are split into equal row chunks.  tiles.fwd_rows caps
    // the chunk height further (smaller LDS footprint, more resident blocks).
    const int V = p->g.n_det_v;
    const int word = (kind == FWD_WEIGHT) ? 8 : 4;
    int max_rows = (64 * 1024 - 16 * kFwdWaves - 16) / (kBandStride * word) - 2;
    if (p->tiles.fwd_rows > 0) max_rows = std::min(max_rows, (int)p->tiles.fwd_rows);
    const int n_chunks = (V + max_rows - 1) / max_rows;
    return (V + n_chunks - 1) / n_chunks;
}

int backward_launch(xr_plan* p, int kind, const float* proj, int code_base, const int32_t* d_codes, int n_views,
                    float* vol, float* tmp, float* factor, float alpha, int overwrite, unsigned* vmax_out) {
    xr_context* ctx = p->ctx;
    BwdArgs a{};
    a.d = p->dims;
    a.b = p->bas;
    a.vg = p->d_vg;
    a.views = d_codes;
    a.n_views = n_views;
    a.proj = proj;
    a.code_base = code_base;
    a.vol = vol;
    a.tmp = tmp;
    a.factor = factor;
    a.chan_stride = (int64_t)p->g.n_vox * p->g.n_vox * p->g.n_vox;
    a.g2 = p->d_g2;
    a.xw = p->d_xw;
    a.alpha = alpha;
    a.overwrite = overwrite;
    a.vmax_out = (kind == BWD_MLEM || kind == BWD_ART) ? vmax_out : nullptr;
    a.zpl = p->tiles.bwd_zpl;
    a.gx = p->tiles.bwd_gx;
    bool axis = true;
    for (int c = 0; c < p->n_cond; c++) axis = axis && p->axis_cond[c];
    if ((kind == BWD_XTT_RAW || kind == BWD_XTT_MLEM || kind == BWD_XTT_ART) && p->d_orient) {
        a.coef = p->d_orient;  // backwardOrth: per-voxel rotation, general path
        axis = false;
    }
    return timed(ctx, 1, [&] { return launch_backward(a, kind, axis, ctx->stream); });
}

int ensure_norm(xr_plan* p) {
    if (p->d_norm) return XR_OK;
    xr_context* ctx = p->ctx;
    const int P = p->g.n_proj;
    const int64_t lenD = (int64_t)p->g.n_det_u * p->g.n_det_v * P;
    // Both tables are built into buffers of their own and handed to the plan
    // only when complete: after a failure the plan holds neither, and the next
    // call builds them again.
    if (!p->d_bands && p->g.n_vox <= 32767) {  // column bounds packed in 16 bits
        const int n_bands = (p->g.n_det_u + kBandU - 1) / kBandU;
        DevBuf<int32_t> bands;
        XR_HIP(ctx, bands.alloc((size_t)p->n_cond * P * n_bands * p->g.n_vox));
        XR_HIP(ctx, launch_band_rows(p->d_vg, p->dims, p->n_cond * P, n_bands, bands.p, ctx->stream));
        p->d_bands = bands.p;
        bands.p = nullptr;
    }
    DevBuf<float> norm;
    XR_HIP(ctx, norm.alloc((size_t)lenD * p->n_cond));
    std::vector<int32_t> codes((size_t)p->n_cond * P);
    for (size_t i = 0; i < codes.size(); i++) codes[i] = (int32_t)i;
    DevBuf<int32_t> dc;
    XR_HIP(ctx, dc.alloc(codes.size()));
    XR_HIP(ctx, hipMemcpyAsync(dc.p, codes.data(), codes.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    // Splat the bilinear weights alone (factor), then L/(factor + 1e-5) per
    // pixel; the largest factor bounds every pixel sum of later passes.
    DevBuf<float> fmax;
    XR_HIP(ctx, fmax.alloc(1));
    XR_HIP(ctx, hipMemsetAsync(fmax.p, 0, sizeof(float), ctx->stream));
    p->d_fmax = fmax.p;
    const int rc = forward_launch(p, FWD_WEIGHT, nullptr, 0, norm.p, nullptr, 0, dc.p, nullptr, (int)codes.size(),
                                  OUT_NORM_TABLE, nullptr);
    p->d_fmax = nullptr;
    XR_TRY(rc);
    float mf = 0.f;
    XR_HIP(ctx, hipMemcpyAsync(&mf, fmax.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    p->max_factor = mf;
    // The gradient operator has no chord normalisation (its splat carries the
    // factor vs instead, FwdArgs::out_scale): a table of ones.
    if (p->conv == XR_CONV_GRAD) XR_HIP(ctx, launch_fill(norm.p, lenD * p->n_cond, 1.0f, ctx->stream));
    p->d_norm = norm.p;
    norm.p = nullptr;
    return XR_OK;
}

}  // namespace host
}  // namespace xr

using namespace xr::host;

namespace {

// ---------------------------------------------------------------- JSON ----
// Minimal JSON reader for settings.json (objects, arrays, numbers, strings,
// booleans, null).
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& k) const {
        if (kind != OBJ) return nullptr;
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* s;
    const char* e;
    bool ok = true;
    void ws() {
        while (s < e && (*s == ' ' || *s == '\n' || *s == '\r' || *s == '\t')) s++;
    }
    bool lit(const char* w) {
        const size_t n = strlen(w);
        if ((size_t)(e - s) >= n && strncmp(s, w, n) == 0) {
            s += n;
            return true;
        }
        return false;
    }
    std::string str() {
        std::string out;
        if (s >= e || *s != '"') {
            ok = false;
            return out;
        }
        s++;
        while (s < e && *s != '"') {
            if (*s == '\\' && s + 1 < e) {
                s++;
                switch (*s) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        // keep escaped code points as '?': paths and keys here are ASCII
                        out += '?';
                        s += std::min<ptrdiff_t>(4, e - s - 1);
                        break;
                    default: out += *s;
                }
                s++;
            } else {
                out += *s++;
            }
        }
        if (s >= e) ok = false;
        s++;
        return out;
    }
    JVal val() {
        JVal v;
        ws();
        if (s >= e) {
            ok = false;
            return v;
        }
        if (*s == '{') {
            v.kind = JVal::OBJ;
            s++;
            ws();
            if (s < e && *s == '}') {
                s++;
                return v;
            }
            while (ok) {
                ws();
                std::string k = str();
                ws();
                if (s >= e || *s != ':') {
                    ok = false;
                    break;
                }
                s++;
                JVal x = val();
                v.obj.emplace_back(std::move(k), std::move(x));
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == '}') {
                    s++;
                    break;
                }
                ok = false;
            }
        } else if (*s == '[') {
            v.kind = JVal::ARR;
            s++;
            ws();
            if (s < e && *s == ']') {
                s++;
                return v;
            }
            while (ok) {
                v.arr.push_back(val());
                ws();
                if (s < e && *s == ',') {
                    s++;
                    continue;
                }
                if (s < e && *s == ']') {
                    s++;
                    break;
                }
                ok = false;
            }
        } else if (*s == '"') {
            v.kind = JVal::STR;
            v.str = str();
        } else if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JVal::BOOL;
        } else if (lit("null")) {
            v.kind = JVal::NUL;
        } else {
            char* end = nullptr;
            v.kind = JVal::NUM;
            v.num = strtod(s, &end);
            if (end == s) ok = false;
            s = end;
        }
        return v;
    }
};

bool jnum(const JVal* v, double& out) {
    if (!v || v->kind != JVal::NUM) return false;
    out = v->num;
    return true;
}
bool jarr(const JVal* v, std::vector<double>& out) {
    if (!v || v->kind != JVal::ARR) return false;
    out.clear();
    for (auto& x : v->arr) {
        if (x.kind != JVal::NUM) return false;
        out.push_back(x.num);
    }
    return true;
}

// ------------------------------------------------------------- phantom ----
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" {

// ------------------------------------------------------------ context ----

int xr_context_create(xr_context** out, int device) {
    if (!out) return XR_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return XR_ERR_HIP;
    if (device < 0 || device >= n) return XR_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return XR_ERR_HIP;
    if (find_scratch(device) != hipSuccess) return XR_ERR_HIP;
    auto* c = new (std::nothrow) xr_context();
    if (!c) return XR_ERR_ALLOC;
    c->device = device;
    if (device_malloc((void**)&c->d_loss, 2 * sizeof(double)) != hipSuccess ||
        device_malloc((void**)&c->d_scal, 2 * sizeof(float)) != hipSuccess) {
        if (c->d_loss) (void)hipFree(c->d_loss);
        delete c;
        return XR_ERR_ALLOC;
    }
    const float one[2] = {1.0f, 0.0f};
    if (hipMemcpy(c->d_scal, one, sizeof(one), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(c->d_loss);
        (void)hipFree(c->d_scal);
        delete c;
        return XR_ERR_HIP;
    }
    *out = c;
    return XR_OK;
}

void xr_context_destroy(xr_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->d_codes) (void)hipFree(ctx->d_codes);
    if (ctx->d_flags) (void)hipFree(ctx->d_flags);
    if (ctx->d_loss) (void)hipFree(ctx->d_loss);
    if (ctx->d_scal) (void)hipFree(ctx->d_scal);
    delete ctx;
}

int xr_context_set_stream(xr_context* ctx, void* hip_stream) {
    if (!ctx) return XR_ERR_INVALID;
    ctx->stream = (hipStream_t)hip_stream;
    return XR_OK;
}

int xr_context_synchronize(xr_context* ctx) {
    if (!ctx) return XR_ERR_INVALID;
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_memcpy(xr_context* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return XR_ERR_INVALID;
    if (bytes == 0) return XR_OK;
    if (!dst || !src) return fail(ctx, XR_ERR_INVALID, "NULL copy address");
    XR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

const char* xr_last_error(const xr_context* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int xr_timing_enable(xr_context* ctx, int enable) {
    if (!ctx) return XR_ERR_INVALID;
    ctx->timing = enable != 0;
    return XR_OK;
}

int xr_timing_reset(xr_context* ctx) {
    if (!ctx) return XR_ERR_INVALID;
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& e : ctx->events) {
        ctx->pool.push_back(e.a);
        ctx->pool.push_back(e.b);
    }
    ctx->events.clear();
    return XR_OK;
}

int xr_timing_read(xr_context* ctx, int kind, double* total_ms, int64_t* count) {
    if (!ctx || !total_ms || !count) return XR_ERR_INVALID;
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double t = 0;
    int64_t c = 0;
    for (auto& e : ctx->events) {
        if (e.kind != kind) continue;
        float ms = 0;
        XR_HIP(ctx, hipEventElapsedTime(&ms, e.a, e.b));
        t += ms;
        c++;
    }
    *total_ms = t;
    *count = c;
    return XR_OK;
}

float xr_default_vox_size(float sdd, float sod, float det_size) { return sod * det_size / sdd; }

// --------------------------------------------------------------- plans ----

int xr_plan_create(xr_context* ctx, const xr_geometry* g, const xr_scan* s, int convention, int direction,
                   xr_plan** out) {
    if (!ctx || !out) return XR_ERR_INVALID;
    *out = nullptr;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (convention != XR_CONV_IR && convention != XR_CONV_FDK && convention != XR_CONV_GRAD)
        return fail(ctx, XR_ERR_INVALID, "bad convention");
    if (direction != XR_CW && direction != XR_CCW) return fail(ctx, XR_ERR_INVALID, "bad direction");
    std::unique_ptr<xr_plan> p(new (std::nothrow) xr_plan());
    if (!p) return fail(ctx, XR_ERR_ALLOC, "plan allocation failed");
    p->ctx = ctx;
    p->g = *g;
    p->n_cond = s->n_cond;
    p->n_basis = s->n_basis;
    p->conv = convention;
    p->dir = direction;
    p->cond_R.assign(s->cond_R, s->cond_R + 9 * s->n_cond);
    p->cond_t.assign(s->cond_t, s->cond_t + 3 * s->n_cond);
    p->cond_offset.assign(s->cond_offset, s->cond_offset + 3 * s->n_cond);
    if (s->n_basis > 0) p->basis.assign(s->basis, s->basis + 3 * s->n_basis);
    if (s->grating) p->grating.assign(s->grating, s->grating + 3 * s->n_cond);
    const int P = g->n_proj, C = s->n_cond, K = std::max(1, s->n_basis);
    std::vector<ViewGeom> vg((size_t)C * P);
    std::vector<ChordGeom> cg((size_t)C * P);
    std::vector<float> g2((size_t)C * P * K, 0.f);
    const int KS = kXwStride;
    std::vector<float> xw((size_t)C * P * 5 * KS, 0.f);
    p->axis_cond.assign(C, 1);
    for (int c = 0; c < C; c++)
        for (int v = 0; v < P; v++) {
            ViewGeom& t = vg[(size_t)c * P + v];
            build_view(*p, c, v, t, cg[(size_t)c * P + v]);
            if (t.axis == 0.0f) p->axis_cond[c] = 0;
            for (int k = 0; k < s->n_basis; k++) {
                const double sg = (double)p->basis[3 * k] * t.G[0] + (double)p->basis[3 * k + 1] * t.G[1] +
                                  (double)p->basis[3 * k + 2] * t.G[2];
                g2[((size_t)c * P + v) * K + k] = (float)(sg * sg);
                // untilted XTT weight constants (see BwdArgs::xw)
                const double Sx = p->basis[3 * k], Sy = p->basis[3 * k + 1], Sz = p->basis[3 * k + 2];
                const double sig = std::fabs(sg);
                float* row = &xw[((size_t)c * P + v) * 5 * KS];
                const double vs = g->vox_size;
                row[k] = (float)(sig * Sz);
                row[KS + k] = (float)((double)(float)(sg * sg) * (double)(float)(Sx * Sx + Sy * Sy + Sz * Sz));
                row[2 * KS + k] = (float)(sig * vs * Sx);
                row[3 * KS + k] = (float)(sig * vs * Sy);
                row[4 * KS + k] = (float)(sig * ((double)t.d0[0] * Sx + (double)t.d0[1] * Sy));
            }
        }
    Dims& d = p->dims;
    d.N = g->n_vox;
    d.nz = g->n_vox;
    d.U = g->n_det_u;
    d.V = g->n_det_v;
    d.P = P;
    d.K = s->n_basis;
    // voxel centre x - xc: (N+1)/2 (ir.cu:1881), (N-1)/2 (fdk.cu:132), and N/2 on
    // the gradient grid of side N (ir.cu:459: (2(x - 0.5) - N + 1)/2)
    d.xc = (convention == XR_CONV_IR)    ? (float)(g->n_vox + 1) * 0.5f
           : (convention == XR_CONV_FDK) ? (float)(g->n_vox - 1) * 0.5f
                                         : (float)g->n_vox * 0.5f;
    d.hu = 0.5f * (float)g->n_det_u;
    d.hv = 0.5f * (float)g->n_det_v;
    d.umax = (float)g->n_det_u - 0.55f;
    d.vmax = (float)g->n_det_v - 0.55f;
    d.vs = g->vox_size;
    {
        // ir.cu:1513-1514, evaluated left to right in FP32
        volatile float m = g->sod / g->sdd;
        volatile float den = g->det_size * m * g->det_size * m;
        d.ratio = (g->vox_size * g->vox_size) / den;
    }
    d.sod = g->sod;
    d.cfdk = 2.0f * (float)M_PI / (float)P;
    for (int k = 0; k < s->n_basis; k++) {
        float g2max = 0.f;
        for (size_t i = 0; i < (size_t)C * P; i++) g2max = std::max(g2max, g2[i * K + k]);
        const double s2 = (double)p->basis[3 * k] * p->basis[3 * k] + (double)p->basis[3 * k + 1] * p->basis[3 * k + 1] +
                          (double)p->basis[3 * k + 2] * p->basis[3 * k + 2];
        p->xtt_wsum += (float)(s2 * g2max);
        double gg = 0.0;
        for (size_t i = 0; i < (size_t)C * P; i++)
            gg = std::max(gg, (double)vg[i].G[0] * vg[i].G[0] + (double)vg[i].G[1] * vg[i].G[1] +
                                  (double)vg[i].G[2] * vg[i].G[2]);
        p->xtt_wsum_orth += (float)(s2 * s2 * gg);
    }
    for (int k = 0; k < s->n_basis; k++) {
        for (int i = 0; i < 3; i++) p->bas.S[k][i] = p->basis[3 * k + i];
        p->bas.S2[k] = p->bas.S[k][0] * p->bas.S[k][0] + p->bas.S[k][1] * p->bas.S[k][1] +
                       p->bas.S[k][2] * p->bas.S[k][2];
    }
    XR_HIP(ctx, hipSetDevice(ctx->device));
    XR_HIP(ctx, device_malloc((void**)&p->d_vg, vg.size() * sizeof(ViewGeom)));
    XR_HIP(ctx, device_malloc((void**)&p->d_cg, cg.size() * sizeof(ChordGeom)));
    XR_HIP(ctx, device_malloc((void**)&p->d_g2, g2.size() * sizeof(float)));
    XR_HIP(ctx, device_malloc((void**)&p->d_xw, xw.size() * sizeof(float)));
    XR_HIP(ctx, hipMemcpy(p->d_xw, xw.data(), xw.size() * sizeof(float), hipMemcpyHostToDevice));
    XR_HIP(ctx, hipMemcpy(p->d_vg, vg.data(), vg.size() * sizeof(ViewGeom), hipMemcpyHostToDevice));
    XR_HIP(ctx, hipMemcpy(p->d_cg, cg.data(), cg.size() * sizeof(ChordGeom), hipMemcpyHostToDevice));
    XR_HIP(ctx, hipMemcpy(p->d_g2, g2.data(), g2.size() * sizeof(float), hipMemcpyHostToDevice));
    *out = p.release();
    return XR_OK;
}

void xr_plan_destroy(xr_plan* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->d_vg) (void)hipFree(p->d_vg);
    if (p->d_cg) (void)hipFree(p->d_cg);
    if (p->d_g2) (void)hipFree(p->d_g2);
    if (p->d_xw) (void)hipFree(p->d_xw);
    if (p->d_norm) (void)hipFree(p->d_norm);
    if (p->d_bands) (void)hipFree(p->d_bands);
    delete p;
}

// ------------------------------------------------------ device operators ----

int xr_volume_to_columns(xr_context* ctx, const float* ref_dev, float* col_dev, int n_vox, int channels) {
    if (!ctx || !ref_dev || !col_dev || n_vox <= 0 || channels <= 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_to_internal(ref_dev, col_dev, n_vox, channels, ctx->stream); });
}

int xr_volume_from_columns(xr_context* ctx, const float* col_dev, float* ref_dev, int n_vox, int channels) {
    if (!ctx || !ref_dev || !col_dev || n_vox <= 0 || channels <= 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_from_internal(col_dev, ref_dev, n_vox, channels, ctx->stream); });
}

void xr_test_fail_allocations(int after) { xr::host::g_alloc_budget.store(after < 0 ? -1 : after); }

int xr_plan_set_orientation(xr_plan* plan, const float* coef) {
    if (!plan) return XR_ERR_INVALID;
    if (coef && plan->n_basis <= 0) return fail(plan->ctx, XR_ERR_INVALID, "orientation needs a plan with basis vectors");
    plan->d_orient = coef;
    return XR_OK;
}

static Basis basis3(const float* basis9) {
    Basis b{};
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++) b.S[k][i] = basis9[3 * k + i];
    return b;
}

int xr_orth_normal(xr_context* ctx, const float* basis9, const float* vol, int64_t nvox, float* coef, float* loss,
                   int rot, const float* est) {
    if (!ctx || !basis9 || !vol || !coef || !loss || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    const Basis b = basis3(basis9);
    return timed(ctx, 3, [&] { return launch_orth_normal(vol, coef, loss, nvox, rot, est, b, ctx->stream); });
}

int xr_orth_update_estimate(xr_context* ctx, const float* basis9, const float* vol, int64_t nvox, float* est,
                            int iter) {
    if (!ctx || !basis9 || !vol || !est || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    const Basis b = basis3(basis9);
    return timed(ctx, 3, [&] { return launch_orth_update_est(vol, est, nvox, iter, b, ctx->stream); });
}

int xr_orth_md(xr_context* ctx, const float* vol, const float* coef, int64_t nvox, float* md) {
    if (!ctx || !vol || !coef || !md || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_orth_md(vol, coef, md, nvox, ctx->stream); });
}

int xr_orth_mean_filter(xr_context* ctx, const float* coef_src, const float* vol, int n_vox, float weight,
                        float* coef_dst) {
    if (!ctx || !coef_src || !vol || !coef_dst || n_vox <= 0 || coef_src == coef_dst)
        return fail(ctx, XR_ERR_INVALID, "bad args (the filter needs distinct source and destination)");
    return timed(ctx, 3, [&] { return launch_orth_mean_filter(coef_src, vol, n_vox, weight, coef_dst, ctx->stream); });
}

int xr_volume_downsample(xr_context* ctx, const float* fine, int n_vox, int factor, float* coarse) {
    if (!ctx || !fine || !coarse || factor <= 0 || n_vox <= 0 || n_vox % factor != 0)
        return fail(ctx, XR_ERR_INVALID, "downsample: n_vox %d must be a positive multiple of factor %d", n_vox, factor);
    return timed(ctx, 3, [&] { return launch_vol_down(fine, coarse, n_vox / factor, factor, ctx->stream); });
}

int xr_volume_upsample_add(xr_context* ctx, const float* coarse_a, const float* coarse_b, int n_coarse, int factor,
                           float* fine, float floor_value) {
    if (!ctx || !coarse_a || !coarse_b || !fine || factor <= 0 || n_coarse <= 0 ||
        (int64_t)n_coarse * factor > 32768)
        return fail(ctx, XR_ERR_INVALID, "upsample_add: bad sizes (n_coarse %d, factor %d)", n_coarse, factor);
    return timed(ctx, 3,
                 [&] { return launch_vol_up_add(coarse_a, coarse_b, n_coarse, factor, fine, floor_value, ctx->stream); });
}

static int upload_views(xr_plan* p, const int32_t* views, int n_views, const uint8_t* accumulate) {
    xr_context* ctx = p->ctx;
    const int total = p->n_cond * p->g.n_proj;
    for (int i = 0; i < n_views; i++)
        if (views[i] < 0 || views[i] >= total)
            return fail(ctx, XR_ERR_INVALID, "view code %d out of range [0, %d)", views[i], total);
    XR_TRY(ensure_codes(ctx, (size_t)n_views));
    XR_HIP(ctx, hipMemcpyAsync(ctx->d_codes, views, sizeof(int32_t) * n_views, hipMemcpyHostToDevice, ctx->stream));
    if (accumulate) {
        XR_TRY(ensure_flags(ctx, (size_t)n_views));
        XR_HIP(ctx, hipMemcpyAsync(ctx->d_flags, accumulate, (size_t)n_views, hipMemcpyHostToDevice, ctx->stream));
    }
    return XR_OK;
}

// Plans the voxel-driven projectors serve: the IR convention, and the
// gradient grid of IR::gradReconstruct (scalar volumes only).
static int check_projecting(xr_context* ctx, const xr_plan* plan, int xtt) {
    if (plan->conv == XR_CONV_IR) return XR_OK;
    if (plan->conv == XR_CONV_GRAD && !xtt) return XR_OK;
    return fail(ctx, XR_ERR_INVALID, "projection needs an XR_CONV_IR plan (or XR_CONV_GRAD without XTT)");
}

int xr_project(xr_context* ctx, xr_plan* plan, const float* vol, float* proj, const float* sino,
               const int32_t* views, int n_views, const uint8_t* accumulate, int xtt, int out_mode, double* loss) {
    if (!ctx || !plan || !vol || !proj || !views || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (out_mode < XR_OUT_RAW || out_mode > XR_OUT_RESIDUAL) return fail(ctx, XR_ERR_INVALID, "bad out_mode");
    if ((out_mode == XR_OUT_RATIO || out_mode == XR_OUT_RESIDUAL) && !sino)
        return fail(ctx, XR_ERR_INVALID, "sino required for ratio/residual");
    if (xtt && plan->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT projection needs basis vectors");
    XR_TRY(check_projecting(ctx, plan, xtt));
    if (n_views == 0) {
        if (loss) *loss = 0;
        return XR_OK;
    }
    XR_TRY(upload_views(plan, views, n_views, accumulate));
    if (out_mode != XR_OUT_RAW) XR_TRY(ensure_norm(plan));
    if (loss) XR_HIP(ctx, hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
    const int64_t N3 = (int64_t)plan->g.n_vox * plan->g.n_vox * plan->g.n_vox;
    XR_TRY(forward_launch(plan, xtt ? FWD_XTT : FWD_PLAIN, vol, N3, proj, sino, 0, ctx->d_codes,
                          accumulate ? ctx->d_flags : nullptr, n_views, out_mode, ctx->d_loss));
    if (loss) {
        XR_HIP(ctx, hipMemcpyAsync(loss, ctx->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

int xr_project_noalloc(xr_context* ctx, xr_plan* plan, const float* vol, float* proj, const float* sino,
                       const int32_t* views, int n_views, const uint8_t* accumulate, int xtt, int out_mode,
                       double* loss) {
    if (!ctx || !plan || !vol || !proj || !views || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (out_mode < XR_OUT_RAW || out_mode > XR_OUT_RESIDUAL) return fail(ctx, XR_ERR_INVALID, "bad out_mode");
    if ((out_mode == XR_OUT_RATIO || out_mode == XR_OUT_RESIDUAL) && !sino)
        return fail(ctx, XR_ERR_INVALID, "sino required for ratio/residual");
    if (xtt && plan->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT projection needs basis vectors");
    if (plan->conv != XR_CONV_IR) return fail(ctx, XR_ERR_INVALID, "forward projection uses the IR convention");
    if (n_views > kNoAllocMaxViews)
        return fail(ctx, XR_ERR_INVALID, "xr_project_noalloc takes at most %d views per call (got %d)",
                    kNoAllocMaxViews, n_views);
    if (n_views == 0) {
        if (loss) *loss = 0;
        return XR_OK;
    }
    const int total = plan->n_cond * plan->g.n_proj;
    for (int i = 0; i < n_views; i++)
        if (views[i] < 0 || views[i] >= total)
            return fail(ctx, XR_ERR_INVALID, "view code %d out of range [0, %d)", views[i], total);
    unsigned char* base = scratch_of(ctx->device);
    if (!base) return fail(ctx, XR_ERR_HIP, "no scratch address for device %d", ctx->device);
    int32_t* d_codes = reinterpret_cast<int32_t*>(base + kScratchCodes);
    uint8_t* d_flags = base + kScratchFlags;
    double* d_loss = reinterpret_cast<double*>(base + kScratchLoss);
    float* d_vmax = reinterpret_cast<float*>(base + kScratchVmax);
    XR_HIP(ctx, hipMemcpyAsync(d_codes, views, sizeof(int32_t) * n_views, hipMemcpyHostToDevice, ctx->stream));
    if (accumulate)
        XR_HIP(ctx, hipMemcpyAsync(d_flags, accumulate, (size_t)n_views, hipMemcpyHostToDevice, ctx->stream));
    // the FP64 loss starts at zero: two zero floats are a zero double
    XR_HIP(ctx, launch_fill(reinterpret_cast<float*>(d_loss), 2, 0.0f, ctx->stream));
    const int64_t N3 = (int64_t)plan->g.n_vox * plan->g.n_vox * plan->g.n_vox;
    XR_TRY(forward_launch(plan, xtt ? FWD_XTT : FWD_PLAIN, vol, N3, proj, sino, 0, d_codes,
                          accumulate ? d_flags : nullptr, n_views, out_mode, d_loss, nullptr, d_vmax, true));
    if (loss) {
        XR_HIP(ctx, hipMemcpyAsync(loss, d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

int xr_residual(xr_context* ctx, xr_plan* plan, const float* vol, float* resid, const float* sino,
                const int32_t* views, int n_views, int xtt, double* l1, double* l2) {
    if (!ctx || !plan || !vol || !resid || !sino || !views || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (xtt && plan->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT projection needs basis vectors");
    XR_TRY(check_projecting(ctx, plan, xtt));
    if (l1) *l1 = 0;
    if (l2) *l2 = 0;
    if (n_views == 0) return XR_OK;
    XR_TRY(upload_views(plan, views, n_views, nullptr));
    XR_TRY(ensure_norm(plan));
    XR_HIP(ctx, hipMemsetAsync(ctx->d_loss, 0, 2 * sizeof(double), ctx->stream));
    const int64_t N3 = (int64_t)plan->g.n_vox * plan->g.n_vox * plan->g.n_vox;
    XR_TRY(forward_launch(plan, xtt ? FWD_XTT : FWD_PLAIN, vol, N3, resid, sino, 0, ctx->d_codes, nullptr, n_views,
                          OUT_RESIDUAL_L2, ctx->d_loss));
    double sums[2] = {0, 0};
    XR_HIP(ctx, hipMemcpyAsync(sums, ctx->d_loss, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (l1) *l1 = sums[0];
    if (l2) *l2 = sums[1];
    return XR_OK;
}

int xr_backproject(xr_context* ctx, xr_plan* plan, const float* proj, const int32_t* views, int n_views, int xtt,
                   float* tmp, float* factor) {
    if (!ctx || !plan || !proj || !views || !tmp || !factor || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (xtt && plan->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT backprojection needs basis vectors");
    XR_TRY(check_projecting(ctx, plan, xtt));
    if (n_views == 0) return XR_OK;
    XR_TRY(upload_views(plan, views, n_views, nullptr));
    return backward_launch(plan, xtt ? BWD_XTT_RAW : BWD_RAW, proj, 0, ctx->d_codes, n_views, nullptr, tmp, factor,
                           0.f, 0);
}

int xr_backproject_update(xr_context* ctx, xr_plan* plan, const float* proj, const int32_t* views, int n_views,
                          int xtt, int method, float lambda, int subset_size, float* vol) {
    if (!ctx || !plan || !proj || !views || !vol || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    if (xtt && plan->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT update needs basis vectors");
    XR_TRY(check_projecting(ctx, plan, xtt));
    if (subset_size <= 0) return fail(ctx, XR_ERR_INVALID, "subset_size must be positive");
    XR_TRY(upload_views(plan, views, n_views, nullptr));
    int kind;
    if (xtt)
        kind = (method == XR_METHOD_MLEM) ? BWD_XTT_MLEM : BWD_XTT_ART;
    else
        kind = (method == XR_METHOD_MLEM) ? BWD_MLEM : BWD_ART;
    return backward_launch(plan, kind, proj, 0, ctx->d_codes, n_views, vol, nullptr, nullptr,
                           lambda / (float)subset_size, 0);
}

int xr_fdk_taps(const xr_geometry* g, int filter, float* taps) {
    if (!g || !taps || g->n_det_u <= 0) return XR_ERR_INVALID;
    const float d = g->det_size * (g->sod / g->sdd);
    for (int m = 0; m < g->n_det_u; m++) {
        if (filter == XR_FILTER_REFERENCE) {
            // reconstruct.cu:1697: double product rounded to float, FP32 reciprocal
            taps[m] = 1.0f / (float)(M_PI * M_PI * d * (1.0f - 4.0f * (float)(m * m)));
        } else if (filter == XR_FILTER_RAMLAK) {
            if (m == 0)
                taps[m] = (float)(1.0 / (8.0 * (double)d));
            else if (m % 2 == 1)
                taps[m] = (float)(-1.0 / (2.0 * M_PI * M_PI * (double)m * (double)m * (double)d));
            else
                taps[m] = 0.0f;
        } else if (filter == XR_FILTER_HILBERT) {
            // cuFFTtoProjection + hilbertFiltering (fdk.cu:298-375): a forward
            // real-to-complex FFT of the row, every half-spectrum bin times
            // -i/(U/2+1), an unnormalised complex-to-real inverse.  The inverse
            // drops the imaginary parts of the DC and (even U) Nyquist bins, so the
            // row is circularly convolved with h(m) = sum_{k=1}^{(U-1)/2}
            // 2 sin(2 pi k m / U) / (U/2+1), evaluated here in FP64.
            const int U = g->n_det_u;
            const double scale = 2.0 / (double)(U / 2 + 1);
            double h = 0.0;
            for (int k = 1; k <= (U - 1) / 2; k++) h += std::sin(2.0 * M_PI * (double)((int64_t)k * m % U) / (double)U);
            taps[m] = (float)(scale * h);
        } else {
            return XR_ERR_INVALID;
        }
    }
    return XR_OK;
}

static int fdk_filter(xr_context* ctx, xr_plan* plan, int filter, const float* sino, float* filt,
                      const int32_t* views, int n_views, int code_base) {
    if (!ctx || !plan || !sino || !filt || !views || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (n_views == 0) return XR_OK;
    const xr_geometry& g = plan->g;
    std::vector<float> taps(g.n_det_u);
    if (xr_fdk_taps(&g, filter, taps.data()) != XR_OK) return fail(ctx, XR_ERR_INVALID, "bad filter");
    XR_TRY(upload_views(plan, views, n_views, nullptr));
    DevBuf<float> dw, dt;
    XR_HIP(ctx, dw.alloc((size_t)g.n_det_u * g.n_det_v));
    XR_HIP(ctx, dt.alloc(taps.size()));
    XR_HIP(ctx, hipMemcpyAsync(dt.p, taps.data(), taps.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    // the Hilbert path filters the raw rows (its calcWeight is commented out,
    // reconstruct.cu:1763); the FDK filters take the cosine-weighted rows
    const bool hilbert = filter == XR_FILTER_HILBERT;
    if (!hilbert)
        XR_TRY(timed(ctx, 3,
                     [&] { return launch_fdk_weight(dw.p, g.n_det_u, g.n_det_v, g.sdd, g.det_size, ctx->stream); }));
    XR_TRY(timed(ctx, 2, [&] {
        return launch_fdk_filter(sino, filt, hilbert ? nullptr : dw.p, dt.p, g.n_det_u, g.n_det_v, ctx->d_codes,
                                 n_views, code_base, hilbert ? 1 : 0, ctx->stream);
    }));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));  // dw / dt are freed on return
    return XR_OK;
}

int xr_fdk_filter(xr_context* ctx, xr_plan* plan, int filter, const float* sino, float* filt, const int32_t* views,
                  int n_views) {
    return fdk_filter(ctx, plan, filter, sino, filt, views, n_views, 0);
}

static int fdk_backproject(xr_context* ctx, xr_plan* plan, const float* filt, const int32_t* views, int n_views,
                           int overwrite, float* vol, int code_base) {
    if (!ctx || !plan || !filt || !views || !vol || n_views < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (plan->conv != XR_CONV_FDK) return fail(ctx, XR_ERR_INVALID, "FDK backprojection needs an XR_CONV_FDK plan");
    if (n_views == 0) return XR_OK;
    XR_TRY(upload_views(plan, views, n_views, nullptr));
    return backward_launch(plan, BWD_FDK, filt, code_base, ctx->d_codes, n_views, vol, nullptr, nullptr, 0.f,
                           overwrite);
}

int xr_fdk_backproject(xr_context* ctx, xr_plan* plan, const float* filt, const int32_t* views, int n_views,
                       int overwrite, float* vol) {
    return fdk_backproject(ctx, plan, filt, views, n_views, overwrite, vol, 0);
}

int xr_volume_sqrt(xr_context* ctx, float* vol, int64_t n) {
    if (!ctx || !vol || n < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_sqrt_abs(vol, n, ctx->stream); });
}

int xr_pca(xr_context* ctx, const float* mu, int n_basis, const float* basis_host, int64_t nvox, int64_t chan_stride,
           float* md, float* evalue) {
    if (!ctx || !mu || !basis_host || !md || !evalue || n_basis <= 0 || n_basis > kMaxBasis || nvox < 0 ||
        chan_stride < nvox)
        return fail(ctx, XR_ERR_INVALID, "bad args");
    Basis b{};
    for (int k = 0; k < n_basis; k++)
        for (int i = 0; i < 3; i++) b.S[k][i] = basis_host[3 * k + i];
    return timed(ctx, 3, [&] { return launch_pca(mu, n_basis, b, nvox, chan_stride, md, evalue, ctx->stream); });
}

int xr_pca_host(xr_context* ctx, const float* mu_host, int n_basis, const float* basis_host, int64_t nvox,
                float* md_host, float* evalue_host) {
    if (!ctx || !mu_host || !basis_host || !md_host || !evalue_host || n_basis <= 0 || n_basis > kMaxBasis || nvox < 0)
        return fail(ctx, XR_ERR_INVALID, "bad args");
    if (nvox == 0) return XR_OK;
    DevBuf<float> dmu, dmd, dev;
    XR_HIP(ctx, dmu.alloc((size_t)n_basis * nvox));
    XR_HIP(ctx, dmd.alloc((size_t)3 * nvox));
    XR_HIP(ctx, dev.alloc((size_t)3 * nvox));
    XR_HIP(ctx, hipMemcpyAsync(dmu.p, mu_host, sizeof(float) * n_basis * nvox, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(xr_pca(ctx, dmu.p, n_basis, basis_host, nvox, nvox, dmd.p, dev.p));
    XR_HIP(ctx, hipMemcpyAsync(md_host, dmd.p, sizeof(float) * 3 * nvox, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipMemcpyAsync(evalue_host, dev.p, sizeof(float) * 3 * nvox, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_md_angles(xr_context* ctx, const float* md, float* angle, int64_t nvox) {
    if (!ctx || !md || !angle || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_md_angles(md, angle, nvox, ctx->stream); });
}

int xr_md_normalise(xr_context* ctx, float* md, int64_t nvox, float eps) {
    if (!ctx || !md || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_md_normalise(md, nvox, eps, ctx->stream); });
}

int xr_md_angles_host(xr_context* ctx, const float* md_host, float* angle_host, int64_t nvox) {
    if (!ctx || !md_host || !angle_host || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (nvox == 0) return XR_OK;
    DevBuf<float> dmd, dang;
    XR_HIP(ctx, dmd.alloc((size_t)3 * nvox));
    XR_HIP(ctx, dang.alloc((size_t)2 * nvox));
    XR_HIP(ctx, hipMemcpyAsync(dmd.p, md_host, sizeof(float) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(xr_md_angles(ctx, dmd.p, dang.p, nvox));
    XR_HIP(ctx, hipMemcpyAsync(angle_host, dang.p, sizeof(float) * 2 * nvox, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_md_normalise_host(xr_context* ctx, float* md_host, int64_t nvox, float eps) {
    if (!ctx || !md_host || nvox < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    if (nvox == 0) return XR_OK;
    DevBuf<float> dmd;
    XR_HIP(ctx, dmd.alloc((size_t)3 * nvox));
    XR_HIP(ctx, hipMemcpyAsync(dmd.p, md_host, sizeof(float) * 3 * nvox, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(xr_md_normalise(ctx, dmd.p, nvox, eps));
    XR_HIP(ctx, hipMemcpyAsync(md_host, dmd.p, sizeof(float) * 3 * nvox, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

// -------------------------------------------------------------- solvers ----

// Subset structure of reconstruct.cu:61-65, 89: subset s holds views
// (s + batch*j) % P, j < ceil(P/batch).
static void subset_views(int sub, int batch, int P, std::vector<int>& out) {
    const int ss = (P + batch - 1) / batch;
    out.resize(ss);
    for (int j = 0; j < ss; j++) out[j] = (sub + batch * j) % P;
}

int xr_ir_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                      int epochs, int batch, int direction, int method, float lambda, xr_stats* stats) {
    if (!ctx) return XR_ERR_INVALID;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (!sino || !vol) return fail(ctx, XR_ERR_INVALID, "sinogram / volume arrays missing");
    if (epochs < 0 || batch <= 0 || batch > g->n_proj) return fail(ctx, XR_ERR_INVALID, "bad epochs / batch");
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    xr_plan* plan = nullptr;
    XR_TRY(xr_plan_create(ctx, g, s, XR_CONV_IR, direction, &plan));
    std::unique_ptr<xr_plan, void (*)(xr_plan*)> guard(plan, xr_plan_destroy);
    XR_TRY(ensure_norm(plan));
    const int N = g->n_vox, P = g->n_proj, C = s->n_cond;
    const int64_t N3 = (int64_t)N * N * N, lenD = (int64_t)g->n_det_u * g->n_det_v * P;
    const int ss = (P + batch - 1) / batch;
    DevBuf<float> dref, dvol, dsino, dproj;
    DevBuf<int32_t> dcodes;
    DevBuf<uint8_t> dflags;
    XR_HIP(ctx, dref.alloc(N3));
    XR_HIP(ctx, dvol.alloc(N3));
    XR_HIP(ctx, dsino.alloc(lenD));
    XR_HIP(ctx, dproj.alloc(lenD));
    DevBuf<float> dvmax;
    XR_HIP(ctx, dvmax.alloc(1));
    const size_t per_epoch = (size_t)batch * ss;
    XR_HIP(ctx, dcodes.alloc(per_epoch * std::max(epochs, 1)));
    XR_HIP(ctx, dflags.alloc(per_epoch * std::max(epochs, 1)));
    // One engine per call, continuing across conditions (reconstruct.cu:75-81).
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::mt19937_64 rng;
    std::vector<int32_t> codes(per_epoch * std::max(epochs, 1));
    std::vector<uint8_t> flags(codes.size());
    std::vector<int> sv;
    const float alpha = (method == XR_METHOD_ART) ? lambda / (float)ss : 0.f;
    const int kind = (method == XR_METHOD_MLEM) ? BWD_MLEM : BWD_ART;
    const int out_mode = (method == XR_METHOD_MLEM) ? OUT_RATIO : OUT_RESIDUAL;
    for (int c = 0; c < C; c++) {
        // Views visited earlier in an epoch keep their image (proj is zeroed once
        // per epoch, reconstruct.cu:83), so a repeated view accumulates.
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
        XR_HIP(ctx, hipMemcpyAsync(dref.p, vol[c], sizeof(float) * N3, hipMemcpyHostToDevice, ctx->stream));
        XR_TRY(xr_volume_to_columns(ctx, dref.p, dvol.p, N, 1));
        // One condition at a time: its images hold codes c*P .. c*P+P-1.
        XR_HIP(ctx, hipMemcpyAsync(dsino.p, sino[c], sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
        // max |voxel| for the forward's fixed-point scale: measured from the
        // uploaded volume, then left by each backprojection for the next
        // subset (the update writes every voxel once)
        bool vmax_known = false;
        for (int ep = 0; ep < epochs; ep++) {
            XR_HIP(ctx, hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
            for (int pos = 0; pos < batch; pos++) {
                const size_t off = (size_t)ep * per_epoch + (size_t)pos * ss;
                XR_TRY(forward_launch(plan, FWD_PLAIN, dvol.p, 0, dproj.p, dsino.p, c * P, dcodes.p + off,
                                      dflags.p + off, ss, out_mode, ctx->d_loss, vmax_known ? dvmax.p : nullptr));
                XR_HIP(ctx, launch_fill(dvmax.p, 1, 0.0f, ctx->stream));
                XR_TRY(backward_launch(plan, kind, dproj.p, c * P, dcodes.p + off, ss, dvol.p, nullptr, nullptr, alpha,
                                       0, reinterpret_cast<unsigned*>(dvmax.p)));
                vmax_known = true;
            }
            if (stats && stats->losses) {
                XR_HIP(ctx, hipMemcpyAsync(&stats->losses[(size_t)c * epochs + ep], ctx->d_loss, sizeof(double),
                                           hipMemcpyDeviceToHost, ctx->stream));
            }
        }
        XR_TRY(xr_volume_from_columns(ctx, dvol.p, dref.p, N, 1));
        XR_HIP(ctx, hipMemcpyAsync(vol[c], dref.p, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

int xr_xtt_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                       float* const* md, float* const* evalue, int epochs, int batch, int direction, int method,
                       float lambda, xr_stats* stats) {
    if (!ctx) return XR_ERR_INVALID;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (s->n_basis <= 0) return fail(ctx, XR_ERR_INVALID, "XTT needs basis vectors");
    if (!sino || !vol) return fail(ctx, XR_ERR_INVALID, "sinogram / volume arrays missing");
    if (epochs < 0 || batch <= 0 || batch > g->n_proj) return fail(ctx, XR_ERR_INVALID, "bad epochs / batch");
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    xr_plan* plan = nullptr;
    XR_TRY(xr_plan_create(ctx, g, s, XR_CONV_IR, direction, &plan));
    std::unique_ptr<xr_plan, void (*)(xr_plan*)> guard(plan, xr_plan_destroy);
    XR_TRY(ensure_norm(plan));
    const int N = g->n_vox, P = g->n_proj, C = s->n_cond, K = s->n_basis;
    const int64_t N3 = (int64_t)N * N * N, lenD = (int64_t)g->n_det_u * g->n_det_v * P;
    const int ss = (P + batch - 1) / batch;
    DevBuf<float> dref, dvol, dsino, dproj;
    DevBuf<int32_t> dcodes;
    DevBuf<uint8_t> dflags;
    XR_HIP(ctx, dref.alloc(N3 * std::max(K, 3)));
    XR_HIP(ctx, dvol.alloc(N3 * std::max(K, 6)));
    XR_HIP(ctx, dsino.alloc(lenD * C));
    XR_HIP(ctx, dproj.alloc(lenD * C));
    const size_t per_epoch = (size_t)batch * C * ss;
    XR_HIP(ctx, dcodes.alloc(per_epoch * std::max(epochs, 1)));
    XR_HIP(ctx, dflags.alloc(per_epoch * std::max(epochs, 1)));
    std::vector<int> order(batch);
    for (int i = 0; i < batch; i++) order[i] = i;
    std::vector<int32_t> codes(per_epoch * std::max(epochs, 1));
    std::vector<uint8_t> flags(codes.size());
    std::vector<int> sv;
    for (int ep = 0; ep < epochs; ep++) {
        // XTT re-creates the engine every epoch but shuffles the order left by the
        // previous epoch (reconstruct.cu:843-859).
        std::mt19937_64 rng;
        std::shuffle(order.begin(), order.end(), rng);
        std::vector<char> visited((size_t)C * P, 0);
        for (int pos = 0; pos < batch; pos++) {
            subset_views(order[pos], batch, P, sv);
            for (int c = 0; c < C; c++)
                for (int j = 0; j < ss; j++) {
                    const size_t i = (size_t)ep * per_epoch + (size_t)pos * C * ss + (size_t)c * ss + j;
                    const int code = c * P + sv[j];
                    codes[i] = code;
                    flags[i] = visited[code];
                    visited[code] = 1;
                }
        }
    }
    if (epochs > 0) {
        XR_HIP(ctx, hipMemcpyAsync(dcodes.p, codes.data(), codes.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                   ctx->stream));
        XR_HIP(ctx, hipMemcpyAsync(dflags.p, flags.data(), flags.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    for (int k = 0; k < K; k++)
        XR_HIP(ctx, hipMemcpyAsync(dref.p + k * N3, vol[k], sizeof(float) * N3, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(xr_volume_to_columns(ctx, dref.p, dvol.p, N, K));
    for (int c = 0; c < C; c++)
        XR_HIP(ctx, hipMemcpyAsync(dsino.p + c * lenD, sino[c], sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
    const float alpha = (method == XR_METHOD_ART) ? lambda / (float)ss : 0.f;
    const int kind = (method == XR_METHOD_MLEM) ? BWD_XTT_MLEM : BWD_XTT_ART;
    const int out_mode = (method == XR_METHOD_MLEM) ? OUT_RATIO : OUT_RESIDUAL;
    for (int ep = 0; ep < epochs; ep++) {
        XR_HIP(ctx, hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
        for (int pos = 0; pos < batch; pos++) {
            const size_t off = (size_t)ep * per_epoch + (size_t)pos * C * ss;
            XR_TRY(forward_launch(plan, FWD_XTT, dvol.p, 0, dproj.p, dsino.p, 0, dcodes.p + off, dflags.p + off,
                                  C * ss, out_mode, ctx->d_loss));
            XR_TRY(backward_launch(plan, kind, dproj.p, 0, dcodes.p + off, C * ss, dvol.p, nullptr, nullptr, alpha, 0));
        }
        if (stats && stats->losses)
            XR_HIP(ctx, hipMemcpyAsync(&stats->losses[ep], ctx->d_loss, sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    }
    // voxelSqrt (reconstruct.cu:919-922), sinogram overwrite (:924-925), volumes (:926-927).
    XR_TRY(xr_volume_sqrt(ctx, dvol.p, N3 * K));
    for (int c = 0; c < C; c++)
        XR_HIP(ctx, hipMemcpyAsync(sino[c], dproj.p + c * lenD, sizeof(float) * lenD, hipMemcpyDeviceToHost, ctx->stream));
    XR_TRY(xr_volume_from_columns(ctx, dvol.p, dref.p, N, K));
    for (int k = 0; k < K; k++)
        XR_HIP(ctx, hipMemcpyAsync(vol[k], dref.p + k * N3, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (md) {
        // calcEigenVector over all voxels (reconstruct.cu:935-942), on the device,
        // from the reference-layout volumes already in dref.
        DevBuf<float> dmd, dev;
        XR_HIP(ctx, dmd.alloc(N3 * 3));
        XR_HIP(ctx, dev.alloc(N3 * 3));
        XR_TRY(xr_pca(ctx, dref.p, K, s->basis, N3, N3, dmd.p, dev.p));
        for (int i = 0; i < 3; i++) {
            XR_HIP(ctx, hipMemcpyAsync(md[i], dmd.p + i * N3, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
            if (evalue && evalue[i])
                XR_HIP(ctx, hipMemcpyAsync(evalue[i], dev.p + i * N3, sizeof(float) * N3, hipMemcpyDeviceToHost,
                                           ctx->stream));
        }
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

// XTT::orthTwiceReconstruct (reconstruct.cu:956-1239).  Per outer iteration
// five passes: pass 0 runs iter2 epochs of the orientation-aware operator from
// the stored orientation and keeps the result as the volume; passes 1-4 first
// re-estimate the orientation from that volume with sign choice pass-1
// (calcNormalVectorThreeDirec), reset channel 0 to 0.1 (fillVolume touches
// channel 0 only), run iter2/4 epochs and record the best normal per voxel
// (updateEstimationByCoef).  Every pass ends with voxelSqrt, the main
// directions from the kept volume and stored orientation (convertNormVector)
// and the device orientation reset to the stored one.  After the passes the
// orientation is re-estimated with the recorded sign choices
// (calcNormalVectorThreeDirecWithEst) and stored.  One std::mt19937 shuffles
// the subset order for the whole call (seeded here: the reference seeds it
// from std::random_device); the estimate starts at zero (the reference leaves
// it uninitialised).
int xr_orth_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                        float* const* md, int iter1, int iter2, int batch, int direction, int method, float lambda,
                        uint32_t seed, const char* volume_path, const char* direction_path) {
    if (!ctx) return XR_ERR_INVALID;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (s->n_basis != 3) return fail(ctx, XR_ERR_INVALID, "the orientation-aware model has three basis vectors");
    if (!sino || !vol || !md) return fail(ctx, XR_ERR_INVALID, "sinogram / volume / direction arrays missing");
    if (iter1 < 0 || iter2 < 0 || batch <= 0 || batch > g->n_proj) return fail(ctx, XR_ERR_INVALID, "bad iterations / batch");
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    xr_plan* plan = nullptr;
    XR_TRY(xr_plan_create(ctx, g, s, XR_CONV_IR, direction, &plan));
    std::unique_ptr<xr_plan, void (*)(xr_plan*)> guard(plan, xr_plan_destroy);
    XR_TRY(ensure_norm(plan));
    const int N = g->n_vox, P = g->n_proj, C = s->n_cond;
    const int64_t N3 = (int64_t)N * N * N, lenD = (int64_t)g->n_det_u * g->n_det_v * P;
    const int ss = (P + batch - 1) / batch;
    DevBuf<float> dref, dvol, dkept, dsino, dproj, dcoef, dstore, dest, dloss, dmd;
    DevBuf<int32_t> dcodes;
    DevBuf<uint8_t> dflags;
    XR_HIP(ctx, dref.alloc(N3 * 3));
    XR_HIP(ctx, dvol.alloc(N3 * 3));
    XR_HIP(ctx, dkept.alloc(N3 * 3));
    XR_HIP(ctx, dsino.alloc(lenD * C));
    XR_HIP(ctx, dproj.alloc(lenD * C));
    XR_HIP(ctx, dcoef.alloc(N3 * 2));
    XR_HIP(ctx, dstore.alloc(N3 * 2));
    XR_HIP(ctx, dest.alloc(N3 * 2));
    XR_HIP(ctx, dloss.alloc(N3));
    XR_HIP(ctx, dmd.alloc(N3 * 3));
    const size_t per_epoch = (size_t)batch * C * ss;
    XR_HIP(ctx, dcodes.alloc(per_epoch));
    XR_HIP(ctx, dflags.alloc(per_epoch));
    for (int k = 0; k < 3; k++)
        XR_HIP(ctx, hipMemcpyAsync(dref.p + k * N3, vol[k], sizeof(float) * N3, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(xr_volume_to_columns(ctx, dref.p, dvol.p, N, 3));
    XR_HIP(ctx, hipMemcpyAsync(dkept.p, dvol.p, sizeof(float) * N3 * 3, hipMemcpyDeviceToDevice, ctx->stream));
    for (int c = 0; c < C; c++)
        XR_HIP(ctx, hipMemcpyAsync(dsino.p + c * lenD, sino[c], sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
    // stored orientation (phi 0, cos theta 1: no rotation) and the estimate
    XR_HIP(ctx, launch_fill(dstore.p, N3, 0.0f, ctx->stream));
    XR_HIP(ctx, launch_fill(dstore.p + N3, N3, 1.0f, ctx->stream));
    XR_HIP(ctx, hipMemcpyAsync(dcoef.p, dstore.p, sizeof(float) * N3 * 2, hipMemcpyDeviceToDevice, ctx->stream));
    XR_HIP(ctx, launch_fill(dest.p, N3 * 2, 0.0f, ctx->stream));
    XR_TRY(xr_plan_set_orientation(plan, dcoef.p));
    const float alpha = (method == XR_METHOD_ART) ? lambda / (float)ss : 0.f;
    const int kind = (method == XR_METHOD_MLEM) ? BWD_XTT_MLEM : BWD_XTT_ART;
    const int out_mode = (method == XR_METHOD_MLEM) ? OUT_RATIO : OUT_RESIDUAL;
    std::mt19937 engine(seed);
    std::vector<int> order(batch), sv;
    for (int i = 0; i < batch; i++) order[i] = i;
    std::vector<int32_t> codes(per_epoch);
    std::vector<uint8_t> flags(per_epoch);
    auto save = [&](const char* base, int outer, const float* dev_cols, int ch) -> int {
        // raw float32, reference layout: base + "_sequence<outer>d<i>_NxNxN.raw" (reconstruct.cu:1203-1216)
        XR_TRY(xr_volume_from_columns(ctx, dev_cols, dref.p, N, ch));
        std::vector<float> h(N3);
        for (int i = 0; i < ch; i++) {
            XR_HIP(ctx, hipMemcpyAsync(h.data(), dref.p + i * N3, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
            XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            char name[4096];
            snprintf(name, sizeof(name), "%s_sequence%dd%d_%dx%dx%d.raw", base, outer + 1, i + 1, N, N, N);
            FILE* f = fopen(name, "wb");
            if (!f) return fail(ctx, XR_ERR_IO, "cannot write %s", name);
            const size_t w = fwrite(h.data(), sizeof(float), (size_t)N3, f);
            fclose(f);
            if (w != (size_t)N3) return fail(ctx, XR_ERR_IO, "short write %s", name);
        }
        return XR_OK;
    };
    for (int outer = 0; outer < iter1; outer++) {
        for (int ep1 = 0; ep1 < 5; ep1++) {
            if (ep1 != 0) {
                XR_HIP(ctx, hipMemcpyAsync(dvol.p, dkept.p, sizeof(float) * N3 * 3, hipMemcpyDeviceToDevice, ctx->stream));
                XR_TRY(xr_orth_normal(ctx, s->basis, dvol.p, N3, dcoef.p, dloss.p, ep1 - 1, nullptr));
            }
            XR_HIP(ctx, launch_fill(dvol.p, N3, 0.1f, ctx->stream));  // channel 0 only
            const int epochs = (ep1 == 0) ? iter2 : iter2 / 4;
            for (int ep2 = 0; ep2 < epochs; ep2++) {
                std::shuffle(order.begin(), order.end(), engine);
                std::vector<char> visited((size_t)C * P, 0);
                for (int pos = 0; pos < batch; pos++) {
                    subset_views(order[pos], batch, P, sv);
                    for (int c = 0; c < C; c++)
                        for (int j = 0; j < ss; j++) {
                            const size_t i = (size_t)pos * C * ss + (size_t)c * ss + j;
                            const int code = c * P + sv[j];
                            codes[i] = code;
                            flags[i] = visited[code];
                            visited[code] = 1;
                        }
                }
                XR_HIP(ctx, hipMemcpyAsync(dcodes.p, codes.data(), codes.size() * sizeof(int32_t),
                                           hipMemcpyHostToDevice, ctx->stream));
                XR_HIP(ctx, hipMemcpyAsync(dflags.p, flags.data(), flags.size(), hipMemcpyHostToDevice, ctx->stream));
                for (int pos = 0; pos < batch; pos++) {
                    const size_t off = (size_t)pos * C * ss;
                    XR_TRY(forward_launch(plan, FWD_XTT, dvol.p, 0, dproj.p, dsino.p, 0, dcodes.p + off, dflags.p + off,
                                          C * ss, out_mode, ctx->d_loss));
                    XR_TRY(backward_launch(plan, kind, dproj.p, 0, dcodes.p + off, C * ss, dvol.p, nullptr, nullptr,
                                           alpha, 0));
                }
                // the host copy of the subset lists is rewritten next epoch
                XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            }
            XR_TRY(xr_volume_sqrt(ctx, dvol.p, N3 * 3));
            if (ep1 != 0)
                XR_TRY(xr_orth_update_estimate(ctx, s->basis, dvol.p, N3, dest.p, ep1 - 1));
            else
                XR_HIP(ctx, hipMemcpyAsync(dkept.p, dvol.p, sizeof(float) * N3 * 3, hipMemcpyDeviceToDevice, ctx->stream));
            XR_TRY(xr_orth_md(ctx, dkept.p, dstore.p, N3, dmd.p));
            XR_HIP(ctx, hipMemcpyAsync(dcoef.p, dstore.p, sizeof(float) * N3 * 2, hipMemcpyDeviceToDevice, ctx->stream));
        }
        XR_HIP(ctx, hipMemcpyAsync(dvol.p, dkept.p, sizeof(float) * N3 * 3, hipMemcpyDeviceToDevice, ctx->stream));
        XR_TRY(xr_orth_normal(ctx, s->basis, dvol.p, N3, dcoef.p, dloss.p, 0, dest.p));
        XR_HIP(ctx, hipMemcpyAsync(dstore.p, dcoef.p, sizeof(float) * N3 * 2, hipMemcpyDeviceToDevice, ctx->stream));
        XR_TRY(xr_orth_md(ctx, dkept.p, dstore.p, N3, dmd.p));
        if (direction_path) XR_TRY(save(direction_path, outer, dmd.p, 3));
        if (volume_path) XR_TRY(save(volume_path, outer, dkept.p, 3));
    }
    XR_TRY(xr_plan_set_orientation(plan, nullptr));
    // results: the kept volume, its main directions, the final ratio / residual images
    XR_TRY(xr_volume_from_columns(ctx, dkept.p, dref.p, N, 3));
    for (int k = 0; k < 3; k++)
        XR_HIP(ctx, hipMemcpyAsync(vol[k], dref.p + k * N3, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
    XR_TRY(xr_volume_from_columns(ctx, dmd.p, dref.p, N, 3));
    for (int k = 0; k < 3; k++)
        XR_HIP(ctx, hipMemcpyAsync(md[k], dref.p + k * N3, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_fdk_reconstruct(xr_context* ctx, const xr_geometry* g, const xr_scan* s, float* const* sino, float* const* vol,
                       int direction, int filter) {
    if (!ctx) return XR_ERR_INVALID;
    XR_TRY(check_geometry(ctx, g));
    XR_TRY(check_scan(ctx, s));
    if (!sino || !vol) return fail(ctx, XR_ERR_INVALID, "sinogram / volume arrays missing");
    xr_plan* plan = nullptr;
    XR_TRY(xr_plan_create(ctx, g, s, XR_CONV_FDK, direction, &plan));
    std::unique_ptr<xr_plan, void (*)(xr_plan*)> guard(plan, xr_plan_destroy);
    const int N = g->n_vox, P = g->n_proj, C = s->n_cond;
    const int64_t N3 = (int64_t)N * N * N, lenD = (int64_t)g->n_det_u * g->n_det_v * P;
    DevBuf<float> dsino, dfilt, dvol, dref;
    XR_HIP(ctx, dsino.alloc(lenD));
    XR_HIP(ctx, dfilt.alloc(lenD));
    XR_HIP(ctx, dvol.alloc(N3));
    XR_HIP(ctx, dref.alloc(N3));
    std::vector<int32_t> codes(P);
    for (int c = 0; c < C; c++) {
        for (int n = 0; n < P; n++) codes[n] = c * P + n;
        XR_HIP(ctx, hipMemcpyAsync(dsino.p, sino[c], sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
        XR_TRY(fdk_filter(ctx, plan, filter, dsino.p, dfilt.p, codes.data(), P, c * P));
        XR_TRY(fdk_backproject(ctx, plan, dfilt.p, codes.data(), P, 1, dvol.p, c * P));
        XR_TRY(xr_volume_from_columns(ctx, dvol.p, dref.p, N, 1));
        XR_HIP(ctx, hipMemcpyAsync(vol[c], dref.p, sizeof(float) * N3, hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipMemcpyAsync(sino[c], dfilt.p, sizeof(float) * lenD, hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

// ------------------------------------------------------------- settings ----

int xr_load_settings(const char* json_path, const char* tag, xr_settings* out) {
    if (!json_path || !tag || !out) return XR_ERR_INVALID;
    FILE* f = fopen(json_path, "rb");
    if (!f) return XR_ERR_IO;
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    fclose(f);
    JParser jp{text.data(), text.data() + text.size()};
    const JVal root = jp.val();
    if (!jp.ok || root.kind != JVal::OBJ) return XR_ERR_IO;
    const JVal* rv = root.get("recon_variable");
    const JVal* t = root.get(tag);
    if (!rv || !t || t->kind != JVal::OBJ) return XR_ERR_NOT_FOUND;
    std::memset(out, 0, sizeof(*out));
    auto copy_str = [](const JVal* v, char* dst, size_t cap) {
        if (v && v->kind == JVal::STR) snprintf(dst, cap, "%s", v->str.c_str());
    };
    copy_str(t->get("proj_path"), out->proj_path, sizeof(out->proj_path));
    copy_str(t->get("vol_path"), out->vol_path, sizeof(out->vol_path));
    copy_str(t->get("direc_path"), out->direc_path, sizeof(out->direc_path));
    double x;
    if (!jnum(rv->get("blockSize"), x)) return XR_ERR_NOT_FOUND;
    out->block_size = (int)x;
    if (!jnum(rv->get("vector"), x)) return XR_ERR_NOT_FOUND;
    out->n_basis = (int)x;
    if (!jnum(rv->get("condition"), x)) return XR_ERR_NOT_FOUND;
    out->n_cond = (int)x;
    if (out->n_basis < 0 || out->n_basis > 16 || out->n_cond <= 0 || out->n_cond > 64) return XR_ERR_INVALID;
    struct {
        const char* key;
        float* dst;
    } fl[] = {{"sod", &out->sod}, {"sdd", &out->sdd}, {"det_size", &out->det_size}};
    for (auto& e : fl) {
        if (!jnum(t->get(e.key), x)) return XR_ERR_NOT_FOUND;
        *e.dst = (float)x;
    }
    struct {
        const char* key;
        int32_t* dst;
    } in[] = {{"proj", &out->n_proj}, {"num_det_u", &out->n_det_u}, {"num_det_v", &out->n_det_v},
              {"num_voxel", &out->n_vox}};
    for (auto& e : in) {
        if (!jnum(t->get(e.key), x)) return XR_ERR_NOT_FOUND;
        *e.dst = (int32_t)x;
    }
    if (!jnum(t->get("rot"), x)) return XR_ERR_NOT_FOUND;
    const int rot = (int)x;
    std::vector<double> mR, vT, aT, off, base, index;
    if (!jarr(t->get("matrixRot"), mR) || !jarr(t->get("vecTrans"), vT) || !jarr(t->get("areaTrans"), aT) ||
        !jarr(t->get("offset"), off) || !jarr(rv->get("base"), base))
        return XR_ERR_NOT_FOUND;
    if ((int)base.size() < 3 * out->n_basis || aT.size() < 3 || (int)mR.size() < 9 * rot || (int)vT.size() < 3 * rot ||
        (int)off.size() < 3 * rot)
        return XR_ERR_INVALID;
    for (int i = 0; i < 3 * out->n_basis; i++) out->basis[i] = (float)base[i];
    // vecTrans += areaTrans for every rotation (params.cu:74-78), in FP32 as the reference does.
    std::vector<float> vTf(vT.size());
    for (size_t i = 0; i < vT.size(); i++) vTf[i] = (float)vT[i];
    for (int i = 0; i < rot; i++)
        for (int j = 0; j < 3; j++) vTf[3 * i + j] += (float)aT[j];
    // A missing "sorting" key reads as false (the reference's json lookup would throw).
    const JVal* so = t->get("sorting");
    const bool sorting = so && so->kind == JVal::BOOL && so->b;
    if (sorting) {
        if (!jarr(t->get("index"), index) || (int)index.size() < out->n_cond) return XR_ERR_NOT_FOUND;
    }
    if (!sorting && out->n_cond > rot) return XR_ERR_INVALID;
    for (int c = 0; c < out->n_cond; c++) {
        const int src = sorting ? (int)index[c] : c;
        if (src < 0 || src >= rot) return XR_ERR_INVALID;
        out->load_index[c] = src + 1;  // params.cu:83, 90
        for (int j = 0; j < 9; j++) out->cond_R[9 * c + j] = (float)mR[9 * src + j];
        for (int j = 0; j < 3; j++) {
            out->cond_t[3 * c + j] = vTf[3 * src + j];
            out->cond_offset[3 * c + j] = (float)off[3 * src + j];
        }
    }
    return XR_OK;
}

int xr_settings_geometry(const xr_settings* st, xr_geometry* g, xr_scan* s) {
    if (!st || !g || !s) return XR_ERR_INVALID;
    g->sdd = st->sdd;
    g->sod = st->sod;
    g->det_size = st->det_size;
    g->vox_size = xr_default_vox_size(st->sdd, st->sod, st->det_size);
    g->n_vox = st->n_vox;
    g->n_det_u = st->n_det_u;
    g->n_det_v = st->n_det_v;
    g->n_proj = st->n_proj;
    s->n_cond = st->n_cond;
    s->n_basis = st->n_basis;
    s->cond_R = st->cond_R;
    s->cond_t = st->cond_t;
    s->cond_offset = st->cond_offset;
    s->basis = st->basis;
    s->grating = nullptr;
    return XR_OK;
}

// ------------------------------------------------------------- phantoms ----

int xr_phantom_ellipsoid_params(uint64_t seed, int count, double fov_radius, double centre_frac, double axis_lo,
                                double axis_hi, double value_lo, double value_hi, double* out) {
    if (count < 0 || !out || !(fov_radius > 0) || !(axis_lo > 0) || axis_hi < axis_lo) return XR_ERR_INVALID;
    SplitMix64 r{seed};
    for (int e = 0; e < count; e++) {
        double* q = out + 16 * e;
        double c[3];
        do {
            for (int i = 0; i < 3; i++) c[i] = 2.0 * r.uniform() - 1.0;
        } while (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 1.0);
        for (int i = 0; i < 3; i++) q[i] = c[i] * centre_frac * fov_radius;
        for (int i = 0; i < 3; i++) q[3 + i] = (axis_lo + (axis_hi - axis_lo) * r.uniform()) * fov_radius;
        // uniform rotation from a uniform unit quaternion (Shoemake)
        const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
        const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        const double w = a * std::sin(2 * M_PI * u2), x = a * std::cos(2 * M_PI * u2);
        const double y = b * std::sin(2 * M_PI * u3), z = b * std::cos(2 * M_PI * u3);
        const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w),     2 * (x * z + y * w),
                             2 * (x * y + z * w),     1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                             2 * (x * z - y * w),     2 * (y * z + x * w),     1 - 2 * (x * x + y * y)};
        for (int i = 0; i < 9; i++) q[6 + i] = R[i];
        q[15] = value_lo + (value_hi - value_lo) * r.uniform();
    }
    return XR_OK;
}

int xr_phantom_rasterise(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                         float* vol_ref_dev) {
    if (!ctx || (!params_host && count > 0) || count < 0 || n_vox <= 0 || !vol_ref_dev)
        return fail(ctx, XR_ERR_INVALID, "bad args");
    DevBuf<double> dp;
    XR_HIP(ctx, dp.alloc((size_t)16 * std::max(count, 1)));
    if (count > 0)
        XR_HIP(ctx, hipMemcpyAsync(dp.p, params_host, sizeof(double) * 16 * count, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(timed(ctx, 3, [&] { return launch_ellipsoids(vol_ref_dev, n_vox, vox_size, dp.p, count, ctx->stream); }));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_phantom_cylinder_shell(xr_context* ctx, int n_vox, float vox_size, double r_out, double thickness,
                              double value, float* vol_ref_dev) {
    if (!ctx || n_vox <= 0 || !vol_ref_dev || !(vox_size > 0)) return fail(ctx, XR_ERR_INVALID, "bad args");
    XR_TRY(timed(ctx, 3, [&] {
        return launch_cylinder_shell(vol_ref_dev, n_vox, vox_size, r_out, thickness, value, ctx->stream);
    }));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_phantom_fibre_params(uint64_t seed, int count, int n_axes, double kappa, double fov_radius,
                            double centre_frac, double r_lo, double r_hi, double half_length, double* out,
                            double* axes_out) {
    if (count < 0 || n_axes <= 0 || !out || !axes_out || !(kappa > 0) || !(fov_radius > 0) || !(r_lo > 0) ||
        r_hi < r_lo || !(half_length > 0))
        return XR_ERR_INVALID;
    SplitMix64 r{seed};
    for (int a = 0; a < n_axes; a++) {  // uniform on the sphere
        const double z = 2.0 * r.uniform() - 1.0, ph = 2.0 * M_PI * r.uniform();
        const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        axes_out[3 * a] = s * std::cos(ph);
        axes_out[3 * a + 1] = s * std::sin(ph);
        axes_out[3 * a + 2] = z;
    }
    for (int e = 0; e < count; e++) {
        double* q = out + 8 * e;
        double c[3];
        do {
            for (int i = 0; i < 3; i++) c[i] = 2.0 * r.uniform() - 1.0;
        } while (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 1.0);
        for (int i = 0; i < 3; i++) q[i] = c[i] * centre_frac * fov_radius;
        // von Mises-Fisher on S^2 about mu (Wood 1994): w = 1 + ln(U + (1-U) e^{-2 kappa}) / kappa
        const double* mu = axes_out + 3 * (e % n_axes);
        const double U = r.uniform(), ph = 2.0 * M_PI * r.uniform();
        const double w = 1.0 + std::log(U + (1.0 - U) * std::exp(-2.0 * kappa)) / kappa;
        const double sw = std::sqrt(std::max(0.0, 1.0 - w * w));
        // orthonormal e1, e2 perpendicular to mu
        double t[3] = {0, 0, 0};
        const int m = (std::fabs(mu[0]) <= std::fabs(mu[1]) && std::fabs(mu[0]) <= std::fabs(mu[2])) ? 0
                      : (std::fabs(mu[1]) <= std::fabs(mu[2]))                                           ? 1
                                                                                                          : 2;
        t[m] = 1.0;
        double e1[3] = {mu[1] * t[2] - mu[2] * t[1], mu[2] * t[0] - mu[0] * t[2], mu[0] * t[1] - mu[1] * t[0]};
        const double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (int i = 0; i < 3; i++) e1[i] /= n1;
        const double e2[3] = {mu[1] * e1[2] - mu[2] * e1[1], mu[2] * e1[0] - mu[0] * e1[2],
                              mu[0] * e1[1] - mu[1] * e1[0]};
        for (int i = 0; i < 3; i++) q[3 + i] = w * mu[i] + sw * (std::cos(ph) * e1[i] + std::sin(ph) * e2[i]);
        const double nf = std::sqrt(q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
        for (int i = 0; i < 3; i++) q[3 + i] /= nf;
        q[6] = r_lo + (r_hi - r_lo) * r.uniform();
        q[7] = half_length;
    }
    return XR_OK;
}

int xr_phantom_fibres(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                      const float* basis_host, int n_basis, float* vol_ref_dev) {
    if (!ctx || (!params_host && count > 0) || count < 0 || n_vox <= 0 || !basis_host || n_basis <= 0 ||
        n_basis > kMaxBasis || !vol_ref_dev)
        return fail(ctx, XR_ERR_INVALID, "bad args");
    Basis b{};
    for (int k = 0; k < n_basis; k++)
        for (int i = 0; i < 3; i++) b.S[k][i] = basis_host[3 * k + i];
    DevBuf<double> dp;
    XR_HIP(ctx, dp.alloc((size_t)8 * std::max(count, 1)));
    if (count > 0)
        XR_HIP(ctx, hipMemcpyAsync(dp.p, params_host, sizeof(double) * 8 * count, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(timed(ctx, 3, [&] { return launch_fibres(vol_ref_dev, n_vox, vox_size, dp.p, count, b, n_basis, ctx->stream); }));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_phantom_fibres_scalar(xr_context* ctx, const double* params_host, int count, int n_vox, float vox_size,
                             float value, float* vol_ref_dev) {
    if (!ctx || (!params_host && count > 0) || count < 0 || n_vox <= 0 || !vol_ref_dev)
        return fail(ctx, XR_ERR_INVALID, "bad args");
    DevBuf<double> dp;
    XR_HIP(ctx, dp.alloc((size_t)8 * std::max(count, 1)));
    if (count > 0)
        XR_HIP(ctx, hipMemcpyAsync(dp.p, params_host, sizeof(double) * 8 * count, hipMemcpyHostToDevice, ctx->stream));
    XR_TRY(timed(ctx, 3, [&] {
        return launch_fibres_scalar(vol_ref_dev, n_vox, vox_size, dp.p, count, value, ctx->stream);
    }));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XR_OK;
}

int xr_noise_poisson(xr_context* ctx, float* y_dev, const float* p_dev, int64_t n, double I0, uint64_t seed) {
    if (!ctx || !y_dev || !p_dev || n < 0 || !(I0 > 0)) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_poisson_noise(y_dev, p_dev, n, I0, seed, ctx->stream); });
}

int xr_noise_gaussian(xr_context* ctx, float* y_dev, const float* x_dev, int64_t n, double sigma, uint64_t seed) {
    if (!ctx || !y_dev || !x_dev || n < 0 || sigma < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_gaussian_noise(y_dev, x_dev, n, sigma, seed, ctx->stream); });
}

int xr_sino_clamp(xr_context* ctx, float* sino_dev, int64_t n) {
    if (!ctx || !sino_dev || n < 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    return timed(ctx, 3, [&] { return launch_clamp_min(sino_dev, n, 0.0f, 1e-8f, ctx->stream); });
}

void xr_philox4x32(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {counter[0], counter[1], counter[2], counter[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; i++) out[i] = c[i];
}

}  // extern "C"
