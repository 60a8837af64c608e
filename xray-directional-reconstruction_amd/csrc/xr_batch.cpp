// Batched multi-problem IR solver and projector tile autotuner (C ABI in
// include/xr.h).
//
// A batch holds n independent problems of one method (MLEM or ART): IR
// problems (one condition each) and XTT problems (any number of conditions,
// K scattering-direction coefficients), each with any geometry, view count and
// subset count.  IR problems and XTT problems each run in their own persistent
// launches of a subset step.  A subset step of every problem is two
// persistent launches: k_forward_batch over all (problem, row chunk, band,
// view) blocks and k_backward_batch over all (problem, voxel tile) blocks,
// each drawing work from a self-resetting ticket queue in descending order of
// problem size.  With XR_BATCH_GRAPH the whole epoch loop (zero fills, the
// per-problem max |voxel| reduction, both launches per step) is captured once
// as a HIP graph and replayed by later runs with the same epoch count; only
// the subset lists are uploaded per run, outside the graph.
//
// Each problem follows IR::reconstruct (reconstruct.cu:23-156) exactly as
// xr_ir_reconstruct does for one condition: its own default-seeded
// std::mt19937_64 shuffles the subset order every epoch (reconstruct.cu:75,
// 89), views revisited within an epoch accumulate onto their image
// (reconstruct.cu:83), the forward pass writes the MLEM ratio / ART residual
// and the backprojection applies the update in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <type_traits>
#include <vector>

#include "../../include/xr.h"
#include "xr_host.hpp"
#include "xr_internal.hpp"

using namespace xr;
using namespace xr::host;

namespace {

bool valid_tiles(const xr_tiles& t) {
    const int z = t.bwd_zpl, g = t.bwd_gx;
    return (z == 0 || z == 1 || z == 2 || z == 4) && (g == 0 || g == 1 || g == 2 || g == 4 || g == 8 || g == 16) &&
           t.fwd_rows >= 0 && (t.fwd_scan == 0 || t.fwd_scan == 1) &&
           (t.fwd_split == 0 || t.fwd_split == 1 || t.fwd_split == 2 || t.fwd_split == 4);
}

std::array<int, 5> tune_key(const xr_plan* p) {
    return {p->g.n_vox, p->g.n_det_u, p->g.n_det_v, p->g.n_proj, p->axis_cond.empty() ? 0 : (int)p->axis_cond[0]};
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    hipError_t init() {
        hipError_t e = hipEventCreate(&a);
        if (e != hipSuccess) return e;
        return hipEventCreate(&b);
    }
};

// Subset lists of one problem for `epochs` epochs, laid out [epoch][pos][j]
// with pos < maxb (positions >= batch unused).  `rng` and `order` are the
// problem's engine and subset order, continued from the previous run, so that
// consecutive runs of a and b epochs visit subsets exactly as one
// xr_ir_reconstruct solve of a + b epochs does (reconstruct.cu:75, 89).
void problem_codes(int P, int batch, int maxb, int epochs, std::mt19937_64& rng, std::vector<int>& order,
                   std::vector<int32_t>& codes, std::vector<uint8_t>& flags, bool& any_flag) {
    const int ss = (P + batch - 1) / batch;
    codes.assign((size_t)epochs * maxb * ss, 0);
    flags.assign(codes.size(), 0);
    any_flag = false;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(order.begin(), order.end(), rng);
        std::vector<char> visited(P, 0);
        for (int pos = 0; pos < batch; pos++) {
            for (int j = 0; j < ss; j++) {
                const int v = (order[pos] + batch * j) % P;  // reconstruct.cu:61-65
                const size_t i = ((size_t)ep * maxb + pos) * ss + j;
                codes[i] = v;
                flags[i] = visited[v];
                any_flag = any_flag || visited[v];
                visited[v] = 1;
            }
        }
    }
}

// XTT::reconstruct's subset lists (reconstruct.cu:843-870): a default-seeded
// engine created every epoch shuffles the order left by the previous epoch;
// step pos holds, condition after condition, the views of subset order[pos]
// (codes cond * P + view), laid out [epoch][pos][cond][j].
void problem_codes_xtt(int P, int C, int batch, int maxb, int epochs, std::vector<int>& order,
                       std::vector<int32_t>& codes, std::vector<uint8_t>& flags, bool& any_flag) {
    const int ss = (P + batch - 1) / batch;
    codes.assign((size_t)epochs * maxb * C * ss, 0);
    flags.assign(codes.size(), 0);
    any_flag = false;
    for (int ep = 0; ep < epochs; ep++) {
        std::mt19937_64 rng;
        std::shuffle(order.begin(), order.end(), rng);
        std::vector<char> visited((size_t)C * P, 0);
        for (int pos = 0; pos < batch; pos++)
            for (int c = 0; c < C; c++)
                for (int j = 0; j < ss; j++) {
                    const int code = c * P + (order[pos] + batch * j) % P;
                    const size_t i = (((size_t)ep * maxb + pos) * C + c) * ss + j;
                    codes[i] = code;
                    flags[i] = visited[code];
                    any_flag = any_flag || visited[code];
                    visited[code] = 1;
                }
    }
}

}  // namespace

struct xr_batch {
    xr_context* ctx = nullptr;
    int n = 0, method = 0, flags = 0, maxb = 1;
    float lambda = 0.f;
    hipStream_t stream = nullptr;
    struct Problem {
        xr_plan* plan = nullptr;
        int N = 0, U = 0, V = 0, P = 0, batch = 1, ss = 0;
        int C = 1, K = 0;  // conditions, basis vectors (K > 0: an XTT problem)
        bool xtt() const { return K > 0; }
        int64_t vol_len() const { return (int64_t)N * N * N * (K > 0 ? K : 1); }
        float* vol = nullptr;
        float* sino = nullptr;
        float* proj = nullptr;
        int32_t* codes = nullptr;
        uint8_t* flags = nullptr;
        bool any_flag = false;
        std::mt19937_64 rng;     // subset-order engine (default seed), continued across runs
        std::vector<int> order;  // current subset order
        void restart() {
            rng = std::mt19937_64();
            order.resize(batch);
            std::iota(order.begin(), order.end(), 0);
        }
    };
    std::vector<Problem> pr;
    std::vector<int> order;  // problems by descending work
    int epoch_cap = 0;       // epochs the code buffers hold
    // per subset position: device item tables
    struct Step {
        int n_fwd = 0, n_bwd = 0, n_active = 0;
        int bwd_first[3] = {0, 0, 0}, bwd_count[3] = {0, 0, 0};  // backprojection items per zpl class
        FwdArgs* fwd = nullptr;
        int* fwd_end = nullptr;
        BwdArgs* bwd = nullptr;
        int* bwd_end = nullptr;
        // XTT problems: their own forward and backprojection launches
        int n_fwd_x = 0, n_bwd_x = 0;
        FwdArgs* fwd_x = nullptr;
        int* fwd_x_end = nullptr;
        BwdArgs* bwd_x = nullptr;
        int* bwd_x_end = nullptr;
        const float** segs = nullptr;  // volumes of the active problems (d_vmax order)
        int64_t* seg_len = nullptr;
        int absmax_blocks = 1;  // blocks per segment of the max |voxel| launch
    };
    std::vector<Step> steps;
    bool tables_ok = false;
    float* d_vmax = nullptr;   // [n]
    double* d_loss = nullptr;  // [epoch_cap][n]
    unsigned* d_queue = nullptr;  // [4]
    int fwd_grid = 0, bwd_grid[3] = {0, 0, 0}, ticket_slot = 0;
    int fwd_grid_x = 0, bwd_grid_x = 0, kmax = 0;  // XTT launches; largest basis count
    bool fwd_axis_all = false;  // every problem untilted: the untilted-only forward kernel
    size_t fwd_lds = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_epochs = -1;
    double last_ms = 0.0;

    ~xr_batch() { release(); }
    void free_tables() {
        for (auto& s : steps) {
            if (s.fwd) (void)hipFree(s.fwd);
            if (s.fwd_end) (void)hipFree(s.fwd_end);
            if (s.bwd) (void)hipFree(s.bwd);
            if (s.bwd_end) (void)hipFree(s.bwd_end);
            if (s.segs) (void)hipFree(s.segs);
            if (s.seg_len) (void)hipFree(s.seg_len);
            if (s.fwd_x) (void)hipFree(s.fwd_x);
            if (s.fwd_x_end) (void)hipFree(s.fwd_x_end);
            if (s.bwd_x) (void)hipFree(s.bwd_x);
            if (s.bwd_x_end) (void)hipFree(s.bwd_x_end);
        }
        steps.clear();
        tables_ok = false;
        drop_graph();
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_epochs = -1;
    }
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        free_tables();
        for (auto& p : pr) {
            if (p.vol) (void)hipFree(p.vol);
            if (p.sino) (void)hipFree(p.sino);
            if (p.proj) (void)hipFree(p.proj);
            if (p.codes) (void)hipFree(p.codes);
            if (p.flags) (void)hipFree(p.flags);
            if (p.plan) xr_plan_destroy(p.plan);
        }
        pr.clear();
        if (d_vmax) (void)hipFree(d_vmax);
        if (d_loss) (void)hipFree(d_loss);
        if (d_queue) (void)hipFree(d_queue);
        d_vmax = nullptr;
        d_loss = nullptr;
        d_queue = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

// Grow the subset-list buffers to `epochs` epochs (invalidates the tables).
int ensure_epochs(xr_batch* b, int epochs) {
    xr_context* ctx = b->ctx;
    if (epochs <= b->epoch_cap) return XR_OK;
    b->free_tables();
    b->epoch_cap = 0;  // stays empty if an allocation below fails
    for (auto& p : b->pr) {
        if (p.codes) (void)hipFree(p.codes);
        if (p.flags) (void)hipFree(p.flags);
        p.codes = nullptr;
        p.flags = nullptr;
        const size_t cnt = (size_t)epochs * b->maxb * p.C * p.ss;
        XR_HIP(ctx, device_malloc((void**)&p.codes, cnt * sizeof(int32_t)));
        XR_HIP(ctx, device_malloc((void**)&p.flags, cnt));
    }
    if (b->d_loss) (void)hipFree(b->d_loss);
    b->d_loss = nullptr;
    XR_HIP(ctx, device_malloc((void**)&b->d_loss, sizeof(double) * (size_t)epochs * b->n));
    b->epoch_cap = epochs;
    return XR_OK;
}

// Device item tables of every subset position.  Needs the code buffers.
int build_tables(xr_batch* b) {
    xr_context* ctx = b->ctx;
    if (b->tables_ok) return XR_OK;
    b->free_tables();
    const bool mlem = b->method == XR_METHOD_MLEM;
    const int out_mode = mlem ? OUT_RATIO : OUT_RESIDUAL;
    // Persistent grids: every block resident at once.  The forward block's LDS
    // holds the tallest row chunk of any problem.
    int max_vb = 1;
    bool any_plain = false, any_xtt = false;
    b->kmax = 0;
    for (const auto& p : b->pr) {
        max_vb = std::max(max_vb, std::min(forward_rows(p.plan, p.xtt() ? FWD_XTT : FWD_PLAIN), p.V));
        any_plain = any_plain || !p.xtt();
        any_xtt = any_xtt || p.xtt();
        b->kmax = std::max(b->kmax, p.K);
    }
    hipDeviceProp_t prop;
    XR_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    b->fwd_axis_all = true;
    for (const auto& p : b->pr)
        if (!p.xtt()) b->fwd_axis_all = b->fwd_axis_all && p.plan->axis_cond[0];
    // past the 32-bit band and the reduction slots (two per wave of any kernel)
    b->ticket_slot = fwd_red_slot(max_vb + 2, 4) + 2 * std::max(kFwdWaves, kFwdBatchWaves);
    b->fwd_lds = sizeof(unsigned long long) * ((size_t)b->ticket_slot + 1);
    b->fwd_grid = b->fwd_grid_x = b->bwd_grid_x = 0;
    if (any_plain) {
        const int fpc = batch_blocks_per_cu_forward(FWD_PLAIN, b->fwd_lds, b->fwd_axis_all);
        if (fpc <= 0) return fail(ctx, XR_ERR_HIP, "occupancy query failed for the batched forward kernel");
        b->fwd_grid = fpc * prop.multiProcessorCount;
        for (int c = 0; c < 3; c++) {
            const int bpc = batch_blocks_per_cu_backward(mlem ? BWD_MLEM : BWD_ART, c == 2 ? 4 : c + 1);
            if (bpc <= 0) return fail(ctx, XR_ERR_HIP, "occupancy query failed for the batched backprojection");
            b->bwd_grid[c] = bpc * prop.multiProcessorCount;
        }
    }
    if (any_xtt) {
        const int fpc = batch_blocks_per_cu_forward(FWD_XTT, b->fwd_lds, false, b->kmax);
        const int bpc = batch_blocks_per_cu_backward(mlem ? BWD_XTT_MLEM : BWD_XTT_ART, 1, b->kmax);
        if (fpc <= 0 || bpc <= 0) return fail(ctx, XR_ERR_HIP, "occupancy query failed for the batched XTT kernels");
        b->fwd_grid_x = fpc * prop.multiProcessorCount;
        b->bwd_grid_x = bpc * prop.multiProcessorCount;
    }
    b->steps.resize(b->maxb);
    for (int pos = 0; pos < b->maxb; pos++) {
        std::vector<FwdArgs> fi, fx;
        std::vector<int> fsplit, fxsplit;  // each item's own plan setting of the band split (0: automatic)
        std::vector<int> fe, fxe;
        std::vector<BwdArgs> bz[3], bx;  // plain items by z depth class (zpl 1, 2, 4); XTT items
        std::vector<const float*> segs;
        std::vector<int64_t> lens;
        int fsum = 0, fxsum = 0;
        for (int oi = 0; oi < (int)b->order.size(); oi++) {
            const int k = b->order[oi];
            xr_batch::Problem& p = b->pr[k];
            if (pos >= p.batch) continue;
            xr_plan* pl = p.plan;
            const bool xtt = p.xtt();
            const int64_t N3 = (int64_t)p.N * p.N * p.N;
            FwdArgs a{};
            a.d = pl->dims;
            a.b = pl->bas;
            a.vg = pl->d_vg;
            a.cg = pl->d_cg;
            a.views = p.codes;
            a.accum = p.flags;  // all zero unless a subset wraps past P
            a.vol = p.vol;
            a.vol_cond_stride = 0;
            a.chan_stride = N3;
            a.proj = p.proj;
            a.sino = p.sino;
            a.norm = pl->d_norm;
            a.bands = pl->tiles.fwd_scan ? nullptr : pl->d_bands;
            a.g2 = pl->d_g2;
            a.xw = pl->d_xw;
            a.loss = b->d_loss + k;
            // max |voxel| of problem oi: measured by k_absmax_seg at the first
            // subset step (every problem is active there, slot == oi), then by
            // the previous step's backprojection (which updates every voxel)
            a.vmax = b->d_vmax + oi;
            // as forward_launch
            a.wmul = xtt ? pl->dims.vs * pl->dims.ratio * pl->xtt_wsum * 1.0001f : pl->dims.ratio * 1.0001f;
            a.fsum = pl->max_factor * 1.01f + 1.0f;
            a.code_base = 0;
            a.out_mode = out_mode;
            a.n_bands = (p.U + kBandU - 1) / kBandU;
            a.n_slots = p.C * p.ss;
            a.view_group = kFwdViewGroup;
            a.axis = 1;
            for (int c = 0; c < p.C; c++) a.axis = a.axis && pl->axis_cond[c];
            const int rows = forward_rows(pl, xtt ? FWD_XTT : FWD_PLAIN);
            a.n_chunks = 1;
            for (int v0 = 0; v0 < p.V; v0 += rows) {
                a.v0 = v0;
                a.vb = std::min(rows, p.V - v0);
                if (xtt) {
                    fx.push_back(a);
                    fxsplit.push_back(pl->tiles.fwd_split);
                    fxsum += a.n_bands * a.n_slots;
                } else {
                    fi.push_back(a);
                    fsplit.push_back(pl->tiles.fwd_split);
                    fsum += a.n_bands * a.n_slots;
                }
            }
            BwdArgs w{};
            w.d = pl->dims;
            w.b = pl->bas;
            w.vg = pl->d_vg;
            w.views = p.codes;
            w.n_views = p.C * p.ss;
            w.proj = p.proj;
            w.code_base = 0;
            w.vol = p.vol;
            w.chan_stride = N3;
            w.g2 = pl->d_g2;
            w.xw = pl->d_xw;
            w.alpha = mlem ? 0.f : b->lambda / (float)p.ss;  // reconstruct.cu:124, 911
            w.axis = a.axis;
            w.vmax_out = (pos + 1 < p.batch) ? reinterpret_cast<unsigned*>(b->d_vmax + oi) : nullptr;
            if (!backward_shape(w, xtt ? 0 : pl->tiles.bwd_zpl, pl->tiles.bwd_gx, xtt) || (xtt && w.zpl != 1))
                return fail(ctx, XR_ERR_INVALID, "bad tile shape for problem %d", k);
            if (xtt)
                bx.push_back(w);
            else
                bz[w.zpl == 4 ? 2 : w.zpl - 1].push_back(w);
            segs.push_back(p.vol);
            lens.push_back(p.vol_len());
        }
        // Few items in this step (small subsets): split the bands so that the
        // persistent grid is filled.  A problem whose plan sets the split
        // (xr_plan_set_tiles, autotune) keeps its own.
        const int split = forward_split(fsum, std::max(b->fwd_grid, 1), 0);
        fsum = 0;
        for (size_t i = 0; i < fi.size(); i++) {
            FwdArgs& a = fi[i];
            a.n_sub = fsplit[i] > 0 ? fsplit[i] : split;
            fsum += a.n_bands * a.n_sub * a.n_slots;
            fe.push_back(fsum);
        }
        const int xsplit = forward_split(fxsum, std::max(b->fwd_grid_x, 1), 0);
        fxsum = 0;
        for (size_t i = 0; i < fx.size(); i++) {
            FwdArgs& a = fx[i];
            a.n_sub = fxsplit[i] > 0 ? fxsplit[i] : xsplit;
            fxsum += a.n_bands * a.n_sub * a.n_slots;
            fxe.push_back(fxsum);
        }
        xr_batch::Step& s = b->steps[pos];
        s.n_fwd = (int)fi.size();
        s.n_fwd_x = (int)fx.size();
        s.n_active = (int)segs.size();
        if (s.n_active == 0) continue;
        s.absmax_blocks = absmax_blocks(*std::max_element(lens.begin(), lens.end()), s.n_active);
        // backprojection items by z depth class (zpl 4, 2, 1): one launch and
        // one ticket range per class, so prefix sums restart in each class
        std::vector<BwdArgs> bi;
        std::vector<int> be, bxe;
        for (int c = 2; c >= 0; c--) {
            s.bwd_first[c] = (int)bi.size();
            s.bwd_count[c] = (int)bz[c].size();
            int run = 0;
            for (const BwdArgs& w : bz[c]) {
                bi.push_back(w);
                run += w.n_tiles;
                be.push_back(run);
            }
        }
        s.n_bwd = (int)bi.size();
        s.n_bwd_x = (int)bx.size();
        {
            int run = 0;
            for (const BwdArgs& w : bx) {
                run += w.n_tiles;
                bxe.push_back(run);
            }
        }
        auto upload = [&](auto*& dst, const auto& src) -> int {
            using T = typename std::remove_reference<decltype(src)>::type::value_type;
            if (src.empty()) return XR_OK;
            XR_HIP(ctx, device_malloc((void**)&dst, sizeof(T) * src.size()));
            XR_HIP(ctx, hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
            return XR_OK;
        };
        XR_TRY(upload(s.fwd, fi));
        XR_TRY(upload(s.fwd_end, fe));
        XR_TRY(upload(s.bwd, bi));
        XR_TRY(upload(s.bwd_end, be));
        XR_TRY(upload(s.fwd_x, fx));
        XR_TRY(upload(s.fwd_x_end, fxe));
        XR_TRY(upload(s.bwd_x, bx));
        XR_TRY(upload(s.bwd_x_end, bxe));
        XR_TRY(upload(s.segs, segs));
        XR_TRY(upload(s.seg_len, lens));
    }
    b->tables_ok = true;
    return XR_OK;
}

// Enqueue the epoch loop on the batch stream (eagerly or under capture).
int enqueue_epochs(xr_batch* b, int epochs) {
    xr_context* ctx = b->ctx;
    hipStream_t st = b->stream;
    const bool mlem = b->method == XR_METHOD_MLEM;
    hipError_t e = launch_fill(reinterpret_cast<float*>(b->d_loss), (int64_t)2 * epochs * b->n, 0.0f, st);
    if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "fill failed: %s", hipGetErrorString(e));
    for (int ep = 0; ep < epochs; ep++) {
        for (int pos = 0; pos < b->maxb; pos++) {
            const xr_batch::Step& s = b->steps[pos];
            if (s.n_active == 0) continue;
            const int step = ep * b->maxb + pos;
            // max |voxel| of each problem: read from the volumes at the first
            // step (the caller may have replaced them between runs), then left
            // by each step's backprojection for the next (cleared once the
            // forward has read it)
            if (pos == 0) {
                e = launch_fill(b->d_vmax, b->n, 0.0f, st);
                if (e == hipSuccess)
                    e = launch_absmax_seg(s.segs, s.seg_len, b->d_vmax, s.n_active, s.absmax_blocks, st);
                if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "absmax failed: %s", hipGetErrorString(e));
            }
            e = launch_forward_batch(s.fwd, s.fwd_end, s.n_fwd, b->d_queue, b->ticket_slot, b->fwd_lds, b->fwd_grid,
                                     step, ep * b->n, FWD_PLAIN, b->fwd_axis_all, st);
            if (e == hipSuccess)
                e = launch_forward_batch(s.fwd_x, s.fwd_x_end, s.n_fwd_x, b->d_queue, b->ticket_slot, b->fwd_lds,
                                         b->fwd_grid_x, step, ep * b->n, FWD_XTT, false, st, b->kmax);
            if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "batched forward failed: %s", hipGetErrorString(e));
            if (pos + 1 < b->maxb) {
                e = launch_fill(b->d_vmax, b->n, 0.0f, st);
                if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "fill failed: %s", hipGetErrorString(e));
            }
            // one persistent launch per z depth class present (in-order on the stream)
            for (int c = 2; c >= 0; c--) {
                if (s.bwd_count[c] == 0) continue;
                e = launch_backward_batch(s.bwd + s.bwd_first[c], s.bwd_end + s.bwd_first[c], s.bwd_count[c],
                                          b->d_queue + 2, b->bwd_grid[c], step, mlem ? BWD_MLEM : BWD_ART,
                                          c == 2 ? 4 : c + 1, st);
                if (e != hipSuccess)
                    return fail(ctx, XR_ERR_HIP, "batched backprojection failed: %s", hipGetErrorString(e));
            }
            e = launch_backward_batch(s.bwd_x, s.bwd_x_end, s.n_bwd_x, b->d_queue + 2, b->bwd_grid_x, step,
                                      mlem ? BWD_XTT_MLEM : BWD_XTT_ART, 1, st, b->kmax);
            if (e != hipSuccess)
                return fail(ctx, XR_ERR_HIP, "batched XTT backprojection failed: %s", hipGetErrorString(e));
        }
    }
    return XR_OK;
}

}  // namespace


extern "C" {

// ---------------------------------------------------------------- tiles ----

int xr_plan_set_tiles(xr_plan* plan, const xr_tiles* tiles) {
    if (!plan || !tiles) return XR_ERR_INVALID;
    if (!valid_tiles(*tiles)) return fail(plan->ctx, XR_ERR_INVALID, "invalid tile shape (zpl %d, gx %d, rows %d, scan %d, split %d)",
                                          tiles->bwd_zpl, tiles->bwd_gx, tiles->fwd_rows, tiles->fwd_scan,
                                          tiles->fwd_split);
    plan->tiles = *tiles;
    return XR_OK;
}

int xr_plan_get_tiles(const xr_plan* plan, xr_tiles* tiles) {
    if (!plan || !tiles) return XR_ERR_INVALID;
    BwdArgs a{};
    a.d = plan->dims;
    if (!backward_shape(a, plan->tiles.bwd_zpl, plan->tiles.bwd_gx, false)) return XR_ERR_INVALID;
    tiles->bwd_zpl = a.zpl;
    tiles->bwd_gx = a.gx;
    tiles->fwd_rows = forward_rows(plan, FWD_PLAIN);
    tiles->fwd_scan = plan->tiles.fwd_scan;
    tiles->fwd_split = plan->tiles.fwd_split;
    return XR_OK;
}

// ------------------------------------------------------------ autotuner ----

int xr_autotune(xr_context* ctx, xr_plan* plan, const float* vol_cols, const float* sino, xr_tiles* best,
                double* best_ms) {
    if (!ctx || !plan || !vol_cols || !sino) return fail(ctx, XR_ERR_INVALID, "bad args");
    const auto key = tune_key(plan);
    for (auto& kv : ctx->tune_cache) {
        if (kv.first == key) {
            plan->tiles = kv.second;
            if (best) *best = kv.second;
            if (best_ms) *best_ms = -1.0;
            return XR_OK;
        }
    }
    XR_TRY(ensure_norm(plan));
    const int N = plan->g.n_vox, P = plan->g.n_proj;
    const int64_t N3 = (int64_t)N * N * N, lenD = (int64_t)plan->g.n_det_u * plan->g.n_det_v * P;
    DevBuf<float> proj, vol2;
    DevBuf<int32_t> codes;
    XR_HIP(ctx, proj.alloc(lenD));
    XR_HIP(ctx, vol2.alloc(N3));
    XR_HIP(ctx, codes.alloc(P));
    std::vector<int32_t> hc(P);
    std::iota(hc.begin(), hc.end(), 0);
    XR_HIP(ctx, hipMemcpyAsync(codes.p, hc.data(), sizeof(int32_t) * P, hipMemcpyHostToDevice, ctx->stream));
    XR_HIP(ctx, hipMemcpyAsync(vol2.p, vol_cols, sizeof(float) * N3, hipMemcpyDeviceToDevice, ctx->stream));
    EventPair ev;
    XR_HIP(ctx, ev.init());
    const bool timing = ctx->timing;
    ctx->timing = false;
    const xr_tiles saved = plan->tiles;
    auto time_it = [&](auto&& fn, double& ms) -> int {
        XR_TRY(fn());  // warm-up
        XR_HIP(ctx, hipEventRecord(ev.a, ctx->stream));
        for (int r = 0; r < 2; r++) XR_TRY(fn());
        XR_HIP(ctx, hipEventRecord(ev.b, ctx->stream));
        XR_HIP(ctx, hipEventSynchronize(ev.b));
        float t = 0.f;
        XR_HIP(ctx, hipEventElapsedTime(&t, ev.a, ev.b));
        ms = 0.5 * t;
        return XR_OK;
    };
    int rc = XR_OK;
    xr_tiles win = saved;
    double fbest = 1e300, bbest = 1e300;
    // forward: detector-row chunk heights (LDS footprint vs. column scans)
    {
        std::vector<int> seen;
        for (int rows : {0, 96, 64, 32}) {
            plan->tiles.fwd_rows = rows;
            const int eff = forward_rows(plan, FWD_PLAIN);
            if (std::find(seen.begin(), seen.end(), eff) != seen.end()) continue;
            seen.push_back(eff);
            double ms = 0;
            rc = time_it([&] { return forward_launch(plan, FWD_PLAIN, vol_cols, 0, proj.p, sino, 0, codes.p, nullptr,
                                                     P, OUT_RATIO, ctx->d_loss); },
                         ms);
            if (rc != XR_OK) break;
            if (ms < fbest) {
                fbest = ms;
                win.fwd_rows = rows;
            }
        }
    }
    // backprojection: z depth per lane x column-group layout, MLEM on a copy
    // of the volume with a ratio image of ones (the volume is unchanged).
    if (rc == XR_OK) {
        const hipError_t e = launch_fill(proj.p, lenD, 1.0f, ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, XR_ERR_HIP, "fill failed: %s", hipGetErrorString(e));
    }
    plan->tiles.fwd_rows = win.fwd_rows;
    for (int zpl : {1, 2, 4}) {
        if (rc != XR_OK) break;
        if (zpl > 1 && 64 * zpl > 2 * std::max(N, 64)) continue;
        for (int gx : {2, 4, 8}) {
            plan->tiles.bwd_zpl = zpl;
            plan->tiles.bwd_gx = gx;
            double ms = 0;
            rc = time_it([&] { return backward_launch(plan, BWD_MLEM, proj.p, 0, codes.p, P, vol2.p, nullptr,
                                                      nullptr, 0.f, 0); },
                         ms);
            if (rc != XR_OK) break;
            if (ms < bbest) {
                bbest = ms;
                win.bwd_zpl = zpl;
                win.bwd_gx = gx;
            }
        }
    }
    ctx->timing = timing;
    if (rc != XR_OK) {
        plan->tiles = saved;
        return rc;
    }
    plan->tiles = win;
    ctx->tune_cache.push_back({key, win});
    if (best) *best = win;
    if (best_ms) *best_ms = fbest + bbest;
    return XR_OK;
}

// -------------------------------------------------------------- batches ----

int xr_batch_create(xr_context* ctx, int n, const xr_geometry* geoms, const xr_scan* scans, const int32_t* batch,
                    int direction, int method, float lambda, int flags, xr_batch** out) {
    if (!ctx || !out || !geoms || n <= 0) return fail(ctx, XR_ERR_INVALID, "bad args");
    *out = nullptr;
    if (method != XR_METHOD_MLEM && method != XR_METHOD_ART) return fail(ctx, XR_ERR_INVALID, "method must be MLEM or ART");
    std::unique_ptr<xr_batch> b(new (std::nothrow) xr_batch());
    if (!b) return fail(ctx, XR_ERR_ALLOC, "batch allocation failed");
    b->ctx = ctx;
    b->n = n;
    b->method = method;
    b->lambda = lambda;
    b->flags = flags;
    XR_HIP(ctx, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, Z[3] = {0, 0, 0};
    xr_scan plain{};
    plain.n_cond = 1;
    plain.n_basis = 0;
    plain.cond_R = I;
    plain.cond_t = Z;
    plain.cond_offset = Z;
    b->pr.resize(n);
    for (int i = 0; i < n; i++) {
        xr_batch::Problem& p = b->pr[i];
        const xr_scan* s = scans ? &scans[i] : &plain;
        XR_TRY(check_scan(ctx, s));
        if (s->n_basis == 0 && s->n_cond != 1)
            return fail(ctx, XR_ERR_INVALID, "problem %d: an IR problem of a batch has one condition", i);
        XR_TRY(check_geometry(ctx, &geoms[i]));
        p.N = geoms[i].n_vox;
        p.U = geoms[i].n_det_u;
        p.V = geoms[i].n_det_v;
        p.P = geoms[i].n_proj;
        p.C = s->n_cond;
        p.K = s->n_basis;
        p.batch = batch ? batch[i] : 1;
        if (p.batch <= 0 || p.batch > p.P) return fail(ctx, XR_ERR_INVALID, "problem %d: bad batch %d", i, p.batch);
        p.ss = (p.P + p.batch - 1) / p.batch;
        p.restart();
        b->maxb = std::max(b->maxb, p.batch);
        XR_TRY(xr_plan_create(ctx, &geoms[i], s, XR_CONV_IR, direction, &p.plan));
        XR_TRY(ensure_norm(p.plan));
        const int64_t nv = p.vol_len(), lenD = (int64_t)p.U * p.V * p.P * p.C;
        XR_HIP(ctx, device_malloc((void**)&p.vol, sizeof(float) * nv));
        XR_HIP(ctx, device_malloc((void**)&p.sino, sizeof(float) * lenD));
        XR_HIP(ctx, device_malloc((void**)&p.proj, sizeof(float) * lenD));
        XR_HIP(ctx, hipMemsetAsync(p.vol, 0, sizeof(float) * nv, ctx->stream));
        XR_HIP(ctx, hipMemsetAsync(p.sino, 0, sizeof(float) * lenD, ctx->stream));
    }
    // Longest work first: backprojection cost ~ N^3 P, the dominant term.
    b->order.resize(n);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        const auto& a = b->pr[x];
        const auto& c = b->pr[y];
        return (double)a.vol_len() * a.C * a.ss > (double)c.vol_len() * c.C * c.ss;
    });
    XR_HIP(ctx, device_malloc((void**)&b->d_vmax, sizeof(float) * n));
    XR_HIP(ctx, device_malloc((void**)&b->d_queue, sizeof(unsigned) * 4));
    XR_HIP(ctx, hipMemsetAsync(b->d_queue, 0, sizeof(unsigned) * 4, ctx->stream));
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = b.release();
    return XR_OK;
}

void xr_batch_destroy(xr_batch* b) { delete b; }

int xr_batch_buffers(xr_batch* b, int i, float** vol_cols, float** sino) {
    if (!b || i < 0 || i >= b->n) return XR_ERR_INVALID;
    if (vol_cols) *vol_cols = b->pr[i].vol;
    if (sino) *sino = b->pr[i].sino;
    return XR_OK;
}

xr_plan* xr_batch_plan(xr_batch* b, int i) {
    if (!b || i < 0 || i >= b->n) return nullptr;
    return b->pr[i].plan;
}

int xr_batch_autotune(xr_batch* b) {
    if (!b) return XR_ERR_INVALID;
    xr_context* ctx = b->ctx;
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto& p : b->pr)
        if (!p.xtt()) XR_TRY(xr_autotune(ctx, p.plan, p.vol, p.sino, nullptr, nullptr));
    b->free_tables();
    return XR_OK;
}

int xr_batch_run(xr_batch* b, int epochs, double* losses) {
    if (!b) return XR_ERR_INVALID;
    xr_context* ctx = b->ctx;
    if (epochs < 0) return fail(ctx, XR_ERR_INVALID, "epochs must be >= 0");
    b->last_ms = 0.0;
    if (epochs == 0) return XR_OK;
    if ((b->flags & XR_BATCH_AUTOTUNE) && !b->tables_ok) XR_TRY(xr_batch_autotune(b));
    XR_TRY(ensure_epochs(b, epochs));
    XR_TRY(build_tables(b));
    // The caller's data (made on the context stream) must be complete.
    XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> codes;
    std::vector<uint8_t> fl;
    for (auto& p : b->pr) {
        if (p.xtt())
            problem_codes_xtt(p.P, p.C, p.batch, b->maxb, epochs, p.order, codes, fl, p.any_flag);
        else
            problem_codes(p.P, p.batch, b->maxb, epochs, p.rng, p.order, codes, fl, p.any_flag);
        XR_HIP(ctx, hipMemcpyAsync(p.codes, codes.data(), sizeof(int32_t) * codes.size(), hipMemcpyHostToDevice,
                                   b->stream));
        XR_HIP(ctx, hipMemcpyAsync(p.flags, fl.data(), fl.size(), hipMemcpyHostToDevice, b->stream));
        // host vectors are reused: wait before overwriting them
        XR_HIP(ctx, hipStreamSynchronize(b->stream));
    }
    EventPair ev;
    XR_HIP(ctx, ev.init());
    const bool use_graph = (b->flags & XR_BATCH_GRAPH) != 0;
    if (use_graph && b->graph_epochs != epochs) {
        b->drop_graph();
        XR_HIP(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_epochs(b, epochs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != XR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(ctx, XR_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        XR_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_epochs = epochs;
    }
    XR_HIP(ctx, hipEventRecord(ev.a, b->stream));
    if (use_graph) {
        XR_HIP(ctx, hipGraphLaunch(b->exec, b->stream));
    } else {
        XR_TRY(enqueue_epochs(b, epochs));
    }
    XR_HIP(ctx, hipEventRecord(ev.b, b->stream));
    XR_HIP(ctx, hipEventSynchronize(ev.b));
    float ms = 0.f;
    XR_HIP(ctx, hipEventElapsedTime(&ms, ev.a, ev.b));
    b->last_ms = ms;
    if (losses)
        XR_HIP(ctx, hipMemcpy(losses, b->d_loss, sizeof(double) * (size_t)epochs * b->n, hipMemcpyDeviceToHost));
    return XR_OK;
}

int xr_batch_restart(xr_batch* b) {
    if (!b) return XR_ERR_INVALID;
    for (auto& p : b->pr) p.restart();
    return XR_OK;
}

int xr_batch_elapsed_ms(const xr_batch* b, double* ms) {
    if (!b || !ms) return XR_ERR_INVALID;
    *ms = b->last_ms;
    return XR_OK;
}

int xr_batch_reconstruct(xr_context* ctx, int n, const xr_geometry* geoms, const xr_scan* scans,
                         const int32_t* batch, float* const* sino, float* const* vol, int epochs, int direction,
                         int method, float lambda, int flags, double* losses) {
    if (!ctx || !sino || !vol) return fail(ctx, XR_ERR_INVALID, "bad args");
    xr_batch* b = nullptr;
    XR_TRY(xr_batch_create(ctx, n, geoms, scans, batch, direction, method, lambda, flags, &b));
    std::unique_ptr<xr_batch, void (*)(xr_batch*)> guard(b, xr_batch_destroy);
    int64_t maxlen = 0;
    for (auto& p : b->pr) maxlen = std::max(maxlen, p.vol_len());
    DevBuf<float> ref;
    XR_HIP(ctx, ref.alloc(maxlen));
    for (int i = 0; i < n; i++) {
        auto& p = b->pr[i];
        const int64_t nv = p.vol_len(), lenD = (int64_t)p.U * p.V * p.P * p.C;
        XR_HIP(ctx, hipMemcpyAsync(p.sino, sino[i], sizeof(float) * lenD, hipMemcpyHostToDevice, ctx->stream));
        XR_HIP(ctx, hipMemcpyAsync(ref.p, vol[i], sizeof(float) * nv, hipMemcpyHostToDevice, ctx->stream));
        XR_TRY(xr_volume_to_columns(ctx, ref.p, p.vol, p.N, std::max(p.K, 1)));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    XR_TRY(xr_batch_run(b, epochs, losses));
    for (int i = 0; i < n; i++) {
        auto& p = b->pr[i];
        const int64_t nv = p.vol_len();
        if (p.xtt()) {
            // XTT::reconstruct's ending (reconstruct.cu:919-927): voxelSqrt and the
            // sinogram overwritten with the final ratio / residual images
            XR_TRY(xr_volume_sqrt(ctx, p.vol, nv));
            const int64_t lenD = (int64_t)p.U * p.V * p.P * p.C;
            XR_HIP(ctx, hipMemcpyAsync(sino[i], p.proj, sizeof(float) * lenD, hipMemcpyDeviceToHost, ctx->stream));
        }
        XR_TRY(xr_volume_from_columns(ctx, p.vol, ref.p, p.N, std::max(p.K, 1)));
        XR_HIP(ctx, hipMemcpyAsync(vol[i], ref.p, sizeof(float) * nv, hipMemcpyDeviceToHost, ctx->stream));
        XR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return XR_OK;
}

}  // extern "C"
