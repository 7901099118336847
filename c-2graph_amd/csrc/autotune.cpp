// Autotuner: picks launch shape, work-item chunk size and rounds per hipGraph
// replay for the SSSP kernels, and the push/pull switch point of single-source
// PPR, by timing them on sample sources.
//
// The reference compiles in one shape (20 workgroups x 512 threads,
// core/Util.cuh:31-35).  Here the candidates are timed on the caller's sample
// and the fastest is stored in the handle.  Every setting computes the same
// distances (the tunables change scheduling only), so tuning never changes
// results.  Search: coordinate descent over grid_per_cu_single and chunk
// (single-source) or grid_per_cu and batch_chunk (batched), then
// iters_per_graph, then the relax kernels' workgroup size (block), the
// light-edge bound (off, or 3/4, 1, 3/2, 2 times the default) and, for
// single-source, the
// XCD split; two passes, best of two runs per candidate.
#include "graph.h"
#include <cstring>

using namespace c2g;

namespace {

// Captured launch sequences hold tunables as kernel arguments; a new light
// bound rebuilds the light rows.
int set_tune_locked(Graph* g, const c2g_tune& t) {
    if (memcmp(&g->tune, &t, sizeof t) != 0) {
        if (g->sssp.exec) { (void)hipGraphExecDestroy(g->sssp.exec); g->sssp.exec = nullptr; }
        if (g->batch.exec) { (void)hipGraphExecDestroy(g->batch.exec); g->batch.exec = nullptr; }
    }
    g->tune = t;
    return light_build(g);
}

double time_once(Graph* g, const int32_t* srcs, int32_t n, bool batched, int* rc) {
    const double t0 = now_ms();
    if (batched) {
        *rc = sssp_batch_group(g, srcs, n, nullptr, nullptr, nullptr, nullptr);
    } else {
        for (int32_t i = 0; i < n && *rc == C2G_OK; ++i)
            *rc = sssp_run(g, srcs[i], 0, nullptr, nullptr, nullptr, nullptr, nullptr, false);
    }
    return now_ms() - t0;
}

double measure(Graph* g, const int32_t* srcs, int32_t n, bool batched, int* rc) {
    double best = 1e30;
    for (int r = 0; r < 2 && *rc == C2G_OK; ++r) best = std::min(best, time_once(g, srcs, n, batched, rc));
    return best;
}

double time_ppr(Graph* g, const int32_t* srcs, int32_t n, float alpha, float eps, int* rc) {
    double best = 1e30;
    for (int r = 0; r < 2 && *rc == C2G_OK; ++r) {
        const double t0 = now_ms();
        for (int32_t i = 0; i < n && *rc == C2G_OK; ++i)
            *rc = ppr_single(g, g->stream, srcs[i], alpha, eps, nullptr, nullptr, nullptr);
        best = std::min(best, now_ms() - t0);
    }
    return best;
}

}  // namespace

// Single-source PPR: workgroups per CU (1, 2, 4, 8), then the push/pull
// switch point: the share of all out-edges (in thousandths) above which a
// round's frontier is propagated by a pull sweep over in-edges instead of
// scattered pushes, timed on the caller's sources, alpha and eps.  Both forms
// compute the same round, so the choices change FP32 summation order only.
extern "C" int c2g_autotune_ppr(c2g_graph* h, const int32_t* srcs, int32_t n, float alpha, float eps,
                                c2g_tune* best_out) {
    if (!h || !srcs || n < 1) return fail(C2G_EINVAL, "c2g_autotune_ppr: need at least one sample source");
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    for (int32_t i = 0; i < n; ++i)
        if (srcs[i] < 0 || srcs[i] >= g->V) return fail(C2G_EINVAL, "source %d out of range", srcs[i]);
    const int32_t m = std::min<int32_t>(n, 16);
    c2g_tune best = g->tune;
    best.profile = 0;
    int rc = set_tune_locked(g, best);
    if (rc) return rc;
    (void)time_ppr(g, srcs, 1, alpha, eps, &rc);  // warm: allocations, in-edge rows
    double tbest = time_ppr(g, srcs, m, alpha, eps, &rc);
    for (int gpc : {1, 2, 4, 8}) {
        if (rc) break;
        c2g_tune t = best;
        t.ppr_grid_per_cu = gpc;
        if (!memcmp(&t, &best, sizeof t)) continue;
        if ((rc = set_tune_locked(g, t))) break;
        const double tt = time_ppr(g, srcs, m, alpha, eps, &rc);
        if (rc) break;
        if (tt < tbest * 0.98) { tbest = tt; best = t; }
    }
    if (!rc) rc = set_tune_locked(g, best);
    for (int v : {0, 25, 50, 100, 200, 400, 800}) {
        if (rc) break;
        c2g_tune t = best;
        t.ppr_pull = v;
        if (!memcmp(&t, &best, sizeof t)) continue;
        if ((rc = set_tune_locked(g, t))) break;
        const double tt = time_ppr(g, srcs, m, alpha, eps, &rc);
        if (rc) break;
        if (tt < tbest * 0.98) { tbest = tt; best = t; }
    }
    const int rs = set_tune_locked(g, best);
    if (rc) return rc;
    if (rs) return rs;
    if (best_out) *best_out = best;
    return C2G_OK;
}

extern "C" int c2g_autotune(c2g_graph* h, const int32_t* srcs, int32_t n, int32_t batched, c2g_tune* best_out) {
    if (!h || !srcs || n < 1) return fail(C2G_EINVAL, "c2g_autotune: need at least one sample source");
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    for (int32_t i = 0; i < n; ++i)
        if (srcs[i] < 0 || srcs[i] >= g->V) return fail(C2G_EINVAL, "source %d out of range", srcs[i]);
    const bool b = batched != 0;
    const int32_t m = b ? std::min<int32_t>(n, 64 * 4) : std::min<int32_t>(n, 16);
    c2g_tune best = g->tune;
    best.profile = 0;
    int rc = set_tune_locked(g, best);
    if (rc) return rc;
    (void)time_once(g, srcs, m, b, &rc);  // warm allocations and code objects
    double tbest = measure(g, srcs, m, b, &rc);
    const std::vector<int> grids = {2, 4, 8, 16};
    const std::vector<int> chunks = b ? std::vector<int>{16, 32, 64, 128} : std::vector<int>{16, 32, 64, 128, 256};
    const std::vector<int> iters = {4, 8, 16};
    const std::vector<int> blocks = {128, 256, 512};
    // light-edge bounds around the graph's default (not the current value,
    // so a single-source pass and a batched pass search the same range);
    // below 3/4 of the default, rounds and their spread across source sets
    // grow quickly
    const int L0 = default_light(g);
    std::vector<int> lights = {0};
    if (L0 > 0)
        for (int v : {3 * L0 / 4, L0, 3 * L0 / 2, 2 * L0})
            if (v > 0 && std::find(lights.begin(), lights.end(), v) == lights.end()) lights.push_back(v);
    const std::vector<int> splits = {0, 1};
    const int dims = b ? 5 : 6;  // single-source also tries the XCD split
    for (int pass = 0; pass < 2 && rc == C2G_OK; ++pass) {
        for (int which = 0; which < dims && rc == C2G_OK; ++which) {
            const std::vector<int>& vals = which == 0   ? grids
                                           : which == 1 ? chunks
                                           : which == 2 ? iters
                                           : which == 3 ? blocks
                                           : which == 4 ? lights
                                                        : splits;
            for (int v : vals) {
                c2g_tune t = best;
                if (which == 0) (b ? t.grid_per_cu : t.grid_per_cu_single) = v;
                else if (which == 1) (b ? t.batch_chunk : t.chunk) = v;
                else if (which == 2) t.iters_per_graph = v;
                else if (which == 3) (b ? t.block : t.block_single) = v;
                else if (which == 4) t.light = v;
                else t.xcd_split = v;
                if (!memcmp(&t, &best, sizeof t)) continue;
                if ((rc = set_tune_locked(g, t))) break;
                const double tt = measure(g, srcs, m, b, &rc);
                if (rc) break;
                // the light bound changes rounds, whose count varies between
                // source sets: leave the default unless clearly better
                if (tt < tbest * (which == 4 ? 0.95 : 0.98)) { tbest = tt; best = t; }
            }
            const int rs = set_tune_locked(g, best);
            if (!rc) rc = rs;
        }
    }
    const int rs = set_tune_locked(g, best);
    if (rc) return rc;
    if (rs) return rs;
    if (best_out) *best_out = best;
    return C2G_OK;
}
