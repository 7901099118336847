// Host-core lanes: the reference's CPU worker algorithms, used for the
// C^2graph-PM handoff (c2g_sssp_migrate) and as the query server's fallback
// lane.  They run only where a caller asks for the host lane explicitly.
//
//   host_bf            Bellman_ford (src/sssp/base-perk/SSSP.cpp:88-119), the
//                      variant that resets its flag after each round (the
//                      purn-perk copy never resets it, SSSP.cpp:104, and can
//                      miss re-improvements).  Pruned graphs relax a vertex's
//                      hidden list under the same rule as the device kernel.
//   host_push          forward_push (src/ppr/w-base-perk/PPR.cpp:178-236) with
//                      the degree-scaled threshold r[v] > eps*outdeg(v) and a
//                      zero-initialised flag array (the reference leaves it
//                      uninitialised, PPR.cpp:182).
#include "graph.h"
#include <cstring>

namespace c2g {

int host_bf(const Graph* g, int32_t* dist, const int32_t* frontier, int32_t count,
            const uint8_t* via_hidden, c2g_stats* st) {
    const int32_t V = g->V;
    if (count < 0 || count > V) return fail(C2G_EINVAL, "frontier count %d out of range", count);
    if (g->wl.empty() && g->E > 0) return fail(C2G_ENOTSUP, "graph has no int weights");
    const bool pruned = !g->hvl.empty() && !g->hwl.empty();
    std::vector<int32_t> cur(frontier, frontier + count), nxt;
    std::vector<uint8_t> flag((size_t)V, 0);
    std::vector<uint8_t> hid;
    if (pruned) {
        hid.assign((size_t)V, 0);
        if (via_hidden) memcpy(hid.data(), via_hidden, (size_t)V);
    }
    nxt.reserve(cur.size());
    int64_t relax = 0, impr = 0;
    int rounds = 0;
    const int32_t* vl = g->vl.data();
    const int32_t* el = g->el.data();
    const int32_t* wl = g->wl.data();
    auto relax_edge = [&](int64_t du, int32_t v, int32_t w, bool hidden_edge) {
        int64_t c = du + (int64_t)w;
        ++relax;
        if (c < (int64_t)dist[v]) {
            dist[v] = (int32_t)c;
            ++impr;
            if (hidden_edge) hid[v] = 1;
            if (!flag[v]) { flag[v] = 1; nxt.push_back(v); }
        }
    };
    while (!cur.empty()) {
        ++rounds;
        nxt.clear();
        for (int32_t u : cur) {
            if ((uint32_t)u >= (uint32_t)V) return fail(C2G_EINVAL, "frontier vertex %d out of range", u);
            const int64_t du = dist[u];
            if (du == INT32_MAX) continue;
            for (int32_t e = vl[u]; e < vl[u + 1]; ++e) relax_edge(du, el[e], wl[e], false);
            if (pruned && hid[u])
                for (int32_t e = g->hvl[u]; e < g->hvl[u + 1]; ++e)
                    relax_edge(du, g->hel[e], g->hwl[e], true);
        }
        for (int32_t v : nxt) flag[v] = 0;
        cur.swap(nxt);
    }
    if (st) {
        st->host_iterations = rounds;
        st->edges_relaxed = relax;
        st->improvements = impr;
    }
    return C2G_OK;
}

static int host_push(const Graph* g, float alpha, float eps, float* p, float* r,
                     const int32_t* frontier, int32_t count, c2g_stats* st) {
    const int32_t V = g->V;
    if (g->fl.empty() && g->E > 0) return fail(C2G_ENOTSUP, "graph has no float weights (flist)");
    const int32_t* vl = g->vl.data();
    const int32_t* el = g->el.data();
    const float* fl = g->fl.data();
    std::vector<float> msg((size_t)V, 0.0f);
    std::vector<uint8_t> flag((size_t)V, 0);
    std::vector<int32_t> act(frontier, frontier + count), touched;
    int rounds = 0;
    int64_t pushes = 0;
    while (!act.empty()) {
        ++rounds;
        touched.clear();
        for (int32_t u : act) {
            const float ru = r[u];
            r[u] = 0.0f;
            if (vl[u + 1] - vl[u] <= 0) continue;
            const float m = (1.0f - alpha) * ru;
            for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
                const int32_t v = el[e];
                msg[v] += m * fl[e];
                ++pushes;
                if (!flag[v]) { flag[v] = 1; touched.push_back(v); }
            }
        }
        act.clear();
        for (int32_t v : touched) {
            r[v] += msg[v];
            msg[v] = 0.0f;
            flag[v] = 0;
            if (r[v] > (float)(vl[v + 1] - vl[v]) * eps) {
                p[v] += r[v] * alpha;
                act.push_back(v);
            }
        }
    }
    if (st) {
        st->host_iterations = rounds;
        st->edges_relaxed = pushes;
    }
    return C2G_OK;
}

// Adsorption on host cores (server fallback lane): the Jacobi recurrence of
// the device kernel, FP32, rows in parallel.
int host_ads(const Graph* g, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L, float p_inj,
             float p_cont, int32_t iters, float* y_out) {
    if (g->fl.empty() && g->E > 0) return fail(C2G_ENOTSUP, "graph has no float weights (flist)");
    const int64_t n = (int64_t)g->V * L;
    std::vector<float> y0((size_t)n, 0.0f), yt((size_t)n), yn((size_t)n);
    for (int32_t i = 0; i < nseed; ++i) y0[(size_t)sv[i] * L + sl[i]] = 1.0f;
    yt = y0;
    for (int32_t t = 0; t < iters; ++t) {
        parallel_for(g->V, [&](int64_t lo, int64_t hi, int) {
            std::vector<float> acc((size_t)L);
            for (int64_t v = lo; v < hi; ++v) {
                std::fill(acc.begin(), acc.end(), 0.0f);
                for (int32_t e = g->vl[v]; e < g->vl[v + 1]; ++e) {
                    const float f = g->fl[e];
                    const float* y = &yt[(size_t)g->el[e] * L];
                    for (int32_t l = 0; l < L; ++l) acc[l] += f * y[l];
                }
                for (int32_t l = 0; l < L; ++l) yn[(size_t)v * L + l] = p_inj * y0[(size_t)v * L + l] + p_cont * acc[l];
            }
        });
        yt.swap(yn);
    }
    memcpy(y_out, yt.data(), sizeof(float) * (size_t)n);
    return C2G_OK;
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_host_bellman_ford(const c2g_graph* h, int32_t* dist, const int32_t* frontier,
                          int32_t count, const uint8_t* via_hidden, c2g_stats* st) {
    if (!h || !dist || (count > 0 && !frontier)) return fail(C2G_EINVAL, "null argument");
    if (st) memset(st, 0, sizeof *st);
    double t0 = now_ms();
    int rc = host_bf(&h->g, dist, frontier, count, via_hidden, st);
    if (st) { st->lane = C2G_LANE_HOST; st->ms = now_ms() - t0; }
    return rc;
}

int c2g_host_sssp(const c2g_graph* h, int32_t src, int32_t* dist_out, c2g_stats* st) {
    if (!h || !dist_out) return fail(C2G_EINVAL, "null argument");
    const Graph* g = &h->g;
    if (src < 0 || src >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", src, g->V);
    if (st) memset(st, 0, sizeof *st);
    double t0 = now_ms();
    for (int32_t v = 0; v < g->V; ++v) dist_out[v] = INT32_MAX;
    dist_out[src] = 0;
    std::vector<uint8_t> hid;
    if (!g->hvl.empty()) { hid.assign((size_t)g->V, 0); hid[src] = 1; }
    int rc = host_bf(g, dist_out, &src, 1, hid.empty() ? nullptr : hid.data(), st);
    if (rc) return rc;
    if (st) {
        st->lane = C2G_LANE_HOST;
        st->ms = now_ms() - t0;
        int32_t r = 0;
        int64_t m = 0;
        for (int32_t v = 0; v < g->V; ++v)
            if (dist_out[v] != INT32_MAX) { ++r; m += g->vl[v + 1] - g->vl[v]; }
        st->reached = r;
        st->reached_edges = m;
    }
    return C2G_OK;
}

int c2g_host_forward_push(const c2g_graph* h, float alpha, float eps, float* p, float* r,
                          const int32_t* frontier, int32_t count, c2g_stats* st) {
    if (!h || !p || !r || (count > 0 && !frontier)) return fail(C2G_EINVAL, "null argument");
    if (st) memset(st, 0, sizeof *st);
    double t0 = now_ms();
    int rc = host_push(&h->g, alpha, eps, p, r, frontier, count, st);
    if (st) { st->lane = C2G_LANE_HOST; st->ms = now_ms() - t0; }
    return rc;
}

int c2g_host_ppr(const c2g_graph* h, int32_t src, float alpha, float eps, float* p_out,
                 float* r_out, c2g_stats* st) {
    if (!h || !p_out) return fail(C2G_EINVAL, "null argument");
    const Graph* g = &h->g;
    if (src < 0 || src >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", src, g->V);
    std::vector<float> rr;
    float* r = r_out;
    if (!r) { rr.assign((size_t)g->V, 0.0f); r = rr.data(); }
    for (int32_t v = 0; v < g->V; ++v) { p_out[v] = 0.0f; r[v] = 0.0f; }
    r[src] = 1.0f;
    p_out[src] = alpha;  // src/ppr/base/Main.cu:146-150
    return c2g_host_forward_push(h, alpha, eps, p_out, r, &src, 1, st);
}

}  // extern "C"
