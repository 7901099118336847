// C^2graph-P preprocessing: purn-sssp and purn-ppr.
//
// Makes the same pruning decisions as the reference (src/purn-sssp.cpp:61-352,
// src/purn-ppr.cpp:63-344) and writes byte-identical pruned CSRs; hidden lists
// are written per vertex (hvlist has V+1 entries, the layout the pruned-graph
// loaders index: src/sssp/purn/Graph.h:40, src/sssp/purn/Main.cu:72).  For
// purn-sssp the reference writes hvlist in prune order (one entry per
// processed vertex, src/purn-sssp.cpp:283-284); the processed vertices are
// written to sssp/porder.bin so that layout can be rebuilt.
//
// Algorithm (shared by both; W is int for SSSP, float for PPR):
//   * graph[u]  = distinct (v,w) out-pairs sorted by (v,w), self-loops dropped
//                 (csr_convert, core/util.hpp:80-115);
//     graphT[v] = distinct in-neighbours u, sorted, each with an alive flag
//                 (csr_convert_idx, src/purn-sssp.cpp:40-58); indeg = alive count.
//   * Visit order: SSSP sweeps k = 0..29 and, within each k, i = 0..V-1,
//     taking i when 0 < indeg(i) <= k and i was not taken before (the take
//     happens even if i has no out-edges, which then prunes nothing)
//     (src/purn-sssp.cpp:142-150).  PPR makes one pass i = 0..V-1 taking i
//     when graph[i] is non-empty and 0 < indeg(i) <= 5 (src/purn-ppr.cpp:146-152).
//   * For a taken i with out-edges: roots = alive in-neighbours r with
//     dis_r = w(r,i).  Each out-pair (v,d3) of i with v not a root is hidden
//     when at most one root lacks an edge to v (cnt + 2 > |roots|,
//     src/purn-sssp.cpp:207): v moves to i's hidden list, roots lacking v
//     become alive in-neighbours of v, and i -> v turns dead.  Then every root
//     merges the hidden pairs: SSSP w(r,v) = min(w(r,v), d3 + dis_r)
//     (src/purn-sssp.cpp:239-275); PPR w(r,v) += d3 * 0.8 * dis_r in double,
//     stored as float (src/purn-ppr.cpp:243-259).
// Only the data structures differ from the reference: merges with no hidden
// pairs are skipped (they rebuild the same row) and small merges insert by
// binary search.
#include "common.h"
#include <cstring>
#include <string>

namespace c2g {

int mkdirs(const std::string& d);
int64_t read_vec_i32(const std::string& p, std::vector<int32_t>& v);
int64_t read_vec_f32(const std::string& p, std::vector<float>& v);
int write_vec(const std::string& p, const void* d, int64_t n);

template <class W>
struct Pruner {
    using Pair = std::pair<int32_t, W>;
    int32_t V = 0;
    std::vector<std::vector<Pair>> graph;
    std::vector<std::vector<std::pair<int32_t, uint8_t>>> graphT;
    std::vector<int32_t> indeg;
    std::vector<std::vector<Pair>> hidden;
    std::vector<int32_t> porder;
    int64_t e_in = 0;

    int build(int32_t nv, const int32_t* vl, const int32_t* el, const W* wl) {
        V = nv;
        graph.assign((size_t)V, {});
        graphT.assign((size_t)V, {});
        hidden.assign((size_t)V, {});
        indeg.assign((size_t)V, 0);
        e_in = vl[V];
        // Rows as std::set<pair> would give them: sorted by (v,w), unique, no loops.
        parallel_for(V, [&](int64_t lo, int64_t hi, int) {
            std::vector<Pair> row;
            for (int64_t u = lo; u < hi; ++u) {
                row.clear();
                for (int32_t e = vl[u]; e < vl[u + 1]; ++e)
                    if (el[e] != u) row.emplace_back(el[e], wl[e]);
                std::sort(row.begin(), row.end());
                row.erase(std::unique(row.begin(), row.end()), row.end());
                graph[u] = row;
            }
        });
        // A target listed twice with different weights makes the reference's
        // merges depend on duplicate order; convert2binary and the generator
        // never produce it, so it is rejected.
        for (int32_t u = 0; u < V; ++u)
            for (size_t k = 1; k < graph[u].size(); ++k)
                if (graph[u][k].first == graph[u][k - 1].first)
                    return fail(C2G_EFORMAT, "vertex %d lists target %d twice", u, graph[u][k].first);
        // In-neighbour lists in increasing source order (distinct targets per
        // source): each thread owns a range of targets and scans the rows in
        // source order, so the lists do not depend on the thread count.
        const int T = std::max(1, std::min(host_threads(), 16));
        parallel_for(T, [&](int64_t t0, int64_t t1, int) {
            for (int64_t t = t0; t < t1; ++t) {
                const int32_t lo = (int32_t)((int64_t)V * t / T), hi = (int32_t)((int64_t)V * (t + 1) / T);
                std::vector<int32_t> cnt((size_t)(hi - lo), 0);
                for (int32_t u = 0; u < V; ++u)
                    for (const auto& p : graph[u])
                        if (p.first >= lo && p.first < hi) cnt[p.first - lo]++;
                for (int32_t v = lo; v < hi; ++v) graphT[v].reserve(cnt[v - lo]);
                for (int32_t u = 0; u < V; ++u)
                    for (const auto& p : graph[u])
                        if (p.first >= lo && p.first < hi) graphT[p.first].emplace_back(u, 1);
                for (int32_t v = lo; v < hi; ++v) indeg[v] = (int32_t)graphT[v].size();
            }
        }, T, 1);
        return C2G_OK;
    }

    static bool has_target(const std::vector<Pair>& row, int32_t v) {
        auto it = std::lower_bound(row.begin(), row.end(), v,
                                   [](const Pair& a, int32_t b) { return a.first < b; });
        return it != row.end() && it->first == v;
    }

    template <class Combine, class Fresh>
    static void merge_into(std::vector<Pair>& row, const std::vector<Pair>& add, W dis, Combine comb,
                           Fresh fresh) {
        if (add.empty()) return;
        if (add.size() * 16 < row.size()) {
            for (const auto& h : add) {
                auto it = std::lower_bound(row.begin(), row.end(), h.first,
                                           [](const Pair& a, int32_t b) { return a.first < b; });
                if (it != row.end() && it->first == h.first) it->second = comb(it->second, h.second, dis);
                else row.insert(it, Pair(h.first, fresh(h.second, dis)));
            }
            return;
        }
        std::vector<Pair> out;
        out.reserve(row.size() + add.size());
        size_t a = 0, b = 0;
        while (a < row.size() || b < add.size()) {
            if (a == row.size()) { out.emplace_back(add[b].first, fresh(add[b].second, dis)); ++b; }
            else if (b == add.size()) out.push_back(row[a++]);
            else if (row[a].first == add[b].first) {
                out.emplace_back(row[a].first, comb(row[a].second, add[b].second, dis));
                ++a; ++b;
            } else if (row[a].first > add[b].first) { out.emplace_back(add[b].first, fresh(add[b].second, dis)); ++b; }
            else out.push_back(row[a++]);
        }
        row.swap(out);
    }

    // Prunes vertex i (it has out-edges and 0 < indeg <= cap).
    template <class Combine, class Fresh>
    int process(int32_t i, Combine comb, Fresh fresh, std::vector<int32_t>& roots, std::vector<W>& dis,
                std::vector<Pair>& keep, std::vector<Pair>& hid, std::vector<int32_t>& lacking) {
        roots.clear();
        dis.clear();
        for (const auto& t : graphT[i]) {
            if (!t.second) continue;
            const auto& row = graph[t.first];
            auto it = std::lower_bound(row.begin(), row.end(), i,
                                       [](const Pair& a, int32_t b) { return a.first < b; });
            if (it == row.end() || it->first != i)
                return fail(C2G_EFORMAT, "prune: in-neighbour %d of %d has no edge to it", t.first, i);
            roots.push_back(t.first);
            dis.push_back(it->second);
        }
        const int32_t j = (int32_t)roots.size();
        keep.clear();
        hid.clear();
        for (const auto& e : graph[i]) {
            const int32_t v = e.first;
            if (std::find(roots.begin(), roots.end(), v) != roots.end()) { keep.push_back(e); continue; }
            int32_t cnt = 0;
            lacking.clear();
            for (int32_t k = 0; k < j; ++k) {
                if (has_target(graph[roots[k]], v)) ++cnt;
                else lacking.push_back(roots[k]);
            }
            if (cnt + 2 > j) {
                hid.push_back(e);
                auto& tv = graphT[v];
                for (int32_t r : lacking) {
                    if (r == v) continue;
                    auto it = std::lower_bound(tv.begin(), tv.end(), r,
                                               [](const std::pair<int32_t, uint8_t>& a, int32_t b) { return a.first < b; });
                    if (it == tv.end() || it->first != r) { tv.insert(it, {r, 1}); indeg[v]++; }
                    else if (!it->second) { it->second = 1; indeg[v]++; }
                }
                auto it = std::lower_bound(tv.begin(), tv.end(), i,
                                           [](const std::pair<int32_t, uint8_t>& a, int32_t b) { return a.first < b; });
                if (it == tv.end() || it->first != i)
                    return fail(C2G_EFORMAT, "prune: %d missing from in-list of %d", i, v);
                indeg[v]--;
                it->second = 0;
            } else {
                keep.push_back(e);
            }
        }
        for (int32_t k = 0; k < j; ++k) merge_into(graph[roots[k]], hid, dis[k], comb, fresh);
        graph[i].swap(keep);
        hidden[i] = hid;
        porder.push_back(i);
        return C2G_OK;
    }

    void flatten(std::vector<int32_t>& vl, std::vector<int32_t>& el, std::vector<W>& wl,
                 std::vector<int32_t>& hvl, std::vector<int32_t>& hel, std::vector<W>& hwl) const {
        vl.assign((size_t)V + 1, 0);
        hvl.assign((size_t)V + 1, 0);
        for (int32_t v = 0; v < V; ++v) {
            vl[v + 1] = vl[v] + (int32_t)graph[v].size();
            hvl[v + 1] = hvl[v] + (int32_t)hidden[v].size();
        }
        el.resize(vl[V]);
        wl.resize(vl[V]);
        hel.resize(hvl[V]);
        hwl.resize(hvl[V]);
        parallel_for(V, [&](int64_t lo, int64_t hi, int) {
            for (int64_t v = lo; v < hi; ++v) {
                int32_t o = vl[v];
                for (const auto& p : graph[v]) { el[o] = p.first; wl[o++] = p.second; }
                o = hvl[v];
                for (const auto& p : hidden[v]) { hel[o] = p.first; hwl[o++] = p.second; }
            }
        });
    }
};

static int prune_sssp_core(int32_t V, const int32_t* vl, const int32_t* el, const int32_t* wl,
                           std::vector<int32_t>& ov, std::vector<int32_t>& oe, std::vector<int32_t>& ow,
                           std::vector<int32_t>& hv, std::vector<int32_t>& he, std::vector<int32_t>& hw,
                           std::vector<int32_t>& porder) {
    Pruner<int32_t> P;
    int rc = P.build(V, vl, el, wl);
    if (rc) return rc;
    const int32_t dmax = 29;  // src/purn-sssp.cpp:139
    std::vector<uint8_t> taken((size_t)V, 0);
    std::vector<int32_t> roots, lacking;
    std::vector<int32_t> dis;
    std::vector<std::pair<int32_t, int32_t>> keep, hid;
    auto comb = [](int32_t w, int32_t d3, int32_t dr) { return std::min(w, d3 + dr); };
    auto fresh = [](int32_t d3, int32_t dr) { return d3 + dr; };
    for (int32_t k = 0; k <= dmax; ++k)
        for (int32_t i = 0; i < V; ++i) {
            const int32_t d = P.indeg[i];
            if (d > k || taken[i] || d == 0) continue;
            taken[i] = 1;
            if (!P.graph[i].empty() && d <= dmax)
                if ((rc = P.process(i, comb, fresh, roots, dis, keep, hid, lacking))) return rc;
        }
    P.flatten(ov, oe, ow, hv, he, hw);
    porder = P.porder;
    return C2G_OK;
}

static int prune_ppr_core(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                          std::vector<int32_t>& ov, std::vector<int32_t>& oe, std::vector<float>& ow,
                          std::vector<int32_t>& hv, std::vector<int32_t>& he, std::vector<float>& hw) {
    Pruner<float> P;
    int rc = P.build(V, vl, el, fl);
    if (rc) return rc;
    const int32_t dmax = 5;  // src/purn-ppr.cpp:146
    std::vector<int32_t> roots, lacking;
    std::vector<float> dis;
    std::vector<std::pair<int32_t, float>> keep, hid;
    // (w_rv + (w_iv * 0.8) * w_ri) evaluated in double, stored as float.
    auto comb = [](float w, float d3, float dr) { return (float)((double)w + (double)d3 * 0.8 * (double)dr); };
    auto fresh = [](float d3, float dr) { return (float)((double)d3 * 0.8 * (double)dr); };
    for (int32_t i = 0; i < V; ++i) {
        const int32_t d = P.indeg[i];
        if (!P.graph[i].empty() && d > 0 && d <= dmax)
            if ((rc = P.process(i, comb, fresh, roots, dis, keep, hid, lacking))) return rc;
    }
    P.flatten(ov, oe, ow, hv, he, hw);
    return C2G_OK;
}

template <class T>
static T* to_malloc(const std::vector<T>& v) {
    T* p = (T*)malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_prune_sssp_mem(int32_t V, const int32_t* vlist, const int32_t* elist, const int32_t* wlist,
                       int32_t** out_vlist, int32_t** out_elist, int32_t** out_wlist,
                       int32_t** out_hvlist, int32_t** out_helist, int32_t** out_hwlist,
                       int32_t** out_porder, int64_t* out_nporder) {
    if (V < 0 || !vlist || !elist || !wlist || !out_vlist || !out_elist || !out_wlist || !out_hvlist ||
        !out_helist || !out_hwlist)
        return fail(C2G_EINVAL, "c2g_prune_sssp_mem: null argument");
    std::vector<int32_t> ov, oe, ow, hv, he, hw, po;
    int rc = prune_sssp_core(V, vlist, elist, wlist, ov, oe, ow, hv, he, hw, po);
    if (rc) return rc;
    *out_vlist = to_malloc(ov); *out_elist = to_malloc(oe); *out_wlist = to_malloc(ow);
    *out_hvlist = to_malloc(hv); *out_helist = to_malloc(he); *out_hwlist = to_malloc(hw);
    if (out_porder) *out_porder = to_malloc(po);
    if (out_nporder) *out_nporder = (int64_t)po.size();
    return C2G_OK;
}

int c2g_prune_ppr_mem(int32_t V, const int32_t* vlist, const int32_t* elist, const float* flist,
                      int32_t** out_vlist, int32_t** out_elist, float** out_wlist, int32_t** out_hvlist,
                      int32_t** out_helist, float** out_hwlist) {
    if (V < 0 || !vlist || !elist || !flist || !out_vlist || !out_elist || !out_wlist || !out_hvlist ||
        !out_helist || !out_hwlist)
        return fail(C2G_EINVAL, "c2g_prune_ppr_mem: null argument");
    std::vector<int32_t> ov, oe, hv, he;
    std::vector<float> ow, hw;
    int rc = prune_ppr_core(V, vlist, elist, flist, ov, oe, ow, hv, he, hw);
    if (rc) return rc;
    *out_vlist = to_malloc(ov); *out_elist = to_malloc(oe); *out_wlist = to_malloc(ow);
    *out_hvlist = to_malloc(hv); *out_helist = to_malloc(he); *out_hwlist = to_malloc(hw);
    return C2G_OK;
}

// purn-sssp <ds>: origin/{vlist,elist,wlist} -> sssp/{vlist,elist,wlist,hvlist,helist,hwlist,porder}
int c2g_prune_sssp(const char* ds, int64_t* stats_out) {
    if (!ds) return fail(C2G_EINVAL, "null dataset path");
    const std::string d(ds);
    std::vector<int32_t> vl, el, wl;
    int64_t n;
    if ((n = read_vec_i32(d + "/origin/vlist.bin", vl)) < 0) return (int)n;
    if ((n = read_vec_i32(d + "/origin/elist.bin", el)) < 0) return (int)n;
    if ((n = read_vec_i32(d + "/origin/wlist.bin", wl)) < 0) return (int)n;
    if (vl.empty() || (int64_t)vl.back() != (int64_t)el.size() || wl.size() != el.size())
        return fail(C2G_EFORMAT, "%s/origin: inconsistent vlist/elist/wlist lengths", ds);
    const int32_t V = (int32_t)vl.size() - 1;
    for (int32_t e : el) if ((uint32_t)e >= (uint32_t)V) return fail(C2G_EFORMAT, "elist entry %d out of range", e);
    std::vector<int32_t> ov, oe, ow, hv, he, hw, po;
    int rc = prune_sssp_core(V, vl.data(), el.data(), wl.data(), ov, oe, ow, hv, he, hw, po);
    if (rc) return rc;
    if ((rc = mkdirs(d + "/sssp"))) return rc;
    if ((rc = write_vec(d + "/sssp/vlist.bin", ov.data(), (int64_t)ov.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/elist.bin", oe.data(), (int64_t)oe.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/wlist.bin", ow.data(), (int64_t)ow.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/hvlist.bin", hv.data(), (int64_t)hv.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/helist.bin", he.data(), (int64_t)he.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/hwlist.bin", hw.data(), (int64_t)hw.size()))) return rc;
    if ((rc = write_vec(d + "/sssp/porder.bin", po.data(), (int64_t)po.size()))) return rc;
    if (stats_out) { stats_out[0] = (int64_t)oe.size(); stats_out[1] = (int64_t)he.size(); stats_out[2] = (int64_t)po.size(); }
    return C2G_OK;
}

// purn-ppr <ds>: origin/{vlist,elist,flist} -> ppr/{vlist,elist,wlist,hvlist,helist,hwlist}
int c2g_prune_ppr(const char* ds, int64_t* stats_out) {
    if (!ds) return fail(C2G_EINVAL, "null dataset path");
    const std::string d(ds);
    std::vector<int32_t> vl, el;
    std::vector<float> fl;
    int64_t n;
    if ((n = read_vec_i32(d + "/origin/vlist.bin", vl)) < 0) return (int)n;
    if ((n = read_vec_i32(d + "/origin/elist.bin", el)) < 0) return (int)n;
    if ((n = read_vec_f32(d + "/origin/flist.bin", fl)) < 0) return (int)n;
    if (vl.empty() || (int64_t)vl.back() != (int64_t)el.size() || fl.size() != el.size())
        return fail(C2G_EFORMAT, "%s/origin: inconsistent vlist/elist/flist lengths", ds);
    const int32_t V = (int32_t)vl.size() - 1;
    for (int32_t e : el) if ((uint32_t)e >= (uint32_t)V) return fail(C2G_EFORMAT, "elist entry %d out of range", e);
    std::vector<int32_t> ov, oe, hv, he;
    std::vector<float> ow, hw;
    int rc = prune_ppr_core(V, vl.data(), el.data(), fl.data(), ov, oe, ow, hv, he, hw);
    if (rc) return rc;
    if ((rc = mkdirs(d + "/ppr"))) return rc;
    if ((rc = write_vec(d + "/ppr/vlist.bin", ov.data(), (int64_t)ov.size()))) return rc;
    if ((rc = write_vec(d + "/ppr/elist.bin", oe.data(), (int64_t)oe.size()))) return rc;
    if ((rc = write_vec(d + "/ppr/wlist.bin", ow.data(), (int64_t)ow.size()))) return rc;
    if ((rc = write_vec(d + "/ppr/hvlist.bin", hv.data(), (int64_t)hv.size()))) return rc;
    if ((rc = write_vec(d + "/ppr/helist.bin", he.data(), (int64_t)he.size()))) return rc;
    if ((rc = write_vec(d + "/ppr/hwlist.bin", hw.data(), (int64_t)hw.size()))) return rc;
    if (stats_out) { stats_out[0] = (int64_t)oe.size(); stats_out[1] = (int64_t)he.size(); stats_out[2] = 0; }
    return C2G_OK;
}

}  // extern "C"
