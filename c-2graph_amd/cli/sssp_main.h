// Shared main of sssp-base / sssp-purn / sssp-base-M / sssp-purn-M.
#pragma once
#include "cli_common.h"

// mode: 0 base, 1 purn, 2 base-M, 3 purn-M
inline int sssp_main(int argc, char** argv, int mode, const char* usage) {
    Args a = parse_args(argc, argv, {"rounds"});
    const bool mig = mode >= 2;
    if (a.pos.size() < (mig ? 3u : 2u)) {
        fprintf(stderr, "usage: %s\n", usage);
        return 2;
    }
    const std::string ds = a.pos[0];
    int32_t src = atoi(a.pos[1].c_str());
    const int budget = mig ? atoi(a.pos[2].c_str()) : 0;
    const bool purn = mode == 1 || mode == 3;
    std::cout << ds << std::endl;
    c2g_graph* g = open_graph(ds, purn ? C2G_VARIANT_SSSP_PURN : C2G_VARIANT_ORIGIN, false, (int)a.num("device", -1));
    if (!g) return 1;
    int32_t V;
    int64_t E, HE;
    c2g_graph_info(g, &V, &E, &HE, nullptr);
    const int32_t* vl;
    c2g_graph_host_csr(g, &vl, nullptr, nullptr, nullptr);
    std::vector<int32_t> hv;
    if (purn) {
        int64_t n = c2g_read_i32((ds + "/sssp/hvlist.bin").c_str(), nullptr, 0);
        hv.resize(n);
        c2g_read_i32((ds + "/sssp/hvlist.bin").c_str(), hv.data(), n);
    }
    std::vector<int32_t> indeg;
    if (mode == 2) {
        const int32_t* el;
        c2g_graph_host_csr(g, nullptr, &el, nullptr, nullptr);
        indeg.assign(V, 0);
        for (int64_t e = 0; e < E; ++e) indeg[el[e]]++;
    }
    // Source rules: sssp/base/Main.cu:61-64, sssp/purn/Main.cu:68-71 (without
    // the 1000-hidden-edge limit), sssp/base-perk/Main.cu:94, sssp/purn-perk/Main.cu:89.
    auto ok = [&](int32_t s) {
        const int d = vl[s + 1] - vl[s];
        if (mode == 0 || mode == 1) return d >= 1;
        if (mode == 2) return d >= 10 && indeg[s] >= 10;
        const int h = hv[s + 1] - hv[s];
        return h >= 10 && h <= 1000;
    };
    while (src < V && !ok(src)) ++src;
    if (src >= V) {
        fprintf(stderr, "no source at or after the given id satisfies the source rule\n");
        return 1;
    }
    const int reps = (int)a.num("reps", 10);
    std::vector<int32_t> dist(V);
    c2g_stats st{};
    std::cout << "\n==================== SSSP with FORWARD PUSH starts ============" << std::endl;
    const double t0 = wall_s();
    for (int r = 0; r < reps; ++r) {
        int rc = mig ? c2g_sssp_migrate(g, src, budget, dist.data(), &st) : c2g_sssp(g, src, dist.data(), &st);
        if (rc) return die(rc);
    }
    const double t1 = wall_s();
    std::cout << "\n==================== SSSP with FORWARD PUSH ends ============" << std::endl;
    std::cout << "iteration_num: " << st.iterations << " (host " << st.host_iterations << ")" << std::endl;
    for (int i = 0; i < 10 && i < V; i++) std::cout << i << "\t" << dist[i] << std::endl;
    std::cout << "gpu runtime: " << (t1 - t0) << " seconds" << std::endl;
    std::cout << "源顶点source = " << src << std::endl;
    if (a.has("rounds") && !mig) {
        // per-iteration frontier sizes, the reference's iteration_act_num
        // (sssp/base/Main.cu:87-93 prints them in a commented-out loop)
        std::vector<c2g_round> rs(100000);
        int32_t nr = 0;
        int rc2 = c2g_sssp_rounds(g, src, dist.data(), rs.data(), (int32_t)rs.size(), &nr, &st);
        if (rc2) return die(rc2);
        std::cout << "iteration\tfrontier\tedges\timprovements\tms" << std::endl;
        for (int i = 0; i < nr && i < (int)rs.size(); ++i)
            std::cout << i << "\t" << rs[i].frontier << "\t" << rs[i].edges << "\t" << rs[i].improvements << "\t"
                      << rs[i].ms << std::endl;
    }
    int rc = maybe_write(a, dist.data(), V, false);
    c2g_close(g);
    return rc;
}
