// Shared main of ppr-base / ppr-purn / ppr-base-M / ppr-purn-M.
#pragma once
#include "cli_common.h"
#include <iostream>

// mode: 0 base, 1 purn, 2 base-M, 3 purn-M.  The -M tools read origin/ and
// ppr/ respectively (the reference swaps them, ppr/w-base-perk/Graph.h:83-85,
// ppr/w-purn-perk/Graph.h:72-83).
inline int ppr_main(int argc, char** argv, int mode, const char* usage) {
    Args a = parse_args(argc, argv);
    const bool mig = mode >= 2;
    if (a.pos.size() < (mig ? 3u : 2u)) {
        fprintf(stderr, "usage: %s\n", usage);
        return 2;
    }
    const std::string ds = a.pos[0];
    const int32_t src = atoi(a.pos[1].c_str());
    const int budget = mig ? atoi(a.pos[2].c_str()) : 0;
    const bool purn = mode == 1 || mode == 3;
    // Defaults are the workload's (alpha 0.15, eps 1e-7 * outdeg); the
    // reference's compiled-in values are alpha 0.2 (ppr/base/Main.cu:61).
    const float alpha = (float)a.real("alpha", 0.15);
    const float eps = (float)a.real("eps", 1e-7);
    c2g_graph* g = open_graph(ds, purn ? C2G_VARIANT_PPR_PURN : C2G_VARIANT_ORIGIN, true, (int)a.num("device", -1));
    if (!g) return 1;
    int32_t V;
    c2g_graph_info(g, &V, nullptr, nullptr, nullptr);
    if (src < 0 || src >= V) {
        fprintf(stderr, "source %d out of range\n", src);
        return 1;
    }
    const int reps = (int)a.num("reps", 10);
    std::vector<float> p(V), r(V);
    c2g_stats st{};
    std::cout << "\n==================== PPR with FORWARD PUSH starts ====================" << std::endl;
    const double t0 = wall_s();
    for (int i = 0; i < reps; ++i) {
        int rc = mig ? c2g_ppr_migrate(g, src, alpha, eps, budget, (int)a.num("handoff", 10000), p.data(), r.data(), &st)
                     : c2g_ppr(g, src, alpha, eps, p.data(), r.data(), &st);
        if (rc) return die(rc);
    }
    const double t1 = wall_s();
    std::cout << "source = " << src << "end ------------" << reps << std::endl;
    std::cout << "gpu runtime: " << (t1 - t0) << " seconds" << std::endl;
    std::cout << "iteration_num: " << st.iterations << " (host " << st.host_iterations << ")" << std::endl;
    std::cout << "==================== PPR with FORWARD PUSH ends ====================\n" << std::endl;
    for (int i = 0; i <= 10 && i < V; i++)
        std::cout << i << ".\tpagerank\t " << p[i] << "\tresidual\t" << r[i] << std::endl;
    int rc = maybe_write(a, p.data(), V, true);
    c2g_close(g);
    return rc;
}
