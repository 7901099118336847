// adsorption <dataset_path> [--labels 16] [--frac 0.01] [--iters 20] [--seed 1]
//            [--p_inj .25 --p_cont .70 --p_abnd .05] [--out Y.bin]
// Multi-label adsorption over origin/flist.bin with labels seeded on a random
// fraction of vertices.
#include <iostream>
#include "cli_common.h"
int main(int argc, char** argv) {
    Args a = parse_args(argc, argv);
    if (a.pos.empty()) { fprintf(stderr, "usage: adsorption <dataset_path> [--labels L] [--iters T] [--out file]\n"); return 2; }
    c2g_graph* g = open_graph(a.pos[0], C2G_VARIANT_ORIGIN, true, (int)a.num("device", -1));
    if (!g) return 1;
    int32_t V;
    c2g_graph_info(g, &V, nullptr, nullptr, nullptr);
    const int L = (int)a.num("labels", 16);
    const int n = std::max(1, (int)(V * a.real("frac", 0.01)));
    std::vector<int32_t> sv(n), sl(n);
    uint64_t x = (uint64_t)a.num("seed", 1) * 0x9e3779b97f4a7c15ull + 1;
    for (int i = 0; i < n; ++i) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        sv[i] = (int32_t)(x % (uint64_t)V);
        sl[i] = (int32_t)((x >> 32) % (uint64_t)L);
    }
    std::vector<float> y((size_t)V * L);
    c2g_stats st{};
    const double t0 = wall_s();
    int rc = c2g_adsorption(g, sv.data(), sl.data(), n, L, (float)a.real("p_inj", 0.25), (float)a.real("p_cont", 0.70),
                            (float)a.real("p_abnd", 0.05), (int)a.num("iters", 20), y.data(), &st);
    if (rc) return die(rc);
    std::cout << "gpu runtime: " << (wall_s() - t0) << " seconds" << std::endl;
    for (int i = 0; i < 10 && i < V; ++i) {
        std::cout << i;
        for (int l = 0; l < L && l < 4; ++l) std::cout << "\t" << y[(size_t)i * L + l];
        std::cout << std::endl;
    }
    rc = maybe_write(a, y.data(), (int64_t)V * L, true);
    c2g_close(g);
    return rc;
}
