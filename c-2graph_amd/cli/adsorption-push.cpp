// adsorption-push <dataset_path> <source> [--out file]
// The reference's adsorption kernel (src/adsorption/w-base/Main.cu:24-103):
// a single-source push over origin/flist.bin with alpha 0.2, r[s] = 10 and the
// constant threshold 1/V.  Prints the reference's lines (run time and
// "i.\tpagerank\t p\tresidual\t r" for i = 0..10); --out writes the dump
// "i p r" per line (the reference writes it to a fixed path, Main.cu:411-425).
#include <fstream>
#include <iostream>
#include "cli_common.h"
int main(int argc, char** argv) {
    Args a = parse_args(argc, argv);
    if (a.pos.size() < 2) { fprintf(stderr, "usage: adsorption-push <dataset_path> <source> [--out file]\n"); return 2; }
    c2g_graph* g = open_graph(a.pos[0], C2G_VARIANT_ORIGIN, true, (int)a.num("device", -1));
    if (!g) return 1;
    int32_t V;
    c2g_graph_info(g, &V, nullptr, nullptr, nullptr);
    const int32_t src = atoi(a.pos[1].c_str());
    std::vector<float> p(V), r(V);
    c2g_stats st{};
    std::cout << "\n==================== PPR with FORWARD PUSH starts ====================" << std::endl;
    const double t0 = wall_s();
    int rc = c2g_adsorption_push(g, src, p.data(), r.data(), &st);
    if (rc) return die(rc);
    const double t1 = wall_s();
    std::cout << "source = " << src << "end ------------" << std::endl;
    std::cout << "gpu runtime: " << (t1 - t0) << " seconds" << std::endl;
    std::cout << "iteration_num: " << st.iterations << std::endl;
    std::cout << "==================== PPR with FORWARD PUSH ends ====================\n" << std::endl;
    for (int i = 0; i <= 10 && i < V; ++i)
        std::cout << i << ".\tpagerank\t " << p[i] << "\tresidual\t" << r[i] << std::endl;
    if (a.has("out")) {
        std::ofstream f(a.str("out", ""));
        if (!f) { fprintf(stderr, "cannot write %s\n", a.str("out", "").c_str()); return 1; }
        for (int i = 0; i < V; ++i) f << i << " " << p[i] << " " << r[i] << "\n";
    }
    c2g_close(g);
    return 0;
}
