// purn-sssp <dataset_path> (src/purn-sssp.cpp:61-352)
#include <iostream>
#include "cli_common.h"
int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "Usage: purn-sssp <dataset_path>\n"); return 2; }
    fprintf(stderr, "purn start...\n");
    int64_t st[3];
    const double t0 = wall_s();
    int rc = c2g_prune_sssp(argv[1], st);
    if (rc) return die(rc);
    int64_t E0 = c2g_read_i32((std::string(argv[1]) + "/origin/elist.bin").c_str(), nullptr, 0);
    printf("the ratio is: %f\n", E0 ? 1.0 * st[0] / E0 : 1.0);
    printf("the added gr size is : %lld\n", (long long)st[1]);
    printf("processed vertices: %lld\n", (long long)st[2]);
    fprintf(stderr, "Filter time : %.4f(s)\n", wall_s() - t0);
    return 0;
}
