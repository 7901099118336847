// purn-ppr <dataset_path> (src/purn-ppr.cpp:63-344)
#include <iostream>
#include "cli_common.h"
int main(int argc, char** argv) {
    if (argc != 2) { fprintf(stderr, "Usage: purn-ppr <dataset_path>\n"); return 2; }
    fprintf(stderr, "purn start...\n");
    int64_t st[3];
    const double t0 = wall_s();
    int rc = c2g_prune_ppr(argv[1], st);
    if (rc) return die(rc);
    printf("2st e1 : %lld\n", (long long)st[0]);
    printf("the added gr size is : %lld\n", (long long)st[1]);
    fprintf(stderr, "Filter time : %.4f(s)\n", wall_s() - t0);
    return 0;
}
