// rmat-gen <dataset_path> <scale> [--ef 16] [--seed 1] [--wmax 64] [--symmetric]
//         [--a .57 --b .19 --c .19] [--no-flist]
// Writes a seeded R-MAT graph as <dataset_path>/origin/{vlist,elist,wlist,flist}.bin.
#include "cli_common.h"
int main(int argc, char** argv) {
    Args a = parse_args(argc, argv, {"symmetric", "no-flist"});
    if (a.pos.size() < 2) {
        fprintf(stderr, "usage: rmat-gen <dataset_path> <scale> [--ef 16] [--seed 1] [--wmax 64] [--symmetric]\n");
        return 2;
    }
    int32_t *vl = nullptr, *el = nullptr, *wl = nullptr;
    float* fl = nullptr;
    int64_t E = 0;
    const int scale = atoi(a.pos[1].c_str());
    const bool want_f = !a.has("no-flist");
    int rc = c2g_rmat(scale, (int)a.num("ef", 16), a.real("a", 0.57), a.real("b", 0.19), a.real("c", 0.19),
                      (uint64_t)a.num("seed", 1), a.has("symmetric") ? 1 : 0, (int)a.num("wmax", 64), want_f ? 1 : 0,
                      &vl, &el, &wl, want_f ? &fl : nullptr, &E);
    if (rc) return die(rc);
    rc = c2g_write_dataset(a.pos[0].c_str(), 1 << scale, vl, el, wl, fl);
    if (rc) return die(rc);
    printf("Vertex Number: %d\nEdge Number: %lld\n", 1 << scale, (long long)E);
    c2g_free(vl); c2g_free(el); c2g_free(wl); c2g_free(fl);
    return 0;
}
