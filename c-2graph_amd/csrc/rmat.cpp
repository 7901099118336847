// Seeded R-MAT generator writing the reference's CSR layout.
//
// The reference ingests SNAP adjacency text (utils/convert2binary.cpp:10-80):
// it drops self-loops (:67), sorts and dedups each row (:76-79) and writes
// vlist/elist/wlist/flist (:118-134).  The synthetic workloads instead come
// from this generator, which produces the same layout:
//   * edge k in [0, edge_factor << scale): quadrant recursion over `scale`
//     levels with probabilities (a, b, c, 1-a-b-c); level j of edge k uses
//     32 bits of hash3(seed, 1, k*ceil(scale/2) + j/2);
//   * vertex ids pass through perm(), a seeded bijection of [0, 2^scale)
//     (two rounds of odd-multiply-add and xorshift), so hubs are not all at
//     low ids;
//   * self-loops are dropped, rows sorted ascending, duplicates removed;
//   * int weight  w(u,v) = 1 + floor(hi32(hash3(seed, 2, u<<32|v)) * wmax / 2^32),
//     uniform in [1, wmax] and independent of edge order;
//   * float weight f(u,v) = x(u,v) / sum_row x, x = U(0,1] from hash3(seed, 3, .),
//     summed and divided in double, then rounded to float.
#include "common.h"
#include <cstdlib>
#include <cstring>

namespace c2g {

struct Perm {
    uint32_t mask, shift;
    uint32_t m1, a1, m2, a2;
    Perm(int scale, uint64_t seed) {
        mask = scale >= 32 ? 0xffffffffu : ((1u << scale) - 1u);
        shift = (uint32_t)((scale + 1) / 2);
        uint64_t h = hash3(seed, 7, 0), g = hash3(seed, 7, 1);
        m1 = (uint32_t)h | 1u;
        a1 = (uint32_t)(h >> 32);
        m2 = (uint32_t)g | 1u;
        a2 = (uint32_t)(g >> 32);
    }
    inline uint32_t operator()(uint32_t x) const {
        x = (x * m1 + a1) & mask;
        x ^= x >> shift;
        x = (x * m2 + a2) & mask;
        x ^= x >> shift;
        return x & mask;
    }
};

struct RmatDraw {
    int scale;
    Stream rng;
    uint32_t ta, tab, tabc;
    int draws;
    Perm perm;
    RmatDraw(int s, uint64_t sd, double a, double b, double c)
        : scale(s), rng(sd, 1), perm(s, sd) {
        auto th = [](double p) -> uint32_t {
            double x = p * 4294967296.0;
            if (x >= 4294967295.0) return 0xffffffffu;
            if (x <= 0) return 0;
            return (uint32_t)x;
        };
        ta = th(a);
        tab = th(a + b);
        tabc = th(a + b + c);
        draws = (s + 1) / 2;
    }
    // Returns false for a self-loop.
    inline bool edge(uint64_t k, uint32_t& u, uint32_t& v) const {
        uint32_t uu = 0, vv = 0;
        uint64_t base = k * (uint64_t)draws;
        for (int j = 0; j < scale; j += 2) {
            uint64_t r = rng(base + (uint64_t)(j >> 1));
            for (int h = 0; h < 2 && j + h < scale; ++h) {
                uint32_t x = (uint32_t)(h ? (r >> 32) : r);
                uint32_t bu, bv;
                if (x < ta) { bu = 0; bv = 0; }
                else if (x < tab) { bu = 0; bv = 1; }
                else if (x < tabc) { bu = 1; bv = 0; }
                else { bu = 1; bv = 1; }
                uu = (uu << 1) | bu;
                vv = (vv << 1) | bv;
            }
        }
        u = perm(uu);
        v = perm(vv);
        return u != v;
    }
};

static inline int32_t int_weight(const Stream& s2, uint32_t u, uint32_t v, int32_t wmax) {
    uint64_t h = s2(((uint64_t)u << 32) | v) >> 32;
    return 1 + (int32_t)((h * (uint64_t)wmax) >> 32);
}
static inline double unit_weight(const Stream& s3, uint32_t u, uint32_t v) {
    uint64_t h = s3(((uint64_t)u << 32) | v) >> 11;  // 53 bits
    return (double)(h + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
}

int rmat_build(int scale, int ef, double a, double b, double c, uint64_t seed, int symmetric,
               int32_t wmax, int want_f, int32_t** o_vl, int32_t** o_el, int32_t** o_wl,
               float** o_fl, int64_t* o_E) {
    if (scale < 1 || scale > 30 || ef < 1)
        return fail(C2G_EINVAL, "rmat: scale must be in [1,30] and edge_factor >= 1");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
        return fail(C2G_EINVAL, "rmat: bad quadrant probabilities");
    const int64_t V = (int64_t)1 << scale;
    const int64_t M = (int64_t)ef << scale;
    const int64_t slots = M * (symmetric ? 2 : 1);
    if (slots > INT32_MAX) return fail(C2G_EINVAL, "rmat: %lld edge slots exceed int32 offsets", (long long)slots);
    RmatDraw dr(scale, seed, a, b, c);

    // Pass 1: draw every edge as a key (u << 32 | v), counting keys per
    // (thread, source bucket); buckets are ranges of 2^bshift source ids.
    const int bshift = scale > 12 ? scale - 12 : 0;
    const int64_t NB = V >> bshift;
    const int T = host_threads();
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)std::max<int64_t>(slots, 1));
    uint64_t* sorted = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)std::max<int64_t>(slots, 1));
    if (!keys || !sorted) { free(keys); free(sorted); return fail(C2G_ENOMEM, "rmat: key allocation failed"); }
    std::vector<int64_t> hist((size_t)T * NB, 0), tcount((size_t)T, 0);
    const Stream s2(seed, 2), s3(seed, 3);
    parallel_for(T, [&](int64_t lo, int64_t hi, int) {
        for (int64_t t = lo; t < hi; ++t) {
            const int64_t k0 = M * t / T, k1 = M * (t + 1) / T;
            uint64_t* out = keys + (symmetric ? 2 * k0 : k0);
            int64_t* h = hist.data() + t * NB;
            int64_t n = 0;
            for (int64_t k = k0; k < k1; ++k) {
                uint32_t u, v;
                if (!dr.edge((uint64_t)k, u, v)) continue;
                out[n++] = ((uint64_t)u << 32) | v;
                h[u >> bshift]++;
                if (symmetric) { out[n++] = ((uint64_t)v << 32) | u; h[v >> bshift]++; }
            }
            tcount[t] = n;
        }
    }, T, 1);
    // Bucket offsets, bucket-major then thread.
    std::vector<int64_t> boff((size_t)NB + 1, 0), cur((size_t)T * NB);
    {
        int64_t o = 0;
        for (int64_t bk = 0; bk < NB; ++bk) {
            boff[bk] = o;
            for (int t = 0; t < T; ++t) { cur[(size_t)t * NB + bk] = o; o += hist[(size_t)t * NB + bk]; }
        }
        boff[NB] = o;
    }
    // Pass 2: scatter into buckets (each thread owns its cursor ranges).
    parallel_for(T, [&](int64_t lo, int64_t hi, int) {
        for (int64_t t = lo; t < hi; ++t) {
            const int64_t k0 = M * t / T;
            const uint64_t* in = keys + (symmetric ? 2 * k0 : k0);
            int64_t* cu = cur.data() + t * NB;
            for (int64_t i = 0; i < tcount[t]; ++i) {
                uint64_t key = in[i];
                sorted[cu[(key >> 32) >> bshift]++] = key;
            }
        }
    }, T, 1);
    free(keys);
    // Pass 3: sort + dedup each bucket, count deduplicated out-degrees.
    std::vector<int32_t> cnt((size_t)V, 0);
    std::vector<int64_t> bkeep((size_t)NB, 0);
    parallel_dyn(NB, [&](int64_t lo, int64_t hi, int) {
        for (int64_t bk = lo; bk < hi; ++bk) {
            uint64_t* s0 = sorted + boff[bk];
            uint64_t* s1 = sorted + boff[bk + 1];
            std::sort(s0, s1);
            uint64_t* e = std::unique(s0, s1);
            bkeep[bk] = e - s0;
            for (uint64_t* p = s0; p < e; ++p) cnt[*p >> 32]++;
        }
    });
    int32_t* vl = (int32_t*)malloc(sizeof(int32_t) * ((size_t)V + 1));
    if (!vl) { free(sorted); return fail(C2G_ENOMEM, "rmat: vlist allocation failed"); }
    vl[0] = 0;
    for (int64_t i = 0; i < V; ++i) vl[i + 1] = vl[i] + cnt[i];
    const int64_t E = vl[V];
    int32_t* el = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1));
    int32_t* wl = wmax > 0 ? (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(E, 1)) : nullptr;
    float* fl = want_f ? (float*)malloc(sizeof(float) * (size_t)std::max<int64_t>(E, 1)) : nullptr;
    if (!el || (wmax > 0 && !wl) || (want_f && !fl)) {
        free(sorted); free(vl); free(el); free(wl); free(fl);
        return fail(C2G_ENOMEM, "rmat: edge array allocation failed");
    }
    // Pass 4: write rows (buckets hold whole rows in ascending order).
    parallel_dyn(NB, [&](int64_t lo, int64_t hi, int) {
        for (int64_t bk = lo; bk < hi; ++bk) {
            const uint64_t* s0 = sorted + boff[bk];
            const int64_t n = bkeep[bk];
            if (!n) continue;
            int64_t o = vl[s0[0] >> 32];
            for (int64_t j = 0; j < n; ++j) {
                const uint32_t u = (uint32_t)(s0[j] >> 32), v = (uint32_t)s0[j];
                el[o + j] = (int32_t)v;
                if (wl) wl[o + j] = int_weight(s2, u, v, wmax);
            }
            if (fl) {
                for (int64_t j = 0; j < n;) {
                    const uint32_t u = (uint32_t)(s0[j] >> 32);
                    int64_t j1 = j;
                    double sum = 0;
                    while (j1 < n && (uint32_t)(s0[j1] >> 32) == u) sum += unit_weight(s3, u, (uint32_t)s0[j1++]);
                    for (int64_t q = j; q < j1; ++q)
                        fl[o + q] = (float)(unit_weight(s3, u, (uint32_t)s0[q]) / sum);
                    j = j1;
                }
            }
        }
    });
    free(sorted);
    *o_vl = vl;
    *o_el = el;
    if (o_wl) *o_wl = wl; else free(wl);
    if (o_fl) *o_fl = fl; else free(fl);
    *o_E = E;
    return C2G_OK;
}

}  // namespace c2g

extern "C" int c2g_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                        uint64_t seed, int32_t symmetric, int32_t wmax, int32_t want_f,
                        int32_t** vlist, int32_t** elist, int32_t** wlist, float** flist,
                        int64_t* out_E) {
    if (!vlist || !elist || !out_E) return c2g::fail(C2G_EINVAL, "c2g_rmat: null output");
    if (wmax > 0 && !wlist) return c2g::fail(C2G_EINVAL, "c2g_rmat: wmax > 0 needs wlist");
    if (want_f && !flist) return c2g::fail(C2G_EINVAL, "c2g_rmat: want_f needs flist");
    return c2g::rmat_build(scale, edge_factor, a, b, c, seed, symmetric, wmax, want_f, vlist,
                           elist, wlist, flist, out_E);
}
