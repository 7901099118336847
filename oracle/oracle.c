/*
 * oracle.c — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's traversal arithmetic, used as the checker
 * by tests/, by __graft_entry__.smoke() and as the CPU baseline leg of
 * bench.py.  Nothing in the product library links, loads or calls this file.
 *
 * Functions and what pins each one (the tests named here live in tests/):
 *   orc_dijkstra       exact int32 SSSP (binary-heap Dijkstra, int64 inside).
 *                      The reference's CalcuSSSP (src/sssp/base/Main.cu:109-356)
 *                      converges to the same unique shortest-path distances.
 *                      Pinned by: scipy.sparse.csgraph.dijkstra on random graphs,
 *                      brute-force Bellman-Ford on tiny graphs, hand graphs with
 *                      closed-form distances, and the edge invariants
 *                      (test_oracle.py).
 *   orc_bellman_ford   restatement of the reference CPU lane Bellman_ford
 *                      (src/sssp/base-perk/SSSP.cpp:88-119, the variant that
 *                      resets its flag).  Pinned by equality with orc_dijkstra
 *                      from {s:0},{s} (test_oracle.py).
 *   orc_ppr_power      FP64 power iteration x <- alpha*e_s + (1-alpha) P^T x with
 *                      dangling rows dropped (the reference drops dangling mass:
 *                      src/ppr/base/Main.cu:216-221).  Pinned by a scipy sparse
 *                      direct solve and the directed-cycle closed form.
 *   orc_forward_push   restatement of forward_push (src/ppr/w-base-perk/PPR.cpp:
 *                      178-236) with the degree-scaled threshold r > eps*deg and
 *                      a zero-initialised flag array; seeds r[s]=1, p[s]=alpha as
 *                      src/ppr/base/Main.cu:146-150.  Pinned by the push
 *                      certificate |p - pi|_1 <= sum(r) against orc_ppr_power.
 *   orc_adsorption     FP64 Jacobi of Y_{t+1} = p_inj*Y0 + p_cont*W*Y_t.  Pinned
 *                      by a dense numpy matrix-power evaluation on small graphs.
 *   orc_sssp_certificate  exactness check of an int32 distance vector for
 *                      weights >= 1 (Graph500-style): d[s] = 0; d[v] <= d[u]+w
 *                      on every edge out of a reached u; every reached v != s
 *                      has a tight in-edge from a reached u; unreached vertices
 *                      hold INT32_MAX.  Tight in-edges strictly decrease d, so
 *                      they lead back to s and every finite d[v] is a real path
 *                      length; feasibility makes it no longer than the shortest
 *                      one.  Pinned by: Dijkstra outputs pass, and each single
 *                      perturbation (+1, -1, an unreached vertex marked reached,
 *                      a reached one marked unreached, d[s] != 0) fails with
 *                      its own code (test_oracle.py).
 *   orc_adsorption_push  restatement of the reference's adsorption kernel
 *                      (src/adsorption/w-base/Main.cu: alpha 0.2, r[s] = 10,
 *                      p[s] = 0.8*r[s], constant threshold 1/V).  Pinned by the
 *                      push invariant against orc_ppr_power: with
 *                      p' = p - (0.8-alpha)*10*e_s, |10*pi_s - p'|_1 = sum(r)
 *                      on graphs without dangling vertices, and r <= 1/V.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <math.h>

/* ---------------------------------------------------------------- Dijkstra */
/* Indexed binary min-heap over vertices keyed by int64 distance. */
typedef struct {
    int32_t* heap;  /* heap[i] = vertex   */
    int32_t* pos;   /* pos[v]  = index or -1 */
    int64_t* key;
    int32_t n;
} heap_t;

static void h_up(heap_t* h, int32_t i) {
    int32_t v = h->heap[i];
    int64_t k = h->key[v];
    while (i > 0) {
        int32_t p = (i - 1) >> 1;
        int32_t pv = h->heap[p];
        if (h->key[pv] <= k) break;
        h->heap[i] = pv;
        h->pos[pv] = i;
        i = p;
    }
    h->heap[i] = v;
    h->pos[v] = i;
}
static void h_down(heap_t* h, int32_t i) {
    int32_t v = h->heap[i];
    int64_t k = h->key[v];
    for (;;) {
        int32_t l = 2 * i + 1;
        if (l >= h->n) break;
        int32_t r = l + 1, c = l;
        if (r < h->n && h->key[h->heap[r]] < h->key[h->heap[l]]) c = r;
        if (h->key[h->heap[c]] >= k) break;
        h->heap[i] = h->heap[c];
        h->pos[h->heap[i]] = i;
        i = c;
    }
    h->heap[i] = v;
    h->pos[v] = i;
}

/* dist_out[v] = shortest distance from src, INT32_MAX if unreachable.
 * Returns 0, or -1 if a finite distance does not fit below INT32_MAX. */
int orc_dijkstra(int32_t V, const int32_t* vl, const int32_t* el, const int32_t* wl,
                 int32_t src, int32_t* dist_out) {
    if (src < 0 || src >= V) return -2;
    heap_t h;
    h.heap = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    h.pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    h.key = (int64_t*)malloc(sizeof(int64_t) * (size_t)V);
    if (!h.heap || !h.pos || !h.key) { free(h.heap); free(h.pos); free(h.key); return -3; }
    for (int32_t v = 0; v < V; ++v) { h.key[v] = INT64_MAX; h.pos[v] = -1; }
    h.n = 0;
    h.key[src] = 0;
    h.heap[0] = src; h.pos[src] = 0; h.n = 1;
    int overflow = 0;
    while (h.n > 0) {
        int32_t u = h.heap[0];
        h.pos[u] = -2;                 /* settled */
        h.n--;
        if (h.n > 0) { h.heap[0] = h.heap[h.n]; h.pos[h.heap[0]] = 0; h_down(&h, 0); }
        int64_t du = h.key[u];
        for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
            int32_t v = el[e];
            if (h.pos[v] == -2) continue;
            int64_t c = du + (int64_t)wl[e];
            if (c < h.key[v]) {
                h.key[v] = c;
                if (h.pos[v] < 0) { h.heap[h.n] = v; h.pos[v] = h.n; h.n++; }
                h_up(&h, h.pos[v]);
            }
        }
    }
    for (int32_t v = 0; v < V; ++v) {
        int64_t k = h.key[v];
        if (k == INT64_MAX) dist_out[v] = INT32_MAX;
        else if (k >= INT32_MAX) { dist_out[v] = INT32_MAX; overflow = 1; }
        else dist_out[v] = (int32_t)k;
    }
    free(h.heap); free(h.pos); free(h.key);
    return overflow ? -1 : 0;
}

/* Returns 0 when d is the exact distance vector from src (weights >= 1), else
 * -1 (d[src] != 0 or src out of range), -2 (an edge out of a reached vertex
 * violates d[v] <= d[u] + w), -3 (a reached v != src without a tight in-edge),
 * -4 (a weight below 1), -5 (no memory).  Vertices are split over OpenMP
 * threads; `tight` marks are idempotent byte stores. */
int orc_sssp_certificate(int32_t V, const int32_t* vl, const int32_t* el, const int32_t* wl,
                         int32_t src, const int32_t* d) {
    if (src < 0 || src >= V || d[src] != 0) return -1;
    uint8_t* tight = (uint8_t*)calloc((size_t)V, 1);
    if (!tight) return -5;
    int bad_w = 0, bad_f = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(|: bad_w, bad_f)
    for (int32_t u = 0; u < V; ++u) {
        if (d[u] == INT32_MAX) continue;
        const int64_t du = d[u];
        for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
            const int32_t v = el[e];
            if (wl[e] < 1) { bad_w = 1; continue; }
            const int64_t c = du + (int64_t)wl[e];
            const int64_t dv = d[v];  /* INT32_MAX for an unreached v */
            if (dv > c) bad_f = 1;
            else if (dv == c) {
#pragma omp atomic write
                tight[v] = 1;
            }
        }
    }
    int rc = 0;
    if (bad_w) rc = -4;
    else if (bad_f) rc = -2;
    else
        for (int32_t v = 0; v < V; ++v)
            if (v != src && d[v] != INT32_MAX && !tight[v]) { rc = -3; break; }
    free(tight);
    return rc;
}

/* ---------------------------------------------------------- Bellman-Ford lane
 * src/sssp/base-perk/SSSP.cpp:88-119: serial frontier relaxation; a vertex is
 * appended to the next frontier the first time in a round that it improves;
 * the flag array is reset after each round.  Distances in place; `frontier`
 * holds `count` start vertices.  Returns the number of rounds.  Sums are done
 * in int64 and saturate at INT32_MAX (unreachable). */
int orc_bellman_ford(int32_t V, const int32_t* vl, const int32_t* el, const int32_t* wl,
                     int32_t* dist, const int32_t* frontier, int32_t count,
                     int64_t* relax_out) {
    if (count <= 0) return 0;
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    int32_t* nxt = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    uint8_t* flag = (uint8_t*)calloc((size_t)V, 1);
    int32_t nc = count, nn = 0, rounds = 0;
    int64_t relax = 0;
    memcpy(cur, frontier, sizeof(int32_t) * (size_t)count);
    while (nc > 0) {
        rounds++;
        nn = 0;
        for (int32_t i = 0; i < nc; ++i) {
            int32_t u = cur[i];
            if (vl[u + 1] - vl[u] <= 0) continue;
            int64_t du = dist[u];
            for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
                int32_t v = el[e];
                int64_t c = du + (int64_t)wl[e];
                if (c > INT32_MAX) c = INT32_MAX;
                relax++;
                if (!flag[v] && (int64_t)dist[v] > c) { nxt[nn++] = v; flag[v] = 1; }
                if (c < dist[v]) dist[v] = (int32_t)c;
            }
        }
        for (int32_t i = 0; i < nn; ++i) flag[nxt[i]] = 0;
        int32_t* t = cur; cur = nxt; nxt = t;
        nc = nn;
    }
    free(cur); free(nxt); free(flag);
    if (relax_out) *relax_out = relax;
    return rounds;
}

/* ----------------------------------------------------------- PPR (FP64)
 * x <- alpha*e_s + (1-alpha) * P^T x, P[u][v] = f(u,v), dangling rows zero.
 * Iterates until the L1 change is below tol or max_iter.  Returns iterations. */
int orc_ppr_power(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                  int32_t src, double alpha, double tol, int32_t max_iter, double* x) {
    /* pull form y[v] = alpha*[v==s] + (1-alpha) * sum_{u->v} f(u,v) x[u]
     * over the in-edges (built once), so threads split the rows v */
    const int64_t E = vl[V];
    int64_t* tv = (int64_t*)calloc((size_t)V + 1, sizeof(int64_t));
    int32_t* tu = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    double* tf = (double*)malloc(sizeof(double) * (size_t)(E > 0 ? E : 1));
    double* y = (double*)malloc(sizeof(double) * (size_t)V);
    for (int64_t e = 0; e < E; ++e) tv[el[e] + 1]++;
    for (int32_t v = 0; v < V; ++v) tv[v + 1] += tv[v];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(V > 0 ? V : 1));
    for (int32_t v = 0; v < V; ++v) fill[v] = tv[v];
    for (int32_t u = 0; u < V; ++u)
        for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
            const int64_t k = fill[el[e]]++;
            tu[k] = u;
            tf[k] = (double)fl[e];
        }
    free(fill);
    for (int32_t v = 0; v < V; ++v) x[v] = 0.0;
    x[src] = alpha;  /* first iterate from x=0 */
    int it = 1;
    for (; it < max_iter; ++it) {
        double d = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : d)
        for (int32_t v = 0; v < V; ++v) {
            double acc = 0.0;
            for (int64_t k = tv[v]; k < tv[v + 1]; ++k) acc += tf[k] * x[tu[k]];
            const double yv = (v == src ? alpha : 0.0) + (1.0 - alpha) * acc;
            d += fabs(yv - x[v]);
            y[v] = yv;
        }
        memcpy(x, y, sizeof(double) * (size_t)V);
        if (d < tol) { ++it; break; }
    }
    free(y); free(tv); free(tu); free(tf);
    return it;
}

/* ------------------------------------------------------ forward push lane
 * src/ppr/w-base-perk/PPR.cpp:178-236 with the degree-scaled threshold
 * r[v] > eps*outdeg(v) and a zero-initialised flag.  p, r updated in place
 * from the active list `frontier` (count entries); FP32 state like the
 * reference.  Returns the number of rounds. */
/* credit: the share of an activated residual credited to p (alpha for PPR). */
static int push_rounds(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                       float alpha, float eps, int degscale, float credit, float* p, float* r,
                       const int32_t* frontier, int32_t count, int32_t max_rounds,
                       int64_t* push_out) {
    if (count <= 0) return 0;
    float* msg = (float*)calloc((size_t)V, sizeof(float));
    uint8_t* flag = (uint8_t*)calloc((size_t)V, 1);
    int32_t* act = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    int32_t* tch = (int32_t*)malloc(sizeof(int32_t) * (size_t)V);
    int32_t na = count, nt = 0, rounds = 0;
    int64_t pushes = 0;
    memcpy(act, frontier, sizeof(int32_t) * (size_t)count);
    while (na > 0 && rounds < max_rounds) {
        rounds++;
        nt = 0;
        for (int32_t i = 0; i < na; ++i) {
            int32_t u = act[i];
            float ru = r[u];
            r[u] = 0.0f;
            if (vl[u + 1] - vl[u] <= 0) continue;
            float m = (1.0f - alpha) * ru;
            for (int32_t e = vl[u]; e < vl[u + 1]; ++e) {
                int32_t v = el[e];
                msg[v] += m * fl[e];
                pushes++;
                if (!flag[v]) { tch[nt++] = v; flag[v] = 1; }
            }
        }
        na = 0;
        for (int32_t i = 0; i < nt; ++i) {
            int32_t v = tch[i];
            r[v] += msg[v];
            msg[v] = 0.0f;
            flag[v] = 0;
            if (r[v] > (degscale ? (float)(vl[v + 1] - vl[v]) * eps : eps)) {
                p[v] += r[v] * credit;
                act[na++] = v;
            }
        }
    }
    free(msg); free(flag); free(act); free(tch);
    if (push_out) *push_out = pushes;
    return rounds;
}

int orc_forward_push(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                     float alpha, float eps, float* p, float* r,
                     const int32_t* frontier, int32_t count, int32_t max_rounds,
                     int64_t* push_out) {
    return push_rounds(V, vl, el, fl, alpha, eps, 1, alpha, p, r, frontier, count, max_rounds, push_out);
}

/* The reference's adsorption kernel (src/adsorption/w-base/Main.cu): alpha
 * 0.2 and rmax 1/V (:63-64), r[s] = 10 and p[s] += 0.8*r[s] (:149-152), then
 * rounds of push (1-alpha) r_u f(u,v) (:208-320) and activation r > rmax with
 * p += alpha*r (:351-355), the source active in the first round. */
int orc_adsorption_push(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                        int32_t src, float* p, float* r, int32_t max_rounds, int64_t* push_out) {
    const float alpha = 0.2f, rmax = 1.0f / (float)V;
    for (int32_t v = 0; v < V; ++v) { p[v] = 0.0f; r[v] = 0.0f; }
    r[src] = 10.0f;
    p[src] += 0.8f * r[src];
    return push_rounds(V, vl, el, fl, alpha, rmax, 0, alpha, p, r, &src, 1, max_rounds, push_out);
}

/* The reference's other adsorption variants.  mode 0 with rmax 1/(V+1):
 * adsorption/w-purn (w-base's push on the purn-ppr graph's live CSR, whose
 * hidden-list seeding is commented out, w-purn/Main.cu:67, :157-169,
 * :183-186).  mode 1: adsorption/w-base-perk, whose source gets
 * p[s] += alpha*r[s] (:306-309) and whose activation credits
 * p[v] += (1-alpha)*r[v] (:555-557), rmax 1/V (w-base-perk/Graph.h:117),
 * alpha 0.2 (PPR.cpp:20).  Pinned (test_oracle.py) by the push invariant
 * rewritten for the credit rule: with A = total residual activated,
 * |10*pi_s - alpha*A|_1 = sum(r), where A = p / credit after the source's
 * seeding credit is taken out. */
int orc_adsorption_push_mode(int32_t V, const int32_t* vl, const int32_t* el, const float* fl, int32_t src,
                             int32_t mode, float rmax, float* p, float* r, int32_t max_rounds, int64_t* push_out) {
    const float alpha = 0.2f;
    for (int32_t v = 0; v < V; ++v) { p[v] = 0.0f; r[v] = 0.0f; }
    r[src] = 10.0f;
    p[src] += (mode == 1 ? alpha : 0.8f) * r[src];
    return push_rounds(V, vl, el, fl, alpha, rmax, 0, mode == 1 ? 1.0f - alpha : alpha, p, r, &src, 1, max_rounds,
                       push_out);
}

/* Seeds r[s]=1, p[s]=alpha (src/ppr/base/Main.cu:146-150) then pushes. */
int orc_ppr_push(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                 int32_t src, float alpha, float eps, float* p, float* r,
                 int32_t max_rounds, int64_t* push_out) {
    for (int32_t v = 0; v < V; ++v) { p[v] = 0.0f; r[v] = 0.0f; }
    r[src] = 1.0f;
    p[src] = alpha;
    return orc_forward_push(V, vl, el, fl, alpha, eps, p, r, &src, 1, max_rounds, push_out);
}

/* ------------------------------------------------------ adsorption (FP64)
 * Y0[v][l] = 1 for each seed (v,l); Y_0 = Y0;
 * Y_{t+1}[v][l] = p_inj*Y0[v][l] + p_cont * sum_{e=(v->u)} f(e) * Y_t[u][l].
 * Y is vertex-major V x L.  */
void orc_adsorption(int32_t V, const int32_t* vl, const int32_t* el, const float* fl,
                    const int32_t* seed_v, const int32_t* seed_l, int32_t nseed, int32_t L,
                    double p_inj, double p_cont, int32_t iters, double* Y) {
    size_t n = (size_t)V * (size_t)L;
    double* Y0 = (double*)calloc(n, sizeof(double));
    double* Z = (double*)malloc(sizeof(double) * n);
    for (int32_t i = 0; i < nseed; ++i) Y0[(size_t)seed_v[i] * L + seed_l[i]] = 1.0;
    memcpy(Y, Y0, sizeof(double) * n);
    for (int32_t t = 0; t < iters; ++t) {
        /* rows are independent: threads split them, each row summed in edge order */
#pragma omp parallel for schedule(dynamic, 4096)
        for (int32_t v = 0; v < V; ++v) {
            double* z = Z + (size_t)v * L;
            for (int32_t l = 0; l < L; ++l) z[l] = 0.0;
            for (int32_t e = vl[v]; e < vl[v + 1]; ++e) {
                const double* y = Y + (size_t)el[e] * L;
                double f = (double)fl[e];
                for (int32_t l = 0; l < L; ++l) z[l] += f * y[l];
            }
            const double* y0 = Y0 + (size_t)v * L;
            for (int32_t l = 0; l < L; ++l) z[l] = p_inj * y0[l] + p_cont * z[l];
        }
        memcpy(Y, Z, sizeof(double) * n);
    }
    free(Y0); free(Z);
}
