// Multi-label adsorption label propagation on gfx950 (FP32).
//
// The reference's "adsorption" kernels (src/adsorption/w-base/Main.cu:123-395,
// w-base-perk/Main.cu:270-593) are a single-source push with r[s]=10 and
// p[s]+=0.8*10 (w-base/Main.cu:149-152).  The target workload is the
// published multi-label recurrence (Baluja et al. 2008; Talukdar & Crammer
// 2009) over the same row-normalised flist weights:
//     Y0[v][l] = 1 for each seed (v,l)
//     Y_{t+1}[v] = p_inj * Y0[v] + p_cont * sum_{e=(v->u)} f(e) * Y_t[u]
// (the walk from v continues to u with probability p_cont*f(v,u), stops with
// p_inj emitting v's injected labels, or is abandoned with p_abnd).
//
// Layout: Y is vertex-major [V][L]; a wave holds 64/G vertex rows of G lanes
// (G = L rounded up to 16, 32 or 64), so gathering Y_t[u] for one edge is one
// coalesced 4*G-byte row.  Rows are pre-split into items of <= 256 edges once
// per graph; a row with one item is written directly, split rows accumulate
// with FP32 atomicAdd into the p_inj*Y0 image written by the init kernel.
#include "graph.h"
#include "wave.h"
#include <cstring>
#include <cmath>

namespace c2g {

constexpr int kAdsChunk = 256;

// Ynext = p_inj * Y0 for rows that are split across items (others are
// overwritten), plus the seed injections.
__global__ void ads_init_k(float* __restrict__ Y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        Y[i] = 0.0f;
}

// Rows split across items start from the injected image (their items add
// into it atomically); whole rows are written outright by their item.
__global__ void ads_split_init_k(float* __restrict__ Yn, const float* __restrict__ Yinj,
                                 const int32_t* __restrict__ splitv, int64_t nsplit, int L) {
    const int64_t n = nsplit * L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = (int64_t)splitv[i / L] * L + i % L;
        Yn[x] = Yinj[x];
    }
}

__global__ void ads_seed_k(float* __restrict__ Y, const int32_t* __restrict__ sv, const int32_t* __restrict__ sl,
                           int nseed, int L, float val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nseed) atomicAdd(&Y[(int64_t)sv[i] * L + sl[i]], val);
}

// items: int4(v, e_begin, e_end, split).  Y0 is given as the injected image
// (p_inj already applied) for split rows via the init of Ynext; whole rows
// read it from Yinj.
template <int G>
__global__ void __launch_bounds__(256)
ads_spmm_k(const int2* __restrict__ adjf, const int4* __restrict__ items, int64_t nitems,
           const float* __restrict__ Yt, float* __restrict__ Yn, const float* __restrict__ Yinj, int L,
           float p_cont) {
    constexpr int PER = 64 / G;
    const int lane = lane_id();
    const int sub = lane / G, l = lane % G;
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave * PER; base < nitems; base += nwaves * PER) {
        const int64_t i = base + sub;
        if (i >= nitems) continue;
        const int4 it = items[i];
        float acc = 0.0f;
        int e = it.y;
        // 8 row gathers in flight per lane group
        for (; e + 8 <= it.z; e += 8) {
            int2 a[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = adjf[e + k];
            float y[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) y[k] = l < L ? Yt[(int64_t)a[k].x * L + l] : 0.0f;
#pragma unroll
            for (int k = 0; k < 8; ++k) acc += __int_as_float(a[k].y) * y[k];
        }
        for (; e < it.z; ++e) {
            const int2 a = adjf[e];
            if (l < L) acc += __int_as_float(a.y) * Yt[(int64_t)a.x * L + l];
        }
        if (l < L) {
            const int64_t x = (int64_t)it.x * L + l;
            if (it.w) unsafeAtomicAdd(&Yn[x], p_cont * acc);
            else Yn[x] = Yinj[x] + p_cont * acc;
        }
    }
}

// The same product with 4 labels per lane (float4 rows, L % 4 == 0): a row
// of L labels takes G = L/4 lanes (rounded up to a power of two), so one wave
// gathers 64/G rows per instruction, 16 bytes per lane.  Items are sorted by
// edge count, so the rows of a wave run for about the same number of steps.
template <int G>
__global__ void __launch_bounds__(256)
ads_spmm4_k(const int2* __restrict__ adjf, const int4* __restrict__ items, int64_t nitems,
            const float4* __restrict__ Yt, float4* __restrict__ Yn, const float4* __restrict__ Yinj, int L4,
            float p_cont) {
    constexpr int PER = 64 / G;
    const int lane = lane_id();
    const int sub = lane / G, l = lane % G;
    const bool on = l < L4;
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t base = wave * PER; base < nitems; base += nwaves * PER) {
        const int64_t i = base + sub;
        if (i >= nitems) continue;
        const int4 it = items[i];
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        int e = it.y;
        for (; e + 8 <= it.z; e += 8) {
            int2 a[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = adjf[e + k];
            float4 y[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) y[k] = on ? Yt[(int64_t)a[k].x * L4 + l] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float f = __int_as_float(a[k].y);
                acc.x += f * y[k].x;
                acc.y += f * y[k].y;
                acc.z += f * y[k].z;
                acc.w += f * y[k].w;
            }
        }
        for (; e < it.z; ++e) {
            const int2 a = adjf[e];
            if (on) {
                const float f = __int_as_float(a.y);
                const float4 y = Yt[(int64_t)a.x * L4 + l];
                acc.x += f * y.x;
                acc.y += f * y.y;
                acc.z += f * y.z;
                acc.w += f * y.w;
            }
        }
        if (on) {
            const int64_t x = (int64_t)it.x * L4 + l;
            if (it.w) {
                float* yn = reinterpret_cast<float*>(&Yn[x]);
                unsafeAtomicAdd(&yn[0], p_cont * acc.x);
                unsafeAtomicAdd(&yn[1], p_cont * acc.y);
                unsafeAtomicAdd(&yn[2], p_cont * acc.z);
                unsafeAtomicAdd(&yn[3], p_cont * acc.w);
            } else {
                const float4 b = Yinj[x];
                Yn[x] = make_float4(b.x + p_cont * acc.x, b.y + p_cont * acc.y, b.z + p_cont * acc.z,
                                    b.w + p_cont * acc.w);
            }
        }
    }
}

// out[v][l] = Y[v][c0 + l] for l < L: one query's label block of a batched run.
// With inv set the run used relabelled vertices: row v of the result is row
// inv[v] of Y.
__global__ void ads_cols_k(const float* __restrict__ Y, int Lt, int c0, int L, float* __restrict__ out, int64_t V,
                           const int32_t* __restrict__ inv) {
    const int64_t n = V * L;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i / L;
        const int64_t src = inv ? (int64_t)inv[v] : v;
        out[i] = Y[src * Lt + c0 + (i - v * L)];
    }
}

struct AdsState {
    int4* items = nullptr;
    int64_t nitems = 0;
    float* Y[3] = {nullptr, nullptr, nullptr};  // two iterates + injected image
    int64_t cap = 0;
    int32_t* seeds = nullptr;
    int64_t seedcap = 0;
    int32_t* splitv = nullptr;  // vertices whose row is split across items
    int64_t nsplit = 0;
    HostStage out;              // result copies to the host
    // Relabelled copy of the weighted CSR: vertices renumbered by descending
    // in-degree, so the rows gathered most often share cache lines.
    int relabeled = -1;         // layout the items were built for (-1: none yet)
    int2* radjf = nullptr;      // [E] (new target id, bits of f) in new row order
    int32_t* inv = nullptr;     // [V] old id -> new id (device)
    std::vector<int32_t> h_inv; // the same on the host (seed mapping)
    std::vector<int32_t> h_row; // [V+1] row offsets of the relabelled CSR
};

static std::mutex g_as_mu;
static std::vector<std::pair<const Graph*, AdsState*>> g_as;

static AdsState* astate(Graph* g) {
    std::lock_guard<std::mutex> lk(g_as_mu);
    for (auto& p : g_as)
        if (p.first == g) return p.second;
    auto* s = new AdsState;
    g_as.emplace_back(g, s);
    return s;
}

void ads_release(Graph* g) {
    std::lock_guard<std::mutex> lk(g_as_mu);
    for (size_t i = 0; i < g_as.size(); ++i)
        if (g_as[i].first == g) {
            AdsState* s = g_as[i].second;
            if (s->items) (void)hipFree(s->items);
            for (auto* y : s->Y) if (y) (void)hipFree(y);
            if (s->seeds) (void)hipFree(s->seeds);
            if (s->splitv) (void)hipFree(s->splitv);
            if (s->radjf) (void)hipFree(s->radjf);
            if (s->inv) (void)hipFree(s->inv);
            s->out.release();
            delete s;
            g_as.erase(g_as.begin() + i);
            return;
        }
}

// Vertices renumbered by descending in-degree (ties by id): new vertex n is
// old vertex perm[n]; its row is the old row in the old edge order with
// targets renamed, so every row sums the same terms in the same order.
static int ads_build_relabel(Graph* g, AdsState* s) {
    const int64_t V = g->V, E = g->E;
    std::vector<int64_t> indeg((size_t)V, 0);
    for (int64_t e = 0; e < E; ++e) indeg[(size_t)g->el[e]] += 1;
    std::vector<int32_t> perm((size_t)V);
    for (int64_t v = 0; v < V; ++v) perm[v] = (int32_t)v;
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return indeg[a] > indeg[b]; });
    s->h_inv.assign((size_t)V, 0);
    for (int64_t n = 0; n < V; ++n) s->h_inv[(size_t)perm[n]] = (int32_t)n;
    s->h_row.assign((size_t)V + 1, 0);
    for (int64_t n = 0; n < V; ++n) s->h_row[n + 1] = s->h_row[n] + (g->vl[perm[n] + 1] - g->vl[perm[n]]);
    std::vector<int2> radj((size_t)std::max<int64_t>(E, 1));
    parallel_for(V, [&](int64_t lo, int64_t hi, int) {
        for (int64_t n = lo; n < hi; ++n) {
            const int32_t v = perm[n];
            int64_t o = s->h_row[n];
            for (int32_t e = g->vl[v]; e < g->vl[v + 1]; ++e, ++o) {
                int32_t fb;
                memcpy(&fb, &g->fl[e], sizeof fb);
                radj[(size_t)o] = make_int2(s->h_inv[(size_t)g->el[e]], fb);
            }
        }
    });
    // both arrays complete before s->radjf marks them ready
    int2* ra = nullptr;
    int32_t* iv = nullptr;
    hipError_t e = hipMalloc(&ra, sizeof(int2) * radj.size());
    if (e == hipSuccess) e = hipMemcpy(ra, radj.data(), sizeof(int2) * radj.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&iv, sizeof(int32_t) * (size_t)std::max<int64_t>(V, 1));
    if (e == hipSuccess && V) e = hipMemcpy(iv, s->h_inv.data(), sizeof(int32_t) * V, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (ra) (void)hipFree(ra);
        if (iv) (void)hipFree(iv);
        return hip_fail(e, "adsorption relabelled CSR");
    }
    s->inv = iv;
    s->radjf = ra;
    return C2G_OK;
}

static int ads_prepare(Graph* g, AdsState* s, int L, int nseed) {
    HIP_TRY(hipSetDevice(g->device));
    const int want = g->tune.ads_relabel ? 1 : 0;
    if (want && !s->radjf) {
        int rc = ads_build_relabel(g, s);
        if (rc) return rc;
    }
    if (s->relabeled != want) {
        if (s->items) (void)hipFree(s->items);
        if (s->splitv) (void)hipFree(s->splitv);
        s->items = nullptr;
        s->splitv = nullptr;
        s->relabeled = -1;  // no items until the new ones are in place
        const int32_t* row = want ? s->h_row.data() : g->vl.data();
        std::vector<int4> it;
        std::vector<int32_t> split;
        it.reserve((size_t)g->V + g->E / kAdsChunk + 1);
        for (int32_t v = 0; v < g->V; ++v) {
            const int r0 = row[v], r1 = row[v + 1];
            if (r1 - r0 <= kAdsChunk) {
                it.push_back(make_int4(v, r0, r1, 0));  // whole row (also deg 0: writes p_inj*Y0)
            } else {
                split.push_back(v);
                for (int e = r0; e < r1; e += kAdsChunk) it.push_back(make_int4(v, e, std::min(e + kAdsChunk, r1), 1));
            }
        }
        // Degree binning: longest items first, so each wave's rows take about
        // the same number of steps (stable: equal lengths keep vertex order).
        std::stable_sort(it.begin(), it.end(), [](const int4& a, const int4& b) { return a.z - a.y > b.z - b.y; });
        s->nitems = (int64_t)it.size();
        HIP_TRY(hipMalloc(&s->items, sizeof(int4) * std::max<int64_t>(s->nitems, 1)));
        HIP_TRY(hipMemcpy(s->items, it.data(), sizeof(int4) * s->nitems, hipMemcpyHostToDevice));
        s->nsplit = (int64_t)split.size();
        HIP_TRY(hipMalloc(&s->splitv, sizeof(int32_t) * std::max<int64_t>(s->nsplit, 1)));
        if (s->nsplit)
            HIP_TRY(hipMemcpy(s->splitv, split.data(), sizeof(int32_t) * s->nsplit, hipMemcpyHostToDevice));
        s->relabeled = want;
    }
    const int64_t need = (int64_t)std::max(g->V, 1) * L;
    if (s->cap < need) {
        for (auto*& y : s->Y) { if (y) (void)hipFree(y); y = nullptr; }
        s->cap = 0;  // empty until all three are in place
        for (auto*& y : s->Y) HIP_TRY(hipMalloc(&y, sizeof(float) * need));
        s->cap = need;
    }
    if (s->seedcap < 2 * (int64_t)nseed) {
        if (s->seeds) (void)hipFree(s->seeds);
        s->seeds = nullptr;
        s->seedcap = 0;
        HIP_TRY(hipMalloc(&s->seeds, sizeof(int32_t) * 2 * std::max(nseed, 1)));
        s->seedcap = 2 * (int64_t)std::max(nseed, 1);
    }
    return C2G_OK;
}

// Device bytes the server's adsorption lane (three V x `labels` iterates, the
// work items and, with ads_relabel, the relabelled CSR copy) would add to
// what the handle holds.
int64_t ads_lane_ws_bytes(Graph* g, int labels) {
    AdsState* s = astate(g);
    const int64_t V = std::max<int64_t>(g->V, 1);
    int64_t b = 0;
    if (s->cap < V * labels) b += 3 * 4 * V * labels;
    if (!s->items) b += 16 * (V + g->E / kAdsChunk + 1) + 4 * V;
    if (g->tune.ads_relabel && !s->radjf) b += 8 * g->E + 4 * V;
    b += 8 * (V / 100 + 1) * (64 / std::max(1, labels));  // seed lists of one batched pass
    return b;
}

// Workgroups per CU of the adsorption kernels: gpc > 0 overrides the graph's
// grid_per_cu (the server runs this lane below full occupancy so that the
// latency-bound SSSP and PPR lanes keep room on every CU).
static int ads_grid(const Graph* g, int gpc) {
    return std::max(1, g->num_cus) * (gpc > 0 ? gpc : g->tune.grid_per_cu);
}

template <int G>
static void launch_spmm(Graph* g, hipStream_t strm, AdsState* s, const float* Yt, float* Yn, const float* Yinj,
                        int L, float pc, int gpc) {
    const int grid = ads_grid(g, gpc);
    hipLaunchKernelGGL(ads_spmm_k<G>, dim3(grid), dim3(256), 0, strm, s->relabeled == 1 ? s->radjf : g->d_adjf,
                       s->items, s->nitems, Yt, Yn,
                       Yinj, L, pc);
}

template <int G>
static void launch_spmm4(Graph* g, hipStream_t strm, AdsState* s, const float* Yt, float* Yn, const float* Yinj,
                         int L, float pc, int gpc) {
    const int grid = ads_grid(g, gpc);
    hipLaunchKernelGGL(ads_spmm4_k<G>, dim3(grid), dim3(256), 0, strm, s->relabeled == 1 ? s->radjf : g->d_adjf,
                       s->items, s->nitems,
                       (const float4*)Yt, (float4*)Yn, (const float4*)Yinj, L / 4, pc);
}

static int ads_core(Graph* g, hipStream_t strm, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L,
                    float p_inj, float p_cont, int32_t iters, int* cur_out, c2g_stats* st, int gpc);

static int ads_check(Graph* g, int32_t L, float p_inj, float p_cont, float p_abnd, int32_t iters) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adjf) return fail(C2G_ENOTSUP, "graph has no float weights (flist) for adsorption");
    if (L < 1 || L > 64) return fail(C2G_EINVAL, "labels must be in [1,64]");
    if (iters < 0) return fail(C2G_EINVAL, "iters must be >= 0");
    if (p_inj < 0 || p_cont < 0 || p_abnd < 0 || std::fabs(p_inj + p_cont + p_abnd - 1.0f) > 1e-5f)
        return fail(C2G_EINVAL, "p_inj, p_cont, p_abnd must be >= 0 and sum to 1");
    return C2G_OK;
}

// k queries of L labels each as one run over k*L labels (columns of the
// recurrence never mix); query j's labels are columns [j*L, (j+1)*L) and its
// result is copied out of that block into y_outs[j].
int ads_run_batch(Graph* g, hipStream_t strm, int k, const std::vector<int32_t>* svs,
                  const std::vector<int32_t>* sls, int32_t L, float p_inj, float p_cont, float p_abnd,
                  int32_t iters, float* const* y_outs, int gpc) {
    if (k < 1 || k * L > 64) return fail(C2G_EINVAL, "batched adsorption needs 1 <= k and k*L <= 64");
    int rc = ads_check(g, k * L, p_inj, p_cont, p_abnd, iters);
    if (rc) return rc;
    std::vector<int32_t> sv, sl;
    for (int j = 0; j < k; ++j)
        for (size_t i = 0; i < svs[j].size(); ++i) {
            if (svs[j][i] < 0 || svs[j][i] >= g->V || sls[j][i] < 0 || sls[j][i] >= L)
                return fail(C2G_EINVAL, "seed (%d,%d) out of range", svs[j][i], sls[j][i]);
            sv.push_back(svs[j][i]);
            sl.push_back(sls[j][i] + j * L);
        }
    int cur = 0;
    if ((rc = ads_core(g, strm, sv.data(), sl.data(), (int32_t)sv.size(), k * L, p_inj, p_cont, iters, &cur, nullptr,
                       gpc)))
        return rc;
    AdsState* s = astate(g);
    const int64_t V = g->V;
    const int grid = ads_grid(g, gpc);
    const int32_t* inv = s->relabeled == 1 ? s->inv : nullptr;
    for (int j = 0; j < k; ++j) {
        if (k == 1 && !inv) {
            if ((rc = s->out.copy_out(y_outs[0], s->Y[cur], sizeof(float) * V * L, strm))) return rc;
            break;
        }
        // the injected image (Y[2]) is free once the iterations are done
        hipLaunchKernelGGL(ads_cols_k, dim3(grid), dim3(256), 0, strm, s->Y[cur], k * L, j * L, L, s->Y[2], V, inv);
        HIP_TRY(hipGetLastError());
        if ((rc = s->out.copy_out(y_outs[j], s->Y[2], sizeof(float) * V * L, strm))) return rc;
    }
    return C2G_OK;
}

int ads_run(Graph* g, hipStream_t strm, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L, float p_inj,
                   float p_cont, float p_abnd, int32_t iters, float* y_out, c2g_stats* st) {
    int rc = ads_check(g, L, p_inj, p_cont, p_abnd, iters);
    if (rc) return rc;
    for (int32_t i = 0; i < nseed; ++i)
        if (sv[i] < 0 || sv[i] >= g->V || sl[i] < 0 || sl[i] >= L)
            return fail(C2G_EINVAL, "seed %d = (%d,%d) out of range", i, sv[i], sl[i]);
    int cur = 0;
    if ((rc = ads_core(g, strm, sv, sl, nseed, L, p_inj, p_cont, iters, &cur, st, 0))) return rc;
    AdsState* s = astate(g);
    if (!y_out) return C2G_OK;
    const float* Y = s->Y[cur];
    if (s->relabeled == 1) {  // rows back to the caller's vertex ids, via the free injected image
        hipLaunchKernelGGL(ads_cols_k, dim3(ads_grid(g, 0)), dim3(256), 0, strm, s->Y[cur], L, 0, L, s->Y[2],
                           (int64_t)g->V, s->inv);
        HIP_TRY(hipGetLastError());
        Y = s->Y[2];
    }
    if ((rc = s->out.copy_out(y_out, Y, sizeof(float) * (int64_t)g->V * L, strm))) return rc;
    return C2G_OK;
}

// The recurrence on the device; the result stays in Y[*cur_out].
static int ads_core(Graph* g, hipStream_t strm, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L,
                    float p_inj, float p_cont, int32_t iters, int* cur_out, c2g_stats* st, int gpc) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    AdsState* s = astate(g);
    int rc = ads_prepare(g, s, L, nseed);
    if (rc) return rc;
    const int64_t n = (int64_t)g->V * L;
    const int grid = ads_grid(g, gpc);
    const double t0 = now_ms();
    float* Yinj = s->Y[2];
    // Yinj = p_inj * Y0 ; Y_0 = Y0
    hipLaunchKernelGGL(ads_init_k, dim3(grid), dim3(256), 0, strm, Yinj, n);
    hipLaunchKernelGGL(ads_init_k, dim3(grid), dim3(256), 0, strm, s->Y[0], n);
    if (nseed > 0) {
        // duplicate (v,l) seeds count once: Y0 is an indicator
        std::vector<std::pair<int32_t, int32_t>> u;
        u.reserve(nseed);
        for (int32_t i = 0; i < nseed; ++i) u.emplace_back(s->relabeled == 1 ? s->h_inv[sv[i]] : sv[i], sl[i]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        std::vector<int32_t> hv(u.size()), hl(u.size());
        for (size_t i = 0; i < u.size(); ++i) { hv[i] = u[i].first; hl[i] = u[i].second; }
        const int m = (int)u.size();
        HIP_TRY(hipMemcpyAsync(s->seeds, hv.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, strm));
        HIP_TRY(hipMemcpyAsync(s->seeds + nseed, hl.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, strm));
        HIP_TRY(hipStreamSynchronize(strm));  // hv/hl are pageable host vectors
        hipLaunchKernelGGL(ads_seed_k, dim3((m + 255) / 256), dim3(256), 0, strm, Yinj, s->seeds,
                           s->seeds + nseed, m, L, p_inj);
        hipLaunchKernelGGL(ads_seed_k, dim3((m + 255) / 256), dim3(256), 0, strm, s->Y[0], s->seeds,
                           s->seeds + nseed, m, L, 1.0f);
    }
    int cur = 0;
    for (int t = 0; t < iters; ++t) {
        float* Yt = s->Y[cur];
        float* Yn = s->Y[cur ^ 1];
        // split rows accumulate atomically into p_inj*Y0
        if (s->nsplit)
            hipLaunchKernelGGL(ads_split_init_k, dim3((int)std::min<int64_t>(grid, (s->nsplit * L + 255) / 256)),
                               dim3(256), 0, strm, Yn, Yinj, s->splitv, s->nsplit, L);
        if (L % 4 == 0) {  // 16-byte rows per lane
            if (L <= 4) launch_spmm4<1>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
            else if (L <= 8) launch_spmm4<2>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
            else if (L <= 16) launch_spmm4<4>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
            else if (L <= 32) launch_spmm4<8>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
            else launch_spmm4<16>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
        } else if (L <= 16) launch_spmm<16>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
        else if (L <= 32) launch_spmm<32>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
        else launch_spmm<64>(g, strm, s, Yt, Yn, Yinj, L, p_cont, gpc);
        cur ^= 1;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(strm));
    const double t1 = now_ms();
    *cur_out = cur;
    if (st) {
        st->lane = C2G_LANE_GPU;
        st->iterations = iters;
        st->edges_relaxed = (int64_t)iters * g->E;
        st->items = (int64_t)iters * s->nitems;
        st->ms = t1 - t0;
    }
    return C2G_OK;
}

}  // namespace c2g

using namespace c2g;

extern "C" int c2g_adsorption(c2g_graph* h, const int32_t* seed_v, const int32_t* seed_label, int32_t nseed,
                              int32_t L, float p_inj, float p_cont, float p_abnd, int32_t iters, float* y_out,
                              c2g_stats* st) {
    if (!h || (nseed > 0 && (!seed_v || !seed_label)) || nseed < 0) return fail(C2G_EINVAL, "bad arguments");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return ads_run(&h->g, h->g.stream, seed_v, seed_label, nseed, L, p_inj, p_cont, p_abnd, iters, y_out, st);
}
