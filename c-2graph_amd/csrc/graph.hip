// Graph handle: dataset loading, validation and the one-off upload to HBM.
//
// Loader behaviour follows the reference's Graph constructors
// (src/sssp/base/Graph.h:30-51, src/sssp/purn/Graph.h:33-48,
// src/ppr/base/Graph.h:31-43, src/ppr/purn/Graph.h:34-46): V = len(vlist)-1,
// E = len(elist).  Unlike them, every length is checked and a mismatch is an
// error code (the reference trusts the files).
#include "graph.h"
#include <cstring>

namespace c2g {

int64_t read_vec_i32(const std::string& p, std::vector<int32_t>& v);
int64_t read_vec_f32(const std::string& p, std::vector<float>& v);
bool file_exists(const std::string& p);

static int validate(Graph* g) {
    const int32_t V = g->V;
    if (V < 0) return fail(C2G_EFORMAT, "vlist is empty");
    if ((int64_t)g->vl.size() != (int64_t)V + 1) return fail(C2G_EFORMAT, "vlist length");
    if (g->vl[0] != 0) return fail(C2G_EFORMAT, "vlist[0] = %d, expected 0", g->vl[0]);
    for (int32_t i = 0; i < V; ++i)
        if (g->vl[i + 1] < g->vl[i]) return fail(C2G_EFORMAT, "vlist not monotone at %d", i);
    if ((int64_t)g->vl[V] != (int64_t)g->el.size())
        return fail(C2G_EFORMAT, "vlist[V] = %d but elist has %zu entries", g->vl[V], g->el.size());
    g->E = (int64_t)g->el.size();
    if (!g->wl.empty() && (int64_t)g->wl.size() != g->E)
        return fail(C2G_EFORMAT, "wlist has %zu entries, elist %lld", g->wl.size(), (long long)g->E);
    if (!g->fl.empty() && (int64_t)g->fl.size() != g->E)
        return fail(C2G_EFORMAT, "flist has %zu entries, elist %lld", g->fl.size(), (long long)g->E);
    for (int64_t e = 0; e < g->E; ++e)
        if ((uint32_t)g->el[e] >= (uint32_t)V) return fail(C2G_EFORMAT, "elist[%lld] = %d out of range", (long long)e, g->el[e]);
    for (int64_t e = 0; e < (int64_t)g->wl.size(); ++e)
        if (g->wl[e] < 0) return fail(C2G_EFORMAT, "negative weight at %lld", (long long)e);
    if (!g->hvl.empty()) {
        if ((int64_t)g->hvl.size() != (int64_t)V + 1)
            return fail(C2G_EFORMAT,
                        "hidden vlist has %zu entries, expected V+1 = %lld (a compacted, "
                        "prune-order hvlist is not supported; regenerate with purn-sssp of this library)",
                        g->hvl.size(), (long long)V + 1);
        if (g->hvl[0] != 0) return fail(C2G_EFORMAT, "hvlist[0] != 0");
        for (int32_t i = 0; i < V; ++i)
            if (g->hvl[i + 1] < g->hvl[i]) return fail(C2G_EFORMAT, "hvlist not monotone at %d", i);
        g->HE = g->hvl[V];
        if ((int64_t)g->hel.size() != g->HE) return fail(C2G_EFORMAT, "helist length %zu != hvlist[V] %lld", g->hel.size(), (long long)g->HE);
        for (int64_t e = 0; e < g->HE; ++e)
            if ((uint32_t)g->hel[e] >= (uint32_t)V) return fail(C2G_EFORMAT, "helist entry out of range");
        if (!g->hwl.empty() && (int64_t)g->hwl.size() != g->HE) return fail(C2G_EFORMAT, "hwlist length");
        if (!g->hwf.empty() && (int64_t)g->hwf.size() != g->HE) return fail(C2G_EFORMAT, "hwlist length");
    }
    return C2G_OK;
}

void default_tune(Graph* g) {
    c2g_tune& t = g->tune;
    memset(&t, 0, sizeof t);
    t.block = 256;
    t.grid_per_cu = 8;
    t.grid_per_cu_single = 2;
    t.ppr_pull = 200;
    t.ads_relabel = 1;
    t.chunk = 32;
    t.batch_chunk = 64;
    t.iters_per_graph = 8;
    t.use_graph = 1;
    t.delta = 0;
    t.light = default_light(g);
    t.batch_order = 1;
    t.ppr_grid_per_cu = 1;
    t.batch_relabel = 1;
}

// Light-edge bound: about a sixteenth of the mean weight (on R-MAT with
// uniform weights U[1,255] and U[1,1023] the measured optimum of batched SSSP
// lies at 1/16 to 1/21, of single-source SSSP at 1/8 to 1/21).  0 when almost
// no edge would be light (a vertex would then only wait a round).
int default_light(const Graph* g) {
    const int64_t E = (int64_t)g->wl.size();
    int light = 0;
    if (E > 0) {
        const int nt = std::max(1, std::min(64, host_threads()));
        std::vector<double> part(64, 0.0);
        parallel_for(E, [&](int64_t lo, int64_t hi, int tid) {
            double a = 0;
            for (int64_t i = lo; i < hi; ++i) a += (double)g->wl[i];
            part[tid] += a;
        }, nt, 1 << 20);
        double sum = 0;
        for (double x : part) sum += x;
        const int64_t L = std::min<int64_t>(std::max<int64_t>(1, std::llround(sum / E / 16.0)), INT32_MAX);
        std::vector<int64_t> cnt(64, 0);
        parallel_for(E, [&](int64_t lo, int64_t hi, int tid) {
            int64_t c = 0;
            for (int64_t i = lo; i < hi; ++i) c += g->wl[i] < L;
            cnt[tid] += c;
        }, nt, 1 << 20);
        int64_t nl = 0;
        for (int64_t x : cnt) nl += x;
        light = nl * 100 >= E ? (int)L : 0;
    }
    return light;
}

// Light-edge rows for tune.light (see Graph::d_lrow).  The light entries are
// packed like the live ones (bit 31 of the target: the target has edges) and
// appended after them in a new d_adj; the old buffer is freed.
int light_build(Graph* g) {
    const int L = g->tune.light;
    if (!g->on_device || g->wl.empty() || g->E == 0 || L <= 0) {
        g->light_built = 0;
        return C2G_OK;
    }
    if (L == g->light_built && g->d_lrow) return C2G_OK;
    HIP_TRY(hipSetDevice(g->device));
    const int64_t V = g->V, E = g->E;
    std::vector<int32_t> lr((size_t)V + 1);
    parallel_for(V, [&](int64_t lo, int64_t hi, int) {
        for (int64_t v = lo; v < hi; ++v) {
            int32_t c = 0;
            for (int32_t e = g->vl[v]; e < g->vl[v + 1]; ++e) c += g->wl[e] < L;
            lr[v + 1] = c;
        }
    }, 0, 1 << 14);
    lr[0] = 0;
    for (int64_t v = 0; v < V; ++v) lr[v + 1] += lr[v];
    const int64_t EL = lr[V];
    if (E + EL > INT32_MAX) {
        // edge offsets are int32: such a graph runs without the mode
        g->light_built = 0;
        return C2G_OK;
    }
    std::vector<int2> tail((size_t)std::max<int64_t>(EL, 1));
    parallel_for(V, [&](int64_t lo, int64_t hi, int) {
        for (int64_t v = lo; v < hi; ++v) {
            int64_t k = lr[v];
            for (int32_t e = g->vl[v]; e < g->vl[v + 1]; ++e) {
                if (g->wl[e] >= L) continue;
                const int32_t t = g->el[e];
                const bool nz = g->vl[t + 1] > g->vl[t] || (!g->hvl.empty() && g->hvl[t + 1] > g->hvl[t]);
                tail[k++] = make_int2(nz ? (int32_t)((uint32_t)t | 0x80000000u) : t, g->wl[e]);
            }
        }
    }, 0, 1 << 14);
    for (int64_t v = 0; v <= V; ++v) lr[v] = (int32_t)(lr[v] + E);
    int2* nb = nullptr;
    HIP_TRY(hipMalloc(&nb, sizeof(int2) * (size_t)(E + std::max<int64_t>(EL, 1))));
    hipError_t e1 = hipMemcpy(nb, g->d_adj, sizeof(int2) * (size_t)E, hipMemcpyDeviceToDevice);
    if (e1 == hipSuccess && EL > 0) e1 = hipMemcpy(nb + E, tail.data(), sizeof(int2) * EL, hipMemcpyHostToDevice);
    if (e1 == hipSuccess && !g->d_lrow) e1 = hipMalloc(&g->d_lrow, sizeof(int32_t) * (size_t)(V + 1));
    if (e1 == hipSuccess) e1 = hipMemcpy(g->d_lrow, lr.data(), sizeof(int32_t) * (size_t)(V + 1), hipMemcpyHostToDevice);
    if (e1 != hipSuccess) {
        (void)hipFree(nb);
        g->light_built = 0;
        return hip_fail(e1, "light rows");
    }
    // captured launch sequences hold the old d_adj
    if (g->sssp.exec) { (void)hipGraphExecDestroy(g->sssp.exec); g->sssp.exec = nullptr; }
    if (g->batch.exec) { (void)hipGraphExecDestroy(g->batch.exec); g->batch.exec = nullptr; }
    (void)hipFree(g->d_adj);
    g->d_adj = nb;
    g->EL = EL;
    g->light_built = L;
    return C2G_OK;
}

static int load_dir(Graph* g, const std::string& ds, int variant, const c2g_opts* o) {
    std::string d;
    int64_t n;
    if (variant == C2G_VARIANT_ORIGIN) d = ds + "/origin/";
    else if (variant == C2G_VARIANT_SSSP_PURN) d = ds + "/sssp/";
    else if (variant == C2G_VARIANT_PPR_PURN) d = ds + "/ppr/";
    else return fail(C2G_EINVAL, "unknown variant %d", variant);
    if ((n = read_vec_i32(d + "vlist.bin", g->vl)) < 0) return (int)n;
    if (n < 1) return fail(C2G_EFORMAT, "%svlist.bin is empty", d.c_str());
    g->V = (int32_t)(n - 1);
    if ((n = read_vec_i32(d + "elist.bin", g->el)) < 0) return (int)n;
    if (variant == C2G_VARIANT_PPR_PURN) {
        if ((n = read_vec_f32(d + "wlist.bin", g->fl)) < 0) return (int)n;
    } else {
        if (file_exists(d + "wlist.bin") && (n = read_vec_i32(d + "wlist.bin", g->wl)) < 0) return (int)n;
        if (o && o->load_flist && variant == C2G_VARIANT_ORIGIN) {
            if ((n = read_vec_f32(d + "flist.bin", g->fl)) < 0) return (int)n;
        }
    }
    if (variant != C2G_VARIANT_ORIGIN) {
        if ((n = read_vec_i32(d + "hvlist.bin", g->hvl)) < 0) return (int)n;
        if ((n = read_vec_i32(d + "helist.bin", g->hel)) < 0) return (int)n;
        if (variant == C2G_VARIANT_SSSP_PURN) {
            if ((n = read_vec_i32(d + "hwlist.bin", g->hwl)) < 0) return (int)n;
        } else {
            if ((n = read_vec_f32(d + "hwlist.bin", g->hwf)) < 0) return (int)n;
        }
    }
    g->variant = variant;
    return validate(g);
}

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int device_upload(Graph* g, int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(C2G_EDEVICE, "no HIP device available");
    if (device < 0) device = count - 1;  // the reference picks the last device (Main.cu:30)
    if (device >= count) return fail(C2G_EINVAL, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    g->device = device;
    g->num_cus = prop.multiProcessorCount;
    g->on_device = false;
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    int rc0 = sssp_scratch_init(device);  // this device's copy of the no-alloc scratch
    if (rc0) return rc0;
    const int64_t V = g->V, E = g->E;
    HIP_TRY(hipMalloc(&g->d_row, sizeof(int32_t) * (V + 1)));
    HIP_TRY(hipMemcpy(g->d_row, g->vl.data(), sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice));
    std::vector<int2> tmp;
    // nz[v] = v has live or hidden out-edges.  SSSP edge entries carry it in
    // bit 31 of the target id (vertex ids are < 2^31) so that kernels skip
    // activating dead-end targets without another load.
    std::vector<uint8_t> nz((size_t)V, 0);
    for (int64_t v = 0; v < V; ++v) {
        nz[v] = g->vl[v + 1] > g->vl[v];
        if (!g->hvl.empty() && g->hvl[v + 1] > g->hvl[v]) nz[v] = 1;
    }
    auto pack = [&](const std::vector<int32_t>& e, const void* w, int64_t n, int2** dst,
                    bool flag) -> int {
        HIP_TRY(hipMalloc(dst, sizeof(int2) * (size_t)std::max<int64_t>(n, 1)));
        const int64_t step = 1 << 24;  // pack and copy in 128 MiB slices
        tmp.resize((size_t)std::min<int64_t>(std::max<int64_t>(n, 1), step));
        for (int64_t o = 0; o < n; o += step) {
            int64_t m = std::min<int64_t>(step, n - o);
            const int32_t* ww = (const int32_t*)w;
            parallel_for(m, [&](int64_t lo, int64_t hi, int) {
                for (int64_t i = lo; i < hi; ++i) {
                    const int32_t t = e[o + i];
                    tmp[i] = make_int2(flag && nz[t] ? (int32_t)((uint32_t)t | 0x80000000u) : t, ww[o + i]);
                }
            });
            HIP_TRY(hipMemcpy(*dst + o, tmp.data(), sizeof(int2) * m, hipMemcpyHostToDevice));
        }
        return C2G_OK;
    };
    int rc;
    // An edgeless graph still gets (1-element) edge arrays so every query runs.
    if ((!g->wl.empty() || E == 0) && (rc = pack(g->el, g->wl.data(), E, &g->d_adj, true))) return rc;
    if ((!g->fl.empty() || E == 0) && (rc = pack(g->el, g->fl.data(), E, &g->d_adjf, false))) return rc;
    if (!g->hvl.empty()) {
        HIP_TRY(hipMalloc(&g->d_hrow, sizeof(int32_t) * (V + 1)));
        HIP_TRY(hipMemcpy(g->d_hrow, g->hvl.data(), sizeof(int32_t) * (V + 1), hipMemcpyHostToDevice));
        if (!g->hwl.empty() && (rc = pack(g->hel, g->hwl.data(), g->HE, &g->d_hadj, true))) return rc;
        if (!g->hwf.empty() && (rc = pack(g->hel, g->hwf.data(), g->HE, &g->d_hadjf, false))) return rc;
    }
    (void)align_up;
    g->on_device = true;
    return light_build(g);
}

template <class T>
static void dfree(T*& p) {
    if (p) { (void)hipFree((void*)p); p = nullptr; }
}

void device_release(Graph* g) {
    if (!g->stream && !g->d_row) {
        // host-only handle: per-graph state records a server's workspace
        // estimates may have created (they hold no device memory)
        batch_release(g);
        ppr_release(g);
        ads_release(g);
        serve_release(g);
        qserve_release(g);
        return;
    }
    (void)hipSetDevice(g->device);
    batch_release(g);
    ppr_release(g);
    ads_release(g);
    serve_release(g);
    qserve_release(g);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->sssp.exec) (void)hipGraphExecDestroy(g->sssp.exec);
    if (g->batch.exec) (void)hipGraphExecDestroy(g->batch.exec);
    dfree(g->sssp.dist); dfree(g->sssp.bits); dfree(g->sssp.q[0]);
    dfree(g->sssp.ctrl); dfree(g->sssp.hid);
    if (g->sssp.h_ctrl) (void)hipHostFree(g->sssp.h_ctrl);
    g->sssp.out.release();
    for (auto& e : g->sssp.ev) if (e) (void)hipEventDestroy(e);
    dfree(g->batch.dist); dfree(g->batch.mask[0]); dfree(g->batch.hmask); dfree(g->batch.ph); dfree(g->batch.ctrl);
    if (g->batch.h_ctrl) (void)hipHostFree(g->batch.h_ctrl);
    dfree(g->d_row); dfree(g->d_adj); dfree(g->d_adjf); dfree(g->d_hrow); dfree(g->d_hadj);
    dfree(g->d_hadjf);
    dfree(g->d_split);
    dfree(g->d_hsplit);
    dfree(g->d_lrow);
    g->light_built = 0;
    if (g->stream) (void)hipStreamDestroy(g->stream);
    g->stream = nullptr;
    g->on_device = false;
}

static int finish_open(c2g_graph* h, const c2g_opts* o, c2g_graph** out) {
    default_tune(&h->g);
    if (!(o && o->no_device)) {
        int rc = device_upload(&h->g, o ? o->device : -1);
        if (rc) { device_release(&h->g); delete h; return rc; }
    }
    if (o && o->verbose) {
        printf("INIT--class Graph is constructed\n");
        printf("The vertex num is:%d\n", h->g.V);
        printf("The edge num is:%lld\n", (long long)h->g.E);
    }
    *out = h;
    return C2G_OK;
}

constexpr size_t kStageChunk = (size_t)32 << 20;
static_assert(64 % kColStage == 0, "column chunks must not span lane groups");

// out[k][v] = A[row(v)][l0 + k] for k < m <= kColStage, row(v) = perm[v]
// (v when perm is null): 64-vertex tiles staged through LDS so both the
// column read (32-byte segments) and the row writes are contiguous.
__global__ void __launch_bounds__(256)
extract_cols_k(const uint32_t* __restrict__ A, const int32_t* __restrict__ perm, int64_t V, int l0, int m,
               uint32_t* __restrict__ out) {
    __shared__ uint32_t tile[64][kColStage + 1];
    const int t = threadIdx.x;
    for (int64_t v0 = (int64_t)blockIdx.x * 64; v0 < V; v0 += (int64_t)gridDim.x * 64) {
        for (int i = t; i < 64 * m; i += 256) {
            const int r = i / m, c = i - r * m;
            const int64_t v = v0 + r;
            tile[r][c] = v < V ? A[(perm ? (int64_t)perm[v] : v) * 64 + l0 + c] : 0u;
        }
        __syncthreads();
        for (int i = t; i < 64 * m; i += 256) {
            const int k = i >> 6, r = i & 63;
            const int64_t v = v0 + r;
            if (v < V) out[(int64_t)k * V + v] = tile[r][k];
        }
        __syncthreads();
    }
}

int ColStage::copy_out(Graph* g, hipStream_t strm, const uint32_t* A, int32_t n, void* const* outs,
                       const int32_t* perm) {
    const int64_t V = g->V;
    if (V == 0 || n <= 0) return C2G_OK;
    if (cap < kColStage * V) {
        release();
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipMalloc(&dev[k], sizeof(uint32_t) * kColStage * V));
            HIP_TRY(hipHostMalloc((void**)&host[k], sizeof(uint32_t) * kColStage * V, hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        }
        cap = kColStage * V;
    }
    const int grid = (int)std::min<int64_t>((V + 63) / 64, (int64_t)std::max(1, g->num_cus) * 16);
    // chunks of up to kColStage queries that have a destination
    std::vector<std::pair<int32_t, int>> chunks;
    for (int32_t q0 = 0; q0 < n; q0 += kColStage) {
        const int m = std::min<int32_t>(kColStage, n - q0);
        bool any = false;
        for (int k = 0; k < m; ++k) any |= outs[q0 + k] != nullptr;
        if (any) chunks.emplace_back(q0, m);
    }
    auto host_copy = [&](size_t c) {
        const int32_t q0 = chunks[c].first;
        const int m = chunks[c].second;
        const uint32_t* h = host[c & 1];
        parallel_for((int64_t)m * V, [&](int64_t lo, int64_t hi, int) {
            while (lo < hi) {
                const int64_t k = lo / V, v0 = lo - k * V, len = std::min(hi - lo, V - v0);
                if (outs[q0 + k]) memcpy((uint32_t*)outs[q0 + k] + v0, h + k * V + v0, sizeof(uint32_t) * len);
                lo += len;
            }
        }, 0, (int64_t)1 << 18);
    };
    for (size_t c = 0; c < chunks.size(); ++c) {
        const int32_t q0 = chunks[c].first;
        const int m = chunks[c].second;
        const int gi = q0 / 64, l0 = q0 % 64;  // a chunk never spans two groups
        // buffer c&1 was last used by chunk c-2, whose host copy has finished
        hipLaunchKernelGGL(extract_cols_k, dim3(grid), dim3(256), 0, strm, A + (int64_t)gi * V * 64, perm, V, l0,
                           m, dev[c & 1]);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(host[c & 1], dev[c & 1], sizeof(uint32_t) * V * m, hipMemcpyDeviceToHost, strm));
        HIP_TRY(hipEventRecord(ev[c & 1], strm));
        if (c > 0) {
            HIP_TRY(hipEventSynchronize(ev[(c - 1) & 1]));
            host_copy(c - 1);
        }
    }
    if (!chunks.empty()) {
        HIP_TRY(hipEventSynchronize(ev[(chunks.size() - 1) & 1]));
        host_copy(chunks.size() - 1);
    }
    return C2G_OK;
}

void ColStage::release() {
    for (int k = 0; k < 2; ++k) {
        if (dev[k]) (void)hipFree(dev[k]);
        if (host[k]) (void)hipHostFree(host[k]);
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        dev[k] = nullptr;
        host[k] = nullptr;
        ev[k] = nullptr;
    }
    cap = 0;
}


int HostStage::copy_out(void* dst, const void* src, size_t bytes, hipStream_t strm) {
    if (bytes == 0) return C2G_OK;
    for (int k = 0; k < 2; ++k) {  // each piece on its own: a failed attempt is completed next time
        if (!buf[k]) HIP_TRY(hipHostMalloc((void**)&buf[k], kStageChunk, hipHostMallocDefault));
        if (!ev[k]) HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
    }
    const size_t nchunks = (bytes + kStageChunk - 1) / kStageChunk;
    auto host_copy = [&](size_t c) {
        const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
        const char* from = buf[c & 1];
        char* to = (char*)dst + off;
        parallel_for((int64_t)len, [&](int64_t lo, int64_t hi, int) { memcpy(to + lo, from + lo, hi - lo); },
                     0, (int64_t)1 << 20);
    };
    for (size_t c = 0; c < nchunks; ++c) {
        const size_t off = c * kStageChunk, len = std::min(kStageChunk, bytes - off);
        HIP_TRY(hipMemcpyAsync(buf[c & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, strm));
        HIP_TRY(hipEventRecord(ev[c & 1], strm));
        if (c > 0) {
            HIP_TRY(hipEventSynchronize(ev[(c - 1) & 1]));
            host_copy(c - 1);
        }
    }
    HIP_TRY(hipEventSynchronize(ev[(nchunks - 1) & 1]));
    host_copy(nchunks - 1);
    return C2G_OK;
}

void HostStage::release() {
    for (int k = 0; k < 2; ++k) {
        if (buf[k]) (void)hipHostFree(buf[k]);
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        buf[k] = nullptr;
        ev[k] = nullptr;
    }
}

}  // namespace c2g

using namespace c2g;

__global__ void diag_copy_k(const int4* __restrict__ a, int4* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

extern "C" {

int c2g_diag_copy(int64_t bytes, int32_t reps, double* gbps) {
    if (bytes < 16 || reps < 1) return fail(C2G_EINVAL, "bytes >= 16 and reps >= 1");
    int4 *a = nullptr, *b = nullptr;
    const int64_t n = bytes / 16;
    HIP_TRY(hipMalloc(&a, 16 * n));
    HIP_TRY(hipMalloc(&b, 16 * n));
    HIP_TRY(hipMemset(a, 1, 16 * n));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipDeviceProp_t p;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    const int grid = p.multiProcessorCount * 8;
    hipLaunchKernelGGL(diag_copy_k, dim3(grid), dim3(256), 0, 0, a, b, n);  // warm
    HIP_TRY(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(diag_copy_k, dim3(grid), dim3(256), 0, 0, a, b, n);
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (gbps) *gbps = 2.0 * 16 * n * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return C2G_OK;
}

int c2g_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
    return strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int c2g_open(const char* ds, int32_t variant, const c2g_opts* opts, c2g_graph** out) {
    if (!ds || !out) return fail(C2G_EINVAL, "c2g_open: null argument");
    c2g_graph* h = new (std::nothrow) c2g_graph;
    if (!h) return fail(C2G_ENOMEM, "c2g_open: out of memory");
    int rc = load_dir(&h->g, ds, variant, opts);
    if (rc) { delete h; return rc; }
    return finish_open(h, opts, out);
}

int c2g_graph_from_csr(int32_t V, const int32_t* vlist, const int32_t* elist, const int32_t* wlist,
                       const float* flist, const int32_t* hvlist, const int32_t* helist,
                       const int32_t* hwlist, const float* hwf, int32_t variant,
                       const c2g_opts* opts, c2g_graph** out) {
    if (V < 0 || !vlist || !out) return fail(C2G_EINVAL, "c2g_graph_from_csr: bad arguments");
    int64_t E = vlist[V];
    if (E > 0 && !elist) return fail(C2G_EINVAL, "c2g_graph_from_csr: null elist");
    c2g_graph* h = new (std::nothrow) c2g_graph;
    if (!h) return fail(C2G_ENOMEM, "out of memory");
    Graph& g = h->g;
    g.V = V;
    g.variant = variant;
    g.vl.assign(vlist, vlist + V + 1);
    g.el.assign(elist, elist + E);
    if (wlist) g.wl.assign(wlist, wlist + E);
    if (flist) g.fl.assign(flist, flist + E);
    if (hvlist) {
        int64_t HE = hvlist[V];
        g.hvl.assign(hvlist, hvlist + V + 1);
        if (HE > 0 && !helist) { delete h; return fail(C2G_EINVAL, "hidden lists need helist"); }
        g.hel.assign(helist, helist + HE);
        if (hwlist) g.hwl.assign(hwlist, hwlist + HE);
        if (hwf) g.hwf.assign(hwf, hwf + HE);
    }
    int rc = validate(&g);
    if (rc) { delete h; return rc; }
    return finish_open(h, opts, out);
}

int c2g_close(c2g_graph* h) {
    if (!h) return C2G_OK;
    if (h->g.serving.load()) return fail(C2G_EBUSY, "the handle is owned by a running query server");
    device_release(&h->g);
    delete h;
    return C2G_OK;
}

int c2g_graph_info(const c2g_graph* h, int32_t* V, int64_t* E, int64_t* HE, int32_t* variant) {
    if (!h) return fail(C2G_EINVAL, "null graph");
    if (V) *V = h->g.V;
    if (E) *E = h->g.E;
    if (HE) *HE = h->g.HE;
    if (variant) *variant = h->g.variant;
    return C2G_OK;
}

int c2g_graph_host_csr(const c2g_graph* h, const int32_t** vl, const int32_t** el,
                       const int32_t** wl, const float** fl) {
    if (!h) return fail(C2G_EINVAL, "null graph");
    if (vl) *vl = h->g.vl.data();
    if (el) *el = h->g.el.data();
    if (wl) *wl = h->g.wl.empty() ? nullptr : h->g.wl.data();
    if (fl) *fl = h->g.fl.empty() ? nullptr : h->g.fl.data();
    return C2G_OK;
}

int c2g_set_tune(c2g_graph* h, const c2g_tune* t) {
    if (!h || !t) return fail(C2G_EINVAL, "null argument");
    if (t->block != 128 && t->block != 256 && t->block != 512)
        return fail(C2G_EINVAL, "block (threads per workgroup of the relax kernels) must be 128, 256 or 512");
    if (t->block_single != 0 && t->block_single != 128 && t->block_single != 256 && t->block_single != 512)
        return fail(C2G_EINVAL, "block_single must be 0, 128, 256 or 512");
    if (t->xcd_split != 0 && t->xcd_split != 1) return fail(C2G_EINVAL, "xcd_split must be 0 or 1");
    if (t->grid_per_cu < 1 || t->grid_per_cu > 64) return fail(C2G_EINVAL, "grid_per_cu in [1,64]");
    if (t->grid_per_cu_single < 0 || t->grid_per_cu_single > 64)
        return fail(C2G_EINVAL, "grid_per_cu_single in [0,64]");
    if (t->ppr_pull < 0 || t->ppr_pull > 1000) return fail(C2G_EINVAL, "ppr_pull in [0,1000]");
    if (t->ads_relabel != 0 && t->ads_relabel != 1) return fail(C2G_EINVAL, "ads_relabel must be 0 or 1");
    if (t->chunk < 1 || t->chunk > 4096) return fail(C2G_EINVAL, "chunk in [1,4096]");
    if (t->batch_chunk < 1 || t->batch_chunk > 4096) return fail(C2G_EINVAL, "batch_chunk in [1,4096]");
    if (t->iters_per_graph < 2 || t->iters_per_graph > 256 || (t->iters_per_graph & 1))
        return fail(C2G_EINVAL, "iters_per_graph must be even, in [2,256]");
    if (t->delta < 0) return fail(C2G_EINVAL, "delta >= 0");
    if (t->light < 0) return fail(C2G_EINVAL, "light >= 0");
    if (t->batch_order != 0 && t->batch_order != 1) return fail(C2G_EINVAL, "batch_order must be 0 or 1");
    if (t->ppr_grid_per_cu < 0 || t->ppr_grid_per_cu > 64) return fail(C2G_EINVAL, "ppr_grid_per_cu in [0,64]");
    if (t->batch_relabel != 0 && t->batch_relabel != 1) return fail(C2G_EINVAL, "batch_relabel must be 0 or 1");
    C2G_HANDLE_LOCK(h);
    if (memcmp(&h->g.tune, t, sizeof *t) != 0) {
        // Captured launch sequences hold the old values as kernel arguments.
        if (h->g.sssp.exec) { (void)hipGraphExecDestroy(h->g.sssp.exec); h->g.sssp.exec = nullptr; }
        if (h->g.batch.exec) { (void)hipGraphExecDestroy(h->g.batch.exec); h->g.batch.exec = nullptr; }
    }
    h->g.tune = *t;
    return light_build(&h->g);
}

int c2g_get_tune(const c2g_graph* h, c2g_tune* t) {
    if (!h || !t) return fail(C2G_EINVAL, "null argument");
    *t = h->g.tune;
    return C2G_OK;
}

}  // extern "C"
