// Internal graph handle: host copies of the loaded arrays plus their HBM
// images and the per-query device workspaces.
//
// Replaces the reference's Graph (host vectors; src/sssp/base/Graph.h:16-51)
// and DeviceMemory (cudaMalloc + H2D; src/sssp/base/DeviceMemory.cuh:22-99).
//
// HBM layout (all int32 unless stated):
//   row   [V+1]        CSR offsets (the vlist file)
//   adj   [E] int2     interleaved (dst, w) = (elist[e], wlist[e]): one 8-byte
//                      load per edge instead of two 4-byte streams
//   adjf  [E] int2     (dst, float bits of f) for PPR / adsorption
//   hrow  [V+1], hadj [HE] int2   hidden lists (pruned variants)
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include <string>
#include <memory>
#include <mutex>
#include <atomic>
#include "common.h"

namespace c2g {

// A failed HIP call `what` as a C2G error (ENOMEM for memory, else EDEVICE),
// with the runtime's last-error state consumed like HIP_TRY does.
inline int hip_fail(hipError_t e, const char* what) {
    (void)hipGetLastError();
    return ::c2g::fail(e == hipErrorOutOfMemory ? C2G_ENOMEM : C2G_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

// Failing HIP calls return C2G_ENOMEM for an allocation that does not fit
// (device or pinned host memory) and C2G_EDEVICE otherwise.  The runtime's
// last-error state is consumed on the way out, so the launch checks of a
// later call (hipGetLastError) do not report this failure again.
#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            (void)hipGetLastError();                                                   \
            return ::c2g::fail(_e == hipErrorOutOfMemory ? C2G_ENOMEM : C2G_EDEVICE,  \
                               "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                               __FILE__, __LINE__);                                    \
        }                                                                              \
    } while (0)

// Device-to-host copies into pageable memory through two pinned chunks: the
// DMA of one chunk overlaps the parallel host copy of the previous one.
// One instance per lane or state (not shared between threads).
struct HostStage {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int copy_out(void* dst, const void* src, size_t bytes, hipStream_t strm);
    void release();
};

struct Graph;

// Per-query rows out of a lane-major result array [group][V][64] of 32-bit
// words (query q = lane q%64 of group q/64): a device tile transpose of
// kColStage columns at a time into [kColStage][V], one pinned transfer, and a
// parallel host copy into outs[q] (null rows are skipped).  With `perm`,
// vertex v's entry is read from row perm[v] of A.  Two staging
// buffers: the transpose and transfer of one chunk overlap the host copy of
// the previous one.  One instance per lane or state (not shared between
// threads).
constexpr int kColStage = 8;
struct ColStage {
    uint32_t* dev[2] = {nullptr, nullptr};   // [kColStage][V] device staging, double-buffered
    uint32_t* host[2] = {nullptr, nullptr};  // pinned mirrors
    hipEvent_t ev[2] = {nullptr, nullptr};
    int64_t cap = 0;                         // words per staging buffer
    int copy_out(Graph* g, hipStream_t strm, const uint32_t* A, int32_t n, void* const* outs,
                 const int32_t* perm = nullptr);
    void release();
};

// Device workspace of single-query SSSP.
struct SsspWs {
    int32_t* dist = nullptr;     // [V]
    uint32_t* bits = nullptr;    // activation bitmaps: live [nwp], hidden [nwp] (nwp padded to 64 words)
    int4* q[1] = {nullptr};      // frontier items (u, e_begin, e_end, kind)
    int64_t qcap = 0;
    int64_t* ctrl = nullptr;     // control block, see sssp.hip
    int64_t* h_ctrl = nullptr;   // pinned mirror
    uint8_t* hid = nullptr;      // [V] 1 = distance was set through a hidden edge
    hipGraphExec_t exec = nullptr;
    int exec_iters = 0, exec_grid = 0, exec_chunk = 0, exec_tb = 0, exec_split = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    HostStage out;               // distance copies to the host
};

// Device workspace of batched SSSP (64 queries per group, lane = query);
// the per-round queues and bitmaps live in sssp_batch.hip.
struct BatchWs {
    int32_t* dist = nullptr;     // [group][V][64]
    uint64_t* mask[1] = {nullptr};  // [group][V] lanes pending expansion
    uint64_t* hmask = nullptr;   // [group][V] lanes whose value came through a hidden edge
    uint64_t* ph = nullptr;      // [group][V] lanes whose heavy edges (c2g_tune.light) are pending
    int64_t* ctrl = nullptr;
    int64_t* h_ctrl = nullptr;
    hipGraphExec_t exec = nullptr;
    int exec_iters = 0, exec_groups = 0, exec_grid = 0, exec_chunk = 0, exec_tb = 0;
    int groups = 0;              // number of groups allocated
};

struct Graph {
    int32_t V = 0;
    int64_t E = 0, HE = 0;
    int variant = C2G_VARIANT_ORIGIN;
    std::vector<int32_t> vl, el, wl, hvl, hel, hwl;
    std::vector<float> fl, hwf;

    // device side
    bool on_device = false;
    int device = -1;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    int32_t* d_row = nullptr;
    int2* d_adj = nullptr;
    int2* d_adjf = nullptr;
    int32_t* d_hrow = nullptr;
    int2* d_hadj = nullptr;
    int2* d_hadjf = nullptr;
    // split mode of single-source SSSP: per-vertex offsets of the target
    // partitions of each row, [V+1][8] (built on first use)
    int32_t* d_split = nullptr;
    int32_t* d_hsplit = nullptr;
    // Light-edge rows of SSSP (c2g_tune.light): when built, d_adj holds the
    // E live entries followed by each vertex's edges of weight < light_built,
    // row by row, and d_lrow[v] .. d_lrow[v+1] (absolute indices >= E) is
    // v's light row.
    int32_t* d_lrow = nullptr;
    int32_t light_built = 0;
    int64_t EL = 0;
    SsspWs sssp;
    BatchWs batch;
    c2g_tune tune{};
    std::mutex mu;  // serialises the direct entry points on this handle
    // 1 while a query server owns the handle: its lanes use the handle's
    // stream and workspaces without `mu`, so every direct entry point that
    // would touch the device refuses with C2G_EBUSY until c2g_server_stop.
    std::atomic<int> serving{0};

    bool has_w() const { return !wl.empty() || E == 0; }
    bool has_f() const { return !fl.empty() || E == 0; }
};

int device_upload(Graph* g, int device);
void device_release(Graph* g);
void batch_release(Graph* g);
void ppr_release(Graph* g);
void ads_release(Graph* g);
void serve_release(Graph* g);
void default_tune(Graph* g);
int default_light(const Graph* g);
int ensure_sssp_ws(Graph* g);
// Builds (or drops) the light-edge rows for g->tune.light; captured launch
// sequences are invalidated when d_adj moves.
int light_build(Graph* g);
// The light rows when the SSSP kernels should use them (light > 0, built,
// delta == 0, xcd_split == 0), else null.
inline const int32_t* light_rows(const Graph* g) {
    return g->tune.light > 0 && g->light_built == g->tune.light && g->d_lrow && g->tune.delta == 0 &&
                   g->tune.xcd_split == 0
               ? g->d_lrow
               : nullptr;
}

// Host-side helpers shared by lanes
int host_bf(const Graph* g, int32_t* dist, const int32_t* frontier, int32_t count,
            const uint8_t* via_hidden, c2g_stats* st);

// Looks up `device`'s copy of the static scratch of c2g_sssp_noalloc (the
// device must be current); once per device, at upload.
int sssp_scratch_init(int device);

// Per-round trace of a query (c2g_sssp_rounds / c2g_sssp_batch_rounds):
// rounds run one at a time and each is recorded, up to `cap`.
struct RoundLog {
    c2g_round* out = nullptr;
    int32_t cap = 0;
    int32_t n = 0;
    int64_t prev[4] = {0, 0, 0, 0};  // cumulative frontier, edges, lane edges at the last two snapshots
    int64_t prev2[4] = {0, 0, 0, 0};
    int64_t impr = 0, items = 0;     // cumulative improvements and items at the last snapshot
    // Records the round that just ran from cumulative counters: frontier,
    // edges and lane edges are counted when a round's work is queued (by the
    // previous round), improvements and items when it runs.
    void add(const int64_t* q3, int64_t impr_now, int64_t items_now, double ms) {
        if (n < cap) {
            c2g_round& r = out[n];
            r.frontier = prev[0] - prev2[0];
            r.edges = prev[1] - prev2[1];
            r.lane_edges = prev[2] - prev2[2];
            r.improvements = impr_now - impr;
            r.items = items_now - items;
            r.ms = ms;
        }
        ++n;
        for (int k = 0; k < 3; ++k) { prev2[k] = prev[k]; prev[k] = q3[k]; }
        impr = impr_now;
        items = items_now;
    }
    void start(const int64_t* q3) {
        for (int k = 0; k < 3; ++k) { prev2[k] = 0; prev[k] = q3[k]; }
    }
};

// Used by the server and batched callers.
int sssp_run(Graph* g, int32_t src, int32_t budget, int32_t* dist_out, int32_t* frontier_out,
             int32_t* count_out, uint8_t* hid_out, c2g_stats* st, bool copy_out, RoundLog* log = nullptr);
int sssp_batch_group(Graph* g, const int32_t* srcs, int32_t n, int32_t* dist_out,
                     int64_t* sums, int32_t* reached, c2g_stats* st, RoundLog* log = nullptr);
int ppr_single(Graph* g, hipStream_t strm, int32_t src, float alpha, float eps, float* p_out,
               float* r_out, c2g_stats* st);
// Batched PPR (64 queries per lane group) on `strm`; query q's p (and r)
// land in p_outs[q] (r_outs[q]); null rows or a null r_outs are skipped.
int ppr_batch_to(Graph* g, hipStream_t strm, const int32_t* srcs, int32_t n, float alpha, float eps,
                 float* const* p_outs, float* const* r_outs, c2g_stats* st);
int ads_run(Graph* g, hipStream_t strm, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L,
            float p_inj, float p_cont, float p_abnd, int32_t iters, float* y_out, c2g_stats* st);
int ads_run_batch(Graph* g, hipStream_t strm, int k, const std::vector<int32_t>* svs,
                  const std::vector<int32_t>* sls, int32_t L, float p_inj, float p_cont, float p_abnd,
                  int32_t iters, float* const* y_outs, int grid_per_cu = 0);
int host_ads(const Graph* g, const int32_t* sv, const int32_t* sl, int32_t nseed, int32_t L, float p_inj,
             float p_cont, int32_t iters, float* y_out);

// Server lane of the batched SSSP kernels: one 64-lane group, queries join
// free lanes between rounds and leave when their lane has no pending work.
int bserve_init(Graph* g, hipStream_t strm);
int bserve_admit(Graph* g, hipStream_t strm, int lane, int32_t src);
int bserve_step(Graph* g, hipStream_t strm, uint64_t* lanes_with_work);
int bserve_extract(Graph* g, hipStream_t strm, const int* lanes, int n, int32_t* const* outs);

// Server lane of the single-query kernels run side by side: `slots` queries
// (at most 64), query-major state, items tagged with the slot.  Used while
// few queries are in flight, when the 64-lane rows above would be mostly idle.
int qserve_init(Graph* g, hipStream_t strm, int slots);
// Device bytes each server lane's workspace would add to what the handle
// holds now (0 when it is already allocated at that size); the server
// charges them against its HBM budget before a lane takes work.
int64_t qserve_ws_bytes(Graph* g, int slots);
int64_t bserve_ws_bytes(Graph* g);
int64_t ppr_lane_ws_bytes(Graph* g);
int64_t ads_lane_ws_bytes(Graph* g, int labels);
int qserve_admit(Graph* g, hipStream_t strm, int slot, int32_t src);
int qserve_step(Graph* g, hipStream_t strm, uint64_t* slots_with_work);
int qserve_extract(Graph* g, hipStream_t strm, int slot, int32_t* out);
void qserve_release(Graph* g);

}  // namespace c2g

struct c2g_graph {
    c2g::Graph g;
};

// Opening of every direct entry point that uses a handle's device state:
// serialises callers and refuses while a query server owns the handle.
#define C2G_HANDLE_LOCK(h)                                                               \
    std::lock_guard<std::mutex> lk((h)->g.mu);                                           \
    if ((h)->g.serving.load())                                                           \
        return ::c2g::fail(C2G_EBUSY, "the handle is owned by a running query server "   \
                                      "(call c2g_server_stop first)")
