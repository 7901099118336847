/*
 * c2g.h — C-ABI of the MI355X (gfx950) weighted-traversal library.
 *
 * The reference (lihuaibei7951/C-2graph) has no library interface: its
 * boundary is the dataset directory, the CLI mains, and inside each main a
 * `Graph` loader, a `DeviceMemory` owner and one persistent kernel launch per
 * query.  Every entry point below replaces one of those pieces and cites it.
 *
 * Conventions
 *  - Every function returns 0 (C2G_OK) on success or a negative C2G_E* code.
 *    The message of the last failure on the calling thread is returned by
 *    c2g_last_error().  Nothing in the library calls exit() (the reference
 *    does: core/Util.cuh:17-24, core/io.hpp:27-30).
 *  - Plain pointers and sizes only.  The caller owns every host buffer passed
 *    in or out (sized V, or n*V for batched calls); the library owns all device
 *    memory behind the opaque c2g_graph handle, mirroring DeviceMemory's RAII
 *    (src/sssp/base/DeviceMemory.cuh:22-44, 103-122).
 *  - One handle is used from one host thread at a time.  The query server
 *    (c2g_server_*) owns its own handle-internal streams and host lane.
 *  - Distances are int32; INT32_MAX marks an unreachable vertex (the reference
 *    uses 1999999999, src/sssp/base/Main.cu:121).
 *  - GPU entry points fail with C2G_EDEVICE when no HIP device or kernel image
 *    is usable.  There is no silent CPU fallback: the host-core lane runs only
 *    where a function says so (c2g_sssp_migrate, the server's fallback lane,
 *    c2g_host_*), and c2g_stats.lane reports which lane produced a result.
 */
#ifndef C2G_H
#define C2G_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define C2G_OK        0
#define C2G_EINVAL   (-1)  /* bad argument (source out of range, null ptr) */
#define C2G_EIO      (-2)  /* missing / unreadable / unwritable file        */
#define C2G_EFORMAT  (-3)  /* file lengths inconsistent (E != vlist[V] ...) */
#define C2G_ENOMEM   (-4)  /* host or device allocation failed              */
#define C2G_EDEVICE  (-5)  /* no usable HIP device / kernel launch failed   */
#define C2G_ENOTSUP  (-6)  /* graph lacks the data the call needs           */
#define C2G_EBUSY    (-7)  /* admission refused (server)                    */
#define C2G_ETIMEOUT (-8)  /* lane timeout / cancellation                   */

/* ---- dataset variants (the reference's directory layout) ----------------- */
/* <ds>/origin/{vlist,elist,wlist[,flist]}.bin  utils/convert2binary.cpp:122-125 */
#define C2G_VARIANT_ORIGIN    0
/* <ds>/sssp/{vlist,elist,wlist,hvlist,helist,hwlist}.bin  src/purn-sssp.cpp:72-87,342-344 */
#define C2G_VARIANT_SSSP_PURN 1
/* <ds>/ppr/{vlist,elist,wlist(f32),hvlist,helist,hwlist(f32)}.bin  src/purn-ppr.cpp:74-89,330-335 */
#define C2G_VARIANT_PPR_PURN  2

/* ---- lanes -------------------------------------------------------------- */
#define C2G_LANE_GPU  0
#define C2G_LANE_HOST 1
#define C2G_LANE_MIXED 2   /* GPU head + host-core tail (migration) */

/* ---- query kinds (server) ---------------------------------------------- */
#define C2G_Q_SSSP 0
#define C2G_Q_PPR  1
#define C2G_Q_ADS  2

typedef struct c2g_graph c2g_graph;
typedef struct c2g_server c2g_server;

typedef struct {
    int32_t device;      /* HIP ordinal; -1 = last device, as the reference's
                            cudaSetDevice(deviceCount-1), src/sssp/base/Main.cu:30 */
    int32_t load_flist;  /* ORIGIN variant: also load origin/flist.bin (PPR)   */
    int32_t verbose;     /* print loader lines like Graph.h:45-47              */
    int32_t no_device;   /* 1 = host-only handle (host lanes, CLIs, CPU tests)  */
} c2g_opts;

typedef struct {
    int32_t lane;            /* C2G_LANE_*                                        */
    int32_t iterations;      /* frontier iterations executed on the GPU           */
    int32_t host_iterations; /* frontier iterations executed on the host lane     */
    int32_t reached;         /* vertices with a finite result                     */
    int64_t frontier_total;  /* sum over iterations of |F|                        */
    int64_t edges_relaxed;   /* R: edge relaxations attempted                     */
    int64_t improvements;    /* U: relaxations that lowered a value               */
    int64_t reached_edges;   /* m_s = sum of outdeg over reached vertices (GTEPS) */
    int64_t csr_edges;       /* batched: CSR edges streamed (once for all lanes)  */
    int64_t items;           /* frontier work items processed                     */
    double  ms;              /* wall time of the call's device work               */
    double  kernel_ms;       /* profile runs: summed duration of the relax kernel */
    int32_t pull_rounds;     /* single-source PPR: rounds run as pull sweeps      */
    int32_t pad;
} c2g_stats;

/* One frontier round of a traced query (c2g_sssp_rounds, c2g_sssp_batch_rounds):
 * the reference records the frontier size of every iteration in
 * iteration_act_num (src/sssp/base/Main.cu:345-350). */
typedef struct {
    int64_t frontier;     /* vertices relaxed this round (batched: (vertex, group) pairs) */
    int64_t edges;        /* CSR edges streamed (hidden edges included)                   */
    int64_t lane_edges;   /* (query, edge) relaxations: = edges for one query            */
    int64_t improvements; /* relaxations that lowered a distance (writes issued)         */
    int64_t items;        /* work items                                                   */
    double  ms;           /* device time of the round (relax + expand kernels)           */
} c2g_round;

/* Tunables picked by the autotuner (c2g_autotune) or set by hand. */
typedef struct {
    int32_t block;          /* threads per workgroup of the batched SSSP relax
                               kernel (and the single-source one while
                               block_single is 0): 128, 256 or 512 (the
                               reference compiles in 512, core/Util.cuh:31-35) */
    int32_t grid_per_cu;    /* workgroups per CU for grid-stride kernels       */
    int32_t chunk;          /* max edges per single-query frontier item        */
    int32_t batch_chunk;    /* max edges per batched frontier item             */
    int32_t iters_per_graph;/* iterations captured per hipGraph replay         */
    int32_t use_graph;      /* 1 = replay captured iterations, 0 = plain launches */
    int32_t delta;          /* near/far split step for SSSP (0 = plain Bellman-Ford) */
    int32_t profile;        /* 1 = plain launches with HIP events around each relax
                               kernel; the summed time lands in c2g_stats.kernel_ms */
    int32_t grid_per_cu_single; /* workgroups per CU for the single-source SSSP
                               kernels (0 = grid_per_cu)                      */
    int32_t ppr_pull;       /* single-source PPR push/pull switch: a round whose
                               frontier holds more than ppr_pull/1000 of all
                               out-edges runs as a pull sweep over in-edges
                               (0 = push only; 0..1000)                       */
    int32_t ads_relabel;    /* adsorption on a copy of the weighted CSR with
                               vertices renumbered by descending in-degree
                               (rows gathered most often share cache lines);
                               results come back in the caller's ids: 0 / 1 */
    int32_t block_single;   /* threads per workgroup of the single-source relax
                               kernel (0 = block): 128, 256 or 512            */
    int32_t xcd_split;      /* single-source SSSP: 1 = cut each frontier row by
                               the partition (one eighth of the vertex range)
                               of its targets into eight queues, relaxed by the
                               workgroups of one XCD each, so each XCD's L2
                               holds one eighth of the distances: 0 / 1     */
    int32_t light;          /* SSSP light-edge bound L (0 = off): a vertex whose
                               distance was lowered in the last round relaxes
                               only its edges of weight < L; its other edges
                               wait until a round in which it was not lowered
                               again (batched: in which none of the group's
                               queries lowered it), so they are relaxed with a
                               settled distance.  Same distances, fewer
                               relaxations.  Applies while delta == 0 and
                               xcd_split == 0.  Default: about 1/16 of the
                               mean edge weight, 0 when fewer than 1 % of the
                               edges would be light (e.g. unit weights).  */
    int32_t batch_order;    /* batched SSSP: 1 = assign queries to 64-lane
                               groups in order of their sources' two-hop
                               reach, so that each group's lanes are pending
                               in the same rounds (results in the caller's
                               order); 0 = the caller's order           */
    int32_t ppr_grid_per_cu;/* workgroups per CU of single-source PPR and the
                               adsorption push modes (0 = 1): their rounds
                               are small, and fewer workgroups launch and
                               retire faster (1..64)                      */
    int32_t batch_relabel;  /* batched SSSP: 1 = run on a copy of the graph
                               renumbered so that vertices with an in-edge
                               (the only rows a traversal writes) come first,
                               which packs those rows and the set bits of the
                               expand's bitmaps into the low ids (about 4 %
                               faster on R-MAT); sources and distance rows
                               are mapped at the boundary, results are in
                               the caller's ids either way; 0 = the graph's
                               own numbering                               */
} c2g_tune;

/* ======================================================================= */
/* Errors                                                                   */
/* ======================================================================= */
/* Message of the last failing call on this thread ("" if none). */
const char* c2g_last_error(void);
/* Library version string (no reference counterpart). */
const char* c2g_version(void);
/* 1 if a HIP device with a gfx950 kernel image is usable, else 0 (the
 * reference's mains call cudaGetDeviceCount/cudaSetDevice and exit on failure,
 * src/sssp/base/Main.cu:26-30). */
int c2g_device_available(void);

/* ======================================================================= */
/* Format (host only)                                                       */
/* ======================================================================= */
/* Replaces read_binary2vector<T> (core/io.hpp:25-40): element count is
 * filesize/4.  Returns count >= 0, or a negative error.  If buf is NULL only
 * the count is returned. */
int64_t c2g_read_i32(const char* path, int32_t* buf, int64_t cap);
/* read_binary2vector<float> (core/io.hpp:25-40), for flist / hwlist(f32). */
int64_t c2g_read_f32(const char* path, float* buf, int64_t cap);
/* Headerless little-endian writers (convert2binary.cpp:118-134). */
int c2g_write_i32(const char* path, const int32_t* buf, int64_t n);
/* Float form of the writer (convert2binary.cpp:130-134 writes flist). */
int c2g_write_f32(const char* path, const float* buf, int64_t n);

/* Seeded R-MAT generator (replaces convert2binary's text ingestion for the
 * synthetic workloads, convert2binary.cpp:67,76-79: self-loops dropped, rows
 * sorted, duplicates removed).  Edge k of edge_factor<<scale is drawn by
 * quadrant recursion from a counter-based hash of (seed, k); vertex ids are
 * then passed through a seeded bijection of [0, 2^scale).  Outputs are
 * malloc'ed by the library (release with c2g_free); out_E receives the edge
 * count after dedup.
 *   want_w     : int weights uniform in [1,wmax], a hash of (seed, u, v)
 *   want_f     : row-normalised float weights, U(0,1] / row sum
 *   symmetric=1: add the reverse of every edge before dedup.             */
int c2g_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
             uint64_t seed, int32_t symmetric, int32_t wmax, int32_t want_f,
             int32_t** vlist, int32_t** elist, int32_t** wlist, float** flist,
             int64_t* out_E);

/* Writes <ds>/origin/{vlist,elist,wlist[,flist]}.bin (creates directories),
 * the files convert2binary writes (utils/convert2binary.cpp:118-134). */
int c2g_write_dataset(const char* ds, int32_t V, const int32_t* vlist,
                      const int32_t* elist, const int32_t* wlist, const float* flist);

/* convert2binary <adjlist.txt> <ds> (utils/convert2binary.cpp:86-135): reads
 * "src dst1:dst2:..." lines; weights are (rand()%100)+1 from an unseeded
 * glibc-compatible generator; flist = w / row sum.                      */
int c2g_convert_adjlist(const char* txt, const char* ds);

/* ======================================================================= */
/* Pruning (C^2graph-P), host only                                          */
/* ======================================================================= */
/* purn-sssp <ds> (src/purn-sssp.cpp:61-352).  Reads origin/, writes
 * sssp/{vlist,elist,wlist} with the reference's exact pruning decisions, and
 * sssp/{hvlist,helist,hwlist} per vertex (hvlist has V+1 entries, the layout
 * the loader sssp/purn/Graph.h:40 and sssp/purn/Main.cu:72 index), plus
 * sssp/porder.bin (processed vertices in prune order) so the reference's
 * compacted hidden-list layout can be rebuilt.  stats_out (nullable):
 * [0]=edges kept, [1]=hidden edges, [2]=processed vertices.             */
int c2g_prune_sssp(const char* ds, int64_t* stats_out);
/* purn-ppr <ds> (src/purn-ppr.cpp:63-344): reads origin/{vlist,elist,flist},
 * writes ppr/{vlist,elist,wlist(f32),hvlist,helist,hwlist(f32)}.        */
int c2g_prune_ppr(const char* ds, int64_t* stats_out);
/* In-memory forms of the same two passes (src/purn-sssp.cpp:109-349,
 * src/purn-ppr.cpp:113-335) used by the library and tests: inputs are an
 * origin CSR; outputs are malloc'ed by the library and released with
 * c2g_free.                                                              */
int c2g_prune_sssp_mem(int32_t V, const int32_t* vlist, const int32_t* elist,
                       const int32_t* wlist, int32_t** out_vlist,
                       int32_t** out_elist, int32_t** out_wlist,
                       int32_t** out_hvlist, int32_t** out_helist,
                       int32_t** out_hwlist, int32_t** out_porder,
                       int64_t* out_nporder);
int c2g_prune_ppr_mem(int32_t V, const int32_t* vlist, const int32_t* elist,
                      const float* flist, int32_t** out_vlist,
                      int32_t** out_elist, float** out_wlist,
                      int32_t** out_hvlist, int32_t** out_helist,
                      float** out_hwlist);
/* Releases a buffer the library malloc'ed (no reference counterpart). */
void c2g_free(void* p);

/* ======================================================================= */
/* Graph handle (replaces Graph + DeviceMemory)                              */
/* ======================================================================= */
/* Graph(dir) + DeviceMemory(V,E) + CudaMemcpyGraph (src/sssp/base/Graph.h:30-51,
 * src/sssp/purn/Graph.h:33-48, src/ppr/base/Graph.h:31-43,
 * src/sssp/base/DeviceMemory.cuh:22-65).  Loads the variant's files, checks
 * their lengths (vlist[V]==E, wlist length E, hvlist length V+1; a legacy
 * compacted sssp/hvlist is rejected with C2G_EFORMAT) and uploads the graph
 * to HBM once. */
int c2g_open(const char* ds, int32_t variant, const c2g_opts* opts, c2g_graph** out);
/* Same from caller arrays (copied), i.e. DeviceMemory + CudaMemcpyGraph
 * without the file loader (src/sssp/base/DeviceMemory.cuh:47-65).
 * wlist/flist/hidden arrays are nullable; for the PPR_PURN variant pass
 * flist (live edge floats) and hwf (hidden floats).  */
int c2g_graph_from_csr(int32_t V, const int32_t* vlist, const int32_t* elist,
                       const int32_t* wlist, const float* flist,
                       const int32_t* hvlist, const int32_t* helist,
                       const int32_t* hwlist, const float* hwf,
                       int32_t variant, const c2g_opts* opts, c2g_graph** out);
/* Frees the handle and its device memory (~DeviceMemory,
 * src/sssp/base/DeviceMemory.cuh:41-44, 103-122). */
int c2g_close(c2g_graph* g);
/* vert_num / edge_num of the loaded Graph (src/sssp/base/Graph.h:26-27, 33-34);
 * nullable outputs. */
int c2g_graph_info(const c2g_graph* g, int32_t* V, int64_t* E, int64_t* hidden_E,
                   int32_t* variant);
/* Host copies of the loaded arrays, Graph's csr_v / csr_e / csr_w
 * (src/sssp/base/Graph.h:33-35); nullable outputs, sized V+1 / E. */
int c2g_graph_host_csr(const c2g_graph* g, const int32_t** vlist,
                       const int32_t** elist, const int32_t** wlist,
                       const float** flist);
/* Launch shape and scheduling tunables, which the reference compiles in
 * (THREADS_PER_BLOCK 512, MAX_BLOCKS_NUM 20: core/Util.cuh:31-35).  Setting
 * them drops the captured launch sequences that hold the old values. */
int c2g_set_tune(c2g_graph* g, const c2g_tune* t);
int c2g_get_tune(const c2g_graph* g, c2g_tune* t);

/* ======================================================================= */
/* SSSP                                                                     */
/* ======================================================================= */
/* One query on the GPU.  Replaces CalcuSSSP (src/sssp/base/Main.cu:109-356)
 * on ORIGIN graphs and CalcuSSSP + host seeding (src/sssp/purn/Main.cu:67-90,
 * 136-193) on SSSP_PURN graphs.  Runs to convergence (no 1000-iteration cap).
 * On SSSP_PURN graphs the hidden list of the source, and of every vertex
 * whose distance was lowered through a hidden edge, is relaxed, so the
 * distances equal those of the full graph.  dist_out: int32[V]. */
int c2g_sssp(c2g_graph* g, int32_t src, int32_t* dist_out, c2g_stats* st);
/* Batched multi-source SSSP (no reference counterpart: the reference runs
 * one query per launch, src/sssp/base/Main.cu:59-76): n sources in one
 * captured launch sequence with per-query frontiers kept as a 64-bit lane
 * mask per vertex.  dist_out is
 * n*V int32 (query-major, row q = query q).  Results are bit-identical to
 * n calls of c2g_sssp.  st (nullable) receives whole-batch totals. */
int c2g_sssp_batch(c2g_graph* g, const int32_t* srcs, int32_t n,
                   int32_t* dist_out, c2g_stats* st);
/* Same as c2g_sssp_batch but leaves results on the device and only returns
 * a checksum per query (sum of finite distances, count of reached) — used to
 * time the kernels with inputs resident in HBM.  sums/reached nullable. */
int c2g_sssp_batch_resident(c2g_graph* g, const int32_t* srcs, int32_t n,
                            int64_t* sums, int32_t* reached, c2g_stats* st);
/* n single-source queries, one launch sequence per query (the "per-query"
 * comparison point for batching); same outputs as c2g_sssp_batch_resident. */
int c2g_sssp_each_resident(c2g_graph* g, const int32_t* srcs, int32_t n,
                           int64_t* sums, int32_t* reached, c2g_stats* st);

/* CalcuSSSP with an iteration budget (src/sssp/base-perk/Main.cu:263-497,
 * src/sssp/purn-perk/Main.cu:276-521): stops after `budget` GPU iterations
 * and returns the resumable state (dist[V], frontier[*count], *count).
 * frontier_out has room for V entries. */
int c2g_sssp_budget(c2g_graph* g, int32_t src, int32_t budget, int32_t* dist_out,
                    int32_t* frontier_out, int32_t* count_out, c2g_stats* st);
/* c2g_sssp with a per-round trace (CalcuSSSP's iteration_act_num,
 * src/sssp/base/Main.cu:345-350; printed by the commented-out loop at
 * :87-93).  Rounds run one launch pair at a time with a host read-back after
 * each, so the query is slower than c2g_sssp; the distances are the same.
 * Up to `cap` rounds are written to rounds[]; *nrounds gets the count of all
 * rounds (may exceed cap). */
int c2g_sssp_rounds(c2g_graph* g, int32_t src, int32_t* dist_out, c2g_round* rounds, int32_t cap,
                    int32_t* nrounds, c2g_stats* st);
/* c2g_sssp_batch_resident with the same per-round trace (no reference
 * counterpart: the reference runs one query per launch). */
int c2g_sssp_batch_rounds(c2g_graph* g, const int32_t* srcs, int32_t n, c2g_round* rounds, int32_t cap,
                          int32_t* nrounds, c2g_stats* st);
/* c2g_sssp without any allocation (no reference counterpart: it is for callers
 * that cannot allocate, such as tight serving loops or code that must not
 * call hipMalloc).  The distances land in the caller's DEVICE array
 * d_dist (int32[V], on the handle's device); the traversal's other state (bitmaps,
 * hidden flags, frontier queue, control block) lives in a static 64 MiB
 * device array that each device holds for the life of the process.  Graphs
 * whose state does not fit (about 2V + (E + hidden E)/chunk frontier items
 * of 16 bytes, e.g. up to R-MAT scale 20 at edge factor 16) are refused with
 * C2G_EINVAL.  Rounds run as plain launches on the handle's stream; the call
 * returns when the query has converged.  Every call on one device shares that
 * scratch: calls that overlap in time on one device (from several threads or
 * handles) race and give wrong results.  Distances equal c2g_sssp's. */
int c2g_sssp_noalloc(c2g_graph* g, int32_t src, int32_t* d_dist, c2g_stats* st);
/* GPU head for `budget` iterations, then the host-core lane finishes the
 * frontier (the C^2graph-PM handoff, src/sssp/purn-perk/Main.cu:111-182 with
 * Bellman_ford of src/sssp/base-perk/SSSP.cpp:88-119). */
int c2g_sssp_migrate(c2g_graph* g, int32_t src, int32_t budget, int32_t* dist_out,
                     c2g_stats* st);

/* ======================================================================= */
/* Host-core lane (the reference's CPU worker; fallback lane of the server) */
/* ======================================================================= */
/* Frontier Bellman-Ford from a (dist, frontier) snapshot, dist updated in
 * place (src/sssp/base-perk/SSSP.cpp:88-119, the version that resets its flag).
 * On SSSP_PURN graphs hidden lists are relaxed by the same rule as c2g_sssp. */
int c2g_host_bellman_ford(const c2g_graph* g, int32_t* dist, const int32_t* frontier,
                          int32_t count, const uint8_t* via_hidden, c2g_stats* st);
/* Full single-source query on the host lane: the reference CPU worker's
 * entry (SSSP_CPU, src/sssp/purn-perk/SSSP.h:31-37) from {s:0},{s}. */
int c2g_host_sssp(const c2g_graph* g, int32_t src, int32_t* dist_out, c2g_stats* st);
/* Forward push from (p, r, frontier) (src/ppr/w-base-perk/PPR.cpp:178-236)
 * with threshold r[v] > eps*outdeg(v); flag array zero-initialised. */
int c2g_host_forward_push(const c2g_graph* g, float alpha, float eps, float* p,
                          float* r, const int32_t* frontier, int32_t count,
                          c2g_stats* st);
/* Full single-source push on the host lane (ppr_CPU entry,
 * src/ppr/w-base-perk/PPR.h:38-46) from r[s]=1, p[s]=alpha. */
int c2g_host_ppr(const c2g_graph* g, int32_t src, float alpha, float eps,
                 float* p_out, float* r_out, c2g_stats* st);

/* ======================================================================= */
/* PPR (forward push)                                                        */
/* ======================================================================= */
/* calcuatePPR<float> (src/ppr/base/Main.cu:121-388; pruned seeding
 * src/ppr/purn/Main.cu:78-101,171-207): r[s]=1, p[s]=alpha; each iteration
 * every active u sends (1-alpha)*r[u]*f(u,v) to m[v] and zeroes r[u]; then
 * r[v]+=m[v] and v becomes active (p[v]+=alpha*r[v]) when
 * r[v] > eps*outdeg(v).  FP32 state.  p_out, r_out: float[V] (r nullable). */
int c2g_ppr(c2g_graph* g, int32_t src, float alpha, float eps, float* p_out,
            float* r_out, c2g_stats* st);
/* ppr-base-M / ppr-purn-M (src/ppr/w-purn-perk/Main.cu:119-192, 379): GPU
 * rounds continue while (round < budget) || (|frontier| > handoff_min); the
 * host forward push (src/ppr/w-base-perk/PPR.cpp:178-236) then finishes from
 * the handed-over (p, r, frontier).  handoff_min = 10000 in the reference. */
int c2g_ppr_migrate(c2g_graph* g, int32_t src, float alpha, float eps, int32_t budget,
                    int32_t handoff_min, float* p_out, float* r_out, c2g_stats* st);
/* n sources in one launch sequence, lane = query (no reference counterpart:
 * one calcuatePPR launch per query, src/ppr/base/Main.cu:79-83); p_out/r_out
 * are n*V (query-major). */
int c2g_ppr_batch(c2g_graph* g, const int32_t* srcs, int32_t n, float alpha,
                  float eps, float* p_out, float* r_out, c2g_stats* st);

/* ======================================================================= */
/* Adsorption                                                                */
/* ======================================================================= */
/* The reference's own "adsorption" kernel, kept as a reference mode:
 * calcuatePPR of src/adsorption/w-base/Main.cu:123-395, a single-source push
 * with alpha = 0.2, r[s] = 10, p[s] = 0.8 * r[s]
 * (:63-64, :149-152), messages (1-alpha) r_u f(u,v) (:248) and a constant
 * threshold rmax = 1/V (:351-355).  Rounds run to convergence (the reference
 * caps them at 2000, :186).  p_out, r_out: float[V] (r_out nullable). */
int c2g_adsorption_push(c2g_graph* g, int32_t src, float* p_out, float* r_out, c2g_stats* st);

/* The reference's other adsorption push variants (same seeding r[s] = 10,
 * alpha 0.2, messages (1-alpha) r w):
 *   C2G_ADS_PUSH_BASE  p[s] += 0.8 r[s], activation credits alpha r, rmax 1/V:
 *                      c2g_adsorption_push.  On a PPR_PURN graph this is
 *                      adsorption/w-purn: the push on the live pruned CSR,
 *                      no hidden seeding (commented out in the reference,
 *                      src/adsorption/w-purn/Main.cu:157-169), rmax 1/(V+1)
 *                      (:67).
 *   C2G_ADS_PUSH_PERK  adsorption/w-base-perk: p[s] += alpha r[s]
 *                      (src/adsorption/w-base-perk/Main.cu:306-309),
 *                      activation credits (1-alpha) r (:555-557), rmax 1/V
 *                      (Graph.h:117); runs on either graph variant (the
 *                      reference's loader reads ppr/).                   */
#define C2G_ADS_PUSH_BASE 0
#define C2G_ADS_PUSH_PERK 1
int c2g_adsorption_push_mode(c2g_graph* g, int32_t src, int32_t mode, float* p_out, float* r_out,
                             c2g_stats* st);

/* Multi-label adsorption (reference: adsorption/w-base/Main.cu:123-395 is a
 * single-source push; the multi-label recurrence here follows Baluja et al.
 * 2008 / Talukdar & Crammer 2009):
 *   Y0[v][l] = 1 if (v,l) is a seed, else 0
 *   Y_{t+1}[v] = p_inj*Y0[v] + p_cont * sum_{(v->u)} f(v,u) * Y_t[u]
 * for `iters` Jacobi iterations starting from Y_0 = Y0.  p_abnd is the
 * abandonment probability (the dummy label's share) and is checked to make
 * p_inj+p_cont+p_abnd = 1.  y_out: float[V*L], vertex-major.  L <= 64. */
int c2g_adsorption(c2g_graph* g, const int32_t* seed_v, const int32_t* seed_label,
                   int32_t nseed, int32_t L, float p_inj, float p_cont, float p_abnd,
                   int32_t iters, float* y_out, c2g_stats* st);

/* ======================================================================= */
/* Autotuner                                                                 */
/* ======================================================================= */
/* Times the SSSP kernels over a sweep of block shapes, chunk sizes (the
 * degree at which a row is split across waves) and iterations-per-graph on
 * the given sample sources and stores the fastest setting in the handle
 * (returned in best_out, nullable).  Replaces the reference's compile-time
 * launch constants (core/Util.cuh:31-35) and its fixed block/warp/thread
 * degree tiers (src/sssp/base/Main.cu:197-293). */
int c2g_autotune(c2g_graph* g, const int32_t* srcs, int32_t n, int32_t batched,
                 c2g_tune* best_out);

/* Picks c2g_tune.ppr_pull, the push/pull switch point of single-source PPR,
 * by timing c2g_ppr on the sample sources with the caller's alpha and eps,
 * and stores it in the handle (best_out nullable).  The reference only
 * pushes (calcuatePPR<float>, src/ppr/base/Main.cu:208-320); the pull form
 * computes the same round. */
int c2g_autotune_ppr(c2g_graph* g, const int32_t* srcs, int32_t n, float alpha, float eps,
                     c2g_tune* best_out);

/* Diagnostic (no reference counterpart): a plain device copy of `bytes`
 * bytes repeated `reps` times, used to calibrate HBM byte counters; returns
 * the copy rate in GB/s (read + write bytes over time). */
int c2g_diag_copy(int64_t bytes, int32_t reps, double* gbps);

/* ======================================================================= */
/* Serving loop (continuous batching)                                        */
/* ======================================================================= */
typedef struct {
    int32_t max_batch;        /* max SSSP queries per device batch (<=64)      */
    int32_t max_inflight;     /* max queries admitted at once                  */
    int64_t hbm_reserve;      /* bytes of HBM kept free                        */
    int64_t fake_free_hbm;    /* >0: use this as free HBM (tests)              */
    int32_t host_lane;        /* 1 = enable host-core fallback lane            */
    int32_t device_failed;    /* 1 = treat device as failed (tests)            */
    float   alpha, eps;       /* PPR parameters                                */
    int32_t ads_labels, ads_iters;
    float   p_inj, p_cont, p_abnd;
    int32_t sssp_slots;       /* SSSP queries run side by side in query-major
                                 slots before the 64-lane group takes the
                                 overflow: 0 = 16, <0 = lane group only, <=64 */
    int32_t ads_grid_per_cu;  /* workgroups per CU of the adsorption lane's
                                 kernels, below full occupancy so the SSSP
                                 and PPR lanes share every CU: 0 = the
                                 graph's grid_per_cu, else 1..64             */
    int32_t deadline_ms;      /* >0: a request not done this many ms after its
                                 submission completes with C2G_ETIMEOUT     */
    int32_t fail_inject;      /* tests: C2G_SLANE_* bits of device lanes whose
                                 step reports a device error, once each      */
    int32_t fail_after;       /* tests: successful steps of an injected lane
                                 before its failure                          */
    int32_t dry_device;       /* tests: the device lanes (admission, budget,
                                 slots, failover, cancellation) run with host
                                 computations in place of kernels, so they work
                                 without a GPU; fake_free_hbm is the budget  */
    int32_t dry_step_ms;      /* tests: duration of one dry lane step         */
    int32_t reserved[1];
} c2g_server_opts;

/* Device lanes of the server, as bits (c2g_server_opts.fail_inject,
 * c2g_server_stats.lanes_enabled). */
#define C2G_SLANE_SLOTS 1   /* SSSP query-major slot pool       */
#define C2G_SLANE_GROUP 2   /* SSSP 64-lane group (overflow)    */
#define C2G_SLANE_PPR   4   /* batched PPR lane                 */
#define C2G_SLANE_ADS   8   /* adsorption lane                  */

typedef struct {
    int64_t budget;          /* HBM budget: free HBM at start (or fake_free_hbm)
                                minus hbm_reserve                            */
    int64_t charged;         /* workspace bytes the enabled lanes charged      */
    int64_t lane_bytes[4];   /* charged per lane (slots, group, PPR, ads)      */
    int32_t sssp_slots;      /* slots of the SSSP pool after budget sizing     */
    int32_t lanes_enabled;   /* C2G_SLANE_* bits of device lanes in service    */
    int64_t submitted, completed, failed, cancelled;  /* cancelled includes
                                                         deadline expiries */
    int64_t to_host;         /* requests handed to the host lane              */
    int64_t lane_failures;   /* device-lane errors the server recovered from  */
    int32_t queued;          /* waiting for a lane                            */
    int32_t running;         /* admitted to a lane, not finished              */
} c2g_server_stats;

typedef struct {
    int64_t id;
    int32_t kind;     /* C2G_Q_* */
    int32_t status;   /* 0 done, <0 error */
    int32_t lane;     /* C2G_LANE_* */
    int32_t pad;
    double  t_submit_ms, t_admit_ms, t_done_ms;  /* server clock */
} c2g_result_info;

/* Starts the lanes over an open graph.  The server replaces the reference's
 * CPU/GPU collaboration loop (src/sssp/purn-perk/Main.cu:87-246: two GPU
 * streams and one CPU worker thread serving one repeated query) with
 * continuously batched lanes per query kind plus the host-core lane.
 *
 * Admission by device memory: at start every device lane's workspace is
 * charged against the HBM budget in the order SSSP slot pool (its slot count
 * halved until it fits), SSSP lane group, PPR lane, adsorption lane; a lane
 * that does not fit is not started.  Requests of a kind with no device lane
 * go to the host lane when it is enabled and are refused with C2G_EBUSY
 * otherwise; the others wait for a free slot, lane or pass.
 *
 * Lane failure: when a device lane's step fails, its running requests move
 * to the host lane (or finish with the error when the host lane is off), the
 * lane is reset and keeps serving; a lane whose reset fails is retired.
 *
 * While the server runs it owns the handle: every other entry point on it
 * returns C2G_EBUSY until c2g_server_stop. */
int c2g_server_start(c2g_graph* g, const c2g_server_opts* o, c2g_server** out);
/* Submit one query; returns its id (>= 0) or a negative error (C2G_EBUSY when
 * admission is refused).  For C2G_Q_ADS, src is the seed-set index.  The
 * reference hands work to its CPU worker through a pinned ring buffer
 * (initQueue / outQueue, src/sssp/purn-perk/SSSP.h:13-37); here requests
 * enter per-kind queues admitted by free HBM. */
int64_t c2g_server_submit(c2g_server* s, int32_t kind, int32_t src);
/* Poll a finished query: 1 = done (info filled, result copied into out when
 * non-null: int32[V] for SSSP, float[V] for PPR, float[V*L] for adsorption),
 * 0 = not done, <0 error.  wait_ms < 0 blocks. */
int c2g_server_poll(c2g_server* s, int64_t id, int32_t wait_ms, void* out,
                    c2g_result_info* info);
/* Cancels request `id` (no reference counterpart: the reference's CPU worker
 * has no cancellation, src/sssp/purn-perk/SSSP.cpp:140-159).  A queued
 * request finishes at once, a running one at its lane's next step boundary;
 * either way its poll reports C2G_ETIMEOUT.  Returns 1 if the request was
 * cancelled, 0 if it had already finished. */
int c2g_server_cancel(c2g_server* s, int64_t id);
/* Counters and admission state of a running server. */
int c2g_server_get_stats(c2g_server* s, c2g_server_stats* out);
/* Stops the lanes, cancels outstanding requests and frees the server (the
 * reference joins its CPU worker thread, src/sssp/purn-perk/Main.cu:255). */
int c2g_server_stop(c2g_server* s);

#ifdef __cplusplus
}
#endif
#endif /* C2G_H */
