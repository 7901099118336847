// Single-source SSSP on gfx950: frontier relaxation with int32 atomicMin.
//
// Replaces CalcuSSSP (src/sssp/base/Main.cu:109-356) and, for pruned graphs,
// CalcuSSSP + host seeding (src/sssp/purn/Main.cu:67-90, 136-193).
//
// Reference behaviour kept: label-correcting relaxation `cand = dis[u] + w;
// if dis[v] > cand: atomicMin(dis[v], cand), v active` with dis[u] re-read at
// relax time (Main.cu:188, 217-221), frontier = vertices improved in the
// previous round, stop when the frontier is empty.  Changed on purpose: no
// 1000-round cap (Main.cu:168), INT32_MAX as the unreached value (Main.cu:121
// uses 1999999999), no software grid barrier (Main.cu:143-147, unsafe without
// co-residency), no O(V) isactive rescan per round (Main.cu:313-340).
//
// One round is two launches:
//   * `sssp_relax_k` walks the round's work items int4(u, e_begin, e_end,
//     kind | slot << 1) (kind 1 = hidden-list edges).  A vertex of out-degree
//     d gives ceil(d/chunk) items, so hubs are split across waves (the
//     reference's block/warp/thread tiers, Main.cu:197-293, become one
//     balanced path).  A wave takes 64 items, scans their edge counts (wave64
//     shuffle scan) and walks the concatenated edge list; lane k finds its
//     owner item with a 6-step shuffle search.  Edge (dst,w) pairs are one
//     8-byte load from the interleaved `adj` array.  An improving lane issues
//     atomicMin without waiting for the old value and marks v in the
//     activation bitmap with a fire-and-forget atomicOr.
//   * `sssp_expand_k` compacts the bitmap into the next round's items (wave
//     ballot + scan, one queue atomic per wave) and clears it.
//   * The last workgroup of each kernel publishes the counters, so rounds are
//     plain back-to-back launches (captured in a hipGraph); there is no grid
//     barrier.
// A slot pool runs several queries side by side with the same kernels: slot s
// owns dist[s*V..], its own bitmap words and hidden flags (server lane).
//
// Light-edge mode (c2g_tune.light = L > 0, delta 0, no split): a vertex
// lowered in the last round relaxes only its light row (edges of weight < L,
// kept after the E live entries of `adj`) and is marked in the pending
// bitmap; the first round in which it is not lowered again relaxes its full
// row and clears the mark.  Heavy edges are then relaxed with a settled
// distance rather than after every improvement: about half the relaxations
// on R-MAT with uniform weights, for a few more rounds.
//
// Pruned graphs: a vertex's hidden list is relaxed when the vertex is the
// source or when its distance was ever lowered through a hidden edge
// (flag `hid`).  Hidden edges stand for real paths, so extra relaxation can
// never lower a distance below the true one.
#include "graph.h"
#include "wave.h"
#include <cstring>
#include <type_traits>

namespace c2g {

enum : int {
    C_CUR = 0,    // items in the current round's queue
    C_ITER = 2,   // rounds with work
    C_FRONT = 4,  // sum of frontier vertices over rounds
    C_EDGES = 5,  // edges queued for relaxation (R)
    C_IMPR = 6,   // improvement writes issued (U): candidate below the value read
    C_OVF = 7,    // queue overflow guard tripped (must stay 0)
    C_HEDGES = 8, // hidden edges queued
    C_ITEMS = 10, // items processed
    C_THR = 11,   // delta-stepping: current bucket bound (INT32_MAX = plain rounds)
    C_FMIN = 12,  // smallest deferred distance this round
    C_NEAR = 13,  // vertices expanded this round
    C_FAR = 14,   // 1 while deferred vertices remain
    C_SLOTS = 15, // slot pools: bit s set when slot s has items queued for the next round
    C_QN = 16,    // split mode: items of partition p for the next round at C_QN + p (8 entries)
    C_N = 24      // fields; the block is kCtrlWords long (completion tree after them)
};
static_assert(C_N <= kCtrlTree, "control fields overlap the completion tree");

// Split mode (c2g_tune.xcd_split): the edges of a frontier row are cut by
// the partition of their target, p = v / ceil(V/8), into eight queues, and
// the relax workgroups with blockIdx % 8 == p take queue p.  Workgroups are
// handed to the eight XCDs round-robin, so each XCD's L2 gathers only its
// eighth of `dist`.  Rows are sorted by target, so a row's edges into
// partition p are one contiguous range: split[8u + p] is its first edge and
// split[8u + 8] = split[8(u+1)] the end of the row.
constexpr int kParts = 8;

// Device edge entries carry a flag in bit 31 of the target id: set when the
// target has out-edges (or hidden edges), i.e. when activating it can lead
// anywhere.  Vertex ids are < 2^31, so the bit is otherwise unused.
constexpr int kIdMask = 0x7fffffff;
constexpr int kSteps = 4;  // 64-edge steps per lane whose loads are in flight together

__global__ void sssp_init_k(int32_t* __restrict__ dist, uint32_t* __restrict__ bits,
                            uint32_t* __restrict__ hbits, uint8_t* __restrict__ hid, int64_t V,
                            uint32_t* __restrict__ pbits) {
    const int64_t nw = (V + 31) >> 5;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * blockDim.x) {
        dist[i] = INT32_MAX;
        if (hid) hid[i] = 0;
        if (i < nw) {
            bits[i] = 0;
            if (hbits) hbits[i] = 0;
            if (pbits) pbits[i] = 0;
        }
    }
}

// Queues the source: live items and (pruned graphs) hidden items.  In split
// mode (split non-null) thread p < 8 writes partition p's items of the source
// into queue p (q + p * qseg).
__global__ void sssp_seed_k(const int32_t* __restrict__ row, const int32_t* __restrict__ hrow,
                            int32_t* __restrict__ dist, uint8_t* __restrict__ hid,
                            int4* __restrict__ q, int64_t* __restrict__ ctrl, int32_t src,
                            int chunk, int delta, const int32_t* __restrict__ split,
                            const int32_t* __restrict__ hsplit, int64_t qseg) {
    const int r0 = row[src], r1 = row[src + 1];
    const int nl = (r1 - r0 + chunk - 1) / chunk;
    int h0 = 0, h1 = 0, nh = 0;
    if (hrow) { h0 = hrow[src]; h1 = hrow[src + 1]; nh = (h1 - h0 + chunk - 1) / chunk; }
    if (split) {
        __shared__ int64_t s_n[kParts];
        if (threadIdx.x < kParts) {
            const int p = threadIdx.x;
            int4* qp = q + p * qseg;
            int64_t k = 0;
            for (int b = split[(int64_t)src * 8 + p], e = split[(int64_t)src * 8 + p + 1]; b < e; b += chunk)
                qp[k++] = make_int4(src, b, min(b + chunk, e), 0);
            if (hsplit)
                for (int b = hsplit[(int64_t)src * 8 + p], e = hsplit[(int64_t)src * 8 + p + 1]; b < e; b += chunk)
                    qp[k++] = make_int4(src, b, min(b + chunk, e), 1);
            s_n[p] = k;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            dist[src] = 0;
            if (hid) hid[src] = 1;
            for (int i = 0; i < C_N; ++i) ctrl[i] = 0;
            done_tree_clear(ctrl);
            ctrl[C_THR] = delta > 0 ? delta : INT32_MAX;
            ctrl[C_FMIN] = INT32_MAX;
            int64_t tot = 0;
            for (int p = 0; p < kParts; ++p) { ctrl[C_QN + p] = s_n[p]; tot += s_n[p]; }
            ctrl[C_CUR] = tot;
            ctrl[C_FRONT] = 1;
            ctrl[C_EDGES] = r1 - r0;
            ctrl[C_HEDGES] = h1 - h0;
        }
        return;
    }
    for (int j = threadIdx.x; j < nl + nh; j += blockDim.x) {
        if (j < nl) {
            int b = r0 + j * chunk;
            q[j] = make_int4(src, b, min(b + chunk, r1), 0);
        } else {
            int b = h0 + (j - nl) * chunk;
            q[j] = make_int4(src, b, min(b + chunk, h1), 1);
        }
    }
    if (threadIdx.x == 0) {
        dist[src] = 0;
        if (hid) hid[src] = 1;  // the source's hidden list is always relaxed
        for (int i = 0; i < C_N; ++i) ctrl[i] = 0;
        done_tree_clear(ctrl);
        ctrl[C_THR] = delta > 0 ? delta : INT32_MAX;
        ctrl[C_FMIN] = INT32_MAX;
        ctrl[C_CUR] = nl + nh;
        ctrl[C_FRONT] = 1;
        ctrl[C_EDGES] = r1 - r0;
        ctrl[C_HEDGES] = h1 - h0;
    }
}

// Items are int4(u, e_begin, e_end, kind | slot << 1): a single query is
// slot 0; a slot pool runs several queries side by side, slot s owning
// dist[s*V..], hid[s*V..] and the bitmap words [s*nwp..] of each map.
// BLOCK: threads per workgroup (128, 256 or 512; c2g_tune.block).
// SPLIT: the workgroup takes queue p = blockIdx % 8 of the split mode (the
// grid is a multiple of 8).
template <bool PRUNED, int BLOCK, bool SPLIT>
__global__ void __launch_bounds__(BLOCK)
sssp_relax_k(const int2* __restrict__ adj, const int2* __restrict__ hadj,
             int32_t* __restrict__ dist, uint32_t* __restrict__ bits,
             uint32_t* __restrict__ hbits, uint8_t* __restrict__ hid,
             const int4* __restrict__ q0, int64_t* __restrict__ ctrl, int64_t V, int64_t nwp, int64_t qseg) {
    const int part = SPLIT ? (int)(blockIdx.x & (kParts - 1)) : 0;
    const int64_t n = SPLIT ? ctrl[C_QN + part] : ctrl[C_CUR];
    const int4* __restrict__ q = q0 + (SPLIT ? part * qseg : 0);
    const int lane = lane_id();
    const int64_t wave = SPLIT ? (int64_t)(blockIdx.x >> 3) * (BLOCK >> 6) +
                                     __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)
                               : wave_index();
    const int64_t nwaves = SPLIT ? (int64_t)(gridDim.x >> 3) * (BLOCK >> 6) : ((int64_t)gridDim.x * BLOCK) >> 6;
    int64_t impr = 0;

    for (int64_t base = wave * kWave; base < n; base += nwaves * kWave) {
        const int64_t i = base + lane;
        int eb = 0, cnt = 0, du = INT32_MAX, kw = 0;
        if (i < n) {
            const int4 it = q[i];
            eb = it.y;
            cnt = it.z - it.y;
            kw = it.w;
            du = dist[(int64_t)(kw >> 1) * V + it.x];
        }
        int tot;
        const int pre = wave_excl_scan(cnt, &tot);
        // kSteps 64-edge steps per pass: every lane first loads its edges and
        // the targets' current distances (all in flight together), then
        // compares and issues the improving atomics.
        for (int s = 0; s < tot; s += kSteps * kWave) {
            int2 a[kSteps];
            int du_j[kSteps], cur[kSteps], kd[kSteps];
            int64_t sl[kSteps];
#pragma unroll
            for (int j = 0; j < kSteps; ++j) {
                const int k = s + j * kWave + lane;
                const int o = wave_owner(pre, k);
                const int e = __shfl(eb, o, 64) + (k - __shfl(pre, o, 64));
                const int w = __shfl(kw, o, 64);
                du_j[j] = __shfl(du, o, 64);
                kd[j] = PRUNED ? (w & 1) : 0;
                sl[j] = w >> 1;
                a[j] = make_int2(0, 0);
                if (k < tot) a[j] = (PRUNED && kd[j]) ? hadj[e] : adj[e];
            }
#pragma unroll
            for (int j = 0; j < kSteps; ++j) {
                const bool ok = s + j * kWave + lane < tot && du_j[j] <= INT32_MAX - a[j].y;
                cur[j] = ok ? dist[sl[j] * V + (a[j].x & kIdMask)] : INT32_MIN;
            }
#pragma unroll
            for (int j = 0; j < kSteps; ++j) {
                if (cur[j] == INT32_MIN) continue;
                const int v = a[j].x & kIdMask;
                const int cand = du_j[j] + a[j].y;
                if (cand >= cur[j]) continue;
                // No wait for the old value: a lower dist[v] written since the
                // read above comes from this round, whose writer marked the same
                // activation bit.
                atomicMin(&dist[sl[j] * V + v], cand);
                ++impr;
                if (a[j].x < 0) atomicOr(&bits[sl[j] * nwp + (v >> 5)], 1u << (v & 31));
                if (PRUNED) {
                    uint8_t* h = hid + sl[j] * V;
                    if (kd[j]) h[v] = 1;
                    if (kd[j] || h[v]) atomicOr(&hbits[sl[j] * nwp + (v >> 5)], 1u << (v & 31));
                }
            }
        }
    }
    impr = wave_sum64(impr);
    // workgroup total, then one global atomic per workgroup
    __shared__ unsigned long long s_impr;
    if (threadIdx.x == 0) s_impr = 0;
    __syncthreads();
    if (lane == 0 && impr) atomicAdd(&s_impr, (unsigned long long)impr);
    // The last workgroup to finish retires the queue (expand refills it).
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_impr) atomicAdd((unsigned long long*)&ctrl[C_IMPR], s_impr);
        if (wg_last(ctrl + kCtrlTree)) {
            int64_t ntot = n;
            if (SPLIT) {
                ntot = 0;
                for (int p = 0; p < kParts; ++p) { ntot += ctrl[C_QN + p]; ctrl[C_QN + p] = 0; }
            }
            if (ntot > 0) ctrl[C_ITER] += 1;
            ctrl[C_ITEMS] += ntot;
            ctrl[C_CUR] = 0;
            ctrl[C_SLOTS] = 0;
        }
    }
}

// Calls f(v, b, e) for each set bit of m (vertex base + bit) with v's row
// [r[v], r[v+1]), four vertices at a time: their offset loads are issued
// together instead of one dependent pair per vertex.
template <class F>
__device__ __forceinline__ void walk_rows4(uint32_t m, int64_t base, const int32_t* __restrict__ r, F&& f) {
    while (m) {
        int v[4], b[4], e[4];
        bool ok[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ok[k] = m != 0;
            v[k] = ok[k] ? (int)(base + __builtin_ctz(m)) : 0;
            if (ok[k]) m &= m - 1;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            b[k] = ok[k] ? r[v[k]] : 0;
            e[k] = ok[k] ? r[v[k] + 1] : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (ok[k]) f(v[k], b[k], e[k]);
    }
}

// Compacts the activation bitmaps into the next round's items: each lane owns
// one 32-vertex word, counts the items of its set bits (ceil(deg/chunk) per
// live list, the same for hidden lists), the wave scans the counts and takes
// its queue range with a single atomicAdd.  Delta-stepping (delta > 0): only
// vertices with dist < ctrl[C_THR] are expanded; the others keep their bit
// for a later round, and when a round expands nothing the last workgroup
// moves the bound to the bucket of the smallest deferred distance.
// Light-edge mode (lrow and pbits non-null; delta 0, not split): a set bit
// queues the vertex's light row and moves the bit to pbits; a pbits bit
// without a bits bit queues the full row and is cleared.  Pending bits keep
// ctrl[C_FAR] (and the slot's C_SLOTS bit) set.
template <bool PRUNED, bool SPLIT>
__global__ void __launch_bounds__(256)
sssp_expand_k(const int32_t* __restrict__ row, const int32_t* __restrict__ hrow,
              uint32_t* __restrict__ bits, uint32_t* __restrict__ hbits, int4* __restrict__ q,
              int64_t* __restrict__ ctrl, int64_t V, int chunk, int64_t qcap,
              const int32_t* __restrict__ dist, int delta, int slots, int64_t nwp,
              const int32_t* __restrict__ split, const int32_t* __restrict__ hsplit,
              const int32_t* __restrict__ lrow, uint32_t* __restrict__ pbits) {
    // NP queues: one, or in split mode one per target partition (queue p at
    // q + p * qcap, qcap items each)
    constexpr int NP = SPLIT ? kParts : 1;
    const int lane = lane_id();
    const int64_t nw = (V + 31) >> 5;
    const int64_t nwr = (int64_t)slots * nwp;  // nwp is a multiple of 64: a wave stays in one slot
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int32_t T = (int32_t)ctrl[C_THR];
    const bool lt = !SPLIT && lrow != nullptr && pbits != nullptr;
    int64_t front = 0, edges = 0, hedges = 0, near_n = 0;
    int32_t fmin = INT32_MAX;
    // Edge range of partition p of v's live (or hidden) list.
    auto range = [&](bool hidden, int v, int p, int& b, int& e) {
        if (SPLIT) {
            const int32_t* sp = (hidden ? hsplit : split) + (int64_t)v * kParts;
            b = sp[p];
            e = sp[p + 1];
        } else {
            const int32_t* r = hidden ? hrow : row;
            b = r[v];
            e = r[v + 1];
        }
    };
    // The loop's trip count is the same for every thread of a workgroup, so
    // its barriers are reached by all of them.
    __shared__ int64_t s_tot[16][NP];
    __shared__ int64_t s_base[NP];
    // workgroup totals of the round's counters: one global atomic each per
    // workgroup (per-wave atomics on the control block's two cache lines
    // serialised the kernel)
    __shared__ unsigned long long s_acc[5];  // front, edges, hidden edges, near, slot bits
    __shared__ int s_fmin;
    if (threadIdx.x < 5) s_acc[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_fmin = INT32_MAX;
    __syncthreads();
    const int wid = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    uint64_t slotbits = 0;  // slots with items or pending vertices (wave-uniform)
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < nwr; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        const bool inr = i < nwr;
        const int slot = inr ? (int)(i / nwp) : 0;
        const int64_t wi = inr ? i - (int64_t)slot * nwp : nw;
        uint32_t wl = 0, wh = 0, wp = 0;
        if (wi < nw) {
            wl = bits[i];
            if (PRUNED) wh = hbits[i];
            if (lt) wp = pbits[i];
        }
        // light mode: full rows of the pending vertices not lowered again
        const uint32_t hv = lt ? wp & ~wl : 0u;
        if (lt && wi < nw && wp != wl) pbits[i] = wl;
        if (lt && wl) fmin = 0;  // pending vertices remain (ctrl[C_FAR])
        // Split by distance: near bits are expanded now, far bits stay.
        uint32_t nl = wl, nh = wh;
        if (delta > 0 && (wl | wh)) {
            nl = nh = 0;
            for (uint32_t m = wl | wh; m; m &= m - 1) {
                const int b = __builtin_ctz(m);
                const int32_t d = dist[(int64_t)slot * V + wi * 32 + b];
                if (d < T) { nl |= wl & (1u << b); nh |= wh & (1u << b); }
                else fmin = min(fmin, d);
            }
        }
        if (wi < nw) {
            if (wl != (wl & ~nl)) bits[i] = wl & ~nl;
            if (PRUNED && wh != (wh & ~nh)) hbits[i] = wh & ~nh;
        }
        const int tag = slot << 1;
        int cnt[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) cnt[p] = 0;
        if constexpr (!SPLIT) {
            auto count = [&](int, int b, int e) {
                front += 1;
                edges += e - b;
                cnt[0] += (e - b + chunk - 1) / chunk;
            };
            walk_rows4(nl, wi * 32, lt ? lrow : row, count);
            walk_rows4(hv, wi * 32, row, count);
            if (PRUNED)
                walk_rows4(nh, wi * 32, hrow, [&](int, int b, int e) {
                    hedges += e - b;
                    cnt[0] += (e - b + chunk - 1) / chunk;
                });
        } else {
        for (uint32_t m = nl; m; m &= m - 1) {
            const int v = (int)(wi * 32 + __builtin_ctz(m));
            front += 1;
            if (lt) {
                edges += lrow[v + 1] - lrow[v];
                cnt[0] += (lrow[v + 1] - lrow[v] + chunk - 1) / chunk;
                continue;
            }
            edges += row[v + 1] - row[v];
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                int b, e;
                range(false, v, p, b, e);
                cnt[p] += (e - b + chunk - 1) / chunk;
            }
        }
        for (uint32_t m = hv; m; m &= m - 1) {
            const int v = (int)(wi * 32 + __builtin_ctz(m));
            front += 1;
            edges += row[v + 1] - row[v];
            cnt[0] += (row[v + 1] - row[v] + chunk - 1) / chunk;
        }
        if (PRUNED)
            for (uint32_t m = nh; m; m &= m - 1) {
                const int v = (int)(wi * 32 + __builtin_ctz(m));
                hedges += hrow[v + 1] - hrow[v];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    int b, e;
                    range(true, v, p, b, e);
                    cnt[p] += (e - b + chunk - 1) / chunk;
                }
            }
        }
        near_n += __popc(nl | nh | hv);
        int off[NP], tot[NP];
        int any = 0;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            off[p] = wave_excl_scan(cnt[p], &tot[p]);
            any |= tot[p];
        }
        // one queue reservation per workgroup and queue: wave totals meet in LDS
        if (lane == 0)
            for (int p = 0; p < NP; ++p) s_tot[wid][p] = tot[p];
        __syncthreads();
        if ((int)threadIdx.x < NP) {
            const int p = threadIdx.x;
            int64_t sum = 0;
            for (int w = 0; w < nwv; ++w) sum += s_tot[w][p];
            s_base[p] = sum > 0 ? (int64_t)atomicAdd((unsigned long long*)&ctrl[SPLIT ? C_QN + p : C_CUR],
                                                     (unsigned long long)sum)
                                : 0;
        }
        __syncthreads();
        int64_t pos[NP];
        bool room = true;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            int64_t base = s_base[p];
            for (int w = 0; w < wid; ++w) base += s_tot[w][p];
            room = room && base + tot[p] <= qcap;
            pos[p] = base + off[p];
        }
        __syncthreads();  // s_tot and s_base are rewritten in the next iteration
        // a slot with pending vertices (light mode) still has work
        const bool pend = __ballot(lt && wl != 0) != 0;
        if (any || pend) slotbits |= 1ull << slot;
        if (any == 0) continue;
        if (!room) {
            if (lane == 0) atomicOr((unsigned long long*)&ctrl[C_OVF], 1ull);
            continue;
        }
        if constexpr (!SPLIT) {
            // The wave writes its items together, one vertex per lane per
            // step (each lane takes the next vertex of its own word), item j
            // of a step by lane j % 64: consecutive items from consecutive
            // lanes, where per-lane runs scattered every store.  The wave's
            // items fill [wave base, wave base + total) in step order.
            int64_t pw = pos[0] - off[0];  // the wave's base (uniform)
            auto emit = [&](uint32_t mm, const int32_t* __restrict__ r, int kind) {
                while (__ballot(mm != 0)) {
                    int v = 0, b = 0, e = 0;
                    if (mm) {
                        v = (int)(wi * 32 + __builtin_ctz(mm));
                        mm &= mm - 1;
                        b = r[v];
                        e = r[v + 1];
                    }
                    const int c = e > b ? (e - b + chunk - 1) / chunk : 0;
                    int ctot;
                    const int cpre = wave_excl_scan(c, &ctot);
                    for (int j0 = 0; j0 < ctot; j0 += kWave) {
                        const int j = j0 + lane;
                        const int o = wave_owner(cpre, min(j, ctot - 1));
                        const int kk = j - __shfl(cpre, o, 64);
                        const int vo = __shfl(v, o, 64), bo = __shfl(b, o, 64), eo = __shfl(e, o, 64);
                        if (j < ctot) {
                            const int x = bo + kk * chunk;
                            q[pw + j] = make_int4(vo, x, min(x + chunk, eo), tag | kind);
                        }
                    }
                    pw += ctot;
                }
            };
            emit(nl, lt ? lrow : row, 0);
            emit(hv, row, 0);
            if (PRUNED) emit(nh, hrow, 1);
            continue;
        }
        for (uint32_t m = nl; m; m &= m - 1) {
            const int v = (int)(wi * 32 + __builtin_ctz(m));
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                int b0, e;
                range(false, v, p, b0, e);
                for (int b = b0; b < e; b += chunk) q[p * qcap + pos[p]++] = make_int4(v, b, min(b + chunk, e), tag);
            }
        }
        for (uint32_t m = hv; m; m &= m - 1) {
            const int v = (int)(wi * 32 + __builtin_ctz(m));
            for (int b = row[v], e = row[v + 1]; b < e; b += chunk) q[pos[0]++] = make_int4(v, b, min(b + chunk, e), tag);
        }
        if (PRUNED)
            for (uint32_t m = nh; m; m &= m - 1) {
                const int v = (int)(wi * 32 + __builtin_ctz(m));
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    int b0, e;
                    range(true, v, p, b0, e);
                    for (int b = b0; b < e; b += chunk)
                        q[p * qcap + pos[p]++] = make_int4(v, b, min(b + chunk, e), tag | 1);
                }
            }
    }
    front = wave_sum64(front);
    edges = wave_sum64(edges);
    if (PRUNED) hedges = wave_sum64(hedges);
    near_n = wave_sum64(near_n);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) fmin = min(fmin, __shfl_xor(fmin, d, 64));
    if (lane == 0) {
        if (front) atomicAdd(&s_acc[0], (unsigned long long)front);
        if (edges) atomicAdd(&s_acc[1], (unsigned long long)edges);
        if (hedges) atomicAdd(&s_acc[2], (unsigned long long)hedges);
        if (near_n) atomicAdd(&s_acc[3], (unsigned long long)near_n);
        if (slotbits) atomicOr(&s_acc[4], (unsigned long long)slotbits);
        if (fmin != INT32_MAX) atomicMin(&s_fmin, fmin);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_acc[0] | s_acc[2]) {
            atomicAdd((unsigned long long*)&ctrl[C_FRONT], s_acc[0]);
            atomicAdd((unsigned long long*)&ctrl[C_EDGES], s_acc[1]);
            if (PRUNED) atomicAdd((unsigned long long*)&ctrl[C_HEDGES], s_acc[2]);
        }
        if (s_acc[3]) atomicAdd((unsigned long long*)&ctrl[C_NEAR], s_acc[3]);
        if (s_acc[4]) atomicOr((unsigned long long*)&ctrl[C_SLOTS], s_acc[4]);
        if (s_fmin != INT32_MAX) atomicMin((long long*)&ctrl[C_FMIN], (long long)s_fmin);
        wait_vmem();  // this workgroup's counts are performed before it counts off
        if (wg_last(ctrl + kCtrlTree)) {
            const int64_t fm = (int64_t)atomicAdd((unsigned long long*)&ctrl[C_FMIN], 0ull);
            const int64_t nearc = (int64_t)atomicAdd((unsigned long long*)&ctrl[C_NEAR], 0ull);
            ctrl[C_FAR] = fm != INT32_MAX ? 1 : 0;
            if (delta > 0 && nearc == 0 && fm != INT32_MAX)
                ctrl[C_THR] = min((int64_t)INT32_MAX, (fm / delta + 1) * delta);
            ctrl[C_FMIN] = INT32_MAX;
            ctrl[C_NEAR] = 0;
            // items past a queue's end were dropped: stop, report
            bool ovf = false;
            int64_t nn = 0;
            for (int p = 0; p < NP; ++p) {
                const int64_t np = (int64_t)atomicAdd((unsigned long long*)&ctrl[SPLIT ? C_QN + p : C_CUR], 0ull);
                ovf = ovf || np > qcap;
                nn += np;
            }
            if (ovf) {
                ctrl[C_OVF] = 1;
                for (int p = 0; p < NP; ++p) ctrl[SPLIT ? C_QN + p : C_CUR] = 0;
                nn = 0;
            }
            ctrl[C_CUR] = nn;
        }
    }
}

// Checksum of a distance vector: sum of finite values and count of them.
// Checksum of a distance vector (element stride `stride`): sum of finite
// values, count of them, and the out-degree sum of reached vertices (m_s).
__global__ void dist_checksum_k(const int32_t* __restrict__ row, const int32_t* __restrict__ dist,
                                int64_t V, int64_t stride, int64_t* __restrict__ sum,
                                int32_t* __restrict__ cnt, int64_t* __restrict__ medges) {
    int64_t s = 0, c = 0, m = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * blockDim.x) {
        int32_t d = dist[i * stride];
        if (d != INT32_MAX) { s += d; c += 1; m += row[i + 1] - row[i]; }
    }
    s = wave_sum64(s);
    c = wave_sum64(c);
    m = wave_sum64(m);
    if (lane_id() == 0 && (s | c)) {
        atomicAdd((unsigned long long*)sum, (unsigned long long)s);
        atomicAdd((int*)cnt, (int)c);
        atomicAdd((unsigned long long*)medges, (unsigned long long)m);
    }
}

// ------------------------------------------------------------------ host
// Words per activation bitmap of one slot, padded to whole waves.
static int64_t bitmap_words(int64_t V) { return (((V + 31) / 32 + 1) + 63) & ~(int64_t)63; }

// Queue capacity of one partition in split mode: each vertex is activated at
// most once per round and has at most ceil(n/chunk) <= n/chunk + 1 items per
// partition of each list.
static int64_t split_qseg(const Graph* g) {
    const int chunk = std::max(1, g->tune.chunk);
    return 2 * std::max<int64_t>(g->V, 1) + (g->E + g->HE) / chunk + 64;
}

// The per-vertex partition offsets of the live (and hidden) rows, built on
// the host the first time split mode runs: split[8u + p] = first edge of row
// u whose target lies in partition p (targets sorted, p = v / ceil(V/8)),
// split[8V] = end of the last row.
static int ensure_split(Graph* g) {
    if (g->d_split && (!g->d_hrow || g->d_hsplit)) return C2G_OK;
    const int64_t V = g->V;
    const int64_t pw = (V + kParts - 1) / kParts;
    auto build = [&](const std::vector<int32_t>& row, const std::vector<int32_t>& el, int32_t** dst) -> int {
        std::vector<int32_t> sp((size_t)(V + 1) * kParts);
        parallel_for(V, [&](int64_t lo, int64_t hi, int) {
            for (int64_t u = lo; u < hi; ++u) {
                const int32_t* b = el.data() + row[u];
                const int32_t* e = el.data() + row[u + 1];
                sp[u * kParts] = row[u];
                // a row not sorted by target (a caller's CSR may not be) goes
                // whole into partition 0: the eight ranges must tile the row
                const bool sorted = std::is_sorted(b, e);
                for (int p = 1; p < kParts; ++p)
                    sp[u * kParts + p] =
                        sorted ? (int32_t)(std::lower_bound(b, e, (int32_t)std::min<int64_t>(p * pw, INT32_MAX)) -
                                           el.data())
                               : row[u + 1];
            }
        }, 0, 1 << 14);
        for (int p = 0; p < kParts; ++p) sp[V * kParts + p] = row[V];
        HIP_TRY(hipMalloc(dst, sizeof(int32_t) * sp.size()));
        HIP_TRY(hipMemcpy(*dst, sp.data(), sizeof(int32_t) * sp.size(), hipMemcpyHostToDevice));
        return C2G_OK;
    };
    HIP_TRY(hipSetDevice(g->device));
    int rc;
    if (!g->d_split && (rc = build(g->vl, g->el, &g->d_split))) return rc;
    if (g->d_hrow && !g->d_hsplit && (rc = build(g->hvl, g->hel, &g->d_hsplit))) return rc;
    return C2G_OK;
}

static bool split_mode(const Graph* g) { return g->tune.xcd_split != 0; }

int ensure_sssp_ws(Graph* g) {
    SsspWs& w = g->sssp;
    const int64_t V = g->V;
    const int chunk = std::max(1, g->tune.chunk);
    // Each vertex is activated at most once per round for its live list and
    // once for its hidden list: at most ceil(d/chunk) <= d/chunk + 1 items each.
    // Split mode: kParts queues of split_qseg items.
    const int64_t need = split_mode(g) ? kParts * split_qseg(g) : 2 * V + (g->E + g->HE) / chunk + 64;
    if (split_mode(g)) {
        int rc = ensure_split(g);
        if (rc) return rc;
    }
    const int64_t nwp = bitmap_words(V);
    HIP_TRY(hipSetDevice(g->device));
    // Each piece is checked on its own: after a failed allocation the next
    // call allocates what is missing instead of using it.
    if (!w.dist) HIP_TRY(hipMalloc(&w.dist, sizeof(int32_t) * std::max<int64_t>(V, 1)));
    // live, hidden and pending (light mode) activation bitmaps
    if (!w.bits) HIP_TRY(hipMalloc(&w.bits, sizeof(uint32_t) * nwp * 3));
    if (!w.ctrl) HIP_TRY(hipMalloc(&w.ctrl, sizeof(int64_t) * kCtrlWords));
    if (!w.ev[0]) HIP_TRY(hipEventCreate(&w.ev[0]));
    if (!w.ev[1]) HIP_TRY(hipEventCreate(&w.ev[1]));
    if (!w.h_ctrl) HIP_TRY(hipHostMalloc((void**)&w.h_ctrl, sizeof(int64_t) * C_N, hipHostMallocDefault));
    if (g->d_hrow && !w.hid) HIP_TRY(hipMalloc(&w.hid, std::max<int64_t>(V, 1)));
    if (w.qcap < need) {
        if (w.exec) { (void)hipGraphExecDestroy(w.exec); w.exec = nullptr; }
        if (w.q[0]) { (void)hipFree(w.q[0]); w.q[0] = nullptr; }
        w.qcap = 0;
        HIP_TRY(hipMalloc(&w.q[0], sizeof(int4) * need));
        w.qcap = need;
    }
    return C2G_OK;
}

static int relax_grid(const Graph* g) {
    const int per_cu = g->tune.grid_per_cu_single > 0 ? g->tune.grid_per_cu_single : g->tune.grid_per_cu;
    const int grid = std::max(1, g->num_cus) * per_cu;
    return split_mode(g) ? (grid + kParts - 1) / kParts * kParts : grid;  // split: a multiple of 8
}

static bool is_pruned(const Graph* g) { return g->d_hrow != nullptr && g->d_hadj != nullptr; }

// Device buffers of one single-query traversal: the handle's workspace, or
// the caller's distance array plus the static scratch (c2g_sssp_noalloc).
struct SsspBufs {
    int32_t* dist;     // [V]
    uint32_t* bits;    // live [nwp], then hidden [nwp], then pending (light mode) [nwp]
    uint8_t* hid;      // [V] (pruned graphs)
    int4* q;           // frontier items
    int64_t qcap;      // items per queue (split mode: per partition)
    int64_t* ctrl;     // [kCtrlWords]
    const int32_t* split = nullptr;   // split mode: partition offsets (null: one queue)
    const int32_t* hsplit = nullptr;
};

// The workspace's buffers; split mode when the handle's tunables ask for it
// and `allow_split` (budgeted runs read back one queue).
static SsspBufs ws_bufs(Graph* g, bool allow_split = true) {
    SsspWs& w = g->sssp;
    SsspBufs b{w.dist, w.bits, w.hid, w.q[0], w.qcap, w.ctrl};
    if (allow_split && split_mode(g) && g->d_split) {
        b.split = g->d_split;
        b.hsplit = g->d_hsplit;
        b.qcap = split_qseg(g);
    }
    return b;
}

// One relax launch with the tuned workgroup size; hbits/hid null on full
// graphs.  Items may belong to several slots (slot pools).
// qseg > 0: split mode, kParts queues of qseg items (grid a multiple of 8).
static void launch_relax(Graph* g, hipStream_t strm, int grid, int32_t* dist, uint32_t* bits, uint32_t* hbits,
                         uint8_t* hid, const int4* q, int64_t* ctrl, int64_t nwp, int64_t qseg = 0) {
    auto go = [&](auto pruned, auto block, auto split) {
        constexpr bool P = decltype(pruned)::value;
        constexpr int B = decltype(block)::value;
        constexpr bool S = decltype(split)::value;
        hipLaunchKernelGGL((sssp_relax_k<P, B, S>), dim3(grid), dim3(B), 0, strm, g->d_adj,
                           P ? g->d_hadj : (const int2*)nullptr, dist, bits, hbits, hid, q, ctrl, (int64_t)g->V, nwp,
                           qseg);
    };
    auto tb = [&](auto pruned, auto split) {
        switch (g->tune.block_single > 0 ? g->tune.block_single : g->tune.block) {
            case 128: go(pruned, std::integral_constant<int, 128>(), split); break;
            case 512: go(pruned, std::integral_constant<int, 512>(), split); break;
            default: go(pruned, std::integral_constant<int, 256>(), split); break;
        }
    };
    auto sp = [&](auto pruned) {
        if (qseg > 0) tb(pruned, std::true_type());
        else tb(pruned, std::false_type());
    };
    if (hbits) sp(std::true_type());
    else sp(std::false_type());
}

// Launches k rounds (relax + expand) on `b`.  With relax_ms non-null
// (profiling), the workspace's HIP events bracket each relax kernel and the
// time is accumulated.
static int launch_rounds_on(Graph* g, const SsspBufs& b, int k, double* relax_ms) {
    SsspWs& w = g->sssp;
    const bool pruned = is_pruned(g);
    const int grid = relax_grid(g);
    const int64_t nwp = bitmap_words(g->V);
    uint32_t* bits = b.bits;
    uint32_t* hbits = bits + nwp;
    const int egrid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (nwp + 255) / 256));
    const bool split = b.split != nullptr;
    const int32_t* lrow = split ? nullptr : light_rows(g);
    uint32_t* pbits = lrow ? bits + 2 * nwp : nullptr;
    for (int r = 0; r < k; ++r) {
        if (relax_ms) HIP_TRY(hipEventRecord(w.ev[0], g->stream));
        launch_relax(g, g->stream, grid, b.dist, bits, pruned ? hbits : nullptr, pruned ? b.hid : nullptr, b.q, b.ctrl,
                     nwp, split ? b.qcap : 0);
        if (relax_ms) HIP_TRY(hipEventRecord(w.ev[1], g->stream));
        auto expand = [&](auto pr, auto sp) {
            constexpr bool P = decltype(pr)::value;
            constexpr bool S = decltype(sp)::value;
            hipLaunchKernelGGL((sssp_expand_k<P, S>), dim3(egrid), dim3(256), 0, g->stream, g->d_row,
                               P ? g->d_hrow : (const int32_t*)nullptr, bits, P ? hbits : (uint32_t*)nullptr, b.q,
                               b.ctrl, (int64_t)g->V, g->tune.chunk, b.qcap, (const int32_t*)b.dist, g->tune.delta, 1,
                               nwp, b.split, P ? b.hsplit : (const int32_t*)nullptr, lrow, pbits);
        };
        if (pruned) {
            if (split) expand(std::true_type(), std::true_type());
            else expand(std::true_type(), std::false_type());
        } else {
            if (split) expand(std::false_type(), std::true_type());
            else expand(std::false_type(), std::false_type());
        }
        if (relax_ms) {
            HIP_TRY(hipEventSynchronize(w.ev[1]));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
            *relax_ms += ms;
        }
    }
    HIP_TRY(hipGetLastError());
    return C2G_OK;
}

static int launch_rounds(Graph* g, int k, double* relax_ms = nullptr, bool allow_split = true) {
    return launch_rounds_on(g, ws_bufs(g, allow_split), k, relax_ms);
}

// Clears `b` for a query from src and queues the source's items.
static void launch_seed(Graph* g, const SsspBufs& b, int32_t src) {
    const bool pruned = is_pruned(g);
    const int64_t nwp = bitmap_words(g->V);
    const int tb = 256, gridv = (int)std::min<int64_t>((g->V + tb - 1) / tb, 65535);
    hipLaunchKernelGGL(sssp_init_k, dim3(std::max(gridv, 1)), dim3(tb), 0, g->stream, b.dist, b.bits,
                       pruned ? b.bits + nwp : (uint32_t*)nullptr, pruned ? b.hid : (uint8_t*)nullptr,
                       (int64_t)g->V, b.bits + 2 * nwp);
    hipLaunchKernelGGL(sssp_seed_k, dim3(1), dim3(256), 0, g->stream, g->d_row,
                       pruned ? g->d_hrow : (const int32_t*)nullptr, b.dist,
                       pruned ? b.hid : (uint8_t*)nullptr, b.q, b.ctrl, src, g->tune.chunk, g->tune.delta, b.split,
                       pruned ? b.hsplit : (const int32_t*)nullptr, b.qcap);
}

// A hipGraph holding `iters_per_graph` rounds.
static int ensure_exec(Graph* g) {
    SsspWs& w = g->sssp;
    const int K = g->tune.iters_per_graph;
    if (w.exec && w.exec_iters == K && w.exec_grid == relax_grid(g) && w.exec_chunk == g->tune.chunk &&
        w.exec_tb == (g->tune.block_single > 0 ? g->tune.block_single : g->tune.block) &&
        w.exec_split == g->tune.xcd_split)
        return C2G_OK;
    if (w.exec) { (void)hipGraphExecDestroy(w.exec); w.exec = nullptr; }
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_rounds(g, K);
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc) {
        if (e == hipSuccess) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&w.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        w.exec = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    w.exec_iters = K;
    w.exec_grid = relax_grid(g);
    w.exec_chunk = g->tune.chunk;
    w.exec_tb = g->tune.block_single > 0 ? g->tune.block_single : g->tune.block;
    w.exec_split = g->tune.xcd_split;
    return C2G_OK;
}

static void fill_stats(c2g_stats* st, const int64_t* c, double ms, int lane) {
    if (!st) return;
    st->lane = lane;
    st->iterations = (int32_t)c[C_ITER];
    st->frontier_total = c[C_FRONT];
    st->edges_relaxed = c[C_EDGES] + c[C_HEDGES];
    st->improvements = c[C_IMPR];
    st->items = c[C_ITEMS];
    st->ms = ms;
}

static int64_t reached_edges(const Graph* g, const int32_t* dist) {
    int64_t m = 0;
    for (int32_t v = 0; v < g->V; ++v)
        if (dist[v] != INT32_MAX) m += g->vl[v + 1] - g->vl[v];
    return m;
}

// Runs one query on the device.  budget > 0 stops after `budget` rounds and
// leaves the remaining frontier in frontier_out (vertices, deduplicated).
int sssp_run(Graph* g, int32_t src, int32_t budget, int32_t* dist_out, int32_t* frontier_out,
             int32_t* count_out, uint8_t* hid_out, c2g_stats* st, bool copy_out, RoundLog* log) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adj) return fail(C2G_ENOTSUP, "graph has no int weights (wlist) for SSSP");
    if (src < 0 || src >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", src, g->V);
    int rc = ensure_sssp_ws(g);
    if (rc) return rc;
    SsspWs& w = g->sssp;
    const bool pruned = is_pruned(g);
    const int64_t nwp = bitmap_words(g->V);
    uint32_t* bits = w.bits;
    const double t0 = now_ms();
    const bool allow_split = budget <= 0;  // a budgeted run reads its queue back as one list
    launch_seed(g, ws_bufs(g, allow_split), src);
    HIP_TRY(hipGetLastError());
    const bool prof = g->tune.profile != 0;
    const bool one = prof || log;  // one round per host check
    const bool use_graph = g->tune.use_graph && budget <= 0 && !one;
    if (use_graph && (rc = ensure_exec(g))) return rc;
    auto snap = [&](int64_t* q3) {
        q3[0] = w.h_ctrl[C_FRONT];
        q3[1] = w.h_ctrl[C_EDGES] + w.h_ctrl[C_HEDGES];
        q3[2] = q3[1];
    };
    if (log) {
        HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, sizeof(int64_t) * C_N, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        int64_t q3[3];
        snap(q3);
        log->start(q3);
    }
    int done_rounds = 0;
    double relax_ms = 0;
    for (;;) {
        int k = one ? 1 : g->tune.iters_per_graph;
        if (budget > 0) k = std::min(k, budget - done_rounds);
        if (k <= 0) break;
        if (log) HIP_TRY(hipEventRecord(w.ev[0], g->stream));
        if (use_graph) HIP_TRY(hipGraphLaunch(w.exec, g->stream));
        else if ((rc = launch_rounds(g, k, prof ? &relax_ms : nullptr, allow_split))) return rc;
        if (log) HIP_TRY(hipEventRecord(w.ev[1], g->stream));
        done_rounds += k;
        HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, sizeof(int64_t) * C_N, hipMemcpyDeviceToHost,
                               g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (w.h_ctrl[C_OVF]) return fail(C2G_EDEVICE, "frontier queue overflow (internal)");
        if (log) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, w.ev[0], w.ev[1]));
            int64_t q3[3];
            snap(q3);
            log->add(q3, w.h_ctrl[C_IMPR], w.h_ctrl[C_ITEMS], ms);
        }
        if (w.h_ctrl[C_CUR] == 0 && w.h_ctrl[C_FAR] == 0) break;
    }
    const double t1 = now_ms();
    int64_t c[C_N];
    memcpy(c, w.h_ctrl, sizeof c);
    if (copy_out && dist_out) {
        const int rc = w.out.copy_out(dist_out, w.dist, sizeof(int32_t) * g->V, g->stream);
        if (rc) return rc;
    }
    if (count_out) {
        // Remaining frontier (budgeted runs): distinct vertices of the current queue.
        int64_t n = c[C_CUR];
        std::vector<int4> items((size_t)n);
        if (n) HIP_TRY(hipMemcpy(items.data(), w.q[0], sizeof(int4) * n, hipMemcpyDeviceToHost));
        int32_t m = 0;
        std::vector<uint8_t> seen((size_t)g->V, 0);
        for (auto& it : items)
            if (!seen[it.x]) { seen[it.x] = 1; if (frontier_out) frontier_out[m] = it.x; ++m; }
        // vertices deferred by delta-stepping, or whose full row is pending
        // (light mode), are still pending
        const int64_t nwd = (g->V + 31) / 32;
        std::vector<uint32_t> bw((size_t)nwd * 3, 0u);
        HIP_TRY(hipMemcpy(bw.data(), bits, sizeof(uint32_t) * nwd, hipMemcpyDeviceToHost));
        if (pruned) HIP_TRY(hipMemcpy(bw.data() + nwd, bits + nwp, sizeof(uint32_t) * nwd, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bw.data() + 2 * nwd, bits + 2 * nwp, sizeof(uint32_t) * nwd, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < (int64_t)bw.size(); ++k)
            for (uint32_t x = bw[k]; x; x &= x - 1) {
                const int32_t v = (int32_t)((k % nwd) * 32 + __builtin_ctz(x));
                if (!seen[v]) { seen[v] = 1; if (frontier_out) frontier_out[m] = v; ++m; }
            }
        *count_out = m;
        if (hid_out && pruned)
            HIP_TRY(hipMemcpy(hid_out, w.hid, g->V, hipMemcpyDeviceToHost));
    }
    fill_stats(st, c, t1 - t0, C2G_LANE_GPU);
    if (st) st->kernel_ms = relax_ms;
    if (st && copy_out && dist_out) {
        int32_t r = 0;
        for (int32_t v = 0; v < g->V; ++v) r += dist_out[v] != INT32_MAX;
        st->reached = r;
        st->reached_edges = reached_edges(g, dist_out);
    }
    return C2G_OK;
}

// ------------------------------------------------------------ slot pool
// Clears slot s before a query joins it.
__global__ void slot_reset_k(int32_t* __restrict__ dist, uint8_t* __restrict__ hid, uint32_t* __restrict__ bits,
                             uint32_t* __restrict__ hbits, int64_t V, int64_t nw, uint32_t* __restrict__ pbits) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < V; i += (int64_t)gridDim.x * blockDim.x) {
        dist[i] = INT32_MAX;
        if (hid) hid[i] = 0;
        if (i < nw) {
            bits[i] = 0;
            if (hbits) hbits[i] = 0;
            pbits[i] = 0;
        }
    }
}

// Queues the source's items (live list, and hidden list on pruned graphs) at
// the end of the pending queue, tagged with the slot.
__global__ void slot_seed_k(const int32_t* __restrict__ row, const int32_t* __restrict__ hrow,
                            int32_t* __restrict__ dist, uint8_t* __restrict__ hid, int4* __restrict__ q,
                            int64_t* __restrict__ ctrl, int64_t qcap, int slot, int32_t src, int chunk) {
    __shared__ int64_t base;
    const int r0 = row[src], r1 = row[src + 1];
    const int nl = (r1 - r0 + chunk - 1) / chunk;
    int h0 = 0, h1 = 0, nh = 0;
    if (hrow) { h0 = hrow[src]; h1 = hrow[src + 1]; nh = (h1 - h0 + chunk - 1) / chunk; }
    if (threadIdx.x == 0) {
        dist[src] = 0;
        if (hid) hid[src] = 1;  // the source's hidden list is always relaxed
        base = (int64_t)atomicAdd((unsigned long long*)&ctrl[C_CUR], (unsigned long long)(nl + nh));
        if (base + nl + nh > qcap) ctrl[C_OVF] = 1;
    }
    __syncthreads();
    if (base + nl + nh > qcap) return;
    const int tag = slot << 1;
    for (int j = threadIdx.x; j < nl + nh; j += blockDim.x) {
        if (j < nl) {
            const int b = r0 + j * chunk;
            q[base + j] = make_int4(src, b, min(b + chunk, r1), tag);
        } else {
            const int b = h0 + (j - nl) * chunk;
            q[base + j] = make_int4(src, b, min(b + chunk, h1), tag | 1);
        }
    }
}

struct SlotPool {
    int slots = 0;
    int64_t nwp = 0, qcap = 0;
    int32_t* dist = nullptr;   // [slots][V]
    uint32_t* bits = nullptr;  // live [slots][nwp], then hidden [slots][nwp], then pending [slots][nwp]
    uint8_t* hid = nullptr;    // [slots][V] (pruned graphs)
    int4* q = nullptr;
    int64_t* ctrl = nullptr;
    int64_t* h_ctrl = nullptr;
    HostStage out;
};
static std::mutex g_sp_mu;
static std::vector<std::pair<const Graph*, SlotPool*>> g_sp;

static SlotPool* spool(Graph* g) {
    std::lock_guard<std::mutex> lk(g_sp_mu);
    for (auto& p : g_sp)
        if (p.first == g) return p.second;
    auto* s = new SlotPool;
    g_sp.emplace_back(g, s);
    return s;
}

void qserve_release(Graph* g) {
    std::lock_guard<std::mutex> lk(g_sp_mu);
    for (size_t i = 0; i < g_sp.size(); ++i)
        if (g_sp[i].first == g) {
            SlotPool* s = g_sp[i].second;
            for (void* p : {(void*)s->dist, (void*)s->bits, (void*)s->hid, (void*)s->q, (void*)s->ctrl})
                if (p) (void)hipFree(p);
            if (s->h_ctrl) (void)hipHostFree(s->h_ctrl);
            s->out.release();
            delete s;
            g_sp.erase(g_sp.begin() + i);
            return;
        }
}

// Device bytes of a slot pool of `slots` queries on this graph.
static int64_t qserve_bytes(const Graph* g, int slots) {
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.chunk);
    const int64_t qcap = (int64_t)slots * (2 * V + (g->E + g->HE) / chunk) + 64;
    return 4 * slots * V + 12 * slots * bitmap_words(V) + (g->hvl.empty() ? 0 : slots * V) + 16 * qcap +
           8 * kCtrlWords;
}

// Device bytes a slot pool of `slots` would add to what the handle holds.
int64_t qserve_ws_bytes(Graph* g, int slots) {
    if (slots <= 0) return 0;
    std::lock_guard<std::mutex> lk(g_sp_mu);
    for (auto& p : g_sp)
        if (p.first == g && p.second->slots == slots) return 0;
    return qserve_bytes(g, slots);
}

// Sizes the pool for `slots` queries (a pool of another size is freed
// first) and clears its control block and bitmaps, so a pool left by an
// earlier server, stopped with queries in flight, starts empty.
int qserve_init(Graph* g, hipStream_t strm, int slots) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adj) return fail(C2G_ENOTSUP, "graph has no int weights (wlist) for SSSP");
    if (slots < 1 || slots > 64) return fail(C2G_EINVAL, "slot pool size must be in [1,64]");
    SlotPool* s = spool(g);
    HIP_TRY(hipSetDevice(g->device));
    const int64_t needq = (int64_t)slots * (2 * std::max<int64_t>(g->V, 1) + (g->E + g->HE) / std::max(1, g->tune.chunk)) + 64;
    if (s->slots == slots && s->qcap >= needq) {
        HIP_TRY(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 3 * slots * s->nwp, strm));
        HIP_TRY(hipMemsetAsync(s->ctrl, 0, sizeof(int64_t) * kCtrlWords, strm));
        HIP_TRY(hipStreamSynchronize(strm));
        return C2G_OK;
    }
    s->slots = 0;  // unusable until every piece below is in place
    // pieces left by an earlier attempt that failed part way
    for (void* p : {(void*)s->dist, (void*)s->bits, (void*)s->hid, (void*)s->q, (void*)s->ctrl})
        if (p) (void)hipFree(p);
    if (s->h_ctrl) (void)hipHostFree(s->h_ctrl);
    s->dist = nullptr; s->bits = nullptr; s->hid = nullptr; s->q = nullptr; s->ctrl = nullptr; s->h_ctrl = nullptr;
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.chunk);
    s->nwp = bitmap_words(V);
    // per slot and round: each vertex at most once per list, ceil(d/chunk) <= d/chunk + 1 items
    s->qcap = (int64_t)slots * (2 * V + (g->E + g->HE) / chunk) + 64;
    HIP_TRY(hipMalloc(&s->dist, sizeof(int32_t) * slots * V));
    HIP_TRY(hipMalloc(&s->bits, sizeof(uint32_t) * 3 * slots * s->nwp));
    HIP_TRY(hipMemsetAsync(s->bits, 0, sizeof(uint32_t) * 3 * slots * s->nwp, strm));
    if (g->d_hrow) HIP_TRY(hipMalloc(&s->hid, (size_t)slots * V));
    HIP_TRY(hipMalloc(&s->q, sizeof(int4) * s->qcap));
    HIP_TRY(hipMalloc(&s->ctrl, sizeof(int64_t) * kCtrlWords));
    HIP_TRY(hipMemsetAsync(s->ctrl, 0, sizeof(int64_t) * kCtrlWords, strm));
    HIP_TRY(hipHostMalloc((void**)&s->h_ctrl, sizeof(int64_t) * C_N, hipHostMallocDefault));
    HIP_TRY(hipStreamSynchronize(strm));
    s->slots = slots;
    return C2G_OK;
}

int qserve_admit(Graph* g, hipStream_t strm, int slot, int32_t src) {
    SlotPool* s = spool(g);
    if (slot < 0 || slot >= s->slots || src < 0 || src >= g->V) return fail(C2G_EINVAL, "bad slot/source");
    const bool pruned = is_pruned(g);
    const int64_t V = g->V;
    const int grid = (int)std::min<int64_t>((V + 255) / 256, (int64_t)std::max(1, g->num_cus) * 8);
    hipLaunchKernelGGL(slot_reset_k, dim3(std::max(grid, 1)), dim3(256), 0, strm, s->dist + slot * V,
                       pruned ? s->hid + slot * V : (uint8_t*)nullptr, s->bits + slot * s->nwp,
                       pruned ? s->bits + (s->slots + slot) * s->nwp : (uint32_t*)nullptr, V, (V + 31) / 32,
                       s->bits + (2 * s->slots + slot) * s->nwp);
    hipLaunchKernelGGL(slot_seed_k, dim3(1), dim3(256), 0, strm, g->d_row,
                       pruned ? g->d_hrow : (const int32_t*)nullptr, s->dist + slot * V,
                       pruned ? s->hid + slot * V : (uint8_t*)nullptr, s->q, s->ctrl, s->qcap, slot, src,
                       std::max(1, g->tune.chunk));
    HIP_TRY(hipGetLastError());
    return C2G_OK;
}

int qserve_step(Graph* g, hipStream_t strm, uint64_t* slots_with_work) {
    SlotPool* s = spool(g);
    const bool pruned = is_pruned(g);
    const int64_t V = g->V;
    const int grid = std::max(1, g->num_cus) * g->tune.grid_per_cu;
    const int64_t words = (int64_t)s->slots * s->nwp;
    const int egrid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (words + 255) / 256));
    uint32_t* hb = s->bits + (int64_t)s->slots * s->nwp;
    // light mode needs delta 0, which the pool always runs with
    const int32_t* lrow = g->tune.light > 0 && g->light_built == g->tune.light ? g->d_lrow : nullptr;
    uint32_t* pb = lrow ? s->bits + 2 * (int64_t)s->slots * s->nwp : nullptr;
    launch_relax(g, strm, grid, s->dist, s->bits, pruned ? hb : nullptr, pruned ? s->hid : nullptr, s->q, s->ctrl,
                 s->nwp);
    if (pruned) {
        hipLaunchKernelGGL((sssp_expand_k<true, false>), dim3(egrid), dim3(256), 0, strm, g->d_row, g->d_hrow, s->bits,
                           hb, s->q, s->ctrl, V, std::max(1, g->tune.chunk), s->qcap, (const int32_t*)s->dist, 0,
                           s->slots, s->nwp, (const int32_t*)nullptr, (const int32_t*)nullptr, lrow, pb);
    } else {
        hipLaunchKernelGGL((sssp_expand_k<false, false>), dim3(egrid), dim3(256), 0, strm, g->d_row,
                           (const int32_t*)nullptr, s->bits, (uint32_t*)nullptr, s->q, s->ctrl, V,
                           std::max(1, g->tune.chunk), s->qcap, (const int32_t*)s->dist, 0, s->slots, s->nwp,
                           (const int32_t*)nullptr, (const int32_t*)nullptr, lrow, pb);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(s->h_ctrl, s->ctrl, sizeof(int64_t) * C_N, hipMemcpyDeviceToHost, strm));
    HIP_TRY(hipStreamSynchronize(strm));
    if (s->h_ctrl[C_OVF]) return fail(C2G_EDEVICE, "server slot pool capacity exceeded (internal)");
    *slots_with_work = (uint64_t)s->h_ctrl[C_SLOTS];
    return C2G_OK;
}

int qserve_extract(Graph* g, hipStream_t strm, int slot, int32_t* out) {
    SlotPool* s = spool(g);
    if (slot < 0 || slot >= s->slots) return fail(C2G_EINVAL, "bad slot");
    return s->out.copy_out(out, s->dist + (int64_t)slot * g->V, sizeof(int32_t) * g->V, strm);
}

// ------------------------------------------------------- no-alloc variant
// Scratch of c2g_sssp_noalloc: control block, checksum words, activation
// bitmaps, hidden flags and the frontier queue of one traversal, in one
// file-scope device array.  Each device holds its own copy for the life of
// the process; host code reaches it only through the address that
// hipGetSymbolAddress returned on that device (kept in g_scr), never through
// the array's name.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scr_mu;
static void* g_scr[kMaxDevices];

// Looks up this device's copy of the scratch array; `device` must be current.
// Called once per device when a handle is uploaded (c2g_open and
// c2g_graph_from_csr), before any no-alloc query can run there.
int sssp_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(C2G_EINVAL, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (g_scr[device]) return C2G_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scr[device] = p;
    return C2G_OK;
}

static void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scr_mu);
    return g_scr[device];
}

// Byte layout of the scratch for one graph; returns the total.
struct ScratchLayout {
    size_t ctrl, sums, bits, hid, q, total;
    int64_t qcap;
};
static ScratchLayout scratch_layout(const Graph* g) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    ScratchLayout L{};
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.chunk);
    L.qcap = 2 * V + (g->E + g->HE) / chunk + 64;  // the workspace's own bound (ensure_sssp_ws)
    L.ctrl = 0;
    L.sums = up(sizeof(int64_t) * kCtrlWords);
    L.bits = L.sums + up(sizeof(int64_t) * 4);
    L.hid = L.bits + up(sizeof(uint32_t) * 3 * bitmap_words(V));
    L.q = L.hid + (is_pruned(g) ? up((size_t)V) : 0);
    L.total = L.q + sizeof(int4) * (size_t)L.qcap;
    return L;
}

__global__ void zero_i64_k(int64_t* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// One query into the caller's device array d_dist[V], all other state in the
// scratch.  Plain launches (no captured graph), counters read back into host
// stack memory.
static int sssp_noalloc_run(Graph* g, int32_t src, int32_t* d_dist, c2g_stats* st) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adj) return fail(C2G_ENOTSUP, "graph has no int weights (wlist) for SSSP");
    if (src < 0 || src >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", src, g->V);
    const ScratchLayout L = scratch_layout(g);
    if (L.total > kScratchBytes)
        return fail(C2G_EINVAL,
                    "graph too large for c2g_sssp_noalloc: needs %zu bytes of scratch (V=%d, E=%lld, "
                    "hidden E=%lld, chunk %d), the static scratch holds %zu",
                    L.total, g->V, (long long)g->E, (long long)g->HE, std::max(1, g->tune.chunk), kScratchBytes);
    unsigned char* base = (unsigned char*)scratch_ptr(g->device);
    if (!base) return fail(C2G_EDEVICE, "no-alloc scratch was not set up on device %d", g->device);
    HIP_TRY(hipSetDevice(g->device));
    const SsspBufs b{d_dist, (uint32_t*)(base + L.bits), is_pruned(g) ? (uint8_t*)(base + L.hid) : nullptr,
                     (int4*)(base + L.q), L.qcap, (int64_t*)(base + L.ctrl)};
    int64_t* sums = (int64_t*)(base + L.sums);
    const double t0 = now_ms();
    launch_seed(g, b, src);  // clears the bitmaps, hidden flags and control block
    HIP_TRY(hipGetLastError());
    int64_t c[C_N];
    for (;;) {
        int rc = launch_rounds_on(g, b, std::max(1, g->tune.iters_per_graph), nullptr);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(c, b.ctrl, sizeof c, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (c[C_OVF]) return fail(C2G_EDEVICE, "frontier queue overflow (internal)");
        if (c[C_CUR] == 0 && c[C_FAR] == 0) break;
    }
    const double t1 = now_ms();
    if (st) {
        // sums[0] = sum of finite distances, sums[1] = reached count (int32
        // in its low word), sums[2] = out-degree sum of reached vertices
        hipLaunchKernelGGL(zero_i64_k, dim3(1), dim3(64), 0, g->stream, sums, 4);
        hipLaunchKernelGGL(dist_checksum_k, dim3(std::max(1, g->num_cus) * 4), dim3(256), 0, g->stream, g->d_row,
                           (const int32_t*)d_dist, (int64_t)g->V, (int64_t)1, sums, (int32_t*)(sums + 1), sums + 2);
        HIP_TRY(hipGetLastError());
        int64_t h[4];
        HIP_TRY(hipMemcpyAsync(h, sums, sizeof h, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        fill_stats(st, c, t1 - t0, C2G_LANE_GPU);
        st->reached = (int32_t)h[1];
        st->reached_edges = h[2];
    }
    return C2G_OK;
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_sssp(c2g_graph* h, int32_t src, int32_t* dist_out, c2g_stats* st) {
    if (!h || !dist_out) return fail(C2G_EINVAL, "c2g_sssp: null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return sssp_run(&h->g, src, 0, dist_out, nullptr, nullptr, nullptr, st, true);
}

int c2g_sssp_noalloc(c2g_graph* h, int32_t src, int32_t* d_dist, c2g_stats* st) {
    if (!h || !d_dist) return fail(C2G_EINVAL, "c2g_sssp_noalloc: null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return sssp_noalloc_run(&h->g, src, d_dist, st);
}

int c2g_sssp_rounds(c2g_graph* h, int32_t src, int32_t* dist_out, c2g_round* rounds, int32_t cap,
                    int32_t* nrounds, c2g_stats* st) {
    if (!h || !dist_out || (cap > 0 && !rounds) || cap < 0) return fail(C2G_EINVAL, "c2g_sssp_rounds: bad argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    RoundLog log;
    log.out = rounds;
    log.cap = cap;
    int rc = sssp_run(&h->g, src, 0, dist_out, nullptr, nullptr, nullptr, st, true, &log);
    if (nrounds) *nrounds = log.n;
    return rc;
}

int c2g_sssp_budget(c2g_graph* h, int32_t src, int32_t budget, int32_t* dist_out,
                    int32_t* frontier_out, int32_t* count_out, c2g_stats* st) {
    if (!h || !dist_out || !frontier_out || !count_out) return fail(C2G_EINVAL, "null argument");
    if (budget < 1) return fail(C2G_EINVAL, "budget must be >= 1");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return sssp_run(&h->g, src, budget, dist_out, frontier_out, count_out, nullptr, st, true);
}

int c2g_sssp_migrate(c2g_graph* h, int32_t src, int32_t budget, int32_t* dist_out, c2g_stats* st) {
    if (!h || !dist_out) return fail(C2G_EINVAL, "null argument");
    if (budget < 1) return fail(C2G_EINVAL, "budget must be >= 1");
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    std::vector<int32_t> fr((size_t)g->V);
    std::vector<uint8_t> hid((size_t)g->V, 0);
    int32_t cnt = 0;
    c2g_stats gs{};
    int rc = sssp_run(g, src, budget, dist_out, fr.data(), &cnt, hid.data(), &gs, true);
    if (rc) return rc;
    c2g_stats hs{};
    rc = host_bf(g, dist_out, fr.data(), cnt, g->hvl.empty() ? nullptr : hid.data(), &hs);
    if (rc) return rc;
    if (st) {
        *st = gs;
        st->lane = cnt ? C2G_LANE_MIXED : C2G_LANE_GPU;
        st->host_iterations = hs.host_iterations;
        st->edges_relaxed += hs.edges_relaxed;
        st->improvements += hs.improvements;
        int32_t r = 0;
        for (int32_t v = 0; v < g->V; ++v) r += dist_out[v] != INT32_MAX;
        st->reached = r;
        st->reached_edges = reached_edges(g, dist_out);
    }
    return C2G_OK;
}

int c2g_sssp_each_resident(c2g_graph* h, const int32_t* srcs, int32_t n, int64_t* sums,
                           int32_t* reached, c2g_stats* st) {
    if (!h || !srcs || n < 0) return fail(C2G_EINVAL, "null argument");
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    if (st) memset(st, 0, sizeof *st);
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    int64_t* d_sum = nullptr;
    int32_t* d_cnt = nullptr;
    int64_t* d_m = nullptr;
    HIP_TRY(hipSetDevice(g->device));
    HIP_TRY(hipMalloc(&d_sum, sizeof(int64_t) * (std::max(n, 1) + 1)));
    HIP_TRY(hipMalloc(&d_cnt, sizeof(int32_t) * std::max(n, 1)));
    d_m = d_sum + std::max(n, 1);
    HIP_TRY(hipMemsetAsync(d_sum, 0, sizeof(int64_t) * (std::max(n, 1) + 1), g->stream));
    HIP_TRY(hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * std::max(n, 1), g->stream));
    const double t0 = now_ms();
    int rc = C2G_OK;
    for (int32_t q = 0; q < n && rc == C2G_OK; ++q) {
        c2g_stats s{};
        rc = sssp_run(g, srcs[q], 0, nullptr, nullptr, nullptr, nullptr, &s, false);
        if (rc) break;
        hipLaunchKernelGGL(dist_checksum_k, dim3(1024), dim3(256), 0, g->stream, g->d_row,
                           g->sssp.dist, (int64_t)g->V, (int64_t)1, d_sum + q, d_cnt + q, d_m);
        if (st) {
            st->iterations += s.iterations;
            st->frontier_total += s.frontier_total;
            st->edges_relaxed += s.edges_relaxed;
            st->improvements += s.improvements;
            st->items += s.items;
            st->kernel_ms += s.kernel_ms;
        }
    }
    if (rc == C2G_OK) {
        if (hipStreamSynchronize(g->stream) != hipSuccess) rc = fail(C2G_EDEVICE, "sync failed");
    }
    const double t1 = now_ms();
    if (rc == C2G_OK && sums) (void)hipMemcpy(sums, d_sum, sizeof(int64_t) * n, hipMemcpyDeviceToHost);
    if (rc == C2G_OK && reached) (void)hipMemcpy(reached, d_cnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    int64_t medges = 0;
    if (rc == C2G_OK) (void)hipMemcpy(&medges, d_m, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (st) st->reached_edges = medges;
    (void)hipFree(d_sum);
    (void)hipFree(d_cnt);
    if (st) { st->lane = C2G_LANE_GPU; st->ms = t1 - t0; }
    return rc;
}

}  // extern "C"
