// Batched multi-source SSSP on gfx950: 64 queries per group, one wave lane
// per query.
//
// The reference runs one CalcuSSSP launch per query (src/sssp/base/Main.cu:
// 59-76); here many sources of uneven reach share every launch.  Distances
// are bit-identical to single-source runs: each lane is an independent
// label-correcting SSSP with the same relaxation rule
// (`cand = dist[u]+w; atomicMin; re-activate on improvement`,
// src/sssp/base/Main.cu:217-221).
//
// HBM layout of a group g (queries 64g .. 64g+63):
//   dist  [g][V][64] int32   vertex-major: relaxing edge (u,v) for all active
//                            queries is one coalesced 256-byte row access
//   mask  [g][V]     u64     lanes of v improved this round, not yet expanded
//   hmask [g][V]     u64     lanes whose value at v came through a hidden edge
// Per round (kernels captured in a hipGraph):
//   relax_b  : each wave takes one item (u, e_begin, e_end, g) with its lane
//              mask; edges are loaded 64 at a time lane-parallel, then walked
//              in groups of 8 (8 independent row loads in flight); improved
//              lanes are OR-ed into mask[g][v] lane-parallel and (v,g) is
//              marked in the group's activation bitmap.  Pruned graphs run the
//              same kernel twice per round: over the hidden-list items first
//              (edges from hadj, improved lanes also recorded in hmask), then
//              over the live items with exactly the full graph's code.
//   expand_b : scans the bitmaps; for each active (v,g) takes mask[g][v] (clearing it),
//              splits v's row into items of <= batch_chunk edges, and writes a
//              hidden-list item for the lanes in hmask to the hidden queue
//              (pruned graphs).
// Light-edge mode (c2g_tune.light = L > 0): a vertex whose lanes were lowered
// in the last round relaxes only its light row (edges of weight < L) for
// those lanes, and the lanes join ph[g][v]; once a round passes in which no
// lane lowered v, v's full row is relaxed for the lanes in ph[g][v].  Its
// heavy edges are then relaxed with settled distances instead of after every
// improvement, which cuts relaxations about 3x and fetched distance rows about
// 4x on R-MAT with uniform weights, for a few more rounds.
//   ph    [g][V]     u64     lanes whose full row is pending (light mode)
// Vertices with neither out-edges nor hidden edges are never activated (flag
// bit 31 of the edge target), which skips about half of all activations on
// R-MAT graphs.  Activation is recorded in a per-group bitmap with
// fire-and-forget atomics; expand_b compacts it (one atomicAdd per wave).
#include "graph.h"
#include "wave.h"
#include <cstring>
#include <type_traits>

namespace c2g {

enum : int {
    B_SPARE = 0,
    B_ITEMS = 1,   // items for the next relax round
    B_ITER = 2,    // relax rounds with work
    B_FRONT = 5,   // sum over rounds of activated (v,g)
    B_EDGES = 6,   // CSR edges streamed (per item, once for all lanes)
    B_IMPR = 7,    // improvement writes issued (U): candidate below the value read
    B_OVF = 8,     // capacity guard tripped (must stay 0)
    B_LEDGES = 9,  // (query, edge) relaxation attempts
    B_ITEMS_TOT = 10,  // items processed
    B_LANEOR = 11,     // server lane: lanes of group 0 with pending work after expand
    B_FAR = 12,        // 1 while some lane has deferred (far) work
    B_HITEMS = 13,     // hidden-list items for the next relax round (pruned graphs)
    B_N = 16           // fields; the block is kCtrlWords long (completion tree after them)
};
static_assert(B_N <= kCtrlTree, "control fields overlap the completion tree");

constexpr int kJ = 8;  // edges whose row loads are in flight together
// expand, delta-stepping form: vertices per lane whose loads are in flight
// together.  One, so that the kernel fits 64 VGPRs and runs 8 waves per SIMD
// (two slots need 77 VGPRs, 6 waves; both took the same time within 4 %).
constexpr int kS = 1;
constexpr int kIdMaskB = 0x7fffffff;  // bit 31 of an edge target = target has edges

__global__ void batch_fill_k(int32_t* __restrict__ dist, int64_t n_dist, uint64_t* __restrict__ mask,
                             uint64_t* __restrict__ hmask, int64_t n_mask, uint32_t* __restrict__ abits,
                             int64_t n_bits, uint64_t* __restrict__ ph, uint32_t* __restrict__ kbits) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int4* d4 = reinterpret_cast<int4*>(dist);
    const int64_t n4 = n_dist / 4;
    for (int64_t i = t; i < n4; i += stride) d4[i] = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
    for (int64_t i = t; i < n_mask; i += stride) {
        if (mask) mask[i] = 0;
        if (hmask) hmask[i] = 0;
        if (ph) ph[i] = 0;
    }
    for (int64_t i = t; i < n_bits; i += stride) {
        abits[i] = 0;
        if (kbits) kbits[i] = 0;
    }
}

// Seeds lane q of group q/64 at its source (one thread per query).
__global__ void batch_seed_k(const int32_t* __restrict__ srcs, int n, const int32_t* __restrict__ row,
                             const int32_t* __restrict__ hrow, int32_t* __restrict__ dist,
                             uint64_t* __restrict__ mask, uint64_t* __restrict__ hmask,
                             uint32_t* __restrict__ abits, int64_t* __restrict__ ctrl, int64_t V, int64_t nw) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        for (int i = 0; i < B_N; ++i) ctrl[i] = 0;
        done_tree_clear(ctrl);
    }
    if (q >= n) return;
    const int g = q >> 6, lane = q & 63;
    const int32_t s = srcs[q];
    dist[((int64_t)g * V + s) * 64 + lane] = 0;
    if (hmask) atomicOr((unsigned long long*)&hmask[(int64_t)g * V + s], 1ull << lane);
    const bool has = row[s + 1] > row[s] || (hrow && hrow[s + 1] > hrow[s]);
    if (!has) return;
    atomicOr((unsigned long long*)&mask[(int64_t)g * V + s], 1ull << lane);
    atomicOr(&abits[(int64_t)g * nw + (s >> 5)], 1u << (s & 31));
}

// Relaxes edges j0 .. j0+kJ-1 of a chunk for the lanes in exec.  Lane j of
// `my` holds edge j (readlane ignores exec; every lane of `my` holds a value);
// everything about an edge (target, weight, row address, which lanes
// improved) is wave-uniform and lives in SGPRs, so per lane there is one add,
// one compare, the row load and the atomic.  Distances are compared unsigned:
// a candidate of INT32_MAX or more (unreached source, or past the end of the
// chunk) never beats a stored distance, which is at most INT32_MAX.  TAIL:
// some of the kJ edges may lie at or past `ne`; they read the source's own row
// and relax nothing.
template <bool TAIL>
__device__ __forceinline__ void relax_step_b(const int2 my, int j0, int ne, int u, uint32_t du,
                                             int32_t* __restrict__ dG, uint32_t loff,
                                             uint64_t* __restrict__ bits) {
    gchar* rj[kJ];
    uint32_t cj[kJ], oj[kJ];
#pragma unroll
    for (int jj = 0; jj < kJ; ++jj) {
        const int j = TAIL ? min(j0 + jj, 63) : j0 + jj;  // a lane index
        int v = __builtin_amdgcn_readlane(my.x, j) & kIdMaskB;
        uint32_t w = (uint32_t)__builtin_amdgcn_readlane(my.y, j);
        if (TAIL) {
            const bool valid = j0 + jj < ne;
            v = valid ? v : u;
            w = valid ? w : 0x7fffffffu;
        }
        rj[jj] = sgpr_ptr(dG + (int64_t)v * 64);
        cj[jj] = du + w;
        oj[jj] = (uint32_t)*lane_slot<int32_t>(rj[jj], loff);
    }
    // Fire-and-forget atomicMin: a lane whose candidate beats the value it
    // read marks (v, lane) active.  If another relaxation lowered
    // dist[v][lane] below the candidate since the read, it was made in this
    // round and marked the same bit, so waiting for the old value would change
    // nothing but the latency.
#pragma unroll
    for (int jj = 0; jj < kJ; ++jj) {
        // lanes (in exec) whose candidate is below the stored value: the
        // compare's own mask, so the ballot costs no extra instruction
        const uint64_t b = __builtin_amdgcn_uicmp(cj[jj], oj[jj], 36 /* unsigned < */);
        if (cj[jj] < oj[jj])
            (void)__hip_atomic_fetch_min(lane_slot<int32_t>(rj[jj], loff), (int)cj[jj], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        bits[j0 + jj] = b;
    }
}

// HIDDEN: the items are hidden-list items (pruned graphs): edges come from
// `adj` = hadj and every improved lane is also recorded in hmask.  The live
// items of any graph run with HIDDEN = false, which never touches hmask.
// BLOCK: threads per workgroup (128, 256 or 512; c2g_tune.block).
template <bool HIDDEN, int BLOCK>
__global__ void __launch_bounds__(BLOCK) C2G_OCC8_SGPR
relax_b_k(const int2* __restrict__ adj, int32_t* __restrict__ dist, uint64_t* __restrict__ mask,
          uint64_t* __restrict__ hmask, const int4* __restrict__ items, const uint64_t* __restrict__ imask,
          uint32_t* __restrict__ abits, int64_t* __restrict__ ctrl, int64_t V, int64_t nw) {
    // improved-lane mask of each edge of the current 64-edge chunk, per wave
    // (a last step of kJ edges may run up to kJ-1 entries past the chunk)
    __shared__ uint64_t s_bits[BLOCK / 64][64 + kJ];
    const int64_t n = ctrl[HIDDEN ? B_HITEMS : B_ITEMS];
    const int lane = lane_id();
    // uniform wave index: item headers come through scalar loads, and each
    // target row's base address is an SGPR pair
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    uint64_t* __restrict__ bits = s_bits[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    const uint32_t loff = (uint32_t)lane * 4u;  // this lane's byte offset in a 256-byte row
    int64_t impr = 0;
    __shared__ unsigned long long s_impr;  // workgroup total: one global atomic per workgroup
    if (threadIdx.x == 0) s_impr = 0;
    __syncthreads();

    for (int64_t i = wave; i < n; i += nwaves) {
        const int4 it = items[i];
        const uint64_t m = imask[i];
        const int u = it.x;
        const int g = it.w;
        int32_t* __restrict__ dG = dist + (int64_t)g * V * 64;
        const bool act = (m >> lane) & 1ull;
        for (int e0 = it.y; e0 < it.z; e0 += 64) {
            const int ne = min(64, it.z - e0);
            int2 my = make_int2(0, 0);
            if (lane < ne) my = adj[e0 + lane];
            // The item's pending lanes relax the chunk's edges kJ at a time.
            if (act) {
                const uint32_t du = (uint32_t)dG[(int64_t)u * 64 + lane];
                wait_vmem();  // `my` and du, once per chunk rather than at every loop head
                int j0 = 0;
                for (; j0 + kJ <= ne; j0 += kJ) relax_step_b<false>(my, j0, ne, u, du, dG, loff, bits);
                if (j0 < ne) relax_step_b<true>(my, j0, ne, u, du, dG, loff, bits);
            }
            // Lane l applies edge l's improved lanes: record them at v and
            // mark (v,g) active (both fire-and-forget).
            if (lane < ne) {
                const uint64_t mybits = m ? bits[lane] : 0;
                impr += __popcll(mybits);
                if (mybits) {
                    const int v = my.x & kIdMaskB;
                    if (HIDDEN)
                        atomicOr((unsigned long long*)&hmask[(int64_t)g * V + v], (unsigned long long)mybits);
                    if (my.x < 0) {
                        atomicOr((unsigned long long*)&mask[(int64_t)g * V + v], (unsigned long long)mybits);
                        atomicOr(&abits[(int64_t)g * nw + (v >> 5)], 1u << (v & 31));
                    }
                }
            }
        }
    }
    impr = wave_sum64(impr);
    if (lane == 0 && impr) atomicAdd(&s_impr, (unsigned long long)impr);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_impr) atomicAdd((unsigned long long*)&ctrl[B_IMPR], s_impr);
        if (wg_last(ctrl + kCtrlTree)) {
            // a round counts once: the hidden pass runs first and counts only
            // rounds whose live queue is empty
            if (HIDDEN ? (n > 0 && ctrl[B_ITEMS] == 0) : n > 0) ctrl[B_ITER] += 1;
            ctrl[B_ITEMS_TOT] += n;
            ctrl[HIDDEN ? B_HITEMS : B_ITEMS] = 0;
        }
    }
}

// Expand one round.  For each active (v,g): take the pending lanes m of v.
// With delta > 0 the lanes are split by v's current distance against each
// lane's bucket bound thr[g][q] (delta-stepping): near lanes (dist < thr) are
// expanded into items, far lanes stay pending in mask/abits for a later round.
// Per-lane bookkeeping of the round goes to lor[2g] (lanes with near work),
// lor[2g+1] (lanes with far work) and fmin[g][q] (smallest far distance); the
// last workgroup moves every lane that has far work but no near work to the
// bucket of its smallest far distance.  With delta == 0 every pending lane is
// near (plain Bellman-Ford rounds).  Light-edge mode (lrow non-null, delta
// 0): lowered lanes m of v get items over v's light row [lrow[v], lrow[v+1])
// and join ph[g][v], and v stays active; a v with no lowered lane gets items
// over its full row for the lanes in ph[g][v], which it clears.  DELTA:
// compiled for delta > 0 (the plain and light-edge forms leave the
// delta-stepping code out, which keeps them within 64 VGPRs).  PRUNED:
// compiled for graphs with hidden lists (hrow non-null); the other form
// leaves the hidden-list code out.
template <bool DELTA, bool PRUNED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) C2G_OCC8_SGPR
expand_b_k(const int32_t* __restrict__ row, const int32_t* __restrict__ hrow, uint64_t* __restrict__ mask,
           const uint64_t* __restrict__ hmask, uint32_t* __restrict__ abits, int4* __restrict__ items,
           uint64_t* __restrict__ imask, int4* __restrict__ hitems, uint64_t* __restrict__ himask,
           int64_t* __restrict__ ctrl, int64_t V, int64_t nw, int groups,
           int chunk, int64_t itemcap, int64_t hitemcap, const int32_t* __restrict__ dist, int32_t* __restrict__ thr,
           int32_t* __restrict__ fmin, uint64_t* __restrict__ lor, int delta, const int32_t* __restrict__ lrow,
           uint64_t* __restrict__ ph, uint32_t* __restrict__ kbits, const uint32_t* __restrict__ sbits) {
    extern __shared__ int32_t lds_fmin[];  // [groups * 64] when delta > 0
    if constexpr (!PRUNED) {
        hrow = nullptr;
        hmask = nullptr;
        hitems = nullptr;
        himask = nullptr;
    }
    const int lane = lane_id();
    if (delta > 0) {
        for (int t = threadIdx.x; t < groups * 64; t += blockDim.x) lds_fmin[t] = INT32_MAX;
        __syncthreads();
    }
    int64_t front = 0, edges = 0, ledges = 0;
    // Work unit: a block of 64 bitmap words (2048 vertices of one group; nw
    // is a multiple of 64).  The wave loads the block's words in one step;
    // without delta-stepping it walks the block's set bits 64 at a time (one
    // lane per active vertex, see block_bits), the delta-stepping form walks
    // its nonzero words 2*kS at a time, lanes 0-31 on one word and 32-63 on
    // the other.  Either way masks and row offsets arrive in coalesced loads.
    // Pass 1 counts items, one atomic reserves the item range, pass 2 writes
    // the items and takes the expanded lanes out of the mask.
    const int64_t nblk = (nw * groups) >> 6;
    const int64_t nwav = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int half = lane >> 5, bit = lane & 31;
    // Each wave takes a contiguous range of blocks (mostly of one group): its
    // lane summaries go out once per group, and one atomic reserves the item
    // ranges of all its blocks (pass 1 counts over the whole range, pass 2
    // writes).
    const int64_t per = (nblk + nwav - 1) / nwav;
    const int64_t bb0 = min(nblk, (int64_t)wave_index() * per), bb1 = min(nblk, bb0 + per);
    // Workgroup totals (counters, item reservations, lane summaries of the
    // first two groups the workgroup's blocks touch) go out with one global
    // atomic each per workgroup: per-wave atomics on the same few cache lines
    // serialise in the L2.
    __shared__ unsigned long long s_acc[3];  // front, edges, lane edges
    __shared__ unsigned long long s_lor[2][2];
    __shared__ int64_t s_res[2][8];          // per wave: items, hidden items
    __shared__ uint32_t s_keep[4][64];       // per wave: kept bits of the block's words (pass 2)
    __shared__ int64_t s_base[2];
    const int wid = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int g0 = (int)((min(nblk - 1, (int64_t)blockIdx.x * nwv * per) << 6) / nw);
    if (threadIdx.x < 3) s_acc[threadIdx.x] = 0;
    s_keep[threadIdx.x >> 6][threadIdx.x & 63] = 0;
    if (threadIdx.x < 4) s_lor[threadIdx.x >> 1][threadIdx.x & 1] = 0;
    __syncthreads();
    uint64_t nor = 0, forr = 0;
    int gcur = -1;
    auto flush = [&]() {
        if (gcur < 0) return;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            nor |= (uint64_t)shfl64((int64_t)nor, lane ^ d);
            forr |= (uint64_t)shfl64((int64_t)forr, lane ^ d);
        }
        if (lane == 0 && (nor | forr)) {
            const int k = gcur - g0;
            if (k >= 0 && k < 2) {
                if (nor) atomicOr(&s_lor[k][0], (unsigned long long)nor);
                if (forr) atomicOr(&s_lor[k][1], (unsigned long long)forr);
            } else {
                if (nor) atomicOr((unsigned long long*)&lor[2 * gcur], (unsigned long long)nor);
                if (forr) atomicOr((unsigned long long*)&lor[2 * gcur + 1], (unsigned long long)forr);
            }
        }
        nor = forr = 0;
    };
    int cnt = 0, hcnt = 0;
    // pass 2: the wave's next live item and next hidden item (wave-uniform;
    // pass 1 fixes only the wave's totals, so any lane may count any vertex),
    // except that the delta-stepping form writes hidden items per lane, at
    // this lane's next hidden item
    int64_t p = 0, hp = 0;
    bool room = true;
    // One block without delta-stepping (plain rounds, or light-edge mode
    // with kbits): abits marks vertices lowered in the last round, kbits
    // (light mode) vertices kept for their full row.  The bitmaps alone fix
    // every vertex's edge range, so pass 1 (write false) reads only row
    // offsets (and, on pruned graphs, the masks of lowered vertices for their
    // hidden lists); pass 2 reads the lane masks, writes the items, updates
    // mask, ph and both bitmaps, and records the lane summaries.
    auto block_bits = [&](int64_t blk, bool write) {
        const int64_t w0 = blk << 6;
        const uint32_t da = abits[w0 + lane];
        const uint32_t dk = kbits ? kbits[w0 + lane] : 0u;
        const uint32_t myword = da | dk;
        const uint64_t nzw = __ballot(myword != 0);
        if (!nzw) return;
        const int g = (int)(w0 / nw);
        if (write && g != gcur) {
            flush();
            gcur = g;
        }
        const int64_t vb0 = (w0 - (int64_t)g * nw) * 32;  // first vertex of the block
        if (!write && sbits && !hrow) {
            // Pass 1 from the static bitmaps: this lane's word holds 32
            // vertices; each with a non-empty row has one item, and only the
            // rare rows of more than `chunk` edges need their offsets.
            const int64_t wl = w0 - (int64_t)g * nw + lane;
            const uint32_t sl = kbits ? da : 0u;          // light rows
            const uint32_t sf = kbits ? dk & ~da : da;    // full rows
            if (sl | sf) {
                int c = __popc(sl & sbits[2 * nw + wl]) + __popc(sf & sbits[wl]);
                const int64_t vw = vb0 + (int64_t)lane * 32;
                for (uint32_t m = sl & sbits[3 * nw + wl]; m; m &= m - 1) {
                    const int64_t v = vw + __builtin_ctz(m);
                    c += (lrow[v + 1] - lrow[v] + chunk - 1) / chunk - 1;
                }
                for (uint32_t m = sf & sbits[nw + wl]; m; m &= m - 1) {
                    const int64_t v = vw + __builtin_ctz(m);
                    c += (row[v + 1] - row[v] + chunk - 1) / chunk - 1;
                }
                cnt += c;
            }
            return;
        }
        uint64_t* __restrict__ mg = mask + (int64_t)g * V;
        uint64_t* __restrict__ pg = kbits ? ph + (int64_t)g * V : nullptr;
        // The block's set bits 64 at a time: set bit j of the block (in
        // vertex order) goes to lane j % 64 of step j / 64, so a block whose
        // words hold few bits takes one step per 64 vertices rather than one
        // per pair of nonzero words.  Pass 1 (write false) counts items (only
        // the wave's totals matter); pass 2 writes them at wave-uniform
        // positions, updates mask and ph, and collects kept bits per word in
        // LDS, from where each lane rewrites its own word at the end.
        uint32_t* __restrict__ kw = s_keep[wid];
        int nbit;
        const int bpre = wave_excl_scan(__popc(myword), &nbit);
        for (int s0 = 0; s0 < nbit; s0 += kWave) {
            const int j = min(s0 + lane, nbit - 1);
            const bool on = s0 + lane < nbit;
            const int o = wave_owner(bpre, j);  // the word (lane) holding bit j
            const uint32_t wo = (uint32_t)__shfl((int)myword, o, 64);
            const uint32_t dao = (uint32_t)__shfl((int)da, o, 64);
            const int bp = nth_set_bit(wo, j - __shfl(bpre, o, 64));
            const int v = (int)(vb0 + (int64_t)o * 32 + bp);
            const bool dirty = (dao >> bp) & 1u;
            // light mode: a lowered vertex takes its light row, a kept one
            // its full row
            const bool lt = pg && dirty;
            uint64_t m = 0, h = 0, hm = 0;
            int b = 0, e = 0, h0 = 0, h1 = 0;
            if (on) {
                if (write || (hrow && dirty)) m = mg[v];
                if (write && pg) h = pg[v];
                const int32_t* rr = lt ? lrow : row;
                b = rr[v];
                e = rr[v + 1];
                if (hrow && dirty) {
                    h0 = hrow[v];
                    h1 = hrow[v + 1];
                }
            }
            if (hrow && on && dirty) hm = m & hmask[(int64_t)g * V + v];
            int c = on && e > b ? (e - b + chunk - 1) / chunk : 0;
            int hc = hm && h1 > h0 ? (h1 - h0 + chunk - 1) / chunk : 0;
            if (!write) {
                cnt += c;
                hcnt += hc;
                continue;
            }
            bool keep = false;
            uint64_t lanes = 0;
            if (on) {
                lanes = dirty ? m : h;
                const uint64_t pend = lt ? h | m : 0;
                if (dirty && m) mg[v] = 0;
                if (pg && pend != h) pg[v] = pend;
                keep = pend != 0;
                nor |= lanes;
                forr |= pend;
                if (lanes) front += 1;
                edges += e - b;
                ledges += (int64_t)__popcll(lanes) * (e - b);
                if (hm) {
                    edges += h1 - h0;
                    ledges += (int64_t)__popcll(hm) * (h1 - h0);
                }
            }
            if (!room) c = hc = 0;
            // items written by the whole wave: item jj of the step belongs to
            // the lane whose range of the step's prefix holds it, so a hub
            // row's thousands of items take thousands/64 steps, not one
            // lane's serial loop
            auto emit = [&](int n, int b0, int e0, uint64_t lm, int4* __restrict__ out,
                            uint64_t* __restrict__ outm, int64_t& pos) {
                int ntot;
                const int npre = wave_excl_scan(n, &ntot);
                for (int j0 = 0; j0 < ntot; j0 += kWave) {
                    const int jj = j0 + lane;
                    const int ow = wave_owner(npre, min(jj, ntot - 1));
                    const int kk = jj - __shfl(npre, ow, 64);
                    const int vo = __shfl(v, ow, 64), bo = __shfl(b0, ow, 64), eo = __shfl(e0, ow, 64);
                    const uint64_t lo = (uint64_t)shfl64((int64_t)lm, ow);
                    if (jj < ntot) {
                        const int x = bo + kk * chunk;
                        out[pos + jj] = make_int4(vo, x, min(x + chunk, eo), g);
                        outm[pos + jj] = lo;
                    }
                }
                pos += ntot;
            };
            emit(c, b, e, lanes, items, imask, p);
            if (hrow) emit(hc, h0, h1, hm, hitems, himask, hp);
            if (keep) atomicOr(&kw[o], 1u << bp);
        }
        if (!write) return;
        // the wave's LDS operations complete in order: every kept bit of
        // this block is in kw before a lane reads its word
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t nk = kw[lane];
        kw[lane] = 0u;
        // every lowered vertex is consumed; kbits keeps the vertices whose
        // full row is pending
        if (da) abits[w0 + lane] = 0u;
        if (kbits && nk != dk) kbits[w0 + lane] = nk;
    };
    // One block with delta-stepping: pass 1 (write false) counts items and
    // records lane summaries, pass 2 (write true) writes items and updates
    // the masks.
    auto block_delta = [&](int64_t blk, bool write) {
        const int64_t w0 = blk << 6;
        const uint32_t myword = abits[w0 + lane];
        const uint64_t nzw = __ballot(myword != 0);
        if (!nzw) return;
        const int g = (int)(w0 / nw);
        if (!write && g != gcur) {
            flush();
            gcur = g;
        }
        const int64_t vb0 = (w0 - (int64_t)g * nw) * 32;  // first vertex of the block
        uint64_t* __restrict__ mg = mask + (int64_t)g * V;
        const int32_t* dG = dist + (int64_t)g * V * 64;
        const int32_t* tg = thr + g * 64;
        // Near lanes of v (and far lanes, recorded once in pass 1).
        auto split = [&](int v, uint64_t m, bool record) -> uint64_t {
            if (delta <= 0) return m;
            uint64_t near = 0;
            for (int c = 0; c < 16; ++c) {
                const uint64_t mc = (m >> (4 * c)) & 0xFull;
                if (!mc) continue;
                const int4 d4 = *reinterpret_cast<const int4*>(dG + (int64_t)v * 64 + 4 * c);
                const int dv[4] = {d4.x, d4.y, d4.z, d4.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (!((mc >> k) & 1ull)) continue;
                    const int q = 4 * c + k;
                    if (dv[k] < tg[q]) near |= 1ull << q;
                    else if (record) atomicMin(&lds_fmin[g * 64 + q], dv[k]);
                }
            }
            return near;
        };
        // The live items of v: edge range [b, e) and lanes; near = the
        // lowered lanes expanded now (their hidden lists go with them),
        // pend = lanes v keeps for a later bucket.  m and r0/r1 are v's
        // loads, issued for all slots of a step before any is used.
        struct Work { int b, e; uint64_t lanes, near, pend; };
        auto work = [&](int v, uint64_t m, int r0, int r1, bool record) -> Work {
            const uint64_t near = split(v, m, record);
            return Work{r0, r1, near, near, m & ~near};
        };
        // A step takes the next 2*kS nonzero words: slot s of lanes 0-31
        // walks word 2s, of lanes 32-63 word 2s+1.
        struct Step {
            int k[2 * kS];
            uint32_t w[2 * kS];
            bool has[2 * kS];
        };
        auto next_step = [&](uint64_t& rest) -> Step {
            Step st;
#pragma unroll
            for (int j = 0; j < 2 * kS; ++j) {
                st.has[j] = rest != 0;
                st.k[j] = st.has[j] ? __builtin_ctzll(rest) : 0;
                if (st.has[j]) rest &= rest - 1;
                st.w[j] = st.has[j] ? (uint32_t)__builtin_amdgcn_readlane((int)myword, st.k[j]) : 0u;
            }
            return st;
        };
        if (!write) for (uint64_t rest = nzw; rest;) {
            const Step st = next_step(rest);
            bool on[kS];
            int vv[kS], r0[kS], r1[kS];
            uint64_t mm[kS];
#pragma unroll
            for (int sl = 0; sl < kS; ++sl) {
                const uint32_t word = half ? st.w[2 * sl + 1] : st.w[2 * sl];
                on[sl] = (word >> bit) & 1u;
                vv[sl] = (int)(vb0 + (int64_t)(half ? st.k[2 * sl + 1] : st.k[2 * sl]) * 32 + bit);
                mm[sl] = 0;
                r0[sl] = r1[sl] = 0;
                if (on[sl]) {
                    const int v = vv[sl];
                    mm[sl] = mg[v];
                    r0[sl] = row[v];
                    r1[sl] = row[v + 1];
                }
            }
#pragma unroll
            for (int sl = 0; sl < kS; ++sl) {
                if (!on[sl]) continue;
                const int v = vv[sl];
                const Work wk = work(v, mm[sl], r0[sl], r1[sl], true);
                forr |= wk.pend;
                if (!wk.lanes) continue;
                nor |= wk.lanes;
                if (wk.e > wk.b) cnt += (wk.e - wk.b + chunk - 1) / chunk;
                if (hrow && wk.near) {
                    const uint64_t hm = wk.near & hmask[(int64_t)g * V + v];
                    if (hm) {
                        const int h0 = hrow[v], h1 = hrow[v + 1];
                        hcnt += (h1 - h0 + chunk - 1) / chunk;
                        edges += h1 - h0;
                        ledges += (int64_t)__popcll(hm) * (h1 - h0);
                    }
                }
                front += 1;
                edges += wk.e - wk.b;
                ledges += (int64_t)__popcll(wk.lanes) * (wk.e - wk.b);
            }
        }
        if (write) for (uint64_t rest = nzw; rest;) {
            const Step st = next_step(rest);
            bool on[kS];
            int vv[kS], r0[kS], r1[kS];
            uint64_t mm[kS];
#pragma unroll
            for (int sl = 0; sl < kS; ++sl) {
                const uint32_t word = half ? st.w[2 * sl + 1] : st.w[2 * sl];
                on[sl] = (word >> bit) & 1u;
                vv[sl] = (int)(vb0 + (int64_t)(half ? st.k[2 * sl + 1] : st.k[2 * sl]) * 32 + bit);
                mm[sl] = 0;
                r0[sl] = r1[sl] = 0;
                if (on[sl]) {
                    const int v = vv[sl];
                    mm[sl] = mg[v];
                    r0[sl] = row[v];
                    r1[sl] = row[v + 1];
                }
            }
#pragma unroll
            for (int sl = 0; sl < kS; ++sl) {
                bool keep = false;  // v stays active: lanes left for a later bucket
                Work wk{0, 0, 0, 0, 0};
                int c = 0;  // this lane's live items in the slot
                const int v = vv[sl];
                if (on[sl]) {
                    const uint64_t m = mm[sl];
                    wk = work(v, m, r0[sl], r1[sl], false);
                    mg[v] = wk.pend;
                    keep = wk.pend != 0;
                    if (wk.lanes && room) {
                        if (wk.e > wk.b) c = (wk.e - wk.b + chunk - 1) / chunk;
                        if (hrow && wk.near) {
                            const uint64_t hm = wk.near & hmask[(int64_t)g * V + v];
                            if (hm) {
                                const int h0 = hrow[v], h1 = hrow[v + 1];
                                for (int e = h0; e < h1; e += chunk, ++hp) {
                                    hitems[hp] = make_int4(v, e, min(e + chunk, h1), g);
                                    himask[hp] = hm;
                                }
                            }
                        }
                    }
                }
                // The slot's live items, written by the whole wave: item j
                // belongs to the lane whose range of the slot's prefix holds
                // it, so a hub row's thousands of items take thousands/64
                // steps instead of one lane's serial loop.
                int ctot;
                const int cpre = wave_excl_scan(c, &ctot);
                for (int j0 = 0; j0 < ctot; j0 += kWave) {
                    const int j = j0 + lane;
                    const int o = wave_owner(cpre, min(j, ctot - 1));
                    const int k = j - __shfl(cpre, o, 64);
                    const int vo = __shfl(v, o, 64), bo = __shfl(wk.b, o, 64), eo = __shfl(wk.e, o, 64);
                    const uint64_t lo = (uint64_t)shfl64((int64_t)wk.lanes, o);
                    if (j < ctot) {
                        const int e = bo + k * chunk;
                        items[p + j] = make_int4(vo, e, min(e + chunk, eo), g);
                        imask[p + j] = lo;
                    }
                }
                p += ctot;
                // the slot's two words keep only the bits of vertices with
                // lanes left
                const uint64_t kb = __ballot(keep);
                if (lane == 0 && st.has[2 * sl] && (uint32_t)kb != st.w[2 * sl]) abits[w0 + st.k[2 * sl]] = (uint32_t)kb;
                if (lane == 32 && st.has[2 * sl + 1] && (uint32_t)(kb >> 32) != st.w[2 * sl + 1])
                    abits[w0 + st.k[2 * sl + 1]] = (uint32_t)(kb >> 32);
            }
        }
    };
    auto block = [&](int64_t blk, bool write) {
        if constexpr (DELTA) block_delta(blk, write);
        else block_bits(blk, write);
    };
    for (int64_t blk = bb0; blk < bb1; ++blk) block(blk, false);
    // Item ranges: wave totals meet in LDS, one atomic per queue reserves
    // the workgroup's range, each wave takes its part in wave order.
    int tot, htot = 0;
    (void)wave_excl_scan(cnt, &tot);  // the wave's live items (any lane may hold any count)
    const int hoff = hrow ? wave_excl_scan(hcnt, &htot) : 0;
    if (lane == 0) {
        s_res[0][wid] = tot;
        s_res[1][wid] = htot;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        int64_t sum = 0;
        for (int w = 0; w < nwv; ++w) sum += s_res[k][w];
        s_base[k] = sum > 0 ? (int64_t)atomicAdd((unsigned long long*)&ctrl[k ? B_HITEMS : B_ITEMS],
                                                 (unsigned long long)sum)
                            : 0;
    }
    __syncthreads();
    int64_t base = s_base[0], hbase = s_base[1];
    for (int w = 0; w < wid; ++w) {
        base += s_res[0][w];
        hbase += s_res[1][w];
    }
    room = base + tot <= itemcap && (!hrow || hbase + htot <= hitemcap);
    if (!room && lane == 0) atomicOr((unsigned long long*)&ctrl[B_OVF], 1ull);
    p = base;
    hp = hbase + (DELTA ? hoff : 0);  // the bitmap walk writes hidden items per wave, the delta form per lane
    for (int64_t blk = bb0; blk < bb1; ++blk) block(blk, true);
    flush();  // lane bookkeeping of the last group
    front = wave_sum64(front);
    edges = wave_sum64(edges);
    ledges = wave_sum64(ledges);
    if (lane == 0 && front) {
        atomicAdd(&s_acc[0], (unsigned long long)front);
        atomicAdd(&s_acc[1], (unsigned long long)edges);
        atomicAdd(&s_acc[2], (unsigned long long)ledges);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_acc[0]) {
        atomicAdd((unsigned long long*)&ctrl[B_FRONT], s_acc[0]);
        atomicAdd((unsigned long long*)&ctrl[B_EDGES], s_acc[1]);
        atomicAdd((unsigned long long*)&ctrl[B_LEDGES], s_acc[2]);
    }
    if (threadIdx.x < 4) {
        const int k = threadIdx.x >> 1, f = threadIdx.x & 1;
        if (s_lor[k][f] && g0 + k < groups)
            atomicOr((unsigned long long*)&lor[2 * (g0 + k) + f], s_lor[k][f]);
    }
    if (delta > 0) {
        __syncthreads();
        for (int t = threadIdx.x; t < groups * 64; t += blockDim.x)
            if (lds_fmin[t] != INT32_MAX) atomicMin(&fmin[t], lds_fmin[t]);
    }
    // The last workgroup closes the round: bucket moves and lane summaries.
    // Every wave's atomics are performed before the workgroup counts off.
    __shared__ int is_last;
    wait_vmem();
    __syncthreads();
    if (threadIdx.x == 0) is_last = wg_last(ctrl + kCtrlTree);
    __syncthreads();
    if (!is_last) return;
    for (int t = threadIdx.x; t < groups * 64; t += blockDim.x) {
        const int gg = t >> 6, q = t & 63;
        const uint64_t nb = (uint64_t)atomicOr((unsigned long long*)&lor[2 * gg], 0ull);
        const uint64_t fb = (uint64_t)atomicOr((unsigned long long*)&lor[2 * gg + 1], 0ull);
        const int32_t fm = atomicMin(&fmin[t], INT32_MAX);
        if (delta > 0 && !((nb >> q) & 1ull) && ((fb >> q) & 1ull))
            thr[t] = (int32_t)min((int64_t)INT32_MAX, ((int64_t)fm / delta + 1) * delta);
        fmin[t] = INT32_MAX;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t any_far = 0;
        for (int gg = 0; gg < groups; ++gg)
            any_far |= (uint64_t)atomicOr((unsigned long long*)&lor[2 * gg + 1], 0ull);
        ctrl[B_FAR] = any_far ? 1 : 0;
        // group 0: lanes with any pending work (the server lane)
        ctrl[B_LANEOR] = (int64_t)((uint64_t)atomicOr((unsigned long long*)&lor[0], 0ull) |
                                   (uint64_t)atomicOr((unsigned long long*)&lor[1], 0ull));
        for (int gg = 0; gg < groups; ++gg) lor[2 * gg] = lor[2 * gg + 1] = 0;
        int64_t ni = (int64_t)atomicAdd((unsigned long long*)&ctrl[B_ITEMS], 0ull);
        int64_t nh = (int64_t)atomicAdd((unsigned long long*)&ctrl[B_HITEMS], 0ull);
        if (ni > itemcap || nh > hitemcap) { ctrl[B_OVF] = 1; ctrl[B_ITEMS] = 0; ctrl[B_HITEMS] = 0; }
    }
}

// Static per-vertex bitmaps of the batched expand for items of `chunk` edges,
// words [k*nw, (k+1)*nw): k = 0 the vertex's row has at least one item, 1 at
// least two, 2 its light row has at least one, 3 at least two (light words
// zero without light rows).  The expand counts a block's items from them
// with popcounts, reading offsets only for rows of more than `chunk` edges.
__global__ void batch_sbits_k(const int32_t* __restrict__ row, const int32_t* __restrict__ lrow, int64_t V,
                              int64_t nw, int chunk, uint32_t* __restrict__ sbits) {
    const int lane = lane_id();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v0 = wave_index() * 64; v0 < nw * 32; v0 += nwaves * 64) {
        const int64_t v = v0 + lane;
        const int d = v < V ? row[v + 1] - row[v] : 0;
        const int l = lrow && v < V ? lrow[v + 1] - lrow[v] : 0;
        const uint64_t b0 = __ballot(d > 0), b1 = __ballot(d > chunk), b2 = __ballot(l > 0), b3 = __ballot(l > chunk);
        if (lane < 8) {
            const int k = lane >> 1, h = lane & 1;
            const uint64_t b = k == 0 ? b0 : k == 1 ? b1 : k == 2 ? b2 : b3;
            sbits[k * nw + (v0 >> 5) + h] = (uint32_t)(b >> (32 * h));
        }
    }
}

__global__ void batch_thr_init_k(int32_t* __restrict__ thr, int32_t* __restrict__ fmin,
                                 uint64_t* __restrict__ lor, int n, int delta) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        thr[t] = delta > 0 ? delta : INT32_MAX;
        fmin[t] = INT32_MAX;
    }
    if (t < 2 * (n / 64)) lor[t] = 0;
}

// Resets the distance rows of the vertices in live[0..nlive) (those with an
// in-edge, live or hidden) in groups [0, groups): the only rows a traversal
// can write apart from its sources' entries.  16 threads per 256-byte row,
// each writing its 16 bytes of the vertex's row in every group.
__global__ void batch_fill_rows_k(int32_t* __restrict__ dist, const int32_t* __restrict__ live, int64_t nlive,
                                  int groups, int64_t V) {
    const int64_t n = nlive * 16;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int4 inf4 = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += stride) {
        const int64_t v = live[t >> 4];
        const int part = (int)(t & 15);
        for (int g = 0; g < groups; ++g) reinterpret_cast<int4*>(dist + ((int64_t)g * V + v) * 64)[part] = inf4;
    }
}

// Resets the entries the previous traversal seeded: lane q%64 of group q/64
// at srcs[q], for q < n.
__global__ void batch_reset_srcs_k(int32_t* __restrict__ dist, const int32_t* __restrict__ srcs, int n, int64_t V) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dist[((int64_t)(q >> 6) * V + srcs[q]) * 64 + (q & 63)] = INT32_MAX;
}

// Per-lane checksums of groups [g0, g0 + ng), ng <= kCkG: sum of finite
// distances, reached count and out-degree sum of reached vertices (m_s), for
// all 64 lanes of each group.  With `live` non-null only its nrows rows are
// read (the vertices a traversal can reach other than the sources
// themselves); otherwise rows 0 .. nrows-1.  A wave takes two rows per step
// and reads them in every group, so 2*ng row loads are in flight per lane;
// the four waves of a workgroup meet in LDS, and one global atomic per
// (group, lane, counter) and workgroup goes out.  The rows are read once, with
// non-temporal loads (3 % faster than plain loads on R-MAT scale 24).
constexpr int kCkG = 4;
__global__ void __launch_bounds__(256)
batch_checksum_k(const int32_t* __restrict__ row, const int32_t* __restrict__ dist, int64_t V, int g0, int ng,
                 int64_t* __restrict__ sums, int32_t* __restrict__ reached, int64_t* __restrict__ medges,
                 const int32_t* __restrict__ live, int64_t nrows) {
    __shared__ unsigned long long l_s[kCkG][64], l_m[kCkG][64];
    __shared__ unsigned int l_c[kCkG][64];
    for (int t = threadIdx.x; t < kCkG * 64; t += blockDim.x) {
        l_s[t >> 6][t & 63] = 0;
        l_m[t >> 6][t & 63] = 0;
        l_c[t >> 6][t & 63] = 0;
    }
    __syncthreads();
    const int lane = lane_id();
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    int64_t s[kCkG] = {}, m[kCkG] = {};
    int c[kCkG] = {};
    const int32_t* __restrict__ d0 = dist + (int64_t)g0 * V * 64 + lane;
    for (int64_t i0 = wave * 2; i0 < nrows; i0 += nwaves * 2) {
        int64_t v[2];
        int deg[2], d[2][kCkG];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int64_t i = i0 + k;
            v[k] = i < nrows ? (live ? live[i] : i) : -1;
            deg[k] = v[k] >= 0 ? row[v[k] + 1] - row[v[k]] : 0;
#pragma unroll
            for (int gg = 0; gg < kCkG; ++gg)
                d[k][gg] = v[k] >= 0 && gg < ng ? __builtin_nontemporal_load(d0 + ((int64_t)gg * V + v[k]) * 64)
                                                : INT32_MAX;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int gg = 0; gg < kCkG; ++gg)
                if (d[k][gg] != INT32_MAX) {
                    s[gg] += d[k][gg];
                    c[gg] += 1;
                    m[gg] += deg[k];
                }
    }
#pragma unroll
    for (int gg = 0; gg < kCkG; ++gg)
        if (gg < ng && c[gg]) {
            atomicAdd(&l_s[gg][lane], (unsigned long long)s[gg]);
            atomicAdd(&l_m[gg][lane], (unsigned long long)m[gg]);
            atomicAdd(&l_c[gg][lane], (unsigned int)c[gg]);
        }
    __syncthreads();
    for (int t = threadIdx.x; t < ng * 64; t += blockDim.x) {
        const int gg = t >> 6, q = t & 63;
        if (!l_c[gg][q]) continue;
        atomicAdd((unsigned long long*)&sums[(g0 + gg) * 64 + q], l_s[gg][q]);
        atomicAdd(&reached[(g0 + gg) * 64 + q], (int)l_c[gg][q]);
        atomicAdd((unsigned long long*)&medges[(g0 + gg) * 64 + q], l_m[gg][q]);
    }
}

// ------------------------------------------------------------------ host
struct BatchState {
    int groups = 0;
    int64_t nw = 0, itemcap = 0;
    uint32_t* abits = nullptr;  // [groups][nw] activation bitmaps (nw padded to 64 words), then
                                // [groups][nw] kept-vertex bitmaps (light mode)
    int32_t* thr = nullptr;     // [groups][64] bucket bound per lane (delta-stepping)
    int32_t* fmin = nullptr;    // [groups][64] smallest deferred distance this round
    uint64_t* lor = nullptr;    // [groups][2] lanes with near / far work this round
    int4* items = nullptr;
    uint64_t* imask = nullptr;
    int4* hitems = nullptr;     // hidden-list items (pruned graphs)
    uint64_t* himask = nullptr;
    int64_t hitemcap = 0;
    int32_t* srcs = nullptr;
    int64_t* sums = nullptr;    // [64*groups]
    int32_t* reach = nullptr;   // [64*groups]
    int64_t* medges = nullptr;  // [64*groups]
    hipEvent_t ev[2] = {nullptr, nullptr};
    ColStage cols;              // per-query distance rows to the host
    // Rows a traversal can write: vertices with an in-edge (live or hidden).
    // After a traversal that converged, groups [0, clean_groups) hold
    // INT32_MAX in every other row except the entries of the sources in
    // srcs[0..nsrcs) (device copy of the last traversal's sources), and their
    // masks, pending lanes and bitmaps are zero: the next traversal resets
    // only the live rows and those entries.
    int32_t* live = nullptr;
    int64_t nlive = 0;
    std::vector<uint8_t> has_in;  // [V] 1 = vertex has an in-edge (the graph's ids)
    bool live_rl = false;         // the live list is in relabelled ids
    int clean_groups = 0;
    int nsrcs = 0;
    // Relabelled copy of the graph for the batched kernels
    // (c2g_tune.batch_relabel): vertices with an in-edge first, in their
    // order, then the rest.  rrow/radj/rhrow/rhadj mirror d_row/d_adj/
    // d_hrow/d_hadj (radj carries the light rows after its E live entries,
    // rlrow their absolute offsets, as light_build lays out d_adj); rnew[v]
    // is v's new id.
    int32_t* rrow = nullptr;
    int2* radj = nullptr;
    int64_t radj_cap = 0;
    int32_t* rlrow = nullptr;
    int32_t* rhrow = nullptr;
    int2* rhadj = nullptr;
    int32_t* rnew = nullptr;
    std::vector<int32_t> newid, oldid;
    bool rl_base = false;  // rrow, radj's live part, rhrow, rhadj, rnew built
    int rl_light = -1;     // light bound of radj's tail and rlrow (0: none)
    // static expand bitmaps (batch_sbits_k) and the (item size, light bound)
    // they were built for
    uint32_t* sbits = nullptr;
    int sbits_chunk = -1, sbits_light = -1;
    int sbits_rl = -1;
    int exec_rl = -1;  // relabelled mode of the captured round sequence
};

static std::mutex g_bs_mu;
static std::vector<std::pair<const Graph*, BatchState*>> g_bs;

static BatchState* bstate(Graph* g) {
    std::lock_guard<std::mutex> lk(g_bs_mu);
    for (auto& p : g_bs)
        if (p.first == g) return p.second;
    auto* s = new BatchState;
    g_bs.emplace_back(g, s);
    return s;
}

void batch_release(Graph* g) {
    std::lock_guard<std::mutex> lk(g_bs_mu);
    for (size_t i = 0; i < g_bs.size(); ++i)
        if (g_bs[i].first == g) {
            BatchState* s = g_bs[i].second;
            (void)hipFree(s->abits); (void)hipFree(s->items);
            (void)hipFree(s->thr); (void)hipFree(s->fmin); (void)hipFree(s->lor);
            (void)hipFree(s->imask); (void)hipFree(s->srcs); (void)hipFree(s->sums);
            (void)hipFree(s->reach); (void)hipFree(s->medges);
            (void)hipFree(s->hitems); (void)hipFree(s->himask); (void)hipFree(s->live);
            (void)hipFree(s->sbits);
            (void)hipFree(s->rrow); (void)hipFree(s->radj); (void)hipFree(s->rlrow); (void)hipFree(s->rhrow);
            (void)hipFree(s->rhadj); (void)hipFree(s->rnew);
            for (auto& e : s->ev) if (e) (void)hipEventDestroy(e);
            s->cols.release();
            delete s;
            g_bs.erase(g_bs.begin() + i);
            return;
        }
}

// The light bound the batched kernels run with (0: light mode off).
static int light_key(Graph* g) { return light_rows(g) ? g->light_built : 0; }

// Builds (or completes) the relabelled copy for the current light bound.
static int rl_build(Graph* g, BatchState* s) {
    const int64_t V = g->V, E = g->E;
    const int L = light_key(g);
    if (s->rl_base && s->rl_light == L) return C2G_OK;
    if (g->batch.exec) { (void)hipGraphExecDestroy(g->batch.exec); g->batch.exec = nullptr; }
    s->rl_light = -1;  // the light tail and rlrow are invalid until rebuilt below
    std::vector<uint8_t> nz((size_t)V, 0);  // v has live or hidden out-edges (bit 31 of its id in edges)
    for (int64_t v = 0; v < V; ++v)
        nz[v] = g->vl[v + 1] > g->vl[v] || (!g->hvl.empty() && g->hvl[v + 1] > g->hvl[v]);
    // One CSR in new order: row offsets (plus `base`) and entries with
    // mapped targets, keeping the entries `keep` accepts; returns the count.
    auto build = [&](const std::vector<int32_t>& rl, const std::vector<int32_t>& ev, const std::vector<int32_t>& wv,
                     auto keep, int64_t base, std::vector<int32_t>& nrow, std::vector<int2>& nadj) -> int64_t {
        std::vector<int64_t> cnt((size_t)V + 1, 0);
        parallel_for(V, [&](int64_t lo, int64_t hi, int) {
            for (int64_t x = lo; x < hi; ++x) {
                const int32_t v = s->oldid[x];
                int64_t c = 0;
                for (int32_t e = rl[v]; e < rl[v + 1]; ++e) c += keep(e) ? 1 : 0;
                cnt[x + 1] = c;
            }
        }, 0, 1 << 14);
        for (int64_t x = 0; x < V; ++x) cnt[x + 1] += cnt[x];
        nrow.resize((size_t)V + 1);
        for (int64_t x = 0; x <= V; ++x) nrow[x] = (int32_t)(cnt[x] + base);
        nadj.resize((size_t)std::max<int64_t>(cnt[V], 1));
        parallel_for(V, [&](int64_t lo, int64_t hi, int) {
            for (int64_t x = lo; x < hi; ++x) {
                const int32_t v = s->oldid[x];
                int64_t k = cnt[x];
                for (int32_t e = rl[v]; e < rl[v + 1]; ++e) {
                    if (!keep(e)) continue;
                    const int32_t t = ev[e], nt = s->newid[t];
                    nadj[k++] = make_int2(nz[t] ? (int32_t)((uint32_t)nt | 0x80000000u) : nt, wv[e]);
                }
            }
        }, 0, 1 << 14);
        return cnt[V];
    };
    const auto all = [](int32_t) { return true; };
    std::vector<int32_t> nrow;
    std::vector<int2> nadj;
    if (!s->rl_base) {
        // vertices with an in-edge first (the rows a traversal gathers)
        std::vector<uint8_t> hin((size_t)V, 0);
        for (int32_t t : g->el) hin[t] = 1;
        for (int32_t t : g->hel) hin[t] = 1;
        s->oldid.clear();
        s->oldid.reserve((size_t)V);
        for (int64_t v = 0; v < V; ++v) if (hin[v]) s->oldid.push_back((int32_t)v);
        for (int64_t v = 0; v < V; ++v) if (!hin[v]) s->oldid.push_back((int32_t)v);
        s->newid.assign((size_t)V, 0);
        for (int64_t x = 0; x < V; ++x) s->newid[s->oldid[x]] = (int32_t)x;
        build(g->vl, g->el, g->wl, all, 0, nrow, nadj);
        if (!s->rrow) HIP_TRY(hipMalloc(&s->rrow, sizeof(int32_t) * (size_t)(V + 1)));
        HIP_TRY(hipMemcpy(s->rrow, nrow.data(), sizeof(int32_t) * (size_t)(V + 1), hipMemcpyHostToDevice));
        if (!s->rnew) HIP_TRY(hipMalloc(&s->rnew, sizeof(int32_t) * (size_t)std::max<int64_t>(V, 1)));
        HIP_TRY(hipMemcpy(s->rnew, s->newid.data(), sizeof(int32_t) * (size_t)V, hipMemcpyHostToDevice));
        if (g->d_hrow && g->d_hadj) {
            std::vector<int32_t> hrow;
            std::vector<int2> hadj;
            build(g->hvl, g->hel, g->hwl, all, 0, hrow, hadj);
            if (!s->rhrow) HIP_TRY(hipMalloc(&s->rhrow, sizeof(int32_t) * (size_t)(V + 1)));
            HIP_TRY(hipMemcpy(s->rhrow, hrow.data(), sizeof(int32_t) * (size_t)(V + 1), hipMemcpyHostToDevice));
            if (!s->rhadj) HIP_TRY(hipMalloc(&s->rhadj, sizeof(int2) * hadj.size()));
            HIP_TRY(hipMemcpy(s->rhadj, hadj.data(), sizeof(int2) * hadj.size(), hipMemcpyHostToDevice));
        }
    }
    // the light rows (edges of weight < L) after the E live entries
    std::vector<int32_t> lrow;
    std::vector<int2> ladj;
    int64_t EL = 0;
    if (L > 0) EL = build(g->vl, g->el, g->wl, [&](int32_t e) { return g->wl[e] < L; }, E, lrow, ladj);
    const int64_t need = std::max<int64_t>(E + EL, 1);
    if (s->radj_cap < need) {
        int2* nb = nullptr;
        HIP_TRY(hipMalloc(&nb, sizeof(int2) * (size_t)need));
        if (s->rl_base && E > 0) {
            const hipError_t e1 = hipMemcpy(nb, s->radj, sizeof(int2) * (size_t)E, hipMemcpyDeviceToDevice);
            if (e1 != hipSuccess) { (void)hipFree(nb); return hip_fail(e1, "relabelled rows"); }
        }
        (void)hipFree(s->radj);
        s->radj = nb;
        s->radj_cap = need;
    }
    if (!s->rl_base && E > 0) HIP_TRY(hipMemcpy(s->radj, nadj.data(), sizeof(int2) * (size_t)E, hipMemcpyHostToDevice));
    if (EL > 0) HIP_TRY(hipMemcpy(s->radj + E, ladj.data(), sizeof(int2) * (size_t)EL, hipMemcpyHostToDevice));
    if (L > 0) {
        if (!s->rlrow) HIP_TRY(hipMalloc(&s->rlrow, sizeof(int32_t) * (size_t)(V + 1)));
        HIP_TRY(hipMemcpy(s->rlrow, lrow.data(), sizeof(int32_t) * (size_t)(V + 1), hipMemcpyHostToDevice));
    }
    s->rl_base = true;
    s->rl_light = L;
    return C2G_OK;
}

// Whether the batched kernels run on the relabelled copy, and the arrays
// they read (the copy's, or the graph's own).
static bool rl_on(Graph* g) {
    BatchState* s = bstate(g);
    return g->tune.batch_relabel && s->rl_base && s->rl_light == light_key(g);
}
static const int32_t* b_row(Graph* g) { return rl_on(g) ? bstate(g)->rrow : g->d_row; }
static const int2* b_adj(Graph* g) { return rl_on(g) ? bstate(g)->radj : g->d_adj; }
static const int32_t* b_lrow(Graph* g) {
    if (!light_rows(g)) return nullptr;
    return rl_on(g) ? bstate(g)->rlrow : light_rows(g);
}
static const int32_t* b_hrow(Graph* g) { return rl_on(g) ? bstate(g)->rhrow : g->d_hrow; }
static const int2* b_hadj(Graph* g) { return rl_on(g) ? bstate(g)->rhadj : g->d_hadj; }

// Builds the relabelled copy when the tunable asks for it, then puts the
// live list in the ids the kernels will run with (the relabelled ids of the
// vertices with an in-edge are 0..nlive-1).  A change of ids leaves no group
// in a known state.
static int rl_sync(Graph* g) {
    BatchState* s = bstate(g);
    if (g->tune.batch_relabel)
        if (int rc = rl_build(g, s)) return rc;
    const bool on = rl_on(g);
    if (!s->live || on == s->live_rl) return C2G_OK;
    std::vector<int32_t> lv((size_t)s->nlive);
    if (on) {
        for (int64_t k = 0; k < s->nlive; ++k) lv[k] = (int32_t)k;
    } else {
        int64_t k = 0;
        for (int64_t v = 0; v < g->V; ++v)
            if (s->has_in[v]) lv[k++] = (int32_t)v;
    }
    s->clean_groups = 0;
    if (s->nlive > 0) HIP_TRY(hipMemcpy(s->live, lv.data(), sizeof(int32_t) * lv.size(), hipMemcpyHostToDevice));
    s->live_rl = on;
    return C2G_OK;
}

static int ensure_batch(Graph* g, int groups) {
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    HIP_TRY(hipSetDevice(g->device));
    const int64_t V = g->V;
    const int chunk = std::max(1, g->tune.batch_chunk);
    // each piece on its own, so a failed attempt is completed by the next call
    if (!s->ev[0]) HIP_TRY(hipEventCreate(&s->ev[0]));
    if (!s->ev[1]) HIP_TRY(hipEventCreate(&s->ev[1]));
    if (!w.ctrl) {
        HIP_TRY(hipMalloc(&w.ctrl, sizeof(int64_t) * kCtrlWords));
        HIP_TRY(hipMemset(w.ctrl, 0, sizeof(int64_t) * kCtrlWords));
    }
    if (!w.h_ctrl) HIP_TRY(hipHostMalloc((void**)&w.h_ctrl, sizeof(int64_t) * B_N, hipHostMallocDefault));
    // Each (v,g) is activated at most once per round: at most
    // ceil(deg/chunk) <= 1 + deg/chunk live items and as many hidden ones.
    const int64_t needi = groups * (V + g->E / chunk) + 64;
    const bool pruned = g->d_hrow != nullptr;
    const int64_t needh = pruned ? groups * (V + g->HE / chunk) + 64 : 0;
    if (groups > s->groups || s->itemcap < needi || s->hitemcap < needh) {
        if (w.exec) { (void)hipGraphExecDestroy(w.exec); w.exec = nullptr; }
        const int G = std::max(groups, s->groups);
        (void)hipFree(w.dist); (void)hipFree(w.mask[0]); (void)hipFree(w.hmask); (void)hipFree(w.ph);
        w.ph = nullptr;
        (void)hipFree(s->abits); (void)hipFree(s->items); (void)hipFree(s->imask); (void)hipFree(s->srcs);
        (void)hipFree(s->sums); (void)hipFree(s->reach); (void)hipFree(s->medges);
        (void)hipFree(s->thr); (void)hipFree(s->fmin); (void)hipFree(s->lor);
        (void)hipFree(s->hitems); (void)hipFree(s->himask);
        s->hitems = nullptr; s->himask = nullptr; s->hitemcap = 0;
        w.dist = nullptr; w.mask[0] = nullptr; w.hmask = nullptr;
        s->abits = nullptr; s->items = nullptr; s->imask = nullptr; s->srcs = nullptr;
        s->sums = nullptr; s->reach = nullptr; s->medges = nullptr;
        s->thr = nullptr; s->fmin = nullptr; s->lor = nullptr;
        // until every piece is in place the workspace counts as empty, so a
        // failed allocation is retried instead of used
        s->groups = 0; s->itemcap = 0; w.groups = 0; s->clean_groups = 0;
        s->nw = (((V + 31) / 32 + 1) + 63) & ~(int64_t)63;  // whole waves per group
        HIP_TRY(hipMalloc(&s->thr, sizeof(int32_t) * 64 * G));
        HIP_TRY(hipMalloc(&s->fmin, sizeof(int32_t) * 64 * G));
        HIP_TRY(hipMalloc(&s->lor, sizeof(uint64_t) * 2 * G));
        HIP_TRY(hipMalloc(&w.dist, sizeof(int32_t) * (size_t)G * V * 64 + 16));
        HIP_TRY(hipMalloc(&w.mask[0], sizeof(uint64_t) * (size_t)G * V + 8));
        if (g->d_hrow) HIP_TRY(hipMalloc(&w.hmask, sizeof(uint64_t) * (size_t)G * V + 8));
        const int64_t itemcap = G * (V + g->E / chunk) + 64;
        HIP_TRY(hipMalloc(&s->abits, sizeof(uint32_t) * s->nw * G * 2));
        HIP_TRY(hipMemset(s->abits, 0, sizeof(uint32_t) * s->nw * G * 2));
        HIP_TRY(hipMalloc(&s->items, sizeof(int4) * itemcap));
        HIP_TRY(hipMalloc(&s->imask, sizeof(uint64_t) * itemcap));
        if (pruned) {
            const int64_t hcap = G * (V + g->HE / chunk) + 64;
            HIP_TRY(hipMalloc(&s->hitems, sizeof(int4) * hcap));
            HIP_TRY(hipMalloc(&s->himask, sizeof(uint64_t) * hcap));
            s->hitemcap = hcap;
        }
        s->itemcap = itemcap;
        HIP_TRY(hipMalloc(&s->srcs, sizeof(int32_t) * 64 * G));
        HIP_TRY(hipMalloc(&s->sums, sizeof(int64_t) * 64 * G));
        HIP_TRY(hipMalloc(&s->reach, sizeof(int32_t) * 64 * G));
        HIP_TRY(hipMalloc(&s->medges, sizeof(int64_t) * 64 * G));
        s->groups = G;
        w.groups = G;
        w.exec_chunk = -1;
    }
    if (light_rows(g) && !w.ph) {
        if (w.exec) { (void)hipGraphExecDestroy(w.exec); w.exec = nullptr; }
        HIP_TRY(hipMalloc(&w.ph, sizeof(uint64_t) * (size_t)s->groups * V + 8));
        HIP_TRY(hipMemset(w.ph, 0, sizeof(uint64_t) * (size_t)s->groups * V + 8));
    }
    if (!s->live && V > 0) {
        // vertices with an in-edge: on R-MAT about 44 % of V, the only rows
        // besides the sources' that a traversal writes
        s->has_in.assign((size_t)V, 0);
        for (int32_t t : g->el) s->has_in[t] = 1;
        for (int32_t t : g->hel) s->has_in[t] = 1;
        std::vector<int32_t> lv;
        lv.reserve((size_t)V);
        for (int64_t v = 0; v < V; ++v)
            if (s->has_in[v]) lv.push_back((int32_t)v);
        HIP_TRY(hipMalloc(&s->live, sizeof(int32_t) * std::max<size_t>(lv.size(), 1)));
        if (!lv.empty())
            HIP_TRY(hipMemcpy(s->live, lv.data(), sizeof(int32_t) * lv.size(), hipMemcpyHostToDevice));
        s->nlive = (int64_t)lv.size();
        s->live_rl = false;
    }
    return rl_sync(g);
}

// ph when the light-edge mode runs (light rows built and ph allocated).
static uint64_t* light_ph(Graph* g) { return light_rows(g) ? g->batch.ph : nullptr; }
// The kept-vertex bitmaps of the light-edge mode (null when it is off).
static uint32_t* light_kbits(Graph* g) {
    BatchState* s = bstate(g);
    return light_ph(g) ? s->abits + s->nw * s->groups : nullptr;
}

static int relax_grid_b(const Graph* g) { return std::max(1, g->num_cus) * g->tune.grid_per_cu; }

// Builds the static expand bitmaps for the current item size and light
// bound if they are missing or were built for others.
static int ensure_sbits(Graph* g, hipStream_t strm) {
    BatchState* s = bstate(g);
    if (s->nw <= 0 || g->V <= 0) return C2G_OK;
    const int chunk = g->tune.batch_chunk, lb = light_ph(g) ? g->light_built : 0, rl = rl_on(g);
    if (s->sbits && s->sbits_chunk == chunk && s->sbits_light == lb && s->sbits_rl == rl) return C2G_OK;
    if (!s->sbits) HIP_TRY(hipMalloc(&s->sbits, sizeof(uint32_t) * 4 * (size_t)s->nw));
    s->sbits_chunk = s->sbits_light = s->sbits_rl = -1;
    hipLaunchKernelGGL(batch_sbits_k, dim3(std::max(1, g->num_cus) * 4), dim3(256), 0, strm, b_row(g),
                       lb ? b_lrow(g) : (const int32_t*)nullptr, (int64_t)g->V, s->nw, std::max(1, chunk),
                       s->sbits);
    HIP_TRY(hipGetLastError());
    s->sbits_chunk = chunk;
    s->sbits_light = lb;
    s->sbits_rl = rl;
    return C2G_OK;
}

// The static expand bitmaps when they match the current item size and light
// bound, else null (the expand then counts from row offsets).
static const uint32_t* valid_sbits(Graph* g) {
    BatchState* s = bstate(g);
    const int lb = light_ph(g) ? g->light_built : 0;
    return s->sbits && s->sbits_chunk == g->tune.batch_chunk && s->sbits_light == lb && s->sbits_rl == (int)rl_on(g)
               ? s->sbits
               : nullptr;
}

static int expand_b(Graph* g, int groups, hipStream_t strm) {
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
    const int64_t words = s->nw * groups;
    const int egrid = (int)std::max<int64_t>(1, std::min<int64_t>(relax_grid_b(g), (words + 255) / 256));
    const int delta = g->tune.delta;
    const size_t lds = delta > 0 ? sizeof(int32_t) * 64 * groups : 0;
    auto kern = pruned ? (delta > 0 ? expand_b_k<true, true> : expand_b_k<false, true>)
                       : (delta > 0 ? expand_b_k<true, false> : expand_b_k<false, false>);
    hipLaunchKernelGGL(kern, dim3(egrid), dim3(256), lds, strm, b_row(g),
                       pruned ? b_hrow(g) : (const int32_t*)nullptr, w.mask[0],
                       pruned ? w.hmask : (const uint64_t*)nullptr, s->abits, s->items, s->imask, s->hitems,
                       s->himask, w.ctrl, (int64_t)g->V, s->nw, groups, g->tune.batch_chunk, s->itemcap,
                       s->hitemcap, w.dist, s->thr, s->fmin, s->lor, delta,
                       light_ph(g) ? b_lrow(g) : (const int32_t*)nullptr, light_ph(g), light_kbits(g),
                       pruned ? (const uint32_t*)nullptr : valid_sbits(g));
    return C2G_OK;
}

// One round = relax_b + expand_b.  When relax_ms is non-null, events bracket
// the relax kernel and its duration is added to *relax_ms (profiling runs).
// skip_relax: the host knows both queues are empty (a round that only
// expands vertices kept for their full rows), so no relax is launched.
static int launch_round_b(Graph* g, int groups, double* relax_ms, hipStream_t strm = nullptr,
                          bool skip_relax = false) {
    if (!strm) strm = g->stream;
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
    const int grid = relax_grid_b(g);
    if (skip_relax) {
        expand_b(g, groups, strm);
        HIP_TRY(hipGetLastError());
        return C2G_OK;
    }
    if (relax_ms) HIP_TRY(hipEventRecord(s->ev[0], strm));
    auto relax = [&](auto hidden, auto block) {
        constexpr bool H = decltype(hidden)::value;
        constexpr int B = decltype(block)::value;
        hipLaunchKernelGGL((relax_b_k<H, B>), dim3(grid), dim3(B), 0, strm, H ? b_hadj(g) : b_adj(g), w.dist,
                           w.mask[0], H ? w.hmask : (uint64_t*)nullptr, H ? s->hitems : s->items,
                           H ? s->himask : s->imask, s->abits, w.ctrl, (int64_t)g->V, s->nw);
    };
    auto relax_tb = [&](auto hidden) {
        switch (g->tune.block) {
            case 128: relax(hidden, std::integral_constant<int, 128>()); break;
            case 512: relax(hidden, std::integral_constant<int, 512>()); break;
            default: relax(hidden, std::integral_constant<int, 256>()); break;
        }
    };
    if (pruned) relax_tb(std::true_type());  // hidden-list items first (see relax_b_k's round accounting)
    relax_tb(std::false_type());
    if (relax_ms) {
        HIP_TRY(hipEventRecord(s->ev[1], strm));
        HIP_TRY(hipEventSynchronize(s->ev[1]));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
        *relax_ms += ms;
    }
    expand_b(g, groups, strm);
    HIP_TRY(hipGetLastError());
    return C2G_OK;
}

static int ensure_exec_b(Graph* g, int groups) {
    BatchWs& w = g->batch;
    const int K = g->tune.iters_per_graph;
    BatchState* s = bstate(g);
    if (w.exec && w.exec_iters == K && w.exec_grid == relax_grid_b(g) &&
        w.exec_chunk == g->tune.batch_chunk && w.exec_groups == groups && w.exec_tb == g->tune.block &&
        s->exec_rl == (int)rl_on(g))
        return C2G_OK;
    if (w.exec) { (void)hipGraphExecDestroy(w.exec); w.exec = nullptr; }
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = C2G_OK;
    for (int r = 0; r < K && rc == C2G_OK; ++r) rc = launch_round_b(g, groups, nullptr);
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
    w.exec_grid = relax_grid_b(g);
    w.exec_chunk = g->tune.batch_chunk;
    w.exec_groups = groups;  // the captured expand scans this many groups
    w.exec_tb = g->tune.block;
    s->exec_rl = rl_on(g);
    return C2G_OK;
}

// Runs n queries: lane q of group q/64 starts at srcs[q].
static int batch_run(Graph* g, const int32_t* srcs, int32_t n, c2g_stats* st, double* relax_ms,
                     RoundLog* log = nullptr) {
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    const int groups = (n + 63) / 64;
    const int64_t V = g->V;
    const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
    const int fill_grid = std::max(1, g->num_cus) * 16;
    if (int rs = ensure_sbits(g, g->stream)) return rs;
    if (groups <= s->clean_groups && s->live) {
        // after a converged traversal: the last sources' entries and the
        // live rows (masks and bitmaps are already zero; hidden marks are
        // cleared in full)
        if (s->nsrcs > 0)
            hipLaunchKernelGGL(batch_reset_srcs_k, dim3((s->nsrcs + 255) / 256), dim3(256), 0, g->stream, w.dist,
                               s->srcs, s->nsrcs, V);
        hipLaunchKernelGGL(batch_fill_rows_k, dim3(fill_grid), dim3(256), 0, g->stream, w.dist, s->live, s->nlive,
                           groups, V);
        if (pruned)
            hipLaunchKernelGGL(batch_fill_k, dim3(fill_grid), dim3(256), 0, g->stream, (int32_t*)nullptr,
                               (int64_t)0, (uint64_t*)nullptr, w.hmask, (int64_t)groups * V, s->abits, (int64_t)0,
                               (uint64_t*)nullptr, (uint32_t*)nullptr);
    } else {
        hipLaunchKernelGGL(batch_fill_k, dim3(fill_grid), dim3(256), 0, g->stream, w.dist,
                           (int64_t)groups * V * 64, w.mask[0], pruned ? w.hmask : (uint64_t*)nullptr,
                           (int64_t)groups * V, s->abits, s->nw * groups, light_ph(g), light_kbits(g));
    }
    const int clean_before = groups <= s->clean_groups ? s->clean_groups : 0;
    s->clean_groups = 0;  // until this traversal converges
    if (rl_on(g)) {
        // the sources in the relabelled ids (copied before the vector goes)
        std::vector<int32_t> rs((size_t)n);
        for (int32_t q = 0; q < n; ++q) rs[q] = s->newid[srcs[q]];
        HIP_TRY(hipMemcpyAsync(s->srcs, rs.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(s->srcs, srcs, sizeof(int32_t) * n, hipMemcpyHostToDevice, g->stream));
    }
    s->nsrcs = n;
    hipLaunchKernelGGL(batch_thr_init_k, dim3((64 * groups + 255) / 256), dim3(256), 0, g->stream, s->thr, s->fmin,
                       s->lor, 64 * groups, g->tune.delta);
    hipLaunchKernelGGL(batch_seed_k, dim3((n + 255) / 256), dim3(256), 0, g->stream, s->srcs, n, b_row(g),
                       pruned ? b_hrow(g) : (const int32_t*)nullptr, w.dist, w.mask[0],
                       pruned ? w.hmask : (uint64_t*)nullptr, s->abits, w.ctrl, V, s->nw);
    // Turn the seeded activations into the first round's items.
    expand_b(g, groups, g->stream);
    HIP_TRY(hipGetLastError());
    const bool one = relax_ms || log;  // one round per host check
    const bool use_graph = g->tune.use_graph && !one;
    int rc;
    if (use_graph && (rc = ensure_exec_b(g, groups))) return rc;
    auto snap = [&](int64_t* q3) {
        q3[0] = w.h_ctrl[B_FRONT];
        q3[1] = w.h_ctrl[B_EDGES];
        q3[2] = w.h_ctrl[B_LEDGES];
    };
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (one) {
        // the host checks every round: start from the seeded round's counters
        HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, sizeof(int64_t) * B_N, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    if (log) {
        HIP_TRY(hipEventCreate(&ev[0]));
        if (hipEventCreate(&ev[1]) != hipSuccess) { (void)hipEventDestroy(ev[0]); return hip_fail(hipErrorOutOfMemory, "hipEventCreate"); }
        int64_t q3[3];
        snap(q3);
        log->start(q3);
    }
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() { for (int k = 0; k < 2; ++k) if (e[k]) (void)hipEventDestroy(e[k]); }
    } evg{ev};
    for (;;) {
        if (log) HIP_TRY(hipEventRecord(ev[0], g->stream));
        if (use_graph) {
            HIP_TRY(hipGraphLaunch(w.exec, g->stream));
        } else {
            // Profiled runs check after every round so that no empty relax
            // is launched (launch count == rounds with work).
            const int k = one ? 1 : g->tune.iters_per_graph;
            const bool empty = one && w.h_ctrl[B_ITEMS] == 0 && w.h_ctrl[B_HITEMS] == 0;
            for (int r = 0; r < k; ++r)
                if ((rc = launch_round_b(g, groups, relax_ms, nullptr, empty))) return rc;
        }
        if (log) HIP_TRY(hipEventRecord(ev[1], g->stream));
        HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, sizeof(int64_t) * B_N, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (w.h_ctrl[B_OVF]) return fail(C2G_EDEVICE, "batched frontier capacity exceeded (internal)");
        if (log) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
            int64_t q3[3];
            snap(q3);
            log->add(q3, w.h_ctrl[B_IMPR], w.h_ctrl[B_ITEMS_TOT], ms);
        }
        if (w.h_ctrl[B_ITEMS] == 0 && w.h_ctrl[B_HITEMS] == 0 && w.h_ctrl[B_FAR] == 0) break;
    }
    // converged: every mask, pending lane and bitmap of these groups is zero
    s->clean_groups = std::max(clean_before, groups);
    if (st) {
        st->lane = C2G_LANE_GPU;
        st->iterations = (int32_t)w.h_ctrl[B_ITER];
        st->frontier_total = w.h_ctrl[B_FRONT];
        st->edges_relaxed = w.h_ctrl[B_LEDGES];
        st->improvements = w.h_ctrl[B_IMPR];
        st->csr_edges = w.h_ctrl[B_EDGES];
        st->items = w.h_ctrl[B_ITEMS_TOT];
    }
    return C2G_OK;
}

// srcs: the lanes' sources (lane order); a source without in-edges is the
// only row outside the live list a lane can reach (its own, at distance 0).
static int batch_checksums(Graph* g, const int32_t* srcs, int n, int64_t* sums, int32_t* reached,
                           int64_t* medges_total) {
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    const int groups = (n + 63) / 64;
    const int64_t V = g->V;
    HIP_TRY(hipMemsetAsync(s->sums, 0, sizeof(int64_t) * 64 * groups, g->stream));
    HIP_TRY(hipMemsetAsync(s->reach, 0, sizeof(int32_t) * 64 * groups, g->stream));
    HIP_TRY(hipMemsetAsync(s->medges, 0, sizeof(int64_t) * 64 * groups, g->stream));
    for (int gi = 0; gi < groups; gi += kCkG)
        hipLaunchKernelGGL(batch_checksum_k, dim3(std::max(1, g->num_cus) * 8), dim3(256), 0, g->stream, b_row(g),
                           (const int32_t*)w.dist, V, gi, std::min(kCkG, groups - gi), s->sums, s->reach,
                           s->medges, (const int32_t*)s->live, s->live ? s->nlive : V);
    HIP_TRY(hipGetLastError());
    std::vector<int64_t> hs(64 * groups), hm(64 * groups);
    std::vector<int32_t> hr(64 * groups);
    HIP_TRY(hipMemcpyAsync(hs.data(), s->sums, sizeof(int64_t) * 64 * groups, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(hr.data(), s->reach, sizeof(int32_t) * 64 * groups, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(hm.data(), s->medges, sizeof(int64_t) * 64 * groups, hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    int64_t mt = 0;
    for (int q = 0; q < n; ++q) {
        if (s->live && !s->has_in[srcs[q]]) {  // the source's own row, at distance 0
            hr[q] += 1;
            hm[q] += g->vl[srcs[q] + 1] - g->vl[srcs[q]];
        }
        if (sums) sums[q] = hs[q];
        if (reached) reached[q] = hr[q];
        mt += hm[q];
    }
    if (medges_total) *medges_total = mt;
    return C2G_OK;
}

static int check_srcs(const Graph* g, const int32_t* srcs, int32_t n) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adj) return fail(C2G_ENOTSUP, "graph has no int weights (wlist) for SSSP");
    if (n < 1) return fail(C2G_EINVAL, "need at least one source");
    if (n > 64 * 1024) return fail(C2G_EINVAL, "at most 65536 sources per batch");
    if (g->tune.delta > 0 && n > 64 * 256) return fail(C2G_EINVAL, "at most 16384 sources per batch with delta > 0");
    for (int32_t q = 0; q < n; ++q)
        if (srcs[q] < 0 || srcs[q] >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", srcs[q], g->V);
    return C2G_OK;
}

// Lane order of a batch (c2g_tune.batch_order = 1): queries sorted by the
// number of vertices their source reaches in two hops (the sum of its
// out-neighbours' out-degrees), largest first, so that each 64-lane group
// holds queries whose frontiers grow at a similar pace.  A row of `dist` is
// fetched whenever any of its lanes is pending, so lanes that are pending in
// the same rounds share fetches: on the headline's mix (half hubs, half
// out-degree-1 sources) this cuts fetched rows by about 13 %.  perm[k] = the
// caller's index of lane k.
static std::vector<int32_t> lane_order(const Graph* g, const int32_t* srcs, int32_t n) {
    std::vector<int32_t> perm((size_t)n);
    for (int32_t q = 0; q < n; ++q) perm[q] = q;
    if (!g->tune.batch_order || n <= 64 || g->vl.empty()) return perm;
    std::vector<int64_t> reach((size_t)n, 0);
    for (int32_t q = 0; q < n; ++q) {
        const int32_t s = srcs[q];
        for (int32_t e = g->vl[s]; e < g->vl[s + 1]; ++e) {
            const int32_t t = g->el[e];
            reach[q] += g->vl[t + 1] - g->vl[t];
        }
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return reach[a] > reach[b]; });
    return perm;
}

int sssp_batch_group(Graph* g, const int32_t* srcs, int32_t n, int32_t* dist_out, int64_t* sums,
                     int32_t* reached, c2g_stats* st, RoundLog* log) {
    int rc = check_srcs(g, srcs, n);
    if (rc) return rc;
    const int groups = (n + 63) / 64;
    if ((rc = ensure_batch(g, groups))) return rc;
    const std::vector<int32_t> perm = lane_order(g, srcs, n);
    std::vector<int32_t> lsrc((size_t)n);
    for (int32_t k = 0; k < n; ++k) lsrc[k] = srcs[perm[k]];
    const double t0 = now_ms();
    double relax_ms = 0;
    const bool prof = g->tune.profile != 0;
    if ((rc = batch_run(g, lsrc.data(), n, st, prof ? &relax_ms : nullptr, log))) return rc;
    const double t1 = now_ms();
    int64_t medges = 0;
    std::vector<int64_t> lsum((size_t)n);
    std::vector<int32_t> lreach((size_t)n);
    if ((rc = batch_checksums(g, lsrc.data(), n, lsum.data(), lreach.data(), &medges))) return rc;
    int32_t rtot = 0;
    for (int32_t k = 0; k < n; ++k) {
        if (sums) sums[perm[k]] = lsum[k];
        if (reached) reached[perm[k]] = lreach[k];
        rtot += lreach[k];
    }
    if (st) {
        st->ms = t1 - t0;
        st->reached_edges = medges;
        st->kernel_ms = relax_ms;
        st->reached = rtot;
    }
    if (dist_out) {
        // Device layout [g][V][64] -> caller layout [q][V] (lane k is query perm[k]).
        std::vector<void*> outs((size_t)n);
        for (int32_t k = 0; k < n; ++k) outs[k] = dist_out + (int64_t)perm[k] * g->V;
        if ((rc = bstate(g)->cols.copy_out(g, g->stream, (const uint32_t*)g->batch.dist, n, outs.data(),
                                           rl_on(g) ? bstate(g)->rnew : nullptr)))
            return rc;
    }
    return C2G_OK;
}


// ------------------------------------------------------------ server lane
// A query joins lane `lane` of group 0: its distance column was reset to
// INT32_MAX when the lane was last freed (or at init).
__global__ void bserve_admit_k(const int32_t* __restrict__ row, const int32_t* __restrict__ hrow,
                               int32_t* __restrict__ dist, uint64_t* __restrict__ mask, uint64_t* __restrict__ hmask,
                               uint32_t* __restrict__ abits, int32_t* __restrict__ thr, int lane, int32_t s,
                               int delta) {
    dist[(int64_t)s * 64 + lane] = 0;
    thr[lane] = delta > 0 ? delta : INT32_MAX;
    // the source's hidden list is always relaxed (pruned graphs)
    if (hmask) atomicOr((unsigned long long*)&hmask[s], 1ull << lane);
    if (row[s + 1] > row[s] || (hrow && hrow[s + 1] > hrow[s])) {
        atomicOr((unsigned long long*)&mask[s], 1ull << lane);
        atomicOr(&abits[s >> 5], 1u << (s & 31));
    }
}

// Copies the distance columns of `n` lanes into out[k][V] and resets them
// (and, on pruned graphs, the lanes' hidden-edge marks).
__global__ void bserve_extract_k(int32_t* __restrict__ dist, uint64_t* __restrict__ hmask, int64_t V,
                                 uint64_t lanes, const int8_t* __restrict__ slot, const int32_t* __restrict__ perm,
                                 int32_t* __restrict__ out) {
    const int lane = lane_id();
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const bool mine = (lanes >> lane) & 1ull;
    const int k = slot[lane];
    for (int64_t v = wave; v < V; v += nwaves) {
        if (hmask && lane == 0 && (hmask[v] & lanes))
            atomicAnd((unsigned long long*)&hmask[v], ~(unsigned long long)lanes);
        if (!mine) continue;
        const int64_t x = (perm ? (int64_t)perm[v] : v) * 64 + lane;  // row of vertex v
        out[(int64_t)k * V + v] = dist[x];
        dist[x] = INT32_MAX;
    }
}

struct ServeState {
    int32_t* out = nullptr;   // [8][V] device staging
    int32_t* h_out = nullptr; // pinned mirror
    int8_t* slot = nullptr;
};
static std::mutex g_ss_mu;
static std::vector<std::pair<const Graph*, ServeState*>> g_ss;
static ServeState* sstate(Graph* g) {
    std::lock_guard<std::mutex> lk(g_ss_mu);
    for (auto& p : g_ss)
        if (p.first == g) return p.second;
    auto* s = new ServeState;
    g_ss.emplace_back(g, s);
    return s;
}
void serve_release(Graph* g) {
    std::lock_guard<std::mutex> lk(g_ss_mu);
    for (size_t i = 0; i < g_ss.size(); ++i)
        if (g_ss[i].first == g) {
            ServeState* s = g_ss[i].second;
            if (s->out) (void)hipFree(s->out);
            if (s->slot) (void)hipFree(s->slot);
            if (s->h_out) (void)hipHostFree(s->h_out);
            delete s;
            g_ss.erase(g_ss.begin() + i);
            return;
        }
}
constexpr int kServeStage = 8;  // lanes copied out per transfer

// Device bytes the server's 64-lane group (one group of the batched
// workspace plus the extraction staging) would add to what the handle holds.
int64_t bserve_ws_bytes(Graph* g) {
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.batch_chunk);
    const int64_t nw = (((V + 31) / 32 + 1) + 63) & ~(int64_t)63;
    const bool pruned = !g->hvl.empty();
    const int64_t needi = V + g->E / chunk + 64;
    const int64_t needh = pruned ? V + g->HE / chunk + 64 : 0;
    int64_t b = 0;
    {
        std::lock_guard<std::mutex> lk(g_bs_mu);
        bool have = false, rl = false;
        for (auto& p : g_bs)
            if (p.first == g) {
                have = p.second->groups >= 1 && p.second->itemcap >= needi && p.second->hitemcap >= needh;
                rl = p.second->rl_base && p.second->rl_light == (light_rows(g) ? g->light_built : 0);
            }
        if (!have)
            b += 256 * V + 8 * V * (pruned ? 2 : 1) + (light_rows(g) ? 8 * V : 0) + 8 * nw + 24 * (needi + needh) +
                 64 * (4 + 4 + 8 + 8 + 4 + 8) + 8 * kCtrlWords;
        if (g->tune.batch_relabel && !rl)  // the relabelled copy of the graph (at most)
            b += 8 * (g->E + (light_rows(g) ? g->EL : 0)) + 12 * (V + 1) + (pruned ? 4 * (V + 1) + 8 * g->HE : 0);
    }
    {
        std::lock_guard<std::mutex> lk(g_ss_mu);
        bool have = false;
        for (auto& p : g_ss)
            if (p.first == g && p.second->out) have = true;
        if (!have) b += 4 * (int64_t)kServeStage * V + 64;
    }
    return b;
}

int bserve_init(Graph* g, hipStream_t strm) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adj) return fail(C2G_ENOTSUP, "graph has no int weights (wlist) for SSSP");
    int rc = ensure_batch(g, 1);
    if (rc) return rc;
    ServeState* ss = sstate(g);
    if (!ss->out) HIP_TRY(hipMalloc(&ss->out, sizeof(int32_t) * (size_t)kServeStage * std::max(g->V, 1)));
    if (!ss->h_out)
        HIP_TRY(hipHostMalloc((void**)&ss->h_out, sizeof(int32_t) * (size_t)kServeStage * std::max(g->V, 1),
                              hipHostMallocDefault));
    if (!ss->slot) HIP_TRY(hipMalloc(&ss->slot, 64));
    BatchWs& w = g->batch;
    BatchState* s = bstate(g);
    const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
    s->clean_groups = 0;  // the server's lanes leave group 0 in no known state
    HIP_TRY(hipMemsetAsync(w.ctrl, 0, sizeof(int64_t) * kCtrlWords, strm));
    hipLaunchKernelGGL(batch_thr_init_k, dim3(1), dim3(64), 0, strm, s->thr, s->fmin, s->lor, 64, g->tune.delta);
    hipLaunchKernelGGL(batch_fill_k, dim3(std::max(1, g->num_cus) * 16), dim3(256), 0, strm, w.dist,
                       (int64_t)g->V * 64, w.mask[0], pruned ? w.hmask : (uint64_t*)nullptr, (int64_t)g->V,
                       s->abits, s->nw, light_ph(g), light_kbits(g));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(strm));
    return C2G_OK;
}

int bserve_admit(Graph* g, hipStream_t strm, int lane, int32_t src) {
    if (lane < 0 || lane > 63 || src < 0 || src >= g->V) return fail(C2G_EINVAL, "bad lane/source");
    BatchState* s = bstate(g);
    const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
    hipLaunchKernelGGL(bserve_admit_k, dim3(1), dim3(1), 0, strm, b_row(g),
                       pruned ? b_hrow(g) : (const int32_t*)nullptr, g->batch.dist, g->batch.mask[0],
                       pruned ? g->batch.hmask : (uint64_t*)nullptr, s->abits, s->thr, lane,
                       rl_on(g) ? s->newid[src] : src, g->tune.delta);
    HIP_TRY(hipGetLastError());
    return C2G_OK;
}

int bserve_step(Graph* g, hipStream_t strm, uint64_t* lanes_with_work) {
    int rc = ensure_sbits(g, strm);
    if (rc) return rc;
    rc = launch_round_b(g, 1, nullptr, strm);
    if (rc) return rc;
    BatchWs& w = g->batch;
    HIP_TRY(hipMemcpyAsync(w.h_ctrl, w.ctrl, sizeof(int64_t) * B_N, hipMemcpyDeviceToHost, strm));
    HIP_TRY(hipStreamSynchronize(strm));
    if (w.h_ctrl[B_OVF]) return fail(C2G_EDEVICE, "server lane capacity exceeded (internal)");
    *lanes_with_work = (uint64_t)w.h_ctrl[B_LANEOR];
    return C2G_OK;
}

int bserve_extract(Graph* g, hipStream_t strm, const int* lanes, int n, int32_t* const* outs) {
    ServeState* ss = sstate(g);
    const int64_t V = g->V;
    for (int b = 0; b < n; b += kServeStage) {
        const int m = std::min(kServeStage, n - b);
        int8_t slot[64];
        memset(slot, 0, sizeof slot);
        uint64_t lm = 0;
        for (int k = 0; k < m; ++k) { lm |= 1ull << lanes[b + k]; slot[lanes[b + k]] = (int8_t)k; }
        HIP_TRY(hipMemcpyAsync(ss->slot, slot, 64, hipMemcpyHostToDevice, strm));
        const bool pruned = g->d_hrow != nullptr && g->d_hadj != nullptr;
        hipLaunchKernelGGL(bserve_extract_k, dim3(std::max(1, g->num_cus) * 8), dim3(256), 0, strm,
                           g->batch.dist, pruned ? g->batch.hmask : (uint64_t*)nullptr, V, lm, ss->slot,
                           rl_on(g) ? bstate(g)->rnew : (const int32_t*)nullptr, ss->out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(ss->h_out, ss->out, sizeof(int32_t) * V * m, hipMemcpyDeviceToHost, strm));
        HIP_TRY(hipStreamSynchronize(strm));
        parallel_for((int64_t)m * V, [&](int64_t lo, int64_t hi, int) {
            while (lo < hi) {
                const int64_t k = lo / V, v0 = lo - k * V, len = std::min(hi - lo, V - v0);
                if (outs[b + k]) memcpy(outs[b + k] + v0, ss->h_out + k * V + v0, sizeof(int32_t) * len);
                lo += len;
            }
        }, 0, (int64_t)1 << 18);
    }
    return C2G_OK;
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_sssp_batch(c2g_graph* h, const int32_t* srcs, int32_t n, int32_t* dist_out, c2g_stats* st) {
    if (!h || !srcs || !dist_out) return fail(C2G_EINVAL, "null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return sssp_batch_group(&h->g, srcs, n, dist_out, nullptr, nullptr, st);
}

int c2g_sssp_batch_rounds(c2g_graph* h, const int32_t* srcs, int32_t n, c2g_round* rounds, int32_t cap,
                          int32_t* nrounds, c2g_stats* st) {
    if (!h || !srcs || (cap > 0 && !rounds) || cap < 0) return fail(C2G_EINVAL, "c2g_sssp_batch_rounds: bad argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    RoundLog log;
    log.out = rounds;
    log.cap = cap;
    int rc = sssp_batch_group(&h->g, srcs, n, nullptr, nullptr, nullptr, st, &log);
    if (nrounds) *nrounds = log.n;
    return rc;
}

int c2g_sssp_batch_resident(c2g_graph* h, const int32_t* srcs, int32_t n, int64_t* sums,
                            int32_t* reached, c2g_stats* st) {
    if (!h || !srcs) return fail(C2G_EINVAL, "null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return sssp_batch_group(&h->g, srcs, n, nullptr, sums, reached, st);
}

}  // extern "C"
