// Forward-push personalized PageRank on gfx950 (FP32 state).
//
// Replaces calcuatePPR<float> (src/ppr/base/Main.cu:121-388) and its pruned
// variant (src/ppr/purn/Main.cu:78-101, 149-207).  Semantics kept: the source
// starts with r=1 and p=alpha (Main.cu:146-150); each round every active u
// sends (1-alpha)*r[u]*f(u,v) into a message buffer m[v] and its residual
// drops to 0 (Main.cu:208-320); then every touched v adds m[v] into r[v] and
// becomes active, crediting p[v] += alpha*r[v], when r[v] exceeds its
// threshold (Main.cu:339-367).  Dangling mass is dropped (the redirect is
// commented out in the reference, Main.cu:216-221).  Changed on purpose: the
// threshold is eps*outdeg(v) as in the reference CPU lane
// (src/ppr/w-base-perk/PPR.cpp:220) instead of one global rmax, alpha and eps
// are arguments, and rounds run to convergence (no 500-round cap, Main.cu:184).
//
// Single source, one round = two kernels:
//   ppr_push_k   : items (u, e_begin, e_end, bits of r[u]) walked 64 edges at
//                  a time per wave (shuffle owner search as in SSSP); one
//                  hardware FP32 atomicAdd into m[v] per edge, and v marked in
//                  a touched bitmap (instead of the reference's full-V scan).
//                  On dense rounds (the previous expand activated more than
//                  the tunable share of all out-edges) the same kernel runs
//                  the pull form instead: a coalesced sweep over the in-edge
//                  rows, m[v] = (1-alpha) * sum over activated in-neighbours u
//                  of r_u * f(u,v), one segmented wave sum and one add per
//                  row segment.  Activated u are recognised by a round stamp
//                  next to the residual they push.
//   ppr_expand_k : walks the touched bitmap; a touched vertex adds its
//                  message, activates above threshold, and its out-row is cut
//                  into items carrying the residual being pushed (wave scan +
//                  one atomicAdd per wave).
// Pruned graphs (ppr/ variant): the first round's frontier is the source's
// hidden and live neighbours with r = (1-alpha)*w (src/ppr/purn/Main.cu:78-101,
// 193-198), activated unconditionally.
//
// Batched (64 queries per group, lane = query), [g][V][64] float arrays:
//   pprb_push_k     : one item per wave; each edge is one 256-byte FP32
//                     atomicAdd row over the active lanes, plus one touched bit.
//   pprb_activate_k : one wave per touched vertex: per lane r += m, threshold,
//                     credit p, snapshot the pushed residual; lanes -> mask.
//   pprb_emit_k     : compacts touched vertices with a non-empty mask into items.
#include "graph.h"
#include "wave.h"
#include <cstring>

namespace c2g {

enum : int {
    P_CUR = 0, P_ITER = 1, P_FRONT = 4, P_EDGES = 5, P_ITEMS = 6,
    P_OVF = 7,
    P_MODE = 8,    // single source: 1 = the next round runs as a pull sweep
    P_REDGES = 9,  // single source: out-edges of the vertices the running expand activates
    P_PULLS = 10,  // single source: rounds run as pull sweeps
    P_WANT = 11,   // single source: some round qualified for a pull sweep
    P_N = 16       // fields; the block is kCtrlWords long (completion tree after them)
};
static_assert(P_N <= kCtrlTree, "control fields overlap the completion tree");

constexpr int kPullChunk = 64;  // in-edges per pull item
constexpr int kPullWin = 4;     // 64-edge windows of a pull step

__global__ void ppr_init_k(float* __restrict__ p, float* __restrict__ r, float* __restrict__ m,
                           uint32_t* __restrict__ tb, int64_t n) {
    const int64_t nw = (n + 31) >> 5;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        p[i] = 0.0f;
        r[i] = 0.0f;
        m[i] = 0.0f;
        if (i < nw) tb[i] = 0;
    }
}

__device__ __forceinline__ void touch(uint32_t* tb, int v) { atomicOr(&tb[v >> 5], 1u << (v & 31)); }

// Seeds messages for the first round (activated unconditionally by the expand).
__global__ void ppr_seed_k(const int32_t* __restrict__ row, const int2* __restrict__ adjf,
                           const int32_t* __restrict__ hrow, const int2* __restrict__ hadjf,
                           float* __restrict__ p, float* __restrict__ m, uint32_t* __restrict__ tb,
                           int64_t* __restrict__ ctrl, int32_t src, float alpha, int pruned, float res0,
                           float p_src) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < P_N; ++i) ctrl[i] = 0;
        done_tree_clear(ctrl);
    }
    if (!pruned) {
        // residual res0 at the source; its first (forced) activation adds alpha*res0 to p_src
        if (threadIdx.x == 0) {
            m[src] = res0;
            p[src] = p_src;
            touch(tb, src);
        }
        return;
    }
    // Pruned graph: p[s] = alpha, and every hidden then live neighbour h gets
    // r[h] = (1-alpha)*w(s,h) (its p credit alpha*r[h] comes with activation).
    if (threadIdx.x == 0) p[src] = alpha;
    const int h0 = hrow[src], h1 = hrow[src + 1], r0 = row[src], r1 = row[src + 1];
    for (int j = threadIdx.x; j < (h1 - h0) + (r1 - r0); j += blockDim.x) {
        const int2 a = j < h1 - h0 ? hadjf[h0 + j] : adjf[r0 + j - (h1 - h0)];
        atomicAdd(&m[a.x], (1.0f - alpha) * __int_as_float(a.y));
        touch(tb, a.x);
    }
}

// Pushes the queued residuals along their edges: m[v] += (1-alpha) r_u f(u,v)
// (FP32 atomics), and marks v in the touched bitmap for the expand.
// form: 0 = the round's form as the expand chose it; 1 = push rounds only,
// 2 = pull rounds only (a launch of the other form returns at once and
// leaves the round to its partner launch, so push rounds can run on a
// small grid and pull sweeps on a large one).
__global__ void __launch_bounds__(256) C2G_OCC8_SGPR
ppr_push_k(const int2* __restrict__ adjf, float* __restrict__ m, uint32_t* __restrict__ tb,
           const int4* __restrict__ q, int64_t* __restrict__ ctrl, float alpha,
           const int2* __restrict__ tadjf, const int4* __restrict__ titems, int64_t nt,
           const int2* __restrict__ ru, const uint32_t* __restrict__ stamp, int form) {
    const bool pull = ctrl[P_MODE] != 0;
    if ((form == 1 && pull) || (form == 2 && !pull)) return;  // the partner launch's round
    const int64_t n = ctrl[P_CUR];
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = lane_id();
    if (pull) {
        // Pull sweep over the in-edge items (v, e_begin, e_end): the edges of
        // 64 items are walked 64 at a time; lanes of one item form a segment
        // whose sum goes to m[v] with one atomic (an item can continue in the
        // next window, so its parts are added).
        const int S = (int)*stamp;
        for (int64_t base = wave * kWave; base < nt; base += nwaves * kWave) {
            const int64_t i = base + lane;
            int eb = 0, cnt = 0, v = 0;
            if (i < nt) {
                const int4 it = titems[i];
                v = it.x;
                eb = it.y;
                cnt = it.z - it.y;
            }
            int tot;
            const int pre = wave_excl_scan(cnt, &tot);
            // four 64-edge windows per step: their edge loads, then their
            // residual gathers, are in flight together
            for (int s0 = 0; s0 < tot; s0 += kPullWin * kWave) {
                int o[kPullWin], vo[kPullWin];
                int2 a[kPullWin];
                float val[kPullWin];
#pragma unroll
                for (int w = 0; w < kPullWin; ++w) {
                    const int k = s0 + w * kWave + lane;
                    o[w] = wave_owner(pre, k);
                    const int e = __shfl(eb, o[w], 64) + (k - __shfl(pre, o[w], 64));
                    vo[w] = __shfl(v, o[w], 64);
                    a[w] = make_int2(0, 0);
                    if (k < tot) a[w] = tadjf[e];
                }
#pragma unroll
                for (int w = 0; w < kPullWin; ++w) {
                    const int k = s0 + w * kWave + lane;
                    val[w] = 0.0f;
                    if (k < tot) {
                        const int2 x = ru[a[w].x];
                        if (x.x == S) val[w] = __int_as_float(x.y) * __int_as_float(a[w].y);
                    }
                }
#pragma unroll
                for (int w = 0; w < kPullWin; ++w) {
                    const int k = s0 + w * kWave + lane;
                    if (s0 + w * kWave >= tot) break;  // uniform: no edges left in this window
#pragma unroll
                    for (int d = 1; d < kWave; d <<= 1) {
                        const float t = __shfl_up(val[w], d, 64);
                        const int to = __shfl_up(o[w], d, 64);
                        if (lane >= d && to == o[w]) val[w] += t;
                    }
                    const int on = __shfl_down(o[w], 1, 64);
                    const bool last = k < tot && (lane == kWave - 1 || on != o[w] || k + 1 == tot);
                    if (last && val[w] != 0.0f) {
                        unsafeAtomicAdd(&m[vo[w]], (1.0f - alpha) * val[w]);
                        touch(tb, vo[w]);
                    }
                }
            }
        }
    }
    for (int64_t base = wave * kWave; !pull && base < n; base += nwaves * kWave) {
        const int64_t i = base + lane;
        int eb = 0, cnt = 0;
        float msg = 0.0f;
        if (i < n) {
            const int4 it = q[i];
            eb = it.y;
            cnt = it.z - it.y;
            msg = (1.0f - alpha) * __int_as_float(it.w);
        }
        int tot;
        const int pre = wave_excl_scan(cnt, &tot);
        for (int s = 0; s < tot; s += 2 * kWave) {
            const int k0 = s + lane, k1 = s + kWave + lane;
            const int o0 = wave_owner(pre, k0), o1 = wave_owner(pre, k1);
            const int e0 = __shfl(eb, o0, 64) + (k0 - __shfl(pre, o0, 64));
            const int e1 = __shfl(eb, o1, 64) + (k1 - __shfl(pre, o1, 64));
            const float m0 = __shfl(msg, o0, 64), m1 = __shfl(msg, o1, 64);
            int2 a0 = make_int2(0, 0), a1 = make_int2(0, 0);
            if (k0 < tot) a0 = adjf[e0];
            if (k1 < tot) a1 = adjf[e1];
            if (k0 < tot) {
                unsafeAtomicAdd(&m[a0.x], m0 * __int_as_float(a0.y));
                touch(tb, a0.x);
            }
            if (k1 < tot) {
                unsafeAtomicAdd(&m[a1.x], m1 * __int_as_float(a1.y));
                touch(tb, a1.x);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (wg_last(ctrl + kCtrlTree)) {
            if (n > 0) ctrl[P_ITER] += 1;
            if (n > 0 && pull) ctrl[P_PULLS] += 1;
            ctrl[P_ITEMS] += n;
            ctrl[P_CUR] = 0;
        }
    }
}

// Threshold pass over the touched vertices only (the touched bitmap, one
// 32-vertex word per lane): r[v] += m[v]; if r[v] > eps*outdeg(v) (eps
// alone when !degscale; always in the `force` first round), p[v] += credit*r[v]
// (credit = alpha for PPR), r[v] = 0 and v's items join
// the next round carrying the residual to push.  Two passes per word: count
// the items of the activated vertices, take the queue range with one atomic
// per wave, then apply the updates and write the items.  With ru set, each
// activated v also gets ru[v] = (stamp+1, residual) for a pull round, and the
// last workgroup picks the next round's form: pull when the activated
// vertices hold more than pull_edges out-edges and the in-edge rows exist
// (can_pull); a qualifying round without them sets P_WANT, and the host
// builds them at its next check.
__global__ void __launch_bounds__(256) C2G_OCC8_SGPR
ppr_expand_k(const int32_t* __restrict__ row, float* __restrict__ p, float* __restrict__ r,
             float* __restrict__ m, uint32_t* __restrict__ tb, int4* __restrict__ q, int64_t* __restrict__ ctrl,
             int64_t V, float credit, float eps, int chunk, int64_t qcap, int force, int degscale,
             int2* __restrict__ ru, uint32_t* __restrict__ stamp, int64_t pull_edges, int can_pull) {
    const int lane = lane_id();
    const int S1 = ru ? (int)*stamp + 1 : 0;
    const int64_t nw = (V + 31) >> 5;
    // workgroup-wide queue reservations and counter totals: one global
    // atomic each per workgroup rather than per wave
    __shared__ int64_t s_tot[4], s_base;
    __shared__ unsigned long long s_acc[2];
    if (threadIdx.x < 2) s_acc[threadIdx.x] = 0;
    int64_t front = 0, edges = 0;
    // uniform trip count per workgroup (wg_reserve holds barriers)
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < nw; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        uint32_t word = 0;
        if (i < nw) {
            word = tb[i];
            if (word) tb[i] = 0;
        }
        int cnt = 0;
        for (uint32_t b = word; b; b &= b - 1) {
            const int64_t v = i * 32 + __builtin_ctz(b);
            const float mv = m[v];
            if (mv == 0.0f) continue;
            const int d = row[v + 1] - row[v];
            if (force || r[v] + mv > (degscale ? (float)d * eps : eps)) cnt += (d + chunk - 1) / chunk;
        }
        int64_t wend;
        const int64_t pos = wg_reserve(cnt, (unsigned long long*)&ctrl[P_CUR], s_tot, &s_base, &wend);
        const bool room = wend <= qcap;
        if (!room && lane == 0) atomicOr((unsigned long long*)&ctrl[P_OVF], 1ull);
        // One vertex per lane per step (each lane takes the next bit of its
        // own word); the step's items are written by the whole wave, item j
        // by lane j % 64, from the wave's base (lane 0's position) on.
        int64_t pw = shfl64(pos, 0);
        for (uint32_t b = word; __ballot(b != 0);) {
            int64_t v = 0;
            int r0 = 0, r1 = 0, c = 0;
            float rv = 0.0f;
            if (b) {
                v = i * 32 + __builtin_ctz(b);
                b &= b - 1;
                const float mv = m[v];
                if (mv != 0.0f) {
                    m[v] = 0.0f;
                    rv = r[v] + mv;
                    r0 = row[v];
                    r1 = row[v + 1];
                    if (force || rv > (degscale ? (float)(r1 - r0) * eps : eps)) {
                        p[v] += credit * rv;
                        r[v] = 0.0f;
                        if (ru) ru[v] = make_int2(S1, __float_as_int(rv));
                        front += 1;
                        edges += r1 - r0;
                        if (room) c = (r1 - r0 + chunk - 1) / chunk;
                    } else {
                        r[v] = rv;
                    }
                }
            }
            int ctot;
            const int cpre = wave_excl_scan(c, &ctot);
            for (int j0 = 0; j0 < ctot; j0 += kWave) {
                const int j = j0 + lane;
                const int o = wave_owner(cpre, min(j, ctot - 1));
                const int kk = j - __shfl(cpre, o, 64);
                const int vo = __shfl((int)v, o, 64), bo = __shfl(r0, o, 64), eo = __shfl(r1, o, 64);
                const int ro = __shfl(__float_as_int(rv), o, 64);
                if (j < ctot) {
                    const int e = bo + kk * chunk;
                    q[pw + j] = make_int4(vo, e, min(e + chunk, eo), ro);
                }
            }
            pw += ctot;
        }
    }
    front = wave_sum64(front);
    edges = wave_sum64(edges);
    __syncthreads();  // s_acc was zeroed before the loop's first barrier or here
    if (lane == 0 && front) {
        atomicAdd(&s_acc[0], (unsigned long long)front);
        atomicAdd(&s_acc[1], (unsigned long long)edges);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_acc[0]) {
            atomicAdd((unsigned long long*)&ctrl[P_FRONT], s_acc[0]);
            atomicAdd((unsigned long long*)&ctrl[P_EDGES], s_acc[1]);
            atomicAdd((unsigned long long*)&ctrl[P_REDGES], s_acc[1]);
        }
        wait_vmem();  // this workgroup's counts are performed before it counts off
        if (wg_last(ctrl + kCtrlTree)) {
            int64_t c = (int64_t)atomicAdd((unsigned long long*)&ctrl[P_CUR], 0ull);
            if (c > qcap) { ctrl[P_OVF] = 1; ctrl[P_CUR] = 0; }
            const int64_t re = (int64_t)atomicAdd((unsigned long long*)&ctrl[P_REDGES], 0ull);
            const bool want = ru && pull_edges > 0 && re > pull_edges;
            ctrl[P_MODE] = (want && can_pull) ? 1 : 0;
            if (want) ctrl[P_WANT] = 1;
            ctrl[P_REDGES] = 0;
            if (ru) *stamp = (uint32_t)S1;
        }
    }
}

// ---------------------------------------------------------------- batched
__global__ void pprb_fill_k(float* __restrict__ a, int64_t n4, uint32_t* __restrict__ bits, int64_t nb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    float4* a4 = reinterpret_cast<float4*>(a);
    for (int64_t i = t; i < n4; i += stride) a4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t i = t; i < nb; i += stride) bits[i] = 0;
}

// Lane q of group q/64 starts at srcs[q]: m = 1 (or the pruned one-hop seeds).
__global__ void pprb_seed_k(const int32_t* __restrict__ srcs, int n, const int32_t* __restrict__ row,
                            const int2* __restrict__ adjf, const int32_t* __restrict__ hrow,
                            const int2* __restrict__ hadjf, float* __restrict__ P, float* __restrict__ M,
                            uint32_t* __restrict__ tbits, int64_t* __restrict__ ctrl, int64_t V, int64_t nw,
                            float alpha) {
    const int q = blockIdx.x;
    if (q == 0 && threadIdx.x == 0) {
        for (int i = 0; i < P_N; ++i) ctrl[i] = 0;
        done_tree_clear(ctrl);
    }
    if (q >= n) return;
    const int g = q >> 6, lane = q & 63;
    const int s = srcs[q];
    float* Mg = M + (int64_t)g * V * 64;
    if (!hrow) {
        if (threadIdx.x == 0) {
            atomicAdd(&Mg[(int64_t)s * 64 + lane], 1.0f);
            atomicOr(&tbits[(int64_t)g * nw + (s >> 5)], 1u << (s & 31));
        }
        return;
    }
    if (threadIdx.x == 0) P[((int64_t)g * V + s) * 64 + lane] = alpha;
    const int h0 = hrow[s], h1 = hrow[s + 1], r0 = row[s], r1 = row[s + 1];
    for (int j = threadIdx.x; j < (h1 - h0) + (r1 - r0); j += blockDim.x) {
        const int2 a = j < h1 - h0 ? hadjf[h0 + j] : adjf[r0 + j - (h1 - h0)];
        atomicAdd(&Mg[(int64_t)a.x * 64 + lane], (1.0f - alpha) * __int_as_float(a.y));
        atomicOr(&tbits[(int64_t)g * nw + (a.x >> 5)], 1u << (a.x & 31));
    }
}

__global__ void __launch_bounds__(256)
pprb_push_k(const int2* __restrict__ adjf, float* __restrict__ M, const float* __restrict__ RU,
            const int4* __restrict__ items, const uint64_t* __restrict__ imask, uint32_t* __restrict__ tbits,
            int64_t* __restrict__ ctrl, int64_t V, int64_t nw, float alpha) {
    const int64_t n = ctrl[P_CUR];
    const int lane = lane_id();
    const uint32_t loff = (uint32_t)lane * 4u;  // this lane's byte offset in a 256-byte row
    const int64_t wave = wave_index();  // item headers through scalar loads
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = wave; i < n; i += nwaves) {
        const int4 it = items[i];
        const uint64_t mk = imask[i];
        const int g = it.w;
        const bool act = (mk >> lane) & 1ull;
        const float msg = act ? (1.0f - alpha) * RU[((int64_t)g * V + it.x) * 64 + lane] : 0.0f;
        float* Mg = M + (int64_t)g * V * 64;
        for (int e0 = it.y; e0 < it.z; e0 += 64) {
            const int ne = min(64, it.z - e0);
            int2 my = make_int2(0, 0);
            if (lane < ne) my = adjf[e0 + lane];
            if (lane < ne) atomicOr(&tbits[(int64_t)g * nw + (my.x >> 5)], 1u << (my.x & 31));
            // The item's pending lanes push along the chunk's edges.  An
            // edge's target, weight and row address are wave-uniform (SGPRs,
            // the edge read from the lane that loaded it: readlane ignores
            // exec), so per lane there is one multiply and the atomic.
            if (act) {
                gchar* MgS = sgpr_ptr(Mg);  // the group's rows, base in SGPRs
                for (int j = 0; j < ne; ++j) {
                    const int v = __builtin_amdgcn_readlane(my.x, j);
                    const float f = __int_as_float(__builtin_amdgcn_readlane(my.y, j));
                    gchar* row = MgS + (int64_t)v * 256;
                    [[clang::atomic(ignore_denormal_mode)]] {
                        (void)__hip_atomic_fetch_add(lane_slot<float>(row, loff), msg * f, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (wg_last(ctrl + kCtrlTree)) {
            if (n > 0) ctrl[P_ITER] += 1;
            ctrl[P_ITEMS] += n;
            ctrl[P_CUR] = 0;
        }
    }
}

// One wave per touched vertex (walking the touched bitmap word by word).
__global__ void __launch_bounds__(256)
pprb_activate_k(const int32_t* __restrict__ row, float* __restrict__ P, float* __restrict__ R,
                float* __restrict__ M, float* __restrict__ RU, uint64_t* __restrict__ mask,
                const uint32_t* __restrict__ tbits, int64_t V, int64_t nw, int groups, float alpha, float eps,
                int force) {
    const int lane = lane_id();
    const int64_t wave = wave_index();
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t words = nw * groups;
    for (int64_t wi = wave; wi < words; wi += nwaves) {
        uint32_t word = tbits[wi];
        if (!word) continue;
        const int g = (int)(wi / nw);
        const int64_t vb = (wi - (int64_t)g * nw) * 32;
        for (; word; word &= word - 1) {
            const int64_t v = vb + __builtin_ctz(word);
            const int64_t x = ((int64_t)g * V + v) * 64 + lane;
            const float mv = M[x];
            bool a = false;
            if (mv != 0.0f) {
                M[x] = 0.0f;
                const float rv = R[x] + mv;
                const float thr = (float)(row[v + 1] - row[v]) * eps;
                if (force || rv > thr) {
                    P[x] += alpha * rv;
                    R[x] = 0.0f;
                    RU[x] = rv;
                    a = true;
                } else {
                    R[x] = rv;
                }
            }
            const uint64_t b = __ballot(a);
            if (lane == 0) mask[(int64_t)g * V + v] = b;
        }
    }
}

__global__ void __launch_bounds__(256) C2G_OCC8_SGPR
pprb_emit_k(const int32_t* __restrict__ row, uint64_t* __restrict__ mask, uint32_t* __restrict__ tbits,
            int4* __restrict__ items, uint64_t* __restrict__ imask, int64_t* __restrict__ ctrl, int64_t V,
            int64_t nw, int groups, int chunk, int64_t itemcap) {
    const int lane = lane_id();
    __shared__ int64_t s_tot[4], s_base;
    __shared__ unsigned long long s_acc[2];
    if (threadIdx.x < 2) s_acc[threadIdx.x] = 0;
    int64_t front = 0, edges = 0;
    const int64_t n = nw * groups;
    // uniform trip count per workgroup (wg_reserve holds barriers)
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x; i0 < n; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        uint32_t word = 0;
        int g = 0;
        int64_t vb = 0;
        if (i < n) {
            word = tbits[i];
            if (word) tbits[i] = 0;
            g = (int)(i / nw);
            vb = (i - (int64_t)g * nw) * 32;
        }
        int cnt = 0;
        for (uint32_t b = word; b; b &= b - 1) {
            const int64_t v = vb + __builtin_ctz(b);
            if (mask[(int64_t)g * V + v]) {
                const int d = row[v + 1] - row[v];
                cnt += (d + chunk - 1) / chunk;
                front += 1;
                edges += d;
            }
        }
        int64_t wend;
        const int64_t pos = wg_reserve(cnt, (unsigned long long*)&ctrl[P_CUR], s_tot, &s_base, &wend);
        const bool room = wend <= itemcap;
        if (!room && lane == 0) atomicOr((unsigned long long*)&ctrl[P_OVF], 1ull);
        // one vertex per lane per step, the step's items written by the
        // whole wave from the wave's base (lane 0's position) on
        int64_t pw = shfl64(pos, 0);
        for (uint32_t b = word; __ballot(b != 0);) {
            int64_t v = 0;
            int r0 = 0, r1 = 0, c = 0;
            uint64_t mk = 0;
            if (b) {
                v = vb + __builtin_ctz(b);
                b &= b - 1;
                uint64_t* mp = &mask[(int64_t)g * V + v];
                mk = *mp;
                if (mk) {
                    *mp = 0;
                    r0 = row[v];
                    r1 = row[v + 1];
                    if (room) c = (r1 - r0 + chunk - 1) / chunk;
                }
            }
            int ctot;
            const int cpre = wave_excl_scan(c, &ctot);
            for (int j0 = 0; j0 < ctot; j0 += kWave) {
                const int j = j0 + lane;
                const int o = wave_owner(cpre, min(j, ctot - 1));
                const int kk = j - __shfl(cpre, o, 64);
                const int vo = __shfl((int)v, o, 64), bo = __shfl(r0, o, 64), eo = __shfl(r1, o, 64);
                const int go = __shfl(g, o, 64);
                const uint64_t mo = (uint64_t)shfl64((int64_t)mk, o);
                if (j < ctot) {
                    const int e = bo + kk * chunk;
                    items[pw + j] = make_int4(vo, e, min(e + chunk, eo), go);
                    imask[pw + j] = mo;
                }
            }
            pw += ctot;
        }
    }
    front = wave_sum64(front);
    edges = wave_sum64(edges);
    __syncthreads();
    if (lane == 0 && front) {
        atomicAdd(&s_acc[0], (unsigned long long)front);
        atomicAdd(&s_acc[1], (unsigned long long)edges);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_acc[0]) {
            atomicAdd((unsigned long long*)&ctrl[P_FRONT], s_acc[0]);
            atomicAdd((unsigned long long*)&ctrl[P_EDGES], s_acc[1]);
        }
        wait_vmem();  // this workgroup's counts are performed before it counts off
        if (wg_last(ctrl + kCtrlTree)) {
            int64_t c = (int64_t)atomicAdd((unsigned long long*)&ctrl[P_CUR], 0ull);
            if (c > itemcap) { ctrl[P_OVF] = 1; ctrl[P_CUR] = 0; }
        }
    }
}

// ------------------------------------------------------------------ host
struct PprState {
    // single
    float *p = nullptr, *r = nullptr, *m = nullptr;
    uint32_t* tb = nullptr;  // touched bitmap [V/32 + 1]
    int4* q = nullptr;
    int64_t qcap = 0;
    int64_t* ctrl = nullptr;
    int64_t* h_ctrl = nullptr;
    // batched
    int groups = 0;
    float *P = nullptr, *R = nullptr, *M = nullptr, *RU = nullptr;
    uint64_t* mask = nullptr;
    uint32_t* tbits = nullptr;
    int64_t nw = 0;
    int4* items = nullptr;
    uint64_t* imask = nullptr;
    int64_t itemcap = 0;
    int32_t* srcs = nullptr;
    ColStage cols;             // per-query rows out of P and R
    HostStage out;             // single-query p / r copies to the host
    // single-source pull rounds: in-edge rows (u, bits of f(u,v)) cut into
    // items of <= kPullChunk edges, and the stamped residuals
    int2* tadjf = nullptr;
    int4* titems = nullptr;
    int64_t nt = 0;
    int2* ru = nullptr;        // [V] (stamp of the activating expand, bits of the pushed residual)
    uint32_t* stamp = nullptr; // stamp of the last expand (device)
    int64_t stamp_est = 0;     // upper bound of *stamp kept on the host
};

static std::mutex g_ps_mu;
static std::vector<std::pair<const Graph*, PprState*>> g_ps;

static PprState* pstate(Graph* g) {
    std::lock_guard<std::mutex> lk(g_ps_mu);
    for (auto& p : g_ps)
        if (p.first == g) return p.second;
    auto* s = new PprState;
    g_ps.emplace_back(g, s);
    return s;
}

void ppr_release(Graph* g) {
    std::lock_guard<std::mutex> lk(g_ps_mu);
    for (size_t i = 0; i < g_ps.size(); ++i)
        if (g_ps[i].first == g) {
            PprState* s = g_ps[i].second;
            for (void* p : {(void*)s->p, (void*)s->r, (void*)s->m, (void*)s->tb, (void*)s->q, (void*)s->ctrl, (void*)s->P,
                            (void*)s->R, (void*)s->M, (void*)s->RU, (void*)s->mask, (void*)s->tbits,
                            (void*)s->items, (void*)s->imask, (void*)s->srcs, (void*)s->tadjf,
                            (void*)s->titems, (void*)s->ru, (void*)s->stamp})
                if (p) (void)hipFree(p);
            if (s->h_ctrl) (void)hipHostFree(s->h_ctrl);
            s->cols.release();
            s->out.release();
            delete s;
            g_ps.erase(g_ps.begin() + i);
            return;
        }
}

// Device bytes the server's PPR lane (single-source state, one 64-lane group
// of the batched state, its column staging and, when pull rounds are on, the
// in-edge rows) would add to what the handle holds.
int64_t ppr_lane_ws_bytes(Graph* g) {
    PprState* s = pstate(g);
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int64_t nw = (V + 31) / 32 + 1;
    int64_t b = 0;
    if (!s->p) b += 12 * V + 4 * nw + 8 * kCtrlWords;
    const int64_t needq = V + g->E / std::max(1, g->tune.chunk) + 64;
    if (s->qcap < needq) b += 16 * needq;
    const int64_t needi = (V + g->E / std::max(1, g->tune.batch_chunk)) + 64;
    if (s->groups < 1 || s->itemcap < needi) b += 4 * 4 * 64 * V + 8 * V + 4 * nw + 24 * needi + 4 * 64;
    if (s->cols.cap < kColStage * V) b += 2 * 4 * (int64_t)kColStage * V;
    if (g->tune.ppr_pull > 0 && !s->tadjf) b += 8 * g->E + 16 * (V + g->E / 64 + 1) + 8 * V;
    return b;
}

static bool ppr_pruned(const Graph* g) { return g->d_hrow != nullptr && g->d_hadjf != nullptr; }

static int ppr_check(Graph* g, float alpha, float eps) {
    if (!g->on_device) return fail(C2G_EDEVICE, "graph has no device image (no HIP device)");
    if (!g->d_adjf) return fail(C2G_ENOTSUP, "graph has no float weights (flist) for PPR");
    if (!(alpha > 0.0f && alpha < 1.0f)) return fail(C2G_EINVAL, "alpha must be in (0,1)");
    if (!(eps > 0.0f)) return fail(C2G_EINVAL, "eps must be > 0");
    return C2G_OK;
}

static int ensure_ppr_single(Graph* g, PprState* s) {
    HIP_TRY(hipSetDevice(g->device));
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.chunk);
    const int64_t need = V + g->E / chunk + 64;
    // each piece on its own, so a failed attempt is completed by the next call
    if (!s->p) HIP_TRY(hipMalloc(&s->p, sizeof(float) * V));
    if (!s->r) HIP_TRY(hipMalloc(&s->r, sizeof(float) * V));
    if (!s->m) HIP_TRY(hipMalloc(&s->m, sizeof(float) * V));
    if (!s->tb) HIP_TRY(hipMalloc(&s->tb, sizeof(uint32_t) * ((V + 31) / 32 + 1)));
    if (!s->ctrl) HIP_TRY(hipMalloc(&s->ctrl, sizeof(int64_t) * kCtrlWords));
    if (!s->h_ctrl) HIP_TRY(hipHostMalloc((void**)&s->h_ctrl, sizeof(int64_t) * P_N, hipHostMallocDefault));
    if (s->qcap < need) {
        if (s->q) (void)hipFree(s->q);
        s->q = nullptr;
        s->qcap = 0;
        HIP_TRY(hipMalloc(&s->q, sizeof(int4) * need));
        s->qcap = need;
    }
    return C2G_OK;
}

// Stamped residuals for pull rounds (cheap: 8 bytes per vertex); the stamps
// persist across queries and restart long before the 32-bit counter wraps.
static int ensure_pull(Graph* g, PprState* s) {
    const int64_t V = g->V;
    if (!s->ru || !s->stamp) {
        if (!s->ru) HIP_TRY(hipMalloc(&s->ru, sizeof(int2) * (size_t)std::max<int64_t>(V, 1)));
        if (!s->stamp) HIP_TRY(hipMalloc(&s->stamp, sizeof(uint32_t)));
        HIP_TRY(hipMemset(s->ru, 0, sizeof(int2) * (size_t)std::max<int64_t>(V, 1)));
        HIP_TRY(hipMemset(s->stamp, 0, sizeof(uint32_t)));
        s->stamp_est = 0;
    }
    if (s->stamp_est > ((int64_t)1 << 30)) {  // restart the stamps long before they wrap
        HIP_TRY(hipMemset(s->ru, 0, sizeof(int2) * (size_t)std::max<int64_t>(V, 1)));
        HIP_TRY(hipMemset(s->stamp, 0, sizeof(uint32_t)));
        s->stamp_est = 0;
    }
    return C2G_OK;
}

// In-edge rows of the PPR graph for pull rounds, built once per handle on
// the host: thread t owns a range of target vertices and scans every row in
// source order, so each in-row lists its sources in increasing order and the
// layout does not depend on the thread count.
static int ensure_transpose(Graph* g, PprState* s) {
    const int64_t V = g->V, E = g->E;
    if (s->tadjf) return C2G_OK;
    const int32_t* vl = g->vl.data();
    const int32_t* el = g->el.data();
    const float* fl = g->fl.data();
    std::vector<int64_t> trow((size_t)V + 1, 0);
    const int T = std::max(1, std::min(host_threads(), 16));
    parallel_for(T, [&](int64_t t0, int64_t t1, int) {
        for (int64_t t = t0; t < t1; ++t) {
            const int64_t lo = V * t / T, hi = V * (t + 1) / T;
            for (int64_t e = 0; e < E; ++e)
                if (el[e] >= lo && el[e] < hi) trow[(size_t)el[e] + 1] += 1;
        }
    }, T, 1);
    for (int64_t v = 0; v < V; ++v) trow[v + 1] += trow[v];
    std::vector<int2> tadj((size_t)std::max<int64_t>(E, 1));
    parallel_for(T, [&](int64_t t0, int64_t t1, int) {
        for (int64_t t = t0; t < t1; ++t) {
            const int64_t lo = V * t / T, hi = V * (t + 1) / T;
            std::vector<int64_t> pos(trow.begin() + lo, trow.begin() + hi);
            for (int64_t u = 0; u < V; ++u)
                for (int64_t e = vl[u]; e < vl[u + 1]; ++e) {
                    const int32_t v = el[e];
                    if (v < lo || v >= hi) continue;
                    int32_t fb;
                    memcpy(&fb, &fl[e], sizeof fb);
                    tadj[(size_t)pos[v - lo]++] = make_int2((int32_t)u, fb);
                }
        }
    }, T, 1);
    std::vector<int4> items;
    items.reserve((size_t)V + (size_t)(E / kPullChunk) + 1);
    for (int64_t v = 0; v < V; ++v)
        for (int64_t e = trow[v]; e < trow[v + 1]; e += kPullChunk)
            items.push_back(make_int4((int32_t)v, (int32_t)e, (int32_t)std::min<int64_t>(e + kPullChunk, trow[v + 1]), 0));
    // both arrays complete before s->tadjf marks them ready
    int2* ta = nullptr;
    int4* ti = nullptr;
    hipError_t e = hipMalloc(&ta, sizeof(int2) * tadj.size());
    if (e == hipSuccess) e = hipMemcpy(ta, tadj.data(), sizeof(int2) * tadj.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&ti, sizeof(int4) * std::max<size_t>(items.size(), 1));
    if (e == hipSuccess && !items.empty())
        e = hipMemcpy(ti, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (ta) (void)hipFree(ta);
        if (ti) (void)hipFree(ti);
        return hip_fail(e, "PPR in-edge rows");
    }
    s->titems = ti;
    s->nt = (int64_t)items.size();
    s->tadjf = ta;
    return C2G_OK;
}

static const int kPprMaxRounds = 100000;

// Seeding and threshold form of a single-source push: the workload's PPR is
// {threshold eps*outdeg, r0 = 1, p[s] starts at 0}; the reference's adsorption
// push is {constant threshold, r0 = 10, p[s] starts at (1-2*alpha)*r0}.
struct PushSpec {
    int degscale = 1;
    float r0 = 1.0f;
    float p_src = 0.0f;
    float credit = -1.0f;     // share of an activated residual credited to p (< 0: alpha)
    bool hidden_seed = true;  // pruned graphs: seed H(s) and N'(s) (ppr-purn)
};

static int ppr_single_ex(Graph* g, hipStream_t strm, int32_t src, float alpha, float eps, int32_t budget,
                         int32_t handoff_min, float* p_out, float* r_out, std::vector<int32_t>* frontier,
                         std::vector<float>* fr_ru, c2g_stats* st, const PushSpec& spec = PushSpec());

int ppr_single(Graph* g, hipStream_t strm, int32_t src, float alpha, float eps, float* p_out, float* r_out,
                      c2g_stats* st) {
    return ppr_single_ex(g, strm, src, alpha, eps, 0, 0, p_out, r_out, nullptr, nullptr, st);
}

// budget > 0: after `budget` rounds, stop as soon as the frontier holds at
// most handoff_min vertices; the remaining frontier is returned with r[u]
// restored to the residual it was about to push (p already credited).
static int ppr_single_ex(Graph* g, hipStream_t strm, int32_t src, float alpha, float eps, int32_t budget,
                         int32_t handoff_min, float* p_out, float* r_out, std::vector<int32_t>* frontier,
                         std::vector<float>* fr_ru, c2g_stats* st, const PushSpec& spec) {
    int rc = ppr_check(g, alpha, eps);
    if (rc) return rc;
    if (src < 0 || src >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", src, g->V);
    PprState* s = pstate(g);
    if ((rc = ensure_ppr_single(g, s))) return rc;
    const bool pruned = ppr_pruned(g);
    const float credit = spec.credit < 0.0f ? alpha : spec.credit;
    const int64_t V = g->V;
    // push rounds and the expand on the small single-source grid; pull
    // sweeps (dense rounds) on the full grid
    const int grid = std::max(1, g->num_cus) * std::max(1, g->tune.ppr_grid_per_cu);
    const int pgrid = std::max(1, g->num_cus) * g->tune.grid_per_cu;
    const int egrid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (V + 255) / 256));
    const int wgrid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, ((V + 31) / 32 + 255) / 256));
    // Pull rounds: a round whose frontier holds more than ppr_pull/1000 of E
    // out-edges gathers over in-edges instead of scattering.
    const bool use_pull = g->tune.ppr_pull > 0 && g->E > 0;
    if (use_pull && (rc = ensure_pull(g, s))) return rc;
    int can_pull = use_pull && s->tadjf ? 1 : 0;  // in-edge rows are built once a round qualifies
    int2* ru = use_pull ? s->ru : nullptr;
    const int64_t pull_edges = use_pull ? std::max<int64_t>(1, g->E * g->tune.ppr_pull / 1000) : 0;
    const double t0 = now_ms();
    hipLaunchKernelGGL(ppr_init_k, dim3(egrid), dim3(256), 0, strm, s->p, s->r, s->m, s->tb, V);
    const bool hseed = pruned && spec.hidden_seed;
    hipLaunchKernelGGL(ppr_seed_k, dim3(1), dim3(256), 0, strm, g->d_row, g->d_adjf,
                       hseed ? g->d_hrow : (const int32_t*)nullptr, hseed ? g->d_hadjf : (const int2*)nullptr,
                       s->p, s->m, s->tb, s->ctrl, src, alpha, hseed ? 1 : 0, spec.r0, spec.p_src);
    hipLaunchKernelGGL(ppr_expand_k, dim3(wgrid), dim3(256), 0, strm, g->d_row, s->p, s->r, s->m, s->tb, s->q,
                       s->ctrl, V, credit, eps, g->tune.chunk, s->qcap, 1, spec.degscale, ru, s->stamp, pull_edges,
                       can_pull);
    HIP_TRY(hipGetLastError());
    if (use_pull) s->stamp_est += 1;
    int rounds = 0;
    int64_t front_prev = 1;
    for (;;) {
        const int K = budget > 0 ? 1 : g->tune.iters_per_graph;
        for (int k = 0; k < K; ++k) {
            if (can_pull)  // the pull form first: it returns at once in a push round
                hipLaunchKernelGGL(ppr_push_k, dim3(pgrid), dim3(256), 0, strm, g->d_adjf, s->m, s->tb, s->q,
                                   s->ctrl, alpha, s->tadjf, s->titems, s->nt, s->ru, s->stamp, 2);
            hipLaunchKernelGGL(ppr_push_k, dim3(grid), dim3(256), 0, strm, g->d_adjf, s->m, s->tb, s->q,
                               s->ctrl, alpha, s->tadjf, s->titems, can_pull ? s->nt : 0, s->ru, s->stamp,
                               can_pull ? 1 : 0);
            hipLaunchKernelGGL(ppr_expand_k, dim3(wgrid), dim3(256), 0, strm, g->d_row, s->p, s->r, s->m,
                               s->tb, s->q, s->ctrl, V, credit, eps, g->tune.chunk, s->qcap, 0, spec.degscale,
                               ru, s->stamp, pull_edges, can_pull);
        }
        HIP_TRY(hipGetLastError());
        rounds += K;
        if (use_pull) s->stamp_est += K;
        HIP_TRY(hipMemcpyAsync(s->h_ctrl, s->ctrl, sizeof(int64_t) * P_N, hipMemcpyDeviceToHost, strm));
        HIP_TRY(hipStreamSynchronize(strm));
        if (s->h_ctrl[P_OVF]) return fail(C2G_EDEVICE, "PPR frontier queue overflow (internal)");
        if (s->h_ctrl[P_CUR] == 0) break;
        if (use_pull && !can_pull && s->h_ctrl[P_WANT]) {
            // a round was dense enough to pull: build the in-edge rows now;
            // the next expands may then switch (their stamps are in place)
            if ((rc = ensure_transpose(g, s))) return rc;
            can_pull = 1;
        }
        if (budget > 0) {
            const int64_t fsize = s->h_ctrl[P_FRONT] - front_prev;  // vertices activated this round
            front_prev = s->h_ctrl[P_FRONT];
            if (rounds >= budget && fsize <= handoff_min) break;
        }
        if (rounds > kPprMaxRounds) return fail(C2G_ETIMEOUT, "PPR did not converge in %d rounds", kPprMaxRounds);
    }
    if (frontier) {
        // Hand over the queued frontier: distinct u with the residual it carries.
        const int64_t n = s->h_ctrl[P_CUR];
        std::vector<int4> items((size_t)n);
        if (n) HIP_TRY(hipMemcpy(items.data(), s->q, sizeof(int4) * n, hipMemcpyDeviceToHost));
        std::vector<uint8_t> seen((size_t)V, 0);
        frontier->clear();
        fr_ru->clear();
        for (const auto& it : items)
            if (!seen[it.x]) {
                seen[it.x] = 1;
                frontier->push_back(it.x);
                float f;
                memcpy(&f, &it.w, sizeof f);
                fr_ru->push_back(f);
            }
    }
    const double t1 = now_ms();
    if (p_out && (rc = s->out.copy_out(p_out, s->p, sizeof(float) * V, strm))) return rc;
    if (r_out && (rc = s->out.copy_out(r_out, s->r, sizeof(float) * V, strm))) return rc;
    if (st) {
        st->lane = C2G_LANE_GPU;
        st->iterations = (int32_t)s->h_ctrl[P_ITER];
        st->frontier_total = s->h_ctrl[P_FRONT];
        st->edges_relaxed = s->h_ctrl[P_EDGES];
        st->items = s->h_ctrl[P_ITEMS];
        st->pull_rounds = (int32_t)s->h_ctrl[P_PULLS];
        st->ms = t1 - t0;
    }
    return C2G_OK;
}

static int ensure_ppr_batch(Graph* g, PprState* s, int groups) {
    HIP_TRY(hipSetDevice(g->device));
    const int64_t V = std::max<int64_t>(g->V, 1);
    const int chunk = std::max(1, g->tune.batch_chunk);
    const int64_t needi = groups * (V + g->E / chunk) + 64;
    if (!s->ctrl) HIP_TRY(hipMalloc(&s->ctrl, sizeof(int64_t) * kCtrlWords));
    if (!s->h_ctrl) HIP_TRY(hipHostMalloc((void**)&s->h_ctrl, sizeof(int64_t) * P_N, hipHostMallocDefault));
    if (groups > s->groups || s->itemcap < needi) {
        const int G = std::max(groups, s->groups);
        for (void** p : {(void**)&s->P, (void**)&s->R, (void**)&s->M, (void**)&s->RU, (void**)&s->mask,
                         (void**)&s->tbits, (void**)&s->items, (void**)&s->imask, (void**)&s->srcs})
            if (*p) { (void)hipFree(*p); *p = nullptr; }
        s->groups = 0;  // empty until every piece is in place: a failure is retried, never used
        s->itemcap = 0;
        const size_t n = (size_t)G * V * 64;
        HIP_TRY(hipMalloc(&s->P, sizeof(float) * n));
        HIP_TRY(hipMalloc(&s->R, sizeof(float) * n));
        HIP_TRY(hipMalloc(&s->M, sizeof(float) * n));
        HIP_TRY(hipMalloc(&s->RU, sizeof(float) * n));
        HIP_TRY(hipMalloc(&s->mask, sizeof(uint64_t) * (size_t)G * V));
        HIP_TRY(hipMemset(s->mask, 0, sizeof(uint64_t) * (size_t)G * V));
        s->nw = (V + 31) / 32 + 1;
        HIP_TRY(hipMalloc(&s->tbits, sizeof(uint32_t) * s->nw * G));
        s->itemcap = G * (V + g->E / chunk) + 64;
        HIP_TRY(hipMalloc(&s->items, sizeof(int4) * s->itemcap));
        HIP_TRY(hipMalloc(&s->imask, sizeof(uint64_t) * s->itemcap));
        HIP_TRY(hipMalloc(&s->srcs, sizeof(int32_t) * 64 * G));
        s->groups = G;
    }
    return C2G_OK;
}

// Runs n queries (64 lanes per group) to convergence on `strm`; P and R stay
// in HBM as [group][V][64] until the next batched call.
static int pprb_run(Graph* g, hipStream_t strm, const int32_t* srcs, int32_t n, float alpha, float eps,
                    c2g_stats* st) {
    int rc = ppr_check(g, alpha, eps);
    if (rc) return rc;
    if (n < 1 || n > 64 * 1024) return fail(C2G_EINVAL, "need 1..65536 sources");
    for (int32_t q = 0; q < n; ++q)
        if (srcs[q] < 0 || srcs[q] >= g->V) return fail(C2G_EINVAL, "source %d out of range [0,%d)", srcs[q], g->V);
    PprState* s = pstate(g);
    const int groups = (n + 63) / 64;
    if ((rc = ensure_ppr_batch(g, s, groups))) return rc;
    const bool pruned = ppr_pruned(g);
    const int64_t V = g->V;
    const int grid = std::max(1, g->num_cus) * g->tune.grid_per_cu;
    const int64_t words = s->nw * groups;
    const int egrid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, (words + 255) / 256));
    const double t0 = now_ms();
    HIP_TRY(hipMemcpyAsync(s->srcs, srcs, sizeof(int32_t) * n, hipMemcpyHostToDevice, strm));
    const int64_t n4 = (int64_t)groups * V * 64 / 4;
    hipLaunchKernelGGL(pprb_fill_k, dim3(grid), dim3(256), 0, strm, s->P, n4, s->tbits, words);
    hipLaunchKernelGGL(pprb_fill_k, dim3(grid), dim3(256), 0, strm, s->R, n4, s->tbits, (int64_t)0);
    hipLaunchKernelGGL(pprb_fill_k, dim3(grid), dim3(256), 0, strm, s->M, n4, s->tbits, (int64_t)0);
    hipLaunchKernelGGL(pprb_seed_k, dim3(n), dim3(256), 0, strm, s->srcs, n, g->d_row, g->d_adjf,
                       pruned ? g->d_hrow : (const int32_t*)nullptr, pruned ? g->d_hadjf : (const int2*)nullptr,
                       s->P, s->M, s->tbits, s->ctrl, V, s->nw, alpha);
    auto threshold = [&](int force) {
        hipLaunchKernelGGL(pprb_activate_k, dim3(grid), dim3(256), 0, strm, g->d_row, s->P, s->R, s->M,
                           s->RU, s->mask, s->tbits, V, s->nw, groups, alpha, eps, force);
        hipLaunchKernelGGL(pprb_emit_k, dim3(egrid), dim3(256), 0, strm, g->d_row, s->mask, s->tbits,
                           s->items, s->imask, s->ctrl, V, s->nw, groups, g->tune.batch_chunk, s->itemcap);
    };
    threshold(1);
    HIP_TRY(hipGetLastError());
    int rounds = 0;
    for (;;) {
        const int K = g->tune.iters_per_graph;
        for (int k = 0; k < K; ++k) {
            hipLaunchKernelGGL(pprb_push_k, dim3(grid), dim3(256), 0, strm, g->d_adjf, s->M, s->RU, s->items,
                               s->imask, s->tbits, s->ctrl, V, s->nw, alpha);
            threshold(0);
        }
        HIP_TRY(hipGetLastError());
        rounds += K;
        HIP_TRY(hipMemcpyAsync(s->h_ctrl, s->ctrl, sizeof(int64_t) * P_N, hipMemcpyDeviceToHost, strm));
        HIP_TRY(hipStreamSynchronize(strm));
        if (s->h_ctrl[P_OVF]) return fail(C2G_EDEVICE, "batched PPR capacity exceeded (internal)");
        if (s->h_ctrl[P_CUR] == 0) break;
        if (rounds > kPprMaxRounds) return fail(C2G_ETIMEOUT, "PPR did not converge in %d rounds", kPprMaxRounds);
    }
    if (st) {
        st->lane = C2G_LANE_GPU;
        st->iterations = (int32_t)s->h_ctrl[P_ITER];
        st->frontier_total = s->h_ctrl[P_FRONT];
        st->edges_relaxed = s->h_ctrl[P_EDGES];
        st->items = s->h_ctrl[P_ITEMS];
        st->ms = now_ms() - t0;
    }
    return C2G_OK;
}

int ppr_batch_to(Graph* g, hipStream_t strm, const int32_t* srcs, int32_t n, float alpha, float eps,
                 float* const* p_outs, float* const* r_outs, c2g_stats* st) {
    int rc = pprb_run(g, strm, srcs, n, alpha, eps, st);
    if (rc) return rc;
    PprState* s = pstate(g);
    if (p_outs && (rc = s->cols.copy_out(g, strm, (const uint32_t*)s->P, n, (void* const*)p_outs))) return rc;
    if (r_outs && (rc = s->cols.copy_out(g, strm, (const uint32_t*)s->R, n, (void* const*)r_outs))) return rc;
    return C2G_OK;
}

}  // namespace c2g

using namespace c2g;

extern "C" {

int c2g_ppr(c2g_graph* h, int32_t src, float alpha, float eps, float* p_out, float* r_out, c2g_stats* st) {
    if (!h || !p_out) return fail(C2G_EINVAL, "null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    return ppr_single(&h->g, h->g.stream, src, alpha, eps, p_out, r_out, st);
}

int c2g_ppr_migrate(c2g_graph* h, int32_t src, float alpha, float eps, int32_t budget, int32_t handoff_min,
                    float* p_out, float* r_out, c2g_stats* st) {
    if (!h || !p_out) return fail(C2G_EINVAL, "null argument");
    if (budget < 1) return fail(C2G_EINVAL, "budget must be >= 1");
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    if (st) memset(st, 0, sizeof *st);
    std::vector<float> r((size_t)g->V);
    std::vector<int32_t> fr;
    std::vector<float> ru;
    c2g_stats gs{};
    int rc = ppr_single_ex(g, g->stream, src, alpha, eps, budget, std::max(0, handoff_min), p_out, r.data(), &fr,
                           &ru, &gs);
    if (rc) return rc;
    for (size_t i = 0; i < fr.size(); ++i) r[fr[i]] = ru[i];
    c2g_stats hs{};
    if (!fr.empty()) {
        rc = c2g_host_forward_push(h, alpha, eps, p_out, r.data(), fr.data(), (int32_t)fr.size(), &hs);
        if (rc) return rc;
    }
    if (r_out) memcpy(r_out, r.data(), sizeof(float) * g->V);
    if (st) {
        *st = gs;
        st->lane = fr.empty() ? C2G_LANE_GPU : C2G_LANE_MIXED;
        st->host_iterations = hs.host_iterations;
        st->edges_relaxed += hs.edges_relaxed;
    }
    return C2G_OK;
}

int c2g_adsorption_push_mode(c2g_graph* h, int32_t src, int32_t mode, float* p_out, float* r_out, c2g_stats* st) {
    if (!h || !p_out) return fail(C2G_EINVAL, "null argument");
    if (mode != C2G_ADS_PUSH_BASE && mode != C2G_ADS_PUSH_PERK) return fail(C2G_EINVAL, "unknown push mode %d", mode);
    C2G_HANDLE_LOCK(h);
    Graph* g = &h->g;
    if (st) memset(st, 0, sizeof *st);
    // alpha 0.2, r[s] = 10, messages (1-alpha) r w (adsorption/w-base/Main.cu:63-64,
    // :149-152, :248).  BASE: p[s] += 0.8 * r[s] and activations credit
    // p += alpha r (:351-355), rmax 1/V; on a purn-ppr graph this is
    // adsorption/w-purn, the same push on the live pruned CSR with no hidden
    // seeding (its seeding is commented out, w-purn/Main.cu:157-169) and rmax
    // 1/(V+1) (:67).  PERK: p[s] += alpha * r[s] and activations credit
    // p += (1-alpha) r (adsorption/w-base-perk/Main.cu:306-309, :555-557),
    // rmax 1/V (Graph.h:117).
    const float alpha = 0.2f;
    PushSpec spec;
    spec.degscale = 0;
    spec.r0 = 10.0f;
    spec.hidden_seed = false;
    spec.credit = mode == C2G_ADS_PUSH_PERK ? 1.0f - alpha : alpha;
    // the source's forced first activation adds credit*r0 to p_src
    spec.p_src = ((mode == C2G_ADS_PUSH_PERK ? alpha : 0.8f) - spec.credit) * spec.r0;
    const float rmax = 1.0f / (float)(std::max(g->V, 1) + (mode == C2G_ADS_PUSH_BASE && ppr_pruned(g) ? 1 : 0));
    return ppr_single_ex(g, g->stream, src, alpha, rmax, 0, 0, p_out, r_out, nullptr, nullptr, st, spec);
}

int c2g_adsorption_push(c2g_graph* h, int32_t src, float* p_out, float* r_out, c2g_stats* st) {
    return c2g_adsorption_push_mode(h, src, C2G_ADS_PUSH_BASE, p_out, r_out, st);
}

int c2g_ppr_batch(c2g_graph* h, const int32_t* srcs, int32_t n, float alpha, float eps, float* p_out,
                  float* r_out, c2g_stats* st) {
    if (!h || !srcs || !p_out) return fail(C2G_EINVAL, "null argument");
    C2G_HANDLE_LOCK(h);
    if (st) memset(st, 0, sizeof *st);
    if (n < 1 || n > 64 * 1024) return fail(C2G_EINVAL, "need 1..65536 sources");
    const int64_t V = h->g.V;
    std::vector<float*> po((size_t)n), ro(r_out ? (size_t)n : 0);
    for (int32_t q = 0; q < n; ++q) {
        po[q] = p_out + (int64_t)q * V;
        if (r_out) ro[q] = r_out + (int64_t)q * V;
    }
    return ppr_batch_to(&h->g, h->g.stream, srcs, n, alpha, eps, po.data(), r_out ? ro.data() : nullptr, st);
}

}  // extern "C"
