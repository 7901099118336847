// Wave64 primitives for gfx950: replace cub::BlockScan<int,512>::ExclusiveSum
// and __any_sync / __syncthreads_or of the reference kernels
// (src/sssp/base/Main.cu:197, 232, 265, 329).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace c2g {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// Global index of this wave, visibly uniform to the compiler (readfirstlane),
// so loads indexed by it alone (item headers, bitmap words, row offsets)
// become scalar loads into SGPRs.  Assumes blockDim.x is a multiple of 64.
__device__ __forceinline__ int64_t wave_index() {
    return (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Exclusive prefix sum over the 64 lanes of a wave; *total gets the wave sum.
// Every lane of the wave must call it (uniform control flow).
__device__ __forceinline__ int wave_excl_scan(int x, int* total) {
    const int lane = lane_id();
    int v = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    *total = __shfl(v, 63, 64);
    return v - x;
}

// s_waitcnt vmcnt(0): waits for every vector-memory operation this wave has
// in flight.  On gfx950 stores and atomics without return count in vmcnt too,
// and with both kinds pending the compiler can only wait for zero.  A loop
// whose first use of a loaded register sits at its head therefore waits for
// the previous trip's atomics on every trip; waiting once before the loop
// clears that register and leaves the loop head without a wait.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

typedef __attribute__((address_space(1))) char gchar;

// `p` as an SGPR pair in the global address space: a value the compiler can
// no longer fold into per-lane address arithmetic, so a row at `p` plus a
// lane's 32-bit offset costs no 64-bit vector add per row.  `p` must be
// wave-uniform.
__device__ __forceinline__ gchar* sgpr_ptr(const void* p) {
    const uint64_t a = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
    return (gchar*)(((uint64_t)hi << 32) | lo);
}

// This lane's element of the row at `row` (`loff` = its byte offset).
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* lane_slot(gchar* row, uint32_t loff) {
    return (__attribute__((address_space(1))) T*)(row + loff);
}

__device__ __forceinline__ int64_t wave_sum64(int64_t x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        int lo = __shfl_xor((int)(uint32_t)x, d, 64);
        int hi = __shfl_xor((int)(uint32_t)((uint64_t)x >> 32), d, 64);
        x += (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    return x;
}

__device__ __forceinline__ int64_t shfl64(int64_t x, int src) {
    int lo = __shfl((int)(uint32_t)x, src, 64);
    int hi = __shfl((int)(uint32_t)((uint64_t)x >> 32), src, 64);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Largest j in [0,64) with pre_j <= k, where pre is this wave's non-decreasing
// exclusive prefix held one value per lane (six shuffles).
__device__ __forceinline__ int wave_owner(int pre, int k) {
    int o = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) {
        int pj = __shfl(pre, o + st, 64);
        if (pj <= k) o += st;
    }
    return o;
}

// Position of the r-th set bit (r = 0 for the lowest) of w; r < popcount(w).
__device__ __forceinline__ int nth_set_bit(uint32_t w, int r) {
    int pos = 0;
#pragma unroll
    for (int half = 16; half > 0; half >>= 1) {
        const int c = __popc(w & ((1u << half) - 1u));
        if (r >= c) {
            r -= c;
            w >>= half;
            pos += half;
        }
    }
    return pos;
}

// Queue reservation for a whole workgroup with one atomic: each lane asks
// for `cnt` entries; returns the lane's first position, and *wave_end the end
// of its wave's range (for a capacity check).  Wave totals meet in s_tot
// (one entry per wave, at least blockDim.x/64) and the base in *s_base.
// Every thread of the workgroup must call it (it holds barriers).
__device__ __forceinline__ int64_t wg_reserve(int cnt, unsigned long long* ctr, int64_t* s_tot, int64_t* s_base,
                                              int64_t* wave_end) {
    const int lane = lane_id();
    const int wid = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    int tot;
    const int off = wave_excl_scan(cnt, &tot);
    if (lane == 0) s_tot[wid] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t sum = 0;
        for (int w = 0; w < nwv; ++w) sum += s_tot[w];
        *s_base = sum > 0 ? (int64_t)atomicAdd(ctr, (unsigned long long)sum) : 0;
    }
    __syncthreads();
    int64_t base = *s_base;
    for (int w = 0; w < wid; ++w) base += s_tot[w];
    __syncthreads();  // s_tot and s_base may be rewritten by the next call
    *wave_end = base + tot;
    return base + off;
}

// SGPR budget for eight waves per SIMD.  The SIMD's 800 SGPRs are handed out
// in blocks of 16 and the trap handler takes 16 more per wave, so a kernel
// allocated more than 80 runs at most 7 waves per SIMD, although the
// compiler's occupancy estimate (which leaves the trap handler out) still
// says 8.  A latency-bound grid-stride kernel sized for 8 waves per SIMD then
// runs its last workgroups after the first ones finish: one SGPR over the
// limit cost the batched relax 40 % of its time.  Kernels that must hold 8
// waves cap their SGPRs here (the few values over the cap spill to VGPR
// lanes in the prologue).
#define C2G_OCC8_SGPR __attribute__((amdgpu_num_sgpr(80)))

// Last-workgroup detection.  Atomics on one address serialise in the memory
// system, about 11 ns apiece on gfx950: a grid of 2048 workgroups counting
// on a single counter spends about 28 us at the end of every launch, and a
// __threadfence (L2 write-back and invalidate) before each count about 14 us
// more.  Here workgroup b counts on counter b % kDoneSlots (counters 128
// bytes apart, counted in parallel) and the last workgroup of each counter
// on a top counter: about 1 us for 2048 workgroups.  The workgroup that
// completes a counter resets it, so the tree is zero again when the launch
// ends.  No fence: what other workgroups produced reaches the last one only
// through atomics, which it reads back with atomics, and each caller makes
// every wave wait for its own memory operations (wait_vmem) before the
// barrier ahead of the call, so a workgroup's atomics are performed before
// it counts.  Kernel boundaries order everything else.
constexpr int kDoneSlots = 64;
constexpr int kDoneStride = 16;  // int64 words per 128-byte line
// A control block: the kernel's fields in words [0, kCtrlTree), then the tree.
constexpr int kCtrlTree = 32;
constexpr int kCtrlWords = kCtrlTree + (kDoneSlots + 1) * kDoneStride;

// Called by one thread of each workgroup; true in the last workgroup.
__device__ __forceinline__ bool wg_last(int64_t* tree) {
    unsigned long long* c = reinterpret_cast<unsigned long long*>(tree);
    const unsigned nb = gridDim.x, s = blockIdx.x % kDoneSlots;
    const unsigned per = (nb - s + kDoneSlots - 1) / kDoneSlots;  // workgroups counting on s
    const unsigned nslot = nb < (unsigned)kDoneSlots ? nb : (unsigned)kDoneSlots;
    if (atomicAdd(&c[s * kDoneStride], 1ull) != per - 1) return false;
    atomicExch(&c[s * kDoneStride], 0ull);
    if (atomicAdd(&c[kDoneSlots * kDoneStride], 1ull) != nslot - 1) return false;
    atomicExch(&c[kDoneSlots * kDoneStride], 0ull);
    return true;
}

// Zeroes a control block's tree (before its first use).
__device__ __forceinline__ void done_tree_clear(int64_t* ctrl) {
    for (int i = 0; i <= kDoneSlots; ++i) ctrl[kCtrlTree + i * kDoneStride] = 0;
}

}  // namespace c2g
