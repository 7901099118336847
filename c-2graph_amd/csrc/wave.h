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

}  // namespace c2g
