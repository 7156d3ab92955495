// Wave64 / LDS primitives for gfx950 (CDNA4).  All cross-lane traffic goes
// through ds_bpermute; a wavefront is 64 lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace glp {
namespace dev {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t shfl_u32(uint32_t x, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)x);
}
__device__ __forceinline__ int32_t shfl_i32(int32_t x, int src) {
    return __builtin_amdgcn_ds_bpermute(src << 2, x);
}
__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src) {
    uint32_t lo = shfl_u32((uint32_t)x, src);
    uint32_t hi = shfl_u32((uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// Load of data read once per step (CSR neighbours, weights, gathered
// labels): marked nontemporal so it does not evict the label lines that the
// random gathers keep reusing from L2.
template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// Max over the 64 lanes; every lane receives the result.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = max_u64(x, shfl_u64(x, lane ^ d));
    return x;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = min(x, shfl_u32(x, lane ^ d));
    return x;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += shfl_u32(x, lane ^ d);
    return x;
}

// Orders LDS traffic of one wavefront: LDS serves one wave's requests in
// issue order, so a compiler-level barrier is all that is needed between a
// lane's store and another lane's load in the same wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// murmur3 fmix32 finaliser; the slot hash of the LDS and HBM tables (the
// reference's GlobalHT uses the same finaliser, method/hashtable.cuh:19-26).
__device__ __forceinline__ uint32_t fmix32(uint32_t k) {
    k ^= k >> 16;
    k *= 0x85ebca6bu;
    k ^= k >> 13;
    k *= 0xc2b2ae35u;
    k ^= k >> 16;
    return k;
}

__host__ __device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
    if (x <= 1) return 1;
    return 1u << (32 - __builtin_clz(x - 1));
}

}  // namespace dev
}  // namespace glp
