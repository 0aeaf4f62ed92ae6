// wave.hpp -- wave64 primitives for CDNA4 (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pk {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// number of set bits of `mask` in lanes below the calling lane
__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniu(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ float unif(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    uint32_t lo = uniu((uint32_t)v), hi = uniu((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float readlanef(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// butterfly sums inside aligned groups of 2^k lanes (every lane gets the sum)
template <int GROUP>
__device__ __forceinline__ int group_sum(int v) {
#pragma unroll
    for (int m = GROUP / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Halving-tree float sum over aligned groups of 16 lanes: lanes j and j+8,
// then j and j+4, j+2, j+1 (the xor partner holds the same value after each
// step, and IEEE addition commutes, so every lane ends with the tree sum).
__device__ __forceinline__ float tree_sum16(float v) {
    v = v + __shfl_xor(v, 8, kWave);
    v = v + __shfl_xor(v, 4, kWave);
    v = v + __shfl_xor(v, 2, kWave);
    v = v + __shfl_xor(v, 1, kWave);
    return v;
}

__device__ __forceinline__ int wave_min_lane(float v, bool valid, float* out_min) {
    float m = valid ? v : __builtin_inff();
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = fminf(m, __shfl_xor(m, s, kWave));
    uint64_t at = ballot(valid && v == m);
    *out_min = unif(m);
    return at ? (int)__builtin_ctzll(at) : -1;
}

}  // namespace pk
