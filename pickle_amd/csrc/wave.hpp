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

// Device-memory views.  Pointers read from memory (the layer table) are
// generic; loads through them compile to flat instructions, which count
// against the LDS wait counter as well, so every LDS wait would also wait
// for the wave's outstanding HBM gathers.  Casting to the global address
// space gives global_load instructions that only the vector-memory counter
// tracks.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ const gmem<T>* gbl(const T* p) {
    return (const gmem<T>*)p;
}
template <class T>
__device__ __forceinline__ gmem<T>* gbl_rw(T* p) {
    return (gmem<T>*)p;
}
template <class T>
using lmem = __attribute__((address_space(3))) T;
template <class T>
__device__ __forceinline__ const lmem<T>* lds_ptr(const T* p) {
    return (const lmem<T>*)p;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// 16-byte loads from device memory / LDS
__device__ __forceinline__ uint4 ld16_global(const void* p) {
    u32x4 v = *(const gmem<u32x4>*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ld16_lds(const void* p) {
    u32x4 v = *(const lmem<u32x4>*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Cross-lane moves inside rows of 16 lanes by DPP (VALU only, no LDS
// crossbar).  Controls: quad_perm 0x00-0xff, row_ror:n 0x120+n,
// row_half_mirror 0x141.
template <int CTRL>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(dpp<CTRL>(__float_as_int(v)));
}
constexpr int kDppXor1 = 0xb1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4e;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <- lane 7-i within each 8
template <int N>
constexpr int kDppRor = 0x120 + N;    // row_ror:N

// Sums inside aligned groups of 8 lanes (every lane gets the sum): pairs,
// quads, then the two quads of the group meet through the half-row mirror.
__device__ __forceinline__ int sum8(int v) {
    v += dpp<kDppXor1>(v);
    v += dpp<kDppXor2>(v);
    v += dpp<kDppHalfMirror>(v);
    return v;
}

// Halving-tree float sum over aligned groups of 16 lanes: lanes j and j+8,
// then j and j+4, j+2, j+1.  Rotating the row by 8, 4, 2, 1 pairs every
// lane with the same partner set; each partial sum is an addition of two
// values that are bit-identical to the tree's (IEEE addition commutes), so
// every lane ends with the tree sum.
__device__ __forceinline__ float tree_sum16(float v) {
    v = v + dppf<kDppRor<8>>(v);
    v = v + dppf<kDppRor<4>>(v);
    v = v + dppf<kDppRor<2>>(v);
    v = v + dppf<kDppRor<1>>(v);
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
