// brute.hpp -- shared parts of the exhaustive k-nearest-neighbour kernels
// (brute.hip: integer tiles, merge and the entry point; brute_fp64.hip:
// FP64 tiles for f16 / f32 rows).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "distance.hpp"

namespace pk {

constexpr int kBruteK = 32;    // candidates returned per query
constexpr int kIntQ = 128;     // queries per workgroup, integer tiles (4 waves x 32)
constexpr int kFltQ = 64;      // queries per workgroup, FP64 tiles (4 waves x 16)
constexpr int kMaxLists = 64;  // lists per query that one merge wave takes
// workgroups the grids aim for: FP64 tiles, integer tiles (at their
// register counts, about two and four resident per CU)
#ifndef PK_BRUTE_WGS
#define PK_BRUTE_WGS 512
#endif
#ifndef PK_BRUTE_INT_WGS
#define PK_BRUTE_INT_WGS 1024
#endif
constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));

// Sorted ascending insert of (d, id), dropping the last entry.  The caller
// only inserts ids larger than every id in the list, so the new entry goes
// after any equal keys and the list stays sorted by (key, id).  Keys shift
// by min / max (entry j becomes max(kd[j-1], min(kd[j], d))); ids follow
// with one select each, from the same comparisons, no position depending
// on another's result.
template <typename K, int KB>
__device__ __forceinline__ void sorted_insert(K (&kd)[KB], int (&ki)[KB], K d, int id) {
    bool le[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) le[j] = kd[j] <= d;
#pragma unroll
    for (int j = KB - 1; j > 0; j--) {
        const int from = le[j - 1] ? id : ki[j - 1];
        ki[j] = le[j] ? ki[j] : from;
        kd[j] = max(kd[j - 1], min(kd[j], d));
    }
    ki[0] = le[0] ? ki[0] : id;
    kd[0] = min(kd[0], d);
}

__device__ __forceinline__ v4i as_v4i(uint4 u, uint32_t flip) {
    v4i v;
    v.x = (int)(u.x ^ flip);
    v.y = (int)(u.y ^ flip);
    v.z = (int)(u.z ^ flip);
    v.w = (int)(u.w ^ flip);
    return v;
}

// element j of a 16-byte chunk of f32 (4) or f16 (8) values, widened to FP64
template <int DT>
__device__ __forceinline__ double chunk_elem(const uint4& u, int j) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
    if (DT == DT_F32) return (double)__uint_as_float(w[j]);
    const uint16_t bits = (uint16_t)(w[j >> 1] >> (16 * (j & 1)));
    _Float16 x;
    __builtin_memcpy(&x, &bits, 2);
    return (double)(float)x;
}

struct BrutePlan {
    int64_t qblocks, splits, tiles_per_split;
    int nlists;
};

// Splits for the XCD-aware grid: enough for target_wgs workgroups, at
// most kMaxLists / lists_per_split.  From kXcds / 2 splits up, the count is
// rounded up to a multiple of kXcds when the rows allow it, so every XCD
// takes whole splits (brute_block_coords); below that, rounding would more
// than double the lists each query builds.
inline BrutePlan plan_brute_xcd(int64_t nq, int64_t nbase, int q_per_block, int rows_per_tile, int lists_per_split,
                                int64_t target_wgs) {
    BrutePlan p;
    p.qblocks = (nq + q_per_block - 1) / q_per_block;
    const int64_t ntiles = (nbase + rows_per_tile - 1) / rows_per_tile;
    int64_t s = (target_wgs + p.qblocks - 1) / p.qblocks;
    const int64_t smax = kMaxLists / lists_per_split;
    s = s < smax ? s : smax;
    const int64_t up = (s + kXcds - 1) / kXcds * kXcds;
    if (s >= kXcds / 2 && up <= smax && up <= ntiles) s = up;
    s = s < ntiles ? s : ntiles;
    s = s > 1 ? s : 1;
    p.tiles_per_split = (ntiles + s - 1) / s;
    p.splits = (ntiles + p.tiles_per_split - 1) / p.tiles_per_split;
    p.nlists = (int)(p.splits * lists_per_split);
    return p;
}

// Workgroups reach the XCDs in turn (linear id mod kXcds), and each XCD has
// its own L2.  On a one-dimensional grid of qblocks x splits workgroups
// with a split count that is a multiple of kXcds (xcd_map), workgroup L
// takes split (L mod kXcds) + kXcds * (L / kXcds / qblocks): every XCD
// works through whole splits, so all its resident workgroups stream the
// same rows and share them in its L2.  Otherwise the split is L / qblocks.
__device__ __forceinline__ void brute_block_coords(int64_t qblocks, int xcd_map, int64_t& qb, int64_t& split) {
    const int64_t L = blockIdx.x;
    if (xcd_map) {
        const int64_t slot = L / kXcds;
        qb = slot % qblocks;
        split = L % kXcds + kXcds * (slot / qblocks);
    } else {
        qb = L % qblocks;
        split = L / qblocks;
    }
}

// FP64 tiles (brute_fp64.hip): row norms when L2, then the tile kernel on
// the XCD-aware grid of plan p
hipError_t launch_brute_fp64(int dtype, int metric, int nsc, const BrutePlan& p, hipStream_t s,
                             const uint8_t* base, int64_t nbase, int vstride, int dim, int64_t npad,
                             double* nrm, const uint8_t* dq, int64_t nq, double* pd, int32_t* pi);

}  // namespace pk
