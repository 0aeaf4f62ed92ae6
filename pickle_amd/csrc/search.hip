// search.hip -- batched HNSW search, one wavefront per query.
//
// Reference: HNSWGraph::AnnSearch (include/hnsw.hpp:509-541) with
// SlideLayer (:446-494) for the upper layers, SearchLayerFast (:374-444) for
// layer 0 and SelectSimpleExt / extractMinTopK (:118-125,
// include/queue.hpp:265-286) for the result.
//
// A persistent grid: every wave takes the next query index from an atomic
// counter until the batch is drained (uneven ef / k per query balance
// themselves).  Per hop the wave
//   1. pops the closest candidate (heap in LDS, deep levels in device memory),
//   2. reads the degree and the adjacency row, one id per lane (coalesced
//      4-byte loads, both in one round trip),
//   3. tests and marks the ids in the exact visited set (one atomic per id)
//      while it gathers every listed row (one lane group per row, 16-byte
//      loads) and computes the distances, keeping those of new ids,
//   5. applies the reference's accept rule in adjacency order, updating the
//      result and candidate heaps exactly as the reference's binary heaps
//      (each push and result-heap pop is one wave-parallel step).
// Step 5 is cut down with one test: when the result heap is full at the
// start of a chunk, its top only shrinks, so rows at or beyond it are
// rejected without touching the heaps.
#include "search.hpp"

#include "traverse.hpp"

namespace pk {

template <class DF, bool IS_INT>
__global__ __launch_bounds__(kSearchBlock, 4) void k_search(SearchArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    const int gw = blockIdx.x * (kSearchBlock / kWave) + w;
    const bool leader = lane == 0;
    uint8_t* my = smem + (size_t)w * A.lds_per_wave;
    uint64_t* Wl = reinterpret_cast<uint64_t*>(my);
    uint64_t* Cl = Wl + A.wcap;
    int* cl = reinterpret_cast<int*>(Cl + A.ccap);
    float* cd = reinterpret_cast<float*>(cl + kWave);
    uint8_t* qst = reinterpret_cast<uint8_t*>(cd + kWave);

    WaveVisited vis;
    vis.tab = A.vtab + ((size_t)gw << A.vg_bits);
    vis.log = A.vlog + ((size_t)gw << A.vg_bits);
    vis.bits = A.vg_bits;
    vis.mask = (1u << A.vg_bits) - 1;
    vis.count = 0;
    vis.tag = 0;
    uint64_t* cg = A.cglob + (size_t)gw * (size_t)A.cg_cap;

    const int g = lane % DF::LPC;
    for (;;) {
        int qi = 0;
        if (leader) qi = atomicAdd(A.qcounter, 1);
        const int64_t pos = readlane(qi, 0);
        if (pos >= A.nq) break;
        const int64_t q = A.order ? (int64_t)uni(gbl(A.order)[pos]) : pos;
        const uint64_t t_begin = wall_clock64();

        // stage the query in LDS, zero padded to the row stride
        stage_row(qst, A.queries + (size_t)q * (size_t)A.qbytes, A.qbytes, A.vstride, lane);
        DF df;
        df.load_query(qst, g, IS_INT ? (A.vstride >> 4) : A.dim);

        int k = uni(A.k ? A.k[q] : A.k_all);
        int ef = uni(A.ef ? A.ef[q] : A.ef_all);
        int fail = 0;
        if (k < 1 || k > A.k_stride || ef < 1 || ef + 1 > A.wcap) fail = kFailArgs;

        int cur = 0;  // m_ent_id of the top layer (include/hnsw_layer.hpp:70, 186-188)

        // ---- upper layers: SlideLayer (include/hnsw.hpp:446-494) ----
        WorkCount up, l0;
        for (int level = A.enter_level; level >= 1 && !fail; level--) {
            const LayerDev L = A.layers[level];
            cur = greedy_level(df, L, A.vstride, cur, vis, cl, cd, lane, fail, up);
            if (!fail) cur = gbl(L.down)[cur];
        }

        // ---- layer 0: SearchLayerFast (include/hnsw.hpp:374-444) ----
        WaveHeap<true, false> W{Wl, Cl, A.wcap, 0, 0};
        WaveHeap<false, true> C{Cl, cg, A.ccap, A.cg_cap, 0};
        if (!fail) beam_level(df, A.layers[0], A.vstride, cur, ef, W, C, vis, cl, cd, lane, fail, l0);

        // ---- extractMinTopK + external ids (queue.hpp:265-286, hnsw.hpp:118-125) ----
        int32_t* oid = A.out_ids + (size_t)q * A.k_stride;
        float* odist = A.out_dists + (size_t)q * A.k_stride;
        int kk = 0;
        if (!fail) {
            kk = k < W.len ? k : W.len;
            while (W.len > kk) W.pop_par(lane, leader);
            for (int i = 0; i < kk; i++) {
                const uint64_t e = W.top();
                if (leader) {
                    oid[kk - 1 - i] = A.ext0[entry_id(e)];
                    odist[kk - 1 - i] = entry_dist(e);
                }
                W.pop_par(lane, leader);
            }
        }
        for (int i = kk + lane; i < A.k_stride; i += kWave) {
            oid[i] = -1;
            odist[i] = __builtin_inff();
        }
        if (leader) {
            if (A.out_count) A.out_count[q] = fail ? -1 : kk;
            if (fail) atomicOr(A.err, fail);
            if (A.ctr) {
                pk_query_counters c;
                c.hops[0] = l0.hops;
                c.hops[1] = up.hops;
                c.neighbors[0] = l0.neighbors;
                c.neighbors[1] = up.neighbors;
                c.dists[0] = l0.dists;
                c.dists[1] = up.dists;
                c.latency_ticks = (int32_t)(wall_clock64() - t_begin);
                A.ctr[q] = c;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_fill_i32(int* p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class DF, bool INT>
static hipError_t launch_rs(const SearchArgs& a, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((k_search<DF, INT>), dim3(grid), dim3(kSearchBlock), lds, s, a);
    return hipGetLastError();
}

template <int DT, int MET>
static hipError_t launch_dt(const SearchArgs& a, int size_class, int grid, size_t lds, hipStream_t s) {
    constexpr bool INT = DT == DT_U8 || DT == DT_I8;
    if constexpr (INT) {
        if (size_class == 0) return launch_rs<IntDist<DT, MET, 1>, true>(a, grid, lds, s);
        if (size_class == 1) return launch_rs<IntDist<DT, MET, 2>, true>(a, grid, lds, s);
        return launch_rs<IntDist<DT, MET, 4>, true>(a, grid, lds, s);
    } else {
        // classes 8 and 9: the order-free FP16 traversal distance
        if constexpr (DT == DT_F16) {
            if (size_class == kHalfFastClass) return launch_rs<HalfDist<MET, 1>, true>(a, grid, lds, s);
            if (size_class == kHalfFastClass + 1) return launch_rs<HalfDist<MET, 2>, true>(a, grid, lds, s);
        }
        if (size_class == 0) return launch_rs<FloatDist<DT, MET, 8>, false>(a, grid, lds, s);
        return launch_rs<FloatDist<DT, MET, 16>, false>(a, grid, lds, s);
    }
}

int size_class_for(int dtype, int dim, int vstride) {
    if (dtype == DT_U8 || dtype == DT_I8) {
        int chunks = vstride / 16;
        if (chunks <= 8) return 0;
        if (chunks <= 16) return 1;
        if (chunks <= 32) return 2;
        return -1;
    }
    int nmain = dim / 16;
    if (nmain <= 8) return 0;
    if (nmain <= 16) return 1;
    return -1;
}

int result_capacity_for(int ef_max) {
    int cap = (ef_max + 2 + 7) / 8 * 8;
    return cap <= 2 * kWave * kPopSlots ? cap : -1;
}

hipError_t launch_search(const SearchArgs& a, int dtype, int metric, int size_class, int grid, size_t lds,
                         hipStream_t s) {
    hipError_t e = hipSuccess;
    // counters are cleared by kernels, not memset nodes, so the sequence
    // can be captured into a HIP graph and replayed
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, a.qcounter, 0, 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, a.err, 0, 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    switch (dtype * 4 + metric) {
        case DT_U8 * 4 + MET_L2: return launch_dt<DT_U8, MET_L2>(a, size_class, grid, lds, s);
        case DT_U8 * 4 + MET_IP: return launch_dt<DT_U8, MET_IP>(a, size_class, grid, lds, s);
        case DT_I8 * 4 + MET_L2: return launch_dt<DT_I8, MET_L2>(a, size_class, grid, lds, s);
        case DT_I8 * 4 + MET_IP: return launch_dt<DT_I8, MET_IP>(a, size_class, grid, lds, s);
        case DT_F32 * 4 + MET_L2: return launch_dt<DT_F32, MET_L2>(a, size_class, grid, lds, s);
        case DT_F32 * 4 + MET_IP: return launch_dt<DT_F32, MET_IP>(a, size_class, grid, lds, s);
        case DT_F16 * 4 + MET_L2: return launch_dt<DT_F16, MET_L2>(a, size_class, grid, lds, s);
        case DT_F16 * 4 + MET_IP: return launch_dt<DT_F16, MET_IP>(a, size_class, grid, lds, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t search_occupancy(int dtype, int metric, int size_class, size_t lds, int* blocks_per_cu) {
    // every instantiation has the same block size; LDS is the binding limit,
    // the register-limited bound comes from one representative kernel
    const void* fn = nullptr;
    if (dtype == DT_U8 || dtype == DT_I8)
        fn = reinterpret_cast<const void*>(&k_search<IntDist<DT_U8, MET_L2, 1>, true>);
    else
        fn = reinterpret_cast<const void*>(&k_search<FloatDist<DT_F32, MET_L2, 8>, false>);
    (void)metric;
    (void)size_class;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, kSearchBlock, lds);
}

}  // namespace pk
