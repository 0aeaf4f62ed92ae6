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

#include "distance.hpp"
#include "heap.hpp"
#include "visited.hpp"
#include "wave.hpp"

namespace pk {

template <class DF>
struct PassesOf {
    static constexpr int value = 4;
};
template <int DT, int MET>
struct PassesOf<IntDist<DT, MET, 1>> {
    static constexpr int value = 8;
};
template <int DT, int MET>
struct PassesOf<IntDist<DT, MET, 4>> {
    static constexpr int value = 2;
};
template <int DT, int MET>
struct PassesOf<FloatDist<DT, MET, 16>> {
    static constexpr int value = 2;
};

// distances of the `ncand` rows whose ids are in cl[]; results to cd[]
template <class DF>
__device__ void batch_distances(const DF& df, const uint8_t* vec, int vstride, const int* cl, float* cd,
                                int ncand, int lane) {
    constexpr int LPC = DF::LPC;
    constexpr int G = kWave / LPC;
    constexpr int PB = PassesOf<DF>::value;
    const int grp = lane / LPC, g = lane % LPC;
    for (int base = 0; base < ncand; base += G * PB) {
        typename DF::Row rows[PB];
        const uint8_t* rp[PB];
        bool act[PB];
#pragma unroll
        for (int p = 0; p < PB; p++) {
            int ci = base + p * G + grp;
            act[p] = ci < ncand;
            int vid = act[p] ? cl[ci] : 0;
            rp[p] = vec + (size_t)vid * (size_t)vstride;
            df.load_row(rows[p], rp[p], g, act[p]);
        }
#pragma unroll
        for (int p = 0; p < PB; p++) {
            float d = df.finish(rows[p], g, rp[p], act[p]);
            int ci = base + p * G + grp;
            if (act[p] && g == 0) cd[ci] = d;
        }
    }
}

__device__ __forceinline__ int clamp_degree(const LayerDev& L, int v, int deg) {
    int64_t room = L.adj_len - (int64_t)v * L.D;
    if (deg < 0) deg = 0;
    if ((int64_t)deg > room) deg = (int)(room < 0 ? 0 : room);
    return deg;
}

// One hop's row: the degree and the first 64 adjacency slots are loaded
// together (the adjacency allocation has 64 slots of padding past adj_len,
// so the speculative read stays inside it).
struct RowHead {
    int64_t base;
    int deg;
    int first;  // this lane's id among the first 64 slots
};

__device__ __forceinline__ RowHead load_row_head(const LayerDev& L, int v, int lane) {
    RowHead h;
    h.base = (int64_t)v * L.D;
    int d = L.deg[v];
    h.first = L.adj[h.base + lane];
    h.deg = uni(clamp_degree(L, v, d));
    return h;
}

// Expands one chunk of 64 adjacency slots: the visited-set compare-and-swaps
// and the gathers of every valid row go out in the same round trip (rows of
// already-visited ids are fetched and their distances dropped).  Returns the
// new lanes; this lane's id and (if new) distance come back through id / d.
template <class DF>
__device__ uint64_t expand_chunk(const DF& df, const LayerDev& L, int vstride, const RowHead& rh, int b0,
                                 WaveVisited& vis, int* cl, float* cd, int lane, int& fail, int& id, float& d) {
    const int j = b0 + lane;
    const bool inrow = j < rh.deg;
    id = b0 == 0 ? rh.first : (inrow ? L.adj[rh.base + j] : -1);
    const bool valid = inrow && id >= 0 && (int64_t)id < L.n;
    WaveVisited::Pending pend = vis.begin(id, valid);
    const uint64_t vm = ballot(valid);
    const int pos = lanes_below(vm);
    if (valid) cl[pos] = id;
    __builtin_amdgcn_wave_barrier();
    batch_distances(df, L.vec, vstride, cl, cd, __builtin_popcountll(vm), lane);
    __builtin_amdgcn_wave_barrier();
    const uint64_t newm = vis.finish(pend, id, valid, fail);
    d = ((newm >> lane) & 1) ? cd[pos] : __builtin_inff();
    return newm;
}

template <class DF>
__device__ float one_distance(const DF& df, const LayerDev& L, int vstride, int id, int* cl, float* cd,
                              int lane) {
    if (lane == 0) cl[0] = id;
    __builtin_amdgcn_wave_barrier();
    batch_distances(df, L.vec, vstride, cl, cd, 1, lane);
    __builtin_amdgcn_wave_barrier();
    return unif(cd[0]);
}

template <class DF, bool IS_INT>
__global__ __launch_bounds__(kSearchBlock) void k_search(SearchArgs A) {
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
        int64_t q = readlane(qi, 0);
        if (q >= A.nq) break;
        const uint64_t t_begin = wall_clock64();

        // stage the query in LDS, zero padded to the row stride
        const uint8_t* qsrc = A.queries + (size_t)q * (size_t)A.qbytes;
        for (int i = lane; i < A.vstride; i += kWave) qst[i] = i < A.qbytes ? qsrc[i] : 0;
        __builtin_amdgcn_wave_barrier();
        DF df;
        df.load_query(qst, g, IS_INT ? (A.vstride >> 4) : A.dim);

        int k = uni(A.k ? A.k[q] : A.k_all);
        int ef = uni(A.ef ? A.ef[q] : A.ef_all);
        int fail = 0;
        if (k < 1 || k > A.k_stride || ef < 1 || ef + 1 > A.wcap) fail = 2;

        int hop0 = 0, nb0 = 0, ds0 = 0, hopu = 0, nbu = 0, dsu = 0;
        int cur = 0;  // m_ent_id of the top layer (include/hnsw_layer.hpp:70, 186-188)

        // ---- upper layers: SlideLayer (include/hnsw.hpp:446-494) ----
        for (int level = A.enter_level; level >= 1 && !fail; level--) {
            const LayerDev L = A.layers[level];
            float cdist = one_distance(df, L, A.vstride, cur, cl, cd, lane);
            vis.mark(cur, leader, fail);
            bool updated;
            do {
                updated = false;
                const RowHead rh = load_row_head(L, cur, lane);
                hopu++;
                nbu += rh.deg;
                for (int b0 = 0; b0 < rh.deg; b0 += kWave) {
                    int id;
                    float d;
                    uint64_t newm = expand_chunk(df, L, A.vstride, rh, b0, vis, cl, cd, lane, fail, id, d);
                    if (newm == 0) continue;
                    dsu += __builtin_popcountll(newm);
                    float m;
                    int at = wave_min_lane(d, (newm >> lane) & 1, &m);
                    if (at >= 0 && m < cdist) {  // strictly closer, first in row order
                        cur = readlane(id, at);
                        cdist = m;
                        updated = true;
                    }
                }
            } while (updated && !fail);
            vis.clear(lane);
            if (!fail) cur = L.down[cur];
        }

        // ---- layer 0: SearchLayerFast (include/hnsw.hpp:374-444) ----
        WaveHeap<true, false> W{Wl, Cl, A.wcap, 0, 0};
        WaveHeap<false, true> C{Cl, cg, A.ccap, A.cg_cap, 0};
        if (!fail) {
            const LayerDev L = A.layers[0];
            float d0 = one_distance(df, L, A.vstride, cur, cl, cd, lane);
            W.push_par(pack_entry(d0, cur), lane, leader);
            C.push_par(pack_entry(d0, cur), lane, leader);
            vis.mark(cur, leader, fail);
            uint32_t wtop_bits = __float_as_uint(d0);  // bits of W.top()
            while (C.len > 0 && !fail) {
                const uint64_t te = C.top();
                C.pop_deep(lane, leader);
                if (entry_dist(te) > __uint_as_float(wtop_bits)) break;  // hnsw.hpp:403
                const RowHead rh = load_row_head(L, entry_id(te), lane);
                hop0++;
                nb0 += rh.deg;
                for (int b0 = 0; b0 < rh.deg && !fail; b0 += kWave) {
                    int id;
                    float d;
                    uint64_t newm = expand_chunk(df, L, A.vstride, rh, b0, vis, cl, cd, lane, fail, id, d);
                    if (newm == 0) continue;
                    ds0 += __builtin_popcountll(newm);
                    bool isnew = (newm >> lane) & 1;
                    const float wtop = __uint_as_float(wtop_bits);
                    uint64_t cand = W.len >= ef ? ballot(isnew && d < wtop) : newm;
                    while (cand) {
                        int ln = __builtin_ctzll(cand);
                        cand &= cand - 1;
                        const float dv = readlanef(d, ln);
                        if (W.len < ef || __uint_as_float(wtop_bits) > dv) {  // hnsw.hpp:425-431
                            const uint64_t e = pack_entry(dv, readlane(id, ln));
                            W.push_par(e, lane, leader);
                            if (C.full()) {
                                fail = 1;
                                break;
                            }
                            C.push_par(e, lane, leader);
                            if (W.len > ef) W.pop_par(lane, leader);
                            wtop_bits = (uint32_t)W.top();
                        }
                    }
                }
            }
            vis.clear(lane);
        }

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
                c.hops[0] = hop0;
                c.hops[1] = hopu;
                c.neighbors[0] = nb0;
                c.neighbors[1] = nbu;
                c.dists[0] = ds0;
                c.dists[1] = dsu;
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
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, a.qcounter, 0, 1);
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
