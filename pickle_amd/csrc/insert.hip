// insert.hip -- batched HNSW insertion on the GPU.
//
// Reference: HNSWGraph::Insert (include/hnsw.hpp:247-302), SelectHeuristic
// (:70-116), HNSWLayer::NewInID / AddNode / AddEdge (include/hnsw_layer.hpp:
// 168-177, 297-332, 360-465), InitBuffer / Build (include/hnsw.hpp:222-245,
// 497-507) and the level sampler (include/util.hpp:13-64).
//
// A batch of new nodes is inserted in three launches, one wave per node:
//   k_insert_search  descent through the layers above the node's level, then
//                    per level the efC-bounded construction search and the
//                    neighbour heuristic; R and Pruned go to scratch
//   k_insert_rows    AddNode: the node's rows, its vector copy per layer and
//                    the id maps (NewInID's ext <-> internal bookkeeping)
//   k_insert_edges   AddEdge(e, node) for every e in R then Pruned, each under
//                    e's lock (a wave holds one lock at a time)
// Internal ids are assigned by the host in batch order, as NewInID hands
// them out.  The searches of a batch see the graph as it was before the
// batch, so batch sizes grow with the graph (singletons while it is small)
// and a node that opens a new top level is inserted alone.
//
// Rows: by default a row never grows past D entries.  With PK_BUILD_EXACT_ROWS
// the reference's AddEdge is reproduced including its write one past a full
// row (hnsw_layer.hpp:441-447), which with batches of one node makes the
// build identical to the reference's single-threaded Build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../include/pickle_gpu.h"
#include "index_impl.hpp"
#include "traverse.hpp"

namespace pk {

constexpr int kInsBlock = 64;     // one wave per workgroup
constexpr int kRowMax = 512;      // longest row AddEdge handles (exact-rows mode may exceed D)
constexpr int kTopKMax = 128;     // max(2M) supported

struct LayerBuild {
    int32_t* adj;
    int32_t* deg;
    int16_t* heur;
    float* dist;
    int32_t* ext;
    uint8_t* vec;
    uint8_t* has;
    int32_t* down;
    int32_t* lock;
    int64_t adj_len;
    int32_t D;
    int32_t pad_;
};

struct InsertArgs {
    const LayerDev* layers;  // read view (sizes before the batch)
    const LayerBuild* build;
    const uint8_t* base;
    int32_t vstride, qbytes, dim;
    int32_t enter_level;
    int32_t nonempty;  // bit l: layer l had nodes before the batch
    int32_t efc, M;
    int64_t nb;               // nodes in the batch
    const int32_t* node_ext;  // [nb]
    const int32_t* node_lvl;  // [nb]
    const int32_t* node_in;   // [nb][maxl] internal id per level
    int32_t maxl;
    int32_t topk;  // row slots per list in scratch (>= 2M)
    int32_t* r_ids;  // [nb][maxl][2][topk] R then Pruned
    float* r_d;
    int32_t* r_cnt;  // [nb][maxl][2]
    uint64_t* cglob;
    uint32_t* vtab;
    int32_t* vlog;
    int32_t cg_cap, vg_bits;
    int32_t wcap, ccap, lds_per_wave;
    int32_t exact_rows;
    int32_t row_limit;       // longest row AddEdge writes: D (capped rows) up to kRowMax - 1
    int32_t row_cache;       // AddEdge: rows short enough to stage whole in LDS (< row_cache ids)
    int32_t row_cache_long;  // AddEdge on longer rows: kept-set vectors held in LDS
    int64_t rows_after[PK_MAX_LEVELS];  // layer sizes once the batch is in
    int32_t* qcounter;
    int32_t* err;
};

__device__ __forceinline__ int top_inserted_level(const InsertArgs& A, int l) {
    // highest non-empty level the node searches, 0 if none (hnsw.hpp:262-280)
    int top = l < A.enter_level ? l : A.enter_level;
    for (int lev = top; lev >= 1; lev--)
        if ((A.nonempty >> lev) & 1) return lev;
    return 0;
}

__device__ __forceinline__ int next_node(int32_t* counter, int lane) {
    int v = 0;
    if (lane == 0) v = atomicAdd(counter, 1);
    return readlane(v, 0);
}

// ---------------------------------------------------------------------------
// k_insert_search
// ---------------------------------------------------------------------------

template <class DF, bool IS_INT>
__global__ __launch_bounds__(kInsBlock) void k_insert_search(InsertArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const bool leader = lane == 0;
    const int gw = blockIdx.x;
    uint64_t* Wl = reinterpret_cast<uint64_t*>(smem);
    uint64_t* Cl = Wl + A.wcap;
    uint64_t* Ml = Cl + A.ccap;      // heuristic work heap
    uint64_t* Dl = Ml + A.wcap;      // heuristic discarded heap
    int* rid = reinterpret_cast<int*>(Dl + A.wcap);
    float* rd = reinterpret_cast<float*>(rid + kTopKMax);
    int* cl = reinterpret_cast<int*>(rd + kTopKMax);
    float* cd = reinterpret_cast<float*>(cl + kWave);
    uint8_t* qst = reinterpret_cast<uint8_t*>(cd + kTopKMax);
    uint8_t* qst2 = qst + A.vstride;

    WaveVisited vis;
    vis.tab = A.vtab + ((size_t)gw << A.vg_bits);
    vis.log = A.vlog + ((size_t)gw << A.vg_bits);
    vis.bits = A.vg_bits;
    vis.mask = (1u << A.vg_bits) - 1;
    vis.count = 0;
    vis.tag = 0;
    uint64_t* cg = A.cglob + (size_t)gw * (size_t)A.cg_cap;
    const int g = lane % DF::LPC;
    const int qdim = IS_INT ? (A.vstride >> 4) : A.dim;

    for (;;) {
        const int64_t i = next_node(A.qcounter, lane);
        if (i >= A.nb) break;
        const int ext = uni(A.node_ext[i]);
        const int l = uni(A.node_lvl[i]);
        // empty R / Pruned lists for every level until a search fills them
        for (int j = lane; j < 2 * A.maxl; j += kWave) A.r_cnt[(size_t)i * A.maxl * 2 + j] = 0;
        stage_row(qst, A.base + (size_t)ext * A.vstride, A.vstride, A.vstride, lane);
        DF df;
        df.load_query(qst, g, qdim);
        int fail = 0;
        WorkCount wc;
        vis.peak = 0;
        vis.full_at = -1;
        int cur = 0;  // m_ent_id of the top layer
        for (int lev = A.enter_level; lev >= l + 1 && !fail; lev--) {
            if (!((A.nonempty >> lev) & 1)) continue;
            const LayerDev L = A.layers[lev];
            cur = greedy_level(df, L, A.vstride, cur, vis, cl, cd, lane, fail, wc);
            if (!fail) cur = L.down[cur];
        }
        const int top = l < A.enter_level ? l : A.enter_level;
        for (int lev = top; lev >= 0 && !fail; lev--) {
            if (!((A.nonempty >> lev) & 1)) continue;
            const LayerDev L = A.layers[lev];
            const int topk = lev == 0 ? 2 * A.M : A.M;
            WaveHeap<true, false> W{Wl, Cl, A.wcap, 0, 0};
            WaveHeap<false, true> C{Cl, cg, A.ccap, A.cg_cap, 0};
            beam_level(df, L, A.vstride, cur, A.efc, W, C, vis, cl, cd, lane, fail, wc);
            if (fail) break;

            // SelectHeuristic (hnsw.hpp:70-116): W's array is pushed in storage
            // order into a min-heap and taken closest first
            WaveHeap<false, false> Mh{Ml, Cl, A.wcap, 0, 0};
            WaveHeap<false, false> Dh{Dl, Cl, A.wcap, 0, 0};
            for (int t = 0; t < W.len; t++) Mh.push_par(uni64(Wl[t]), lane, leader);
            int nR = 0, nP = 0;
            while (Mh.len > 0 && nR < topk) {
                const uint64_t v1 = Mh.top();
                bool keep = true;
                if (nR > 0) {
                    stage_row(qst2, L.vec + (size_t)entry_id(v1) * A.vstride, A.vstride, A.vstride, lane);
                    DF dv;
                    dv.load_query(qst2, g, qdim);
                    batch_distances(dv, L.vec, A.vstride, rid, cd, nR, lane);
                    __builtin_amdgcn_wave_barrier();
                    bool occl = false;
                    for (int j = lane; j < nR; j += kWave) occl |= entry_dist(v1) > cd[j];
                    keep = ballot(occl) == 0;
                }
                if (keep) {
                    if (leader) {
                        rid[nR] = entry_id(v1);
                        rd[nR] = entry_dist(v1);
                    }
                    nR++;
                } else {
                    Dh.push_par(v1, lane, leader);
                }
                Mh.pop_par(lane, leader);
                __builtin_amdgcn_wave_barrier();
            }
            const size_t slot = ((size_t)i * A.maxl + lev) * 2;
            int32_t* oid = A.r_ids + slot * A.topk;
            float* od = A.r_d + slot * A.topk;
            for (int j = lane; j < nR; j += kWave) {
                oid[j] = rid[j];
                od[j] = rd[j];
            }
            while (Dh.len > 0 && nR + nP < topk) {
                const uint64_t e = Dh.top();
                if (leader) {
                    oid[A.topk + nP] = entry_id(e);
                    od[A.topk + nP] = entry_dist(e);
                }
                nP++;
                Dh.pop_par(lane, leader);
            }
            if (leader) {
                A.r_cnt[slot] = nR;
                A.r_cnt[slot + 1] = nP;
            }
            cur = uni(rid[0]);
            if (lev > 0) cur = L.down[cur];
        }
        if (fail && leader) {
            atomicOr(A.err, fail);
            if (atomicCAS(&A.err[1], 0, 1) == 0) {
                A.err[2] = ext;
                A.err[3] = l;
                A.err[4] = vis.peak;
                A.err[5] = wc.hops;
                A.err[6] = vis.full_at;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// k_insert_rows: NewInID bookkeeping and AddNode (hnsw_layer.hpp:297-332)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kInsBlock) void k_insert_rows(InsertArgs A) {
    const int lane = lane_id();
    const bool leader = lane == 0;
    for (;;) {
        const int64_t i = next_node(A.qcounter, lane);
        if (i >= A.nb) break;
        const int ext = uni(A.node_ext[i]);
        const int l = uni(A.node_lvl[i]);
        const int topl = top_inserted_level(A, l);
        for (int lev = 0; lev <= l; lev++) {
            const LayerBuild B = A.build[lev];
            const int vid = uni(A.node_in[(size_t)i * A.maxl + lev]);
            if (leader) {
                B.ext[vid] = ext;
                if (lev >= 1) B.down[vid] = A.node_in[(size_t)i * A.maxl + lev - 1];
            }
            if (lev > topl) continue;  // a new top level: no AddNode (hnsw.hpp:283-297)
            // vector copy
            const uint4* src = reinterpret_cast<const uint4*>(A.base + (size_t)ext * A.vstride);
            uint4* dst = reinterpret_cast<uint4*>(B.vec + (size_t)vid * A.vstride);
            for (int c = lane; c < A.vstride / 16; c += kWave) dst[c] = src[c];
            // rows: R then Pruned, heuristic = min(D, |R|), degree = min(h + |P|, D)
            const size_t slot = ((size_t)i * A.maxl + lev) * 2;
            const bool searched = lev <= A.enter_level && ((A.nonempty >> lev) & 1);
            const int nR = searched ? min(uni(A.r_cnt[slot]), A.topk) : 0;
            const int nP = searched ? min(uni(A.r_cnt[slot + 1]), A.topk) : 0;
            const int h = nR < B.D ? nR : B.D;
            const int dg = h + nP < B.D ? h + nP : B.D;
            const int32_t* sid = A.r_ids + slot * A.topk;
            const float* sd = A.r_d + slot * A.topk;
            for (int j = lane; j < dg; j += kWave) {
                const int src_j = j < h ? j : A.topk + (j - h);
                B.adj[(size_t)vid * B.D + j] = sid[src_j];
                B.dist[(size_t)vid * B.D + j] = sd[src_j];
            }
            if (leader) {
                B.deg[vid] = dg;
                B.heur[vid] = (int16_t)h;
                B.has[vid] = 1;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// k_insert_edges: AddEdge (hnsw_layer.hpp:360-465) under per-node locks
// ---------------------------------------------------------------------------

// std::lower_bound over a[0, n) (first index whose value is not < x), with
// libstdc++'s probe sequence so an unsorted range behaves the same
__device__ int lower_bound_lds(const float* a, int n, float x) {
    int first = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        const int mid = first + half;
        if (unif(a[mid]) < x) {
            first = mid + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

__device__ bool lock_acquire(int32_t* lk, int lane) {
    for (int spin = 0; spin < (1 << 22); spin++) {
        int got = 0;
        if (lane == 0) got = atomicCAS(lk, 0, 1) == 0;
        if (readlane(got, 0)) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return true;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return false;
}

__device__ void lock_release(int32_t* lk, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) atomicExch(lk, 0);
}

// The suffix rb[offset, dg) in ascending distance order, written to S as
// (distance, slot t or id) entries, when that order is unique: both
// segments [offset, h) and [h, dg) strictly increasing and no distance
// shared between them -- then a min-heap of the entries pops exactly this
// order.  Each entry's rank is its index in its segment plus a binary
// search in the other.  Returns false (S unusable) otherwise.
__device__ bool sorted_suffix(const float* d, const int* id, int offset, int h, int dg, bool by_slot, uint64_t* S,
                              int lane) {
    bool bad = false;
    for (int t = offset + lane; t < dg; t += kWave) {
        const float x = d[t];
        const bool in1 = t < h;
        const int own = in1 ? offset : h;
        const int lo = in1 ? h : offset, hi = in1 ? dg : h;  // the other segment
        if (t > own && !(d[t - 1] < x)) bad = true;
        int first = lo, len = hi - lo;
        while (len > 0) {
            const int half = len >> 1;
            const int mid = first + half;
            if (d[mid] < x) {
                first = mid + 1;
                len = len - half - 1;
            } else {
                len = half;
            }
        }
        if (first < hi && !(d[first] > x)) bad = true;  // a tie (or NaN)
        S[(t - own) + (first - lo)] = pack_entry(x, by_slot ? t : id[t]);
    }
    __builtin_amdgcn_wave_barrier();
    return ballot(bad) == 0;
}

// Is some member of the kept set nb_id[0, n) strictly closer than thr to
// the row loaded as `dv`?  Members with an LDS slot (nb_slot, always a
// prefix of the set) are read from the row cache, the others from device
// memory; the scan stops at the first pass that finds one.
template <class DF>
__device__ __forceinline__ bool kept_closer(const DF& dv, const uint8_t* cache, const uint8_t* vec, int vstride,
                                            const int* nb_id, const int* nb_slot, int in_lds, int n, float thr,
                                            int lane) {
    const int a = n < in_lds ? n : in_lds;
    if (a > 0 && any_closer<DF, true>(dv, cache, vstride, nb_slot, a, thr, lane)) return true;
    return n > a && any_closer<DF, false>(dv, vec, vstride, nb_id + a, n - a, thr, lane);
}

// AddEdge(vid, nid, distance) under vid's lock (hnsw_layer.hpp:360-465).
// Vectors come from an LDS row cache wherever they fit:
//  * short rows (dg + 1 <= A.row_cache): the whole row and the new node are
//    staged once (row entry j in slot j, the new node in slot dg), so the
//    occlusion test and the heuristic re-run read only LDS; the re-run's
//    heaps then hold (distance, slot): their order depends on distances only;
//  * long rows: the kept set's vectors are cached while they fit
//    (A.row_cache_long slots, the rest gathered per step), and the next
//    entry of the re-run -- the heap's next pop, which does not depend on
//    the keep decisions -- is loaded while the current one is compared.
// Every distance keeps the reference's summation order.
template <class DF, bool IS_INT>
__device__ void add_edge(const InsertArgs& A, const LayerBuild& B, int64_t n_rows, int vid, int nid, float distance,
                         int lane, int g, int qdim, int* rb_id, float* rb_d, int* nb_id, float* nb_d, int* nb_slot,
                         float* cd, uint8_t* region, int& fail) {
    const bool leader = lane == 0;
    if (!lock_acquire(&B.lock[vid], lane)) {
        fail = kFailLockTimeout;
        return;
    }
    const size_t rowbase = (size_t)vid * B.D;
    const int vs = A.vstride;
    // a row that grew into the end of the layer's allocation holds what fit
    // (the reference's flat array has no room there either)
    int dg = uni(B.deg[vid]);
    if ((size_t)dg > (size_t)B.adj_len - rowbase) dg = (int)(B.adj_len - rowbase);
    int h = uni((int)B.heur[vid]);
    if (h > dg) h = dg;
    if (dg > kRowMax - 1) {
        fail = kFailRowLimit;
        lock_release(&B.lock[vid], lane);
        return;
    }
    bool bad = false;
    for (int j = lane; j < dg; j += kWave) {
        const int id = B.adj[rowbase + j];
        bad |= id < 0 || (int64_t)id >= n_rows;
        rb_id[j] = id;
        rb_d[j] = B.dist[rowbase + j];
    }
    __builtin_amdgcn_wave_barrier();
    if (ballot(bad)) {
        // a row that reads past its own slots into storage no node has
        // written yet (ids outside the layer): the reference would compute
        // distances to such ids; here the edge is dropped and the row left
        // as it is
        lock_release(&B.lock[vid], lane);
        return;
    }
    const int offset = lower_bound_lds(rb_d, h, distance);

    // region layout.  short rows: row_cache rows, then two heaps of
    // row_cache entries.  long rows: kc kept-set rows, two staging rows
    // (slots kc, kc + 1), then two heaps of kRowMax entries.
    constexpr int kStage = 2;
    const bool cached = dg + 1 <= A.row_cache;
    uint8_t* cache = region;
    const int kc = cached ? A.row_cache : A.row_cache_long;
    const int hcap = cached ? A.row_cache : kRowMax;
    uint64_t* h1 = reinterpret_cast<uint64_t*>(region + (size_t)(cached ? A.row_cache : kc + kStage) * vs);
    uint64_t* h2 = h1 + hcap;

    // the kept set of a re-run starts as the prefix plus the new node; in
    // the long-row layout its entries past kc have no LDS slot (-1)
    const int nid_slot = cached ? dg : (offset < kc ? offset : kc + 1);
    for (int j = lane; j < offset; j += kWave) {
        nb_id[j] = rb_id[j];
        nb_d[j] = rb_d[j];
        nb_slot[j] = j < kc ? j : -1;
    }
    if (leader) {
        nb_id[offset] = nid;
        nb_d[offset] = distance;
        nb_slot[offset] = nid_slot;
    }
    __builtin_amdgcn_wave_barrier();
    // the prefix rows that have slots, and the new node, in one round trip
    const int staged = offset < kc ? offset : kc;
    {
        // ids and slots of what is staged (nb_id / nb_slot [0, staged), then
        // the new node), listed in the heaps' space before the heaps exist
        int* sid = reinterpret_cast<int*>(h1);
        int* sslot = sid + hcap;
        for (int j = lane; j < staged; j += kWave) {
            sid[j] = nb_id[j];
            sslot[j] = nb_slot[j];
        }
        if (leader) {
            sid[staged] = nid;
            sslot[staged] = nid_slot;
        }
        __builtin_amdgcn_wave_barrier();
        stage_rows(cache, B.vec, vs, sid, sslot, 0, staged + 1, lane);
    }
    const int in_lds0 = staged;  // prefix entries with LDS slots

    // is the new node occluded by a closer heuristic neighbour?
    bool insert_n = true;
    if (offset > 0) {
        DF dn;
        dn.load_query(cache + (size_t)nid_slot * vs, g, qdim);
        insert_n = !kept_closer(dn, cache, B.vec, vs, nb_id, nb_slot, in_lds0, offset, distance, lane);
    }

    if (!insert_n) {
        // sorted insert into the pruned part (hnsw_layer.hpp:386-408)
        const int pruned = dg - h;
        const int off = lower_bound_lds(rb_d + h, pruned, distance);
        if (off == pruned && dg < B.D) {
            if (leader) {
                B.adj[rowbase + h + off] = nid;
                B.dist[rowbase + h + off] = distance;
                B.deg[vid] = dg + 1;
            }
        } else if (off < pruned) {
            const int end = dg < B.D ? pruned : pruned - 1;
            for (int j = off + 1 + lane; j <= end; j += kWave) {
                B.adj[rowbase + h + j] = rb_id[h + j - 1];
                B.dist[rowbase + h + j] = rb_d[h + j - 1];
            }
            if (leader) {
                B.adj[rowbase + h + off] = nid;
                B.dist[rowbase + h + off] = distance;
                B.deg[vid] = dg + (dg < B.D);
            }
        }
    } else if (offset >= (A.exact_rows ? A.row_limit : B.D)) {
        // a full row of closer heuristic neighbours: the reference writes
        // slot D (hnsw_layer.hpp:420, 441-447); rows capped at D drop the
        // edge, and so do reference rows at their limit of kRowMax - 1 ids
    } else {
        // re-run the heuristic on the entries from offset on (:409-463).
        // The reference pops them from a min-heap W; when their distances are
        // all distinct that order is the sorted one, built here by a
        // parallel merge of the two sorted segments (`listed`), and the
        // rejected entries then leave Wd in the order they entered it.  With
        // ties the heaps run as the reference's do (their layout orders
        // equal distances).
        if (cached && dg > offset) stage_rows(cache, B.vec, vs, rb_id + offset, nullptr, offset, dg - offset, lane);
        const int n_s = dg - offset;
        uint64_t* S = h1;   // listed: the suffix in pop order
        uint64_t* Sd = h2;  // listed: the rejected entries in order
        const bool listed = sorted_suffix(rb_d, rb_id, offset, h, dg, cached, S, lane);
        WaveHeap<false, false> W{h1, h1, hcap, 0, 0};
        WaveHeap<false, false> Wd{h2, h2, hcap, 0, 0};
        if (!listed)
            for (int t = offset; t < dg; t++) W.push_par(pack_entry(unif(rb_d[t]), cached ? t : uni(rb_id[t])), lane, leader);
        int it = 0, nrej = 0;
        int cur_h = offset + 1;
        int in_lds = cached ? cur_h : (nid_slot < kc ? offset + 1 : in_lds0);  // kept entries with LDS slots
        // long rows: v1's row in staging slot `cur`, the next one loading into the other
        int cur = kc + 1;
        const int cpr = vs >> 4;
        if (!cached && n_s > 0)
            stage_one_row(cache + (size_t)cur * vs, B.vec + (size_t)entry_id(listed ? uni64(S[0]) : W.top()) * vs, vs,
                          lane);
        while (listed ? it < n_s : W.len > 0) {
            uint64_t v1;
            if (listed) {
                v1 = uni64(S[it]);
                it++;
            } else {
                v1 = W.top();
                W.pop_par(lane, leader);
            }
            const bool more = listed ? it < n_s : W.len > 0;
            const int next_key = more ? entry_id(listed ? uni64(S[it]) : W.top()) : 0;
            const int v1_slot = cached ? entry_id(v1) : cur;
            const int v1_id = cached ? uni(rb_id[v1_slot]) : entry_id(v1);
            // long rows: the next entry's row is loaded while this one is compared
            uint4 pf0 = make_uint4(0, 0, 0, 0), pf1 = make_uint4(0, 0, 0, 0);
            const bool prefetch = !cached && more && cpr <= 2 * kWave;
            if (prefetch) {
                const uint8_t* nrow = B.vec + (size_t)next_key * vs;
                if (lane < cpr) pf0 = ld16_global(nrow + 16 * lane);
                if (lane + kWave < cpr) pf1 = ld16_global(nrow + 16 * (lane + kWave));
            }
            DF dv;
            dv.load_query(cache + (size_t)v1_slot * vs, g, qdim);
            const bool occl = kept_closer(dv, cache, B.vec, vs, nb_id, nb_slot, in_lds, cur_h, entry_dist(v1), lane);
            const bool keep = !occl && cur_h < (A.exact_rows ? A.row_limit : B.D);
            if (keep) {
                int slot = v1_slot;
                if (!cached) {
                    // the kept row joins the LDS part of the set while it fits
                    slot = -1;
                    if (in_lds == cur_h && cur_h < kc) {
                        slot = cur_h;
                        for (int c = lane; c < cpr; c += kWave)
                            *reinterpret_cast<uint4*>(cache + (size_t)slot * vs + 16 * c) =
                                ld16_lds(cache + (size_t)v1_slot * vs + 16 * c);
                        in_lds++;
                    }
                } else {
                    in_lds++;
                }
                if (leader) {
                    nb_id[cur_h] = v1_id;
                    nb_d[cur_h] = entry_dist(v1);
                    nb_slot[cur_h] = slot;
                }
                cur_h++;
            } else if (listed) {
                if (leader) Sd[nrej] = v1;
                nrej++;
            } else {
                Wd.push_par(v1, lane, leader);
            }
            if (!cached && more) {
                // the next entry's row into the other staging slot
                cur = cur == kc ? kc + 1 : kc;
                if (prefetch) {
                    if (lane < cpr) *reinterpret_cast<uint4*>(cache + (size_t)cur * vs + 16 * lane) = pf0;
                    if (lane + kWave < cpr) *reinterpret_cast<uint4*>(cache + (size_t)cur * vs + 16 * (lane + kWave)) = pf1;
                } else {
                    stage_one_row(cache + (size_t)cur * vs, B.vec + (size_t)next_key * vs, vs, lane);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        int cur_deg = cur_h;
        for (int r = 0; cur_deg < B.D && (listed ? r < nrej : Wd.len > 0); r++) {
            const uint64_t e = listed ? uni64(Sd[r]) : Wd.top();
            const int e_id = cached ? uni(rb_id[entry_id(e)]) : entry_id(e);
            if (leader) {
                nb_id[cur_deg] = e_id;
                nb_d[cur_deg] = entry_dist(e);
            }
            cur_deg++;
            if (!listed) Wd.pop_par(lane, leader);
        }
        __builtin_amdgcn_wave_barrier();
        // flat row offsets, as the reference's spans: with exact rows a row
        // that outgrows D continues into the next row's slots
        const int wr = rowbase + cur_deg <= (size_t)B.adj_len ? cur_deg : (int)(B.adj_len - rowbase);
        for (int j = lane; j < wr; j += kWave) {
            B.adj[rowbase + j] = nb_id[j];
            B.dist[rowbase + j] = nb_d[j];
        }
        if (leader) {
            B.heur[vid] = (int16_t)(cur_h < wr ? cur_h : wr);
            B.deg[vid] = wr;
        }
    }
    lock_release(&B.lock[vid], lane);
}

template <class DF, bool IS_INT>
__global__ __launch_bounds__(kInsBlock) void k_insert_edges(InsertArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    int* rb_id = reinterpret_cast<int*>(smem);
    float* rb_d = reinterpret_cast<float*>(rb_id + kRowMax);
    int* nb_id = reinterpret_cast<int*>(rb_d + kRowMax);
    float* nb_d = reinterpret_cast<float*>(nb_id + kRowMax);
    float* cd = reinterpret_cast<float*>(nb_d + kRowMax);
    int* nb_slot = reinterpret_cast<int*>(cd + kRowMax);
    // row cache and heaps (see add_edge)
    uint8_t* region = reinterpret_cast<uint8_t*>(nb_slot + kRowMax);
    const int g = lane % DF::LPC;
    const int qdim = IS_INT ? (A.vstride >> 4) : A.dim;
    for (;;) {
        const int64_t i = next_node(A.qcounter, lane);
        if (i >= A.nb) break;
        const int l = uni(A.node_lvl[i]);
        const int topl = top_inserted_level(A, l);
        int fail = 0;
        for (int lev = 0; lev <= topl && !fail; lev++) {
            const bool searched = lev <= A.enter_level && ((A.nonempty >> lev) & 1);
            if (!searched) continue;
            const LayerBuild B = A.build[lev];
            const int64_t n_rows = A.rows_after[lev];
            const int vid = uni(A.node_in[(size_t)i * A.maxl + lev]);
            const size_t slot = ((size_t)i * A.maxl + lev) * 2;
            const int nR = min(uni(A.r_cnt[slot]), A.topk);
            const int nP = min(uni(A.r_cnt[slot + 1]), A.topk);
            const int32_t* sid = A.r_ids + slot * A.topk;
            const float* sd = A.r_d + slot * A.topk;
            for (int t = 0; t < nR + nP && !fail; t++) {
                const int src = t < nR ? t : A.topk + (t - nR);
                add_edge<DF, IS_INT>(A, B, n_rows, uni(sid[src]), vid, unif(sd[src]), lane, g, qdim, rb_id, rb_d,
                                     nb_id, nb_d, nb_slot, cd, region, fail);
            }
        }
        if (fail && lane == 0) atomicOr(A.err, fail);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_fill_i32_ins(int32_t* p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f32_ins(float* p, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class DF, bool INT>
// phase 0: construction searches (R / Pruned lists to scratch only, so the
// phase can run again); phase 1: AddNode rows, then AddEdge
static hipError_t launch_ins(const InsertArgs& a, int phase, int grid, size_t lds_a, size_t lds_c, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32_ins, dim3(1), dim3(64), 0, s, a.qcounter, 0, 1);
    if (phase == 0) {
        hipLaunchKernelGGL((k_insert_search<DF, INT>), dim3(grid), dim3(kInsBlock), lds_a, s, a);
    } else {
        hipLaunchKernelGGL(k_insert_rows, dim3(grid), dim3(kInsBlock), 0, s, a);
        hipLaunchKernelGGL(k_fill_i32_ins, dim3(1), dim3(64), 0, s, a.qcounter, 0, 1);
        hipLaunchKernelGGL((k_insert_edges<DF, INT>), dim3(grid), dim3(kInsBlock), lds_c, s, a);
    }
    return hipGetLastError();
}

template <int DT, int MET>
static hipError_t launch_ins_dt(const InsertArgs& a, int phase, int size_class, int grid, size_t la, size_t lc,
                                hipStream_t s) {
    constexpr bool INT = DT == DT_U8 || DT == DT_I8;
    if constexpr (INT) {
        if (size_class == 0) return launch_ins<IntDist<DT, MET, 1>, true>(a, phase, grid, la, lc, s);
        if (size_class == 1) return launch_ins<IntDist<DT, MET, 2>, true>(a, phase, grid, la, lc, s);
        return launch_ins<IntDist<DT, MET, 4>, true>(a, phase, grid, la, lc, s);
    } else {
        if (size_class == 0) return launch_ins<FloatDist<DT, MET, 8>, false>(a, phase, grid, la, lc, s);
        return launch_ins<FloatDist<DT, MET, 16>, false>(a, phase, grid, la, lc, s);
    }
}

static hipError_t launch_insert(const InsertArgs& a, int phase, int dtype, int metric, int size_class, int grid,
                                size_t la, size_t lc, hipStream_t s) {
    switch (dtype * 4 + metric) {
        case DT_U8 * 4 + MET_L2: return launch_ins_dt<DT_U8, MET_L2>(a, phase, size_class, grid, la, lc, s);
        case DT_U8 * 4 + MET_IP: return launch_ins_dt<DT_U8, MET_IP>(a, phase, size_class, grid, la, lc, s);
        case DT_I8 * 4 + MET_L2: return launch_ins_dt<DT_I8, MET_L2>(a, phase, size_class, grid, la, lc, s);
        case DT_I8 * 4 + MET_IP: return launch_ins_dt<DT_I8, MET_IP>(a, phase, size_class, grid, la, lc, s);
        case DT_F32 * 4 + MET_L2: return launch_ins_dt<DT_F32, MET_L2>(a, phase, size_class, grid, la, lc, s);
        case DT_F32 * 4 + MET_IP: return launch_ins_dt<DT_F32, MET_IP>(a, phase, size_class, grid, la, lc, s);
        case DT_F16 * 4 + MET_L2: return launch_ins_dt<DT_F16, MET_L2>(a, phase, size_class, grid, la, lc, s);
        case DT_F16 * 4 + MET_IP: return launch_ins_dt<DT_F16, MET_IP>(a, phase, size_class, grid, la, lc, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pk

using namespace pk;

namespace {

#define ICHK(expr)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return set_error(PK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int ialloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) != hipSuccess)
        return set_error(PK_ERR_NOMEM, "device allocation failed (" + std::to_string(count * sizeof(T)) + " bytes)");
    return PK_OK;
}

template <class T>
void ifree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return PK_OK;
    hipLaunchKernelGGL(k_fill_i32_ins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    ICHK(hipGetLastError());
    return PK_OK;
}

int fill_f32(float* p, float v, int64_t n, hipStream_t s) {
    if (n <= 0) return PK_OK;
    hipLaunchKernelGGL(k_fill_f32_ins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    ICHK(hipGetLastError());
    return PK_OK;
}

}  // namespace

// the level sampler of include/util.hpp:13-64 (fresh uniform per level)
extern "C" int pk_sample_levels(int64_t n, int M, uint64_t seed, uint8_t* out) {
    if (n < 0 || (n > 0 && !out) || M < 2) return set_error(PK_ERR_ARG, "bad arguments");
    std::vector<double> cum;
    const double r_logm = 1.0 / std::log((double)M);
    double c = 0.0;
    for (;;) {
        const int lev = (int)cum.size();
        const double delta = std::exp(-1.0 * lev / r_logm) * (1 - std::exp(-1.0 / r_logm));
        if (delta < 1e-9) break;
        c += delta;
        cum.push_back(c);
    }
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int64_t i = 0; i < n; i++) {
        int lv = (int)cum.size() - 1;
        for (size_t lev = 0; lev < cum.size(); lev++)
            if (u(rng) < cum[lev]) {
                lv = (int)lev;
                break;
            }
        out[i] = (uint8_t)std::min(lv, PK_MAX_LEVELS - 1);
    }
    return PK_OK;
}

extern "C" int pk_index_init_build(pk_index* h, const uint8_t* levels) {
    if (!h || !levels) return set_error(PK_ERR_ARG, "null argument");
    if (!h->base) return set_error(PK_ERR_STATE, "set the base rows before init_build");
    if (h->graphs_alive || h->servers_alive)
        return set_error(PK_ERR_STATE, "destroy the search graphs and serving loops before init_build");
    if (2 * h->M > kTopKMax) return set_error(PK_ERR_ARG, "M above 64 is not supported by the GPU build");
    h->levels.assign(levels, levels + h->capacity);
    int maxl = 0;
    std::vector<int64_t> cnt(PK_MAX_LEVELS, 0);
    for (int64_t i = 0; i < h->capacity; i++) {
        if (levels[i] >= PK_MAX_LEVELS) return set_error(PK_ERR_ARG, "level above 15");
        maxl = std::max<int>(maxl, levels[i]);
        for (int l = 0; l <= levels[i]; l++) cnt[l]++;
    }
    ICHK(hipStreamSynchronize(h->stream));
    for (int l = 0; l < PK_MAX_LEVELS; l++) pk::free_layer_public(h->layers[l]);
    for (int l = 0; l <= maxl; l++) {
        pk::Layer& L = h->layers[l];
        const int D = l == 0 ? 2 * h->M : h->M;
        const int64_t cap = cnt[l];
        L.cap = cap;
        L.n = 0;
        L.D = D;
        L.adj_len = (cap + 1) * D;  // one spare row: exact-rows mode writes past a full last row
        int rc;
        if ((rc = ialloc(&L.adj, (size_t)L.adj_len + 64)) || (rc = ialloc(&L.deg, (size_t)cap + 1)) ||
            (rc = ialloc(&L.heur, (size_t)cap + 1)) || (rc = ialloc(&L.dist, (size_t)L.adj_len + 64)) ||
            (rc = ialloc(&L.ext, (size_t)cap + 1)) || (rc = ialloc(&L.vec, (size_t)(cap + 1) * h->vstride)) ||
            (rc = ialloc(&L.has, (size_t)cap + 1)) || (rc = ialloc(&L.down, (size_t)cap + 1)) ||
            (rc = ialloc(&L.lock, (size_t)cap + 1)))
            return rc;
        if ((rc = fill_i32(L.adj, -1, L.adj_len + 64, h->stream)) ||
            (rc = fill_f32(L.dist, 3.402823466e+38f, L.adj_len + 64, h->stream)) ||
            (rc = fill_i32(L.deg, 0, cap + 1, h->stream)) || (rc = fill_i32(L.ext, -1, cap + 1, h->stream)) ||
            (rc = fill_i32(L.down, -1, cap + 1, h->stream)) || (rc = fill_i32(L.lock, 0, cap + 1, h->stream)))
            return rc;
        ICHK(hipMemsetAsync(L.heur, 0, (cap + 1) * sizeof(int16_t), h->stream));
        ICHK(hipMemsetAsync(L.has, 0, cap + 1, h->stream));
        ICHK(hipMemsetAsync(L.vec, 0, (size_t)(cap + 1) * h->vstride, h->stream));
        L.loaded = true;
        L.built = true;
    }
    ICHK(hipStreamSynchronize(h->stream));
    h->enter_level = 0;
    h->ready = false;
    h->building = true;
    return pk::publish_layers(h);
}

// One insert phase's kernels on the index stream, directly or (graph) as a
// HIP graph: the launch sequence is captured, and the phase's executable
// graph is instantiated on first use and updated in place with the new
// arguments and grid afterwards (instantiated again if an update is refused).
static int run_insert_phase(pk_index* h, const InsertArgs& a, int phase, int grid, size_t lds_a, size_t lds_c,
                            bool graph) {
    if (!graph) {
        ICHK(launch_insert(a, phase, h->dtype, h->metric, h->size_class, grid, lds_a, lds_c, h->stream));
        return PK_OK;
    }
    hipGraph_t g = nullptr;
    ICHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = launch_insert(a, phase, h->dtype, h->metric, h->size_class, grid, lds_a, lds_c, h->stream);
    const hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_error(PK_ERR_HIP, std::string("insert graph capture: ") +
                                         hipGetErrorString(le != hipSuccess ? le : ce));
    }
    hipGraphExec_t& ex = h->ins.graph[phase];
    if (ex) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(ex, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();  // the refused update leaves no error behind
            (void)hipGraphExecDestroy(ex);
            ex = nullptr;
        }
    }
    hipError_t e = hipSuccess;
    if (!ex) {
        hipGraphExec_t fresh = nullptr;
        e = hipGraphInstantiate(&fresh, g, nullptr, nullptr, 0);
        if (e == hipSuccess) ex = fresh;
    }
    if (e == hipSuccess) e = hipGraphLaunch(ex, h->stream);
    (void)hipGraphDestroy(g);  // the executable graph holds its own copy
    if (e != hipSuccess) {
        if (ex) (void)hipGraphExecDestroy(ex);
        ex = nullptr;
        return set_error(PK_ERR_HIP, std::string("insert graph: ") + hipGetErrorString(e));
    }
    return PK_OK;
}

extern "C" int pk_index_insert(pk_index* h, const int32_t* ext_ids, int64_t n, int max_batch, int flags) {
    if (!h || (n > 0 && !ext_ids)) return set_error(PK_ERR_ARG, "null argument");
    if (flags & ~(PK_BUILD_EXACT_ROWS | PK_BUILD_GRAPH)) return set_error(PK_ERR_ARG, "unknown insert flags");
    const bool graph = (flags & PK_BUILD_GRAPH) != 0;
    if (!h->building) return set_error(PK_ERR_STATE, "call pk_index_init_build first");
    if (max_batch < 1) max_batch = 4096;
    for (int64_t t = 0; t < n; t++)
        if (ext_ids[t] < 0 || ext_ids[t] >= h->capacity || ext_ids[t] >= h->base_rows)
            return set_error(PK_ERR_ARG, "ext id " + std::to_string(ext_ids[t]) + " outside the base rows");
    // visited table: room for the efc-bounded search's visits at a low load;
    // grown (up to 2^20 slots per wave) when a search fills it, since rows
    // that grew past D make some searches touch many more ids
    int vbits = std::max(h->ins.vbits, h->efc <= 64 ? 14 : (h->efc <= 160 ? 15 : 16));
    const int wcap = (h->efc + 2 + 7) / 8 * 8;
    if (wcap > 2 * 64 * kPopSlots) return set_error(PK_ERR_ARG, "efc too large (max 630)");
    const int ccap = 1024;
    size_t lds_a = (size_t)wcap * 8 * 3 + (size_t)ccap * 8 + kTopKMax * 8 + 64 * 4 + kTopKMax * 4 + 2 * h->vstride;
    lds_a = (lds_a + 15) / 16 * 16;
    // AddEdge: six row lists and the staged new node, then a region that
    // holds the LDS row cache with two heaps of its size, or two heaps of
    // kRowMax entries for rows too long for the cache.  The cache takes what
    // keeps the workgroup within 52 KB, three resident per CU (76 rows of
    // float32 d = 128: a row of 2M = 64 ids plus the new node fits).
    //   short rows: row_cache rows + two heaps of row_cache entries
    //   long rows:  row_cache_long kept-set rows + two staging rows + two
    //               heaps of kRowMax entries, in the same space
    const size_t vs = (size_t)h->vstride;
    const size_t lds_c0 = (size_t)kRowMax * 4 * 6;
    const size_t long_heaps = (size_t)kRowMax * 8 * 2;
    constexpr size_t kEdgeLds = 52 * 1024;
    int row_cache = lds_c0 + 2 * (vs + 16) > kEdgeLds
                        ? 0
                        : (int)std::min<size_t>(kRowMax, (kEdgeLds - lds_c0) / (vs + 16));
    if (row_cache < 2) row_cache = 0;
    const size_t short_region = (size_t)row_cache * (vs + 16);
    const size_t stage_rows_long = 2;
    const int row_cache_long =
        short_region > long_heaps + stage_rows_long * vs
            ? (int)std::min<size_t>(kRowMax / 2, (short_region - long_heaps) / vs - stage_rows_long)
            : 0;
    const size_t region = std::max(short_region, long_heaps + (row_cache_long + stage_rows_long) * vs);
    const size_t lds_c = (lds_c0 + region + 15) / 16 * 16;
    int grid_cap = h->num_cus * 8;
    // scratch for the largest batch
    const int maxl_all = [&] {
        int m = 0;
        for (int l = 0; l < PK_MAX_LEVELS; l++)
            if (h->layers[l].built) m = l + 1;
        return m;
    }();
    const int topk = 2 * h->M;
    int rc = pk::ensure_insert_scratch(h, max_batch, maxl_all, topk, grid_cap, vbits);
    if (rc) return rc;
    int32_t diag[8] = {0};
    auto device_flags = [&](int32_t* out) -> int {
        ICHK(hipStreamSynchronize(h->stream));
        ICHK(hipMemcpy(diag, h->ins.err, sizeof(diag), hipMemcpyDeviceToHost));
        *out = diag[0];
        return PK_OK;
    };
    auto diag_text = [&]() {
        return " (flags " + std::to_string(diag[0]) + "; first failing node: ext " + std::to_string(diag[2]) +
               ", level " + std::to_string(diag[3]) + ", " + std::to_string(diag[4]) + " visited ids, " +
               std::to_string(diag[5]) + " hops, table full after " + std::to_string(diag[6]) +
               " inserts, visited table 2^" + std::to_string(vbits) + ", recorded " + std::to_string(diag[1]) + ")";
    };
    ICHK(hipMemsetAsync(h->ins.err, 0, 32, h->stream));

    std::vector<int32_t> b_ext, b_lvl, b_in;
    int64_t t = 0;
    while (t < n) {
        // batch: grows with the graph; a node above the current top level goes alone
        const int64_t size0 = h->layers[0].n;
        int64_t allow = std::max<int64_t>(1, std::min<int64_t>(max_batch, size0 / 8));
        int64_t s = t;
        int batch_top = 0;
        while (s < n && s - t < allow) {
            const int lv = h->levels[ext_ids[s]];
            if (lv > h->enter_level) {
                if (s == t) s++;
                break;
            }
            batch_top = std::max(batch_top, lv);
            s++;
        }
        const int64_t nb = s - t;
        int maxl = 0;
        for (int64_t k = t; k < s; k++) maxl = std::max<int>(maxl, h->levels[ext_ids[k]]);
        maxl += 1;
        b_ext.assign(ext_ids + t, ext_ids + s);
        b_lvl.resize(nb);
        b_in.assign((size_t)nb * maxl, -1);
        int nonempty = 0;
        for (int l = 0; l < PK_MAX_LEVELS; l++)
            if (h->layers[l].built && h->layers[l].n > 0) nonempty |= 1 << l;
        for (int64_t k = 0; k < nb; k++) {
            const int lv = h->levels[b_ext[k]];
            b_lvl[k] = lv;
            for (int l = 0; l <= lv; l++) {
                pk::Layer& L = h->layers[l];
                if (L.n >= L.cap) return set_error(PK_ERR_CAPACITY, "layer " + std::to_string(l) + " is full");
                b_in[(size_t)k * maxl + l] = (int32_t)L.n++;  // NewInID (hnsw_layer.hpp:168-177)
            }
        }
        InsertArgs a{};
        if ((rc = pk::upload_insert_batch(h, b_ext.data(), b_lvl.data(), b_in.data(), nb, maxl))) return rc;
        a.layers = static_cast<const LayerDev*>(h->d_layers);
        a.build = static_cast<const LayerBuild*>(h->d_build);
        a.base = h->base;
        a.vstride = h->vstride;
        a.qbytes = h->qbytes;
        a.dim = h->dim;
        a.enter_level = h->enter_level;
        a.nonempty = nonempty;
        a.efc = h->efc;
        a.M = h->M;
        a.nb = nb;
        a.node_ext = h->ins.d_ext;
        a.node_lvl = h->ins.d_lvl;
        a.node_in = h->ins.d_in;
        a.maxl = maxl;
        a.topk = topk;
        a.r_ids = h->ins.r_ids;
        a.r_d = h->ins.r_d;
        a.r_cnt = h->ins.r_cnt;
        a.cglob = h->ins.cglob;
        a.vtab = h->ins.vtab;
        a.vlog = h->ins.vlog;
        a.cg_cap = std::max(h->cg_cap, 1 << vbits);
        a.vg_bits = vbits;
        a.wcap = wcap;
        a.ccap = ccap;
        a.lds_per_wave = (int)lds_a;
        a.exact_rows = (flags & PK_BUILD_EXACT_ROWS) != 0;
        a.row_limit = h->row_limit > 0 ? std::min(h->row_limit, kRowMax - 1) : kRowMax - 1;
        a.row_cache = row_cache;
        a.row_cache_long = row_cache_long;
        for (int l = 0; l < PK_MAX_LEVELS; l++) a.rows_after[l] = h->layers[l].n;
        a.qcounter = h->ins.counter;
        a.err = h->ins.err;
        int grid = (int)std::min<int64_t>({nb, (int64_t)grid_cap, (int64_t)pk::insert_scratch_waves(h, vbits)});
        for (;;) {
            if ((rc = run_insert_phase(h, a, 0, grid, lds_a, lds_c, graph))) return rc;
            int32_t err = 0;
            if ((rc = device_flags(&err))) return rc;
            if (err == 0) break;
            if (err != kFailVisitedFull || vbits >= 20)
                return set_error(PK_ERR_DEVICE, "insertion search failed on the device" + diag_text());
            vbits++;
            h->ins.vbits = vbits;
            // the larger tables run on fewer waves in the memory already held
            if ((rc = pk::ensure_insert_scratch(h, max_batch, maxl_all, topk, grid_cap, vbits))) return rc;
            grid = (int)std::min<int64_t>({nb, (int64_t)grid_cap, (int64_t)pk::insert_scratch_waves(h, vbits)});
            a.vtab = h->ins.vtab;
            a.vlog = h->ins.vlog;
            a.cglob = h->ins.cglob;
            a.cg_cap = std::max(h->cg_cap, 1 << vbits);
            a.vg_bits = vbits;
            ICHK(hipMemsetAsync(h->ins.err, 0, 32, h->stream));
        }
        if ((rc = run_insert_phase(h, a, 1, grid, lds_a, lds_c, graph))) return rc;
        {
            int32_t err = 0;
            if ((rc = device_flags(&err))) return rc;
            if (err)
                return set_error(PK_ERR_DEVICE, "AddEdge failed on the device (flags " + std::to_string(err) +
                                                    ": 8 = a node lock was held for seconds, 16 = a row reached " +
                                                    std::to_string(kRowMax - 1) + " ids)");
        }
        // the new sizes become visible, then SetEntLevel (hnsw.hpp:146-151):
        // a search launched from another thread that reads the new enter
        // level also finds the layers it starts on published
        if ((rc = pk::publish_layers(h))) return rc;
        int top = h->enter_level;
        for (int64_t k = 0; k < nb; k++) top = std::max(top, b_lvl[k]);
        h->enter_level = top;
        t = s;
    }
    ICHK(hipStreamSynchronize(h->stream));
    int32_t err = 0;
    ICHK(hipMemcpy(&err, h->ins.err, 4, hipMemcpyDeviceToHost));
    if (err) return set_error(PK_ERR_DEVICE, "insertion failed on the device (flags " + std::to_string(err) + ")");
    if (!h->ready && h->layers[0].n > 0) h->ready = true;  // already set while a serving loop searches
    return PK_OK;
}

extern "C" int pk_set_build_row_limit(pk_index* h, int limit) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    if (limit != 0 && (limit < 2 * h->M || limit > kRowMax - 1))
        return set_error(PK_ERR_ARG, "row limit must be 0 (511) or in [2M, 511]");
    h->row_limit = limit;
    return PK_OK;
}
