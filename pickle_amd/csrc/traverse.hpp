// traverse.hpp -- one layer of HNSW traversal, executed by a whole wave.
//
// Shared by the query kernel and the insertion kernels:
//   greedy_level  SlideLayer (include/hnsw.hpp:446-494)
//   beam_level    SearchLayer / SearchLayerFast (include/hnsw.hpp:304-444)
// Both mark visited nodes exactly as the reference (the entry first, every
// adjacency id before its distance is used) and count hops, scanned
// adjacency entries and distance evaluations with the Profiler's
// definitions (include/hnsw.hpp:407-410, 438-442, 488-492).
#pragma once
#include "distance.hpp"
#include "heap.hpp"
#include "search.hpp"
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
template <int MET>
struct PassesOf<HalfDist<MET, 1>> {
    static constexpr int value = 8;
};
template <int MET>
struct PassesOf<HalfDist<MET, 2>> {
    static constexpr int value = 4;
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
        int vid[PB];
        // all candidate ids first (one LDS round trip), then every row
        // gather without a branch between them; lanes past ncand re-read
        // the first candidate's row and their distances are dropped
#pragma unroll
        for (int p = 0; p < PB; p++) {
            const int ci = base + p * G + grp;
            act[p] = ci < ncand;
            vid[p] = cl[act[p] ? ci : 0];
        }
#pragma unroll
        for (int p = 0; p < PB; p++) {
            rp[p] = vec + (size_t)vid[p] * (size_t)vstride;
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

// batch_distances for rows staged in LDS: row ci is rows + slot[ci] * vstride
template <class DF>
__device__ void batch_distances_lds(const DF& df, const uint8_t* rows, int vstride, const int* slot, float* cd,
                                    int ncand, int lane) {
    constexpr int LPC = DF::LPC;
    constexpr int G = kWave / LPC;
    constexpr int PB = PassesOf<DF>::value;
    const int grp = lane / LPC, g = lane % LPC;
    for (int base = 0; base < ncand; base += G * PB) {
        typename DF::Row r[PB];
        const uint8_t* rp[PB];
        bool act[PB];
#pragma unroll
        for (int p = 0; p < PB; p++) {
            const int ci = base + p * G + grp;
            act[p] = ci < ncand;
            rp[p] = rows + (size_t)slot[act[p] ? ci : 0] * (size_t)vstride;
            df.load_row_lds(r[p], rp[p], g);
        }
#pragma unroll
        for (int p = 0; p < PB; p++) {
            float d = df.template finish<true>(r[p], g, rp[p], act[p]);
            int ci = base + p * G + grp;
            if (act[p] && g == 0) cd[ci] = d;
        }
    }
}

// Copies rows ids[0, n) of vec (vstride bytes each, ids valid rows) to the
// LDS rows dst + slot * vstride, slot = slots[j] (or slot_base + j when
// slots is null), with up to 16 16-byte loads in flight per lane.
__device__ __forceinline__ void stage_rows(uint8_t* dst, const uint8_t* vec, int vstride, const int* ids,
                                           const int* slots, int slot_base, int n, int lane) {
    const int cpr = vstride >> 4;  // chunks per row
    const int total = n * cpr;
    constexpr int U = 16;
    for (int c0 = 0; c0 < total; c0 += kWave * U) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + u * kWave + lane;
            const int r = c < total ? c / cpr : 0;
            const int k = c < total ? c - r * cpr : 0;
            v[u] = ld16_global(vec + (size_t)ids[r] * (size_t)vstride + 16 * k);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + u * kWave + lane;
            if (c < total) {
                const int r = c / cpr;
                const int slot = slots ? slots[r] : slot_base + r;
                *reinterpret_cast<uint4*>(dst + (size_t)slot * vstride + 16 * (c - r * cpr)) = v[u];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// True when some row ci < ncand is strictly closer than thr to the query of
// df: rows + slot[ci] * vstride in LDS (FROM_LDS) or vec + id[ci] * vstride
// in device memory.  Rows are compared a pass at a time (G * PB rows) and
// the scan stops at the pass holding the first such row, as the
// reference's loops break at their first occluder.
template <class DF, bool FROM_LDS>
__device__ bool any_closer(const DF& df, const uint8_t* rows, int vstride, const int* list, int ncand, float thr,
                           int lane) {
    constexpr int LPC = DF::LPC;
    constexpr int G = kWave / LPC;
    constexpr int PB = PassesOf<DF>::value;
    const int grp = lane / LPC, g = lane % LPC;
    for (int base = 0; base < ncand; base += G * PB) {
        typename DF::Row r[PB];
        const uint8_t* rp[PB];
        bool act[PB];
#pragma unroll
        for (int p = 0; p < PB; p++) {
            const int ci = base + p * G + grp;
            act[p] = ci < ncand;
            rp[p] = rows + (size_t)list[act[p] ? ci : 0] * (size_t)vstride;
            if (FROM_LDS)
                df.load_row_lds(r[p], rp[p], g);
            else
                df.load_row(r[p], rp[p], g, act[p]);
        }
        bool closer = false;
#pragma unroll
        for (int p = 0; p < PB; p++) {
            const float d = df.template finish<FROM_LDS>(r[p], g, rp[p], act[p]);
            closer |= act[p] && g == 0 && d < thr;
        }
        if (ballot(closer)) return true;
    }
    return false;
}

// One row of vstride bytes (a multiple of 16) from device memory to LDS.
__device__ __forceinline__ void stage_one_row(uint8_t* dst, const uint8_t* src, int vstride, int lane) {
    for (int c = lane; c < (vstride >> 4); c += kWave)
        *reinterpret_cast<uint4*>(dst + 16 * c) = ld16_global(src + 16 * c);
    __builtin_amdgcn_wave_barrier();
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
    int d = gbl(L.deg)[v];
    h.first = gbl(L.adj)[h.base + lane];
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
    id = b0 == 0 ? rh.first : (inrow ? gbl(L.adj)[rh.base + j] : -1);
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

// failure flags (pk_device_error): the visited table (1) or the candidate
// heap (4) ran out of room; a query's k or ef was out of range (2)
constexpr int kFailVisitedFull = 1;
constexpr int kFailArgs = 2;
constexpr int kFailHeapFull = 4;
constexpr int kFailLockTimeout = 8;  // insertion: a node lock was held for seconds
constexpr int kFailRowLimit = 16;    // insertion: a row already holds the longest supported run of ids

struct WorkCount {
    int hops = 0, neighbors = 0, dists = 0;
};

// SlideLayer: greedy descent from `cur`; returns the node it stops at.  The
// row scanned in one pass is that of the node current at the start of the
// pass, and the strictly closest new neighbour (first in row order) wins.
template <class DF>
__device__ int greedy_level(const DF& df, const LayerDev& L, int vstride, int cur, WaveVisited& vis, int* cl,
                            float* cd, int lane, int& fail, WorkCount& wc) {
    const bool leader = lane == 0;
    float cdist = one_distance(df, L, vstride, cur, cl, cd, lane);
    vis.mark(cur, leader, fail);
    bool updated;
    do {
        updated = false;
        const RowHead rh = load_row_head(L, cur, lane);
        wc.hops++;
        wc.neighbors += rh.deg;
        for (int b0 = 0; b0 < rh.deg; b0 += kWave) {
            int id;
            float d;
            uint64_t newm = expand_chunk(df, L, vstride, rh, b0, vis, cl, cd, lane, fail, id, d);
            if (newm == 0) continue;
            wc.dists += __builtin_popcountll(newm);
            float m;
            int at = wave_min_lane(d, (newm >> lane) & 1, &m);
            if (at >= 0 && m < cdist) {
                cur = readlane(id, at);
                cdist = m;
                updated = true;
            }
        }
    } while (updated && !fail);
    vis.clear(lane);
    return cur;
}

// SearchLayer: ef-bounded best-first search from `entry`.  W (the result
// max-heap) and C (the candidate min-heap) must be empty; W holds the
// result afterwards, in the reference's heap layout.
template <class DF, class WH, class CH>
__device__ void beam_level(const DF& df, const LayerDev& L, int vstride, int entry, int ef, WH& W, CH& C,
                           WaveVisited& vis, int* cl, float* cd, int lane, int& fail, WorkCount& wc) {
    const bool leader = lane == 0;
    const float d0 = one_distance(df, L, vstride, entry, cl, cd, lane);
    W.push_par(pack_entry(d0, entry), lane, leader);
    C.push_par(pack_entry(d0, entry), lane, leader);
    vis.mark(entry, leader, fail);
    uint32_t wtop_bits = __float_as_uint(d0);  // bits of W.top()
    while (C.len > 0 && !fail) {
        const uint64_t te = C.top();
        C.pop_deep(lane, leader);
        if (entry_dist(te) > __uint_as_float(wtop_bits)) break;  // hnsw.hpp:403
        const RowHead rh = load_row_head(L, entry_id(te), lane);
        wc.hops++;
        wc.neighbors += rh.deg;
        for (int b0 = 0; b0 < rh.deg && !fail; b0 += kWave) {
            int id;
            float d;
            uint64_t newm = expand_chunk(df, L, vstride, rh, b0, vis, cl, cd, lane, fail, id, d);
            if (newm == 0) continue;
            wc.dists += __builtin_popcountll(newm);
            const bool isnew = (newm >> lane) & 1;
            const float wtop = __uint_as_float(wtop_bits);
            // a full result heap's top only shrinks: rows at or past it are rejected
            uint64_t cand = W.len >= ef ? ballot(isnew && d < wtop) : newm;
            while (cand) {
                const int ln = __builtin_ctzll(cand);
                cand &= cand - 1;
                const float dv = readlanef(d, ln);
                if (W.len < ef || __uint_as_float(wtop_bits) > dv) {  // hnsw.hpp:425-431
                    const uint64_t e = pack_entry(dv, readlane(id, ln));
                    if (C.full()) {
                        fail = kFailHeapFull;
                        break;
                    }
                    // W and C are pushed with one LDS round trip for both
                    // ancestor reads.  W's new root is known without reading
                    // it back: the entry if it rose to the root, else the old
                    // root; after the eviction, what the sift-down put at
                    // the root -- and the evicted-into-root entry (the last
                    // slot's) is known from the push.
                    const auto pw = W.push_read(lane);
                    const auto pc = C.push_read(lane);
                    uint64_t last;
                    const bool rose = W.push_write(pw, e, lane, leader, &last);
                    C.push_write(pc, e, lane, leader);
                    if (W.len > ef)
                        wtop_bits = (uint32_t)W.pop_with(last, lane, leader);
                    else if (rose)
                        wtop_bits = (uint32_t)e;
                }
            }
        }
    }
    vis.clear(lane);
}

// Stages a row of `bytes` (zero padded to vstride) into LDS for use as a
// distance query.
__device__ __forceinline__ void stage_row(uint8_t* dst, const uint8_t* src, int bytes, int vstride, int lane) {
    for (int i = lane; i < vstride; i += kWave) dst[i] = i < bytes ? src[i] : 0;
    __builtin_amdgcn_wave_barrier();
}

}  // namespace pk
