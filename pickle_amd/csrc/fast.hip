// fast.hip -- batched HNSW search, fast mode: sorted beam in registers,
// several expansions per round trip, visited filter in LDS.
//
// Reference: HNSWGraph::AnnSearch (include/hnsw.hpp:509-541): SlideLayer
// (:446-494) above layer 0, SearchLayerFast (:374-444) on layer 0.  This
// mode keeps the reference's accept rule (a neighbour enters the result set
// W when |W| < ef or it is strictly closer than W's farthest entry) and its
// stopping rule (the search ends when no entry of W is left to expand), and
// changes how the work is organised:
//
//   * W is one sorted array of (distance, id, expanded) keys held in
//     registers, 64 keys per register row, RR rows (ef <= 64 * RR).  The
//     reference's separate candidate heap is not needed: a candidate that
//     is not in W is never expanded by the reference either (its distance
//     is at least W's farthest when it is popped, hnsw.hpp:403).
//   * Each round expands the `width` closest unexpanded entries of W at
//     once: their adjacency rows and the rows of every new neighbour are
//     gathered in the same memory round trips.  Width 1 expands in the
//     reference's order; a larger width trades a few extra expansions for
//     fewer dependent round trips per query.
//   * New neighbours that pass the accept rule are sorted with a wave
//     bitonic network and merged into W with bitonic merges of only the
//     tail rows the batch can reach (its smallest key picks 1, 2, 4 or all
//     rows), keeping the ef smallest; up to six accepted neighbours are
//     instead inserted one by one with a DPP row shift.  Duplicates (an id
//     the filter forgot, back while still in W) are removed in registers.
//   * The visited filter is a hash table in LDS of 16-bit entries: a
//     bijective hash of the id (an odd multiplier modulo 2^b, b the bits of
//     the index capacity) names a slot by its high bits and is stored as a
//     tag of its low bits, so a slot and its tag identify one id exactly.
//     A slot is overwritten on collision, so the filter may forget an id,
//     but never takes an id it has not seen for one it has.  A forgotten id
//     is only evaluated again: if it is still in W, the merge
//     places both copies side by side and one is dropped (keeping the
//     expanded mark); if it left W, its distance is at or above W's
//     farthest and the accept rule rejects it.  Results therefore do not
//     depend on the table size, only the work does (on data without ties;
//     with ties, which of two equal-distance ids stays can follow the
//     grouping of the merges, never a returned distance).  No visited state
//     goes to device memory.
//
// Ties between equal distances are ordered by id here, by heap layout in
// the reference; without ties, width 1 returns the reference's results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pickle_gpu.h"
#include "index_impl.hpp"
#include "traverse.hpp"

namespace pk {

constexpr uint64_t kKeyInf = ~0ull;
constexpr int kFastCap = 256;  // most new neighbours gathered before a flush (distances + merge)
constexpr int kInsertMax = 6;  // accepted neighbours inserted one by one rather than sorted and merged

struct FastArgs {
    const LayerDev* layers;
    const int32_t* ext0;
    int32_t enter_level;
    int32_t dim;
    int32_t vstride;
    int32_t qbytes;
    const uint8_t* queries;
    int64_t nq;
    const int32_t* k;
    const int32_t* ef;
    int32_t k_all, ef_all, k_stride;
    int32_t* out_ids;
    float* out_dists;
    int32_t* out_count;
    pk_query_counters* ctr;
    int32_t width;      // entries expanded per round (1..4)
    int32_t slot_bits;  // log2 entries of the LDS visited filter per wave
    int32_t tag_bits;   // id bits kept in an entry (<= 15), b - slot_bits
    uint32_t id_mask;   // 2^b - 1
    int32_t lds_per_wave;
    int32_t list_cap;    // new neighbours gathered before a flush (a multiple of 64, <= kFastCap)
    int32_t max_rounds;  // per query; past it the query fails (flag 1)
    int32_t* qcounter;
    int32_t* err;
    const int32_t* order;  // queue position -> query (null: in order)
};

// ---- keys: order-preserving distance bits, then id, then the pending bit

__device__ __forceinline__ uint32_t dist_key(float d) {
    const uint32_t u = __float_as_uint(d);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_dist(uint64_t k) {
    const uint32_t ku = (uint32_t)(k >> 32);
    return __uint_as_float((ku & 0x80000000u) ? (ku & 0x7fffffffu) : ~ku);
}
// bit 0 set: not expanded yet.  An expanded copy of an id therefore sorts
// before a pending copy of it, and keeping the first of adjacent copies
// keeps the expanded one.
__device__ __forceinline__ uint64_t make_key(float d, int id) {
    return ((uint64_t)dist_key(d) << 32) | ((uint64_t)(uint32_t)id << 1) | 1ull;
}
__device__ __forceinline__ bool pending(uint64_t k) { return (k & 1) && k != kKeyInf; }
__device__ __forceinline__ int key_id(uint64_t k) { return (int)((uint32_t)k >> 1); }

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const int lo = __shfl((int)(uint32_t)v, src, kWave);
    const int hi = __shfl((int)(uint32_t)(v >> 32), src, kWave);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m, kWave);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}
// lane i gets lane i - 1's key, lane 0 gets `first`: a DPP wave shift
// right by one (wave_shr:1, a VALU move on CDNA), the vacated lane 0 keeping
// the old value supplied
__device__ __forceinline__ uint64_t shift_up1(uint64_t v, uint64_t first) {
    constexpr int kWaveShr1 = 0x138;
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)first, (int)(uint32_t)v, kWaveShr1, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(first >> 32), (int)(uint32_t)(v >> 32), kWaveShr1, 0xf,
                                               0xf, false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// lane i gets lane i + 1's key, lane 63 gets `last`: a DPP wave shift
// left by one (wave_shl:1), the vacated lane 63 keeping the value supplied
__device__ __forceinline__ uint64_t shift_down1(uint64_t v, uint64_t last) {
    constexpr int kWaveShl1 = 0x130;
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)last, (int)(uint32_t)v, kWaveShl1, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(last >> 32), (int)(uint32_t)(v >> 32), kWaveShl1, 0xf,
                                               0xf, false);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

// lane l's key, uniform (v_readlane reads any lane, whatever its exec bit)
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}
__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t max64(uint64_t a, uint64_t b) { return a < b ? b : a; }

// ascending sort of one key per lane
__device__ __forceinline__ uint64_t bitonic_sort64(uint64_t x, int lane) {
#pragma unroll
    for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint64_t y = shfl_xor64(x, j);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            x = (lower == up) ? min64(x, y) : max64(x, y);
        }
    }
    return x;
}

// a bitonic sequence of one key per lane, sorted (ascending or descending)
__device__ __forceinline__ uint64_t bitonic_merge64(uint64_t x, int lane, bool ascending) {
#pragma unroll
    for (int j = kWave >> 1; j > 0; j >>= 1) {
        const uint64_t y = shfl_xor64(x, j);
        const bool lower = (lane & j) == 0;
        x = (lower == ascending) ? min64(x, y) : max64(x, y);
    }
    return x;
}

// The sorted beam: position p = 64 * r + lane lives in row r.
template <int RR>
struct Beam {
    uint64_t key[RR];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int r = 0; r < RR; r++) key[r] = kKeyInf;
    }

    // W <- the 64 * RR smallest keys of W and the sorted batch s (one key
    // per lane), when every key of s is above the keys before the last SUB
    // rows: those rows and s hold W's tail, so only they merge -- the
    // smallest 64 of the last row and s, sorted, then one bitonic merge of
    // the SUB rows.  Returns the 64 keys that fall off the end, sorted: the
    // merged sequence continues with them, so a duplicate dropped from W is
    // replaced from there.
    template <int SUB>
    __device__ __forceinline__ uint64_t merge_tail(uint64_t s, int lane) {
        constexpr int R0 = RR - SUB;
        const uint64_t srev = shfl64(s, kWave - 1 - lane);
        if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
        const uint64_t last = key[RR - 1];
        const uint64_t t = min64(last, srev);
        const uint64_t over = bitonic_merge64(max64(last, srev), lane, true);
        if (SUB == 1) {
            key[RR - 1] = bitonic_merge64(t, lane, true);
            return over;
        }
        key[RR - 1] = bitonic_merge64(t, lane, false);  // rows ascending, then this row descending
#pragma unroll
        for (int J = SUB / 2; J >= 1; J >>= 1)
#pragma unroll
            for (int r = R0; r < RR; r++)
                if (((r - R0) & J) == 0) {
                    const uint64_t a = key[r], b = key[r + J];
                    key[r] = min64(a, b);
                    key[r + J] = max64(a, b);
                    if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
                }
        // one row at a time: interleaving the rows' exchanges raises register
        // pressure enough to cost occupancy with eight rows
#pragma unroll
        for (int r = R0; r < RR; r++) {
            key[r] = bitonic_merge64(key[r], lane, true);
            if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
        }
        return over;
    }

    // Over the last SUB rows of W: keeps the first of adjacent copies of one
    // (distance, id) -- the expanded one if any, since it sorts first.  Each
    // dropped copy moves the keys after it one place toward the front, in
    // registers (a DPP shift per row), and the overflow row -- the sorted
    // continuation from merge_tail -- refills W's last position.  Copies
    // appear only when an id the visited filter forgot came back while in
    // W, so the loop usually ends at its first scan.  Returns the copies
    // dropped.
    template <int SUB>
    __device__ int dedup_tail(uint64_t over, int lane) {
        constexpr int R0 = RR - SUB;
        const uint64_t before_tail = R0 > 0 ? readlane64(key[R0 > 0 ? R0 - 1 : 0], kWave - 1) : kKeyInf;
        int dropped = 0;
        for (;;) {
            // the first copy among the tail rows: row prow, lane plane
            int prow = -1, plane = 0;
            uint64_t carry = before_tail;
#pragma unroll
            for (int r = 0; r < SUB; r++) {
                if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
                const uint64_t v = key[R0 + r];
                const uint64_t prev = shift_up1(v, carry);
                const bool dup = v != kKeyInf && prev != kKeyInf && (v >> 1) == (prev >> 1);
                const uint64_t m = ballot(dup);
                if (m != 0 && prow < 0) {
                    prow = r;
                    plane = (int)__builtin_ctzll(m);
                }
                carry = readlane64(v, kWave - 1);
            }
            if (prow < 0) return dropped;
#pragma unroll
            for (int r = 0; r < SUB; r++) {
                if (r < prow) continue;
                if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
                const uint64_t next = r + 1 < SUB ? readlane64(key[R0 + (r + 1 < SUB ? r + 1 : r)], 0)
                                                  : readlane64(over, 0);
                const uint64_t sh = shift_down1(key[R0 + r], next);
                key[R0 + r] = (r == prow && lane < plane) ? key[R0 + r] : sh;
            }
            over = shift_down1(over, kKeyInf);
            dropped++;
        }
    }

    // merge the sorted batch s into W, touching only the tail rows it can
    // reach, then drop duplicates
    __device__ __forceinline__ void merge_dedup(uint64_t s, int lane) {
        const uint64_t smin = readlane64(s, 0);
        // the last key before the last 1, 2, 4 rows
        const uint64_t b1 = RR >= 2 ? readlane64(key[RR >= 2 ? RR - 2 : 0], kWave - 1) : kKeyInf;
        const uint64_t b2 = RR >= 4 ? readlane64(key[RR >= 4 ? RR - 3 : 0], kWave - 1) : kKeyInf;
        const uint64_t b4 = RR >= 8 ? readlane64(key[RR >= 8 ? RR - 5 : 0], kWave - 1) : kKeyInf;
        if (RR >= 2 && smin > b1) {
            dedup_tail<1>(merge_tail<1>(s, lane), lane);
        } else if constexpr (RR >= 4) {
            if (smin > b2) {
                dedup_tail<RR >= 4 ? 2 : 1>(merge_tail<RR >= 4 ? 2 : 1>(s, lane), lane);
            } else if constexpr (RR >= 8) {
                if (smin > b4)
                    dedup_tail<RR >= 8 ? 4 : 1>(merge_tail<RR >= 8 ? 4 : 1>(s, lane), lane);
                else
                    dedup_tail<RR>(merge_tail<RR>(s, lane), lane);
            } else {
                dedup_tail<RR>(merge_tail<RR>(s, lane), lane);
            }
        } else {
            dedup_tail<RR>(merge_tail<RR>(s, lane), lane);
        }
    }

    // Inserts one pending key c at its sorted position, moving the keys
    // after it one place on (the key at position ef - 1 leaves), unless the
    // beam holds the same (distance, id) already or c sorts at or past ef.
    __device__ __forceinline__ void insert1(uint64_t c, int ef, int lane) {
        int pos = 0;
        bool dup = false;
#pragma unroll
        for (int r = 0; r < RR; r++) {
            pos += __builtin_popcountll(ballot(key[r] < c));
            dup |= ballot(key[r] != kKeyInf && (key[r] >> 1) == (c >> 1)) != 0;
        }
        if (dup || pos >= ef) return;
        const int prow = pos >> 6;  // rows before this one do not move
        uint64_t carry = kKeyInf;   // last key of the row before
#pragma unroll
        for (int r = 0; r < RR; r++) {
            if (r < prow) continue;
            const uint64_t last = readlane64(key[r], kWave - 1);
            const uint64_t up = shift_up1(key[r], carry);
            const int p = 64 * r + lane;
            key[r] = p < pos ? key[r] : (p == pos ? c : up);
            carry = last;
            if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
        }
    }

    // positions >= ef leave the beam
    __device__ __forceinline__ void cap(int ef, int lane) {
#pragma unroll
        for (int r = 0; r < RR; r++)
            if (64 * r + lane >= ef) key[r] = kKeyInf;
    }

    __device__ __forceinline__ int count() const {
        int n = 0;
#pragma unroll
        for (int r = 0; r < RR; r++) n += __builtin_popcountll(ballot(key[r] != kKeyInf));
        return n;
    }

    __device__ __forceinline__ uint64_t at(int p) const {
        uint64_t v = kKeyInf;
#pragma unroll
        for (int r = 0; r < RR; r++)
            if (r == (p >> 6)) v = readlane64(key[r], p & 63);
        return v;
    }
};

// LDS visited filter: h = id * odd mod 2^b is a bijection of the ids below
// 2^b; its high bits pick the slot and its low bits, with the top bit set
// (an empty slot holds 0), are the entry.  Test and set: a lane reads its
// slot, then writes its entry.  Two lanes of one instruction with the same
// id both report it new (it is evaluated twice and the beam's merge drops
// the copy); a collision overwrites the older entry.
__device__ __forceinline__ bool filter_new(uint16_t* tab, uint32_t id_mask, int tag_bits, int id) {
    const uint32_t h = ((uint32_t)id * 0x9E3779B1u) & id_mask;
    const uint32_t slot = h >> tag_bits;
    const uint16_t tag = (uint16_t)((h & ((1u << tag_bits) - 1u)) | 0x8000u);
    const uint16_t old = tab[slot];
    tab[slot] = tag;
    return old != tag;
}

// SlideLayer without a visited set (a node seen before cannot be strictly
// closer than the current one): move to the strictly closest neighbour,
// first in row order, until none is closer.
template <class DF>
__device__ int greedy_fast(const DF& df, const LayerDev& L, int vstride, int cur, int* cl, float* cd, int lane,
                           WorkCount& wc) {
    float cdist = one_distance(df, L, vstride, cur, cl, cd, lane);
    wc.dists++;
    for (int guard = 0; guard < (1 << 20); guard++) {
        const RowHead rh = load_row_head(L, cur, lane);
        wc.hops++;
        wc.neighbors += rh.deg;
        int best = cur;
        float bestd = cdist;
        for (int b0 = 0; b0 < rh.deg; b0 += kWave) {
            const int j = b0 + lane;
            const bool inrow = j < rh.deg;
            const int id = b0 == 0 ? rh.first : (inrow ? gbl(L.adj)[rh.base + j] : -1);
            const bool valid = inrow && id >= 0 && (int64_t)id < L.n;
            const uint64_t vm = ballot(valid);
            const int pos = lanes_below(vm);
            if (valid) cl[pos] = id;
            __builtin_amdgcn_wave_barrier();
            const int nc = __builtin_popcountll(vm);
            batch_distances(df, L.vec, vstride, cl, cd, nc, lane);
            __builtin_amdgcn_wave_barrier();
            wc.dists += nc;
            float m;
            const float d = lane < nc ? cd[lane] : __builtin_inff();
            const int at = wave_min_lane(d, lane < nc, &m);
            if (at >= 0 && m < bestd) {
                bestd = m;
                best = readlane(lane < nc ? cl[lane] : 0, at);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (!(bestd < cdist)) break;
        cur = best;
        cdist = bestd;
    }
    return cur;
}

// Distances of the gathered new neighbours cl[0, n), then the accepted
// ones merged into the beam, 64 at a time.
template <class DF, int RR>
__device__ void flush(const DF& df, const LayerDev& L, int vstride, int* cl, float* cd, int n, Beam<RR>& W, int ef,
                      int& len, int lane) {
    batch_distances(df, L.vec, vstride, cl, cd, n, lane);
    __builtin_amdgcn_wave_barrier();
    for (int c0 = 0; c0 < n; c0 += kWave) {
        const int c = c0 + lane;
        uint64_t key = c < n ? make_key(cd[c], cl[c]) : kKeyInf;
        if (len >= ef) {
            const uint32_t worst = (uint32_t)(W.at(ef - 1) >> 32);  // hnsw.hpp:425: top > dist
            if (!((uint32_t)(key >> 32) < worst)) key = kKeyInf;
        }
        uint64_t pass = ballot(key != kKeyInf);
        if (pass == 0) continue;
        if (__builtin_popcountll(pass) <= kInsertMax) {
            // few accepted: one at a time, each against the beam as it stands
            while (pass) {
                const int ln = __builtin_ctzll(pass);
                pass &= pass - 1;
                const uint64_t kc = readlane64(key, ln);
                if (len >= ef && !((uint32_t)(kc >> 32) < (uint32_t)(W.at(ef - 1) >> 32))) continue;
                W.insert1(kc, ef, lane);
                W.cap(ef, lane);  // the key shifted to position ef leaves
                len = W.count();
            }
            continue;
        }
        const uint64_t sorted = bitonic_sort64(key, lane);
        if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
        W.merge_dedup(sorted, lane);
        if constexpr (RR >= 8) __builtin_amdgcn_sched_barrier(0);
        W.cap(ef, lane);
        len = W.count();
    }
    __builtin_amdgcn_wave_barrier();
}

template <class DF, bool IS_INT, int RR>
__global__ __launch_bounds__(256, 4) void k_search_fast(FastArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    uint8_t* my = smem + (size_t)w * A.lds_per_wave;
    uint16_t* vtab = reinterpret_cast<uint16_t*>(my);
    const int nslot = 1 << A.slot_bits;
    int* cl = reinterpret_cast<int*>(vtab + (nslot < 8 ? 8 : nslot));
    float* cd = reinterpret_cast<float*>(cl + A.list_cap);
    uint8_t* qst = reinterpret_cast<uint8_t*>(cd + A.list_cap);
    const bool leader = lane == 0;
    const int g = lane % DF::LPC;

    for (;;) {
        int qi = 0;
        if (leader) qi = atomicAdd(A.qcounter, 1);
        const int64_t pos = readlane(qi, 0);
        if (pos >= A.nq) break;
        const int64_t q = A.order ? (int64_t)uni(gbl(A.order)[pos]) : pos;
        const uint64_t t_begin = wall_clock64();
        stage_row(qst, A.queries + (size_t)q * (size_t)A.qbytes, A.qbytes, A.vstride, lane);
        DF df;
        df.load_query(qst, g, IS_INT ? (A.vstride >> 4) : A.dim);
        const int k = uni(A.k ? A.k[q] : A.k_all);
        const int ef = uni(A.ef ? A.ef[q] : A.ef_all);
        int fail = 0;
        if (k < 1 || k > A.k_stride || ef < 1 || ef > 64 * RR) fail = kFailArgs;

        WorkCount up, l0;
        int cur = 0;  // m_ent_id of the top layer (include/hnsw_layer.hpp:70, 186-188)
        for (int level = A.enter_level; level >= 1 && !fail; level--) {
            const LayerDev L = A.layers[level];
            cur = greedy_fast(df, L, A.vstride, cur, cl, cd, lane, up);
            cur = gbl(L.down)[cur];
        }

        Beam<RR> W;
        W.clear();
        int len = 0;
        if (!fail) {
            // an empty filter for this query
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            for (int i = lane; i < (nslot + 7) / 8; i += kWave)
                reinterpret_cast<u32x4*>(vtab)[i] = u32x4{0u, 0u, 0u, 0u};
            __builtin_amdgcn_wave_barrier();
            const LayerDev L = A.layers[0];
            const float d0 = one_distance(df, L, A.vstride, cur, cl, cd, lane);
            l0.dists++;
            if (leader) {
                W.key[0] = make_key(d0, cur);
                filter_new(vtab, A.id_mask, A.tag_bits, cur);
            }
            len = 1;
            __builtin_amdgcn_wave_barrier();
            int rounds = 0;
            for (;;) {
                // the `width` closest unexpanded entries, marked expanded
                int sel[4];
                int nsel = 0;
#pragma unroll
                for (int r = 0; r < RR; r++) {
                    uint64_t m = ballot(pending(W.key[r]));
                    while (m && nsel < A.width) {
                        const int ln = __builtin_ctzll(m);
                        m &= m - 1;
                        const uint64_t kk = readlane64(W.key[r], ln);
                        if (lane == ln) W.key[r] &= ~1ull;
                        sel[nsel == 0 ? 0 : nsel == 1 ? 1 : nsel == 2 ? 2 : 3] = key_id(kk);
                        nsel++;
                    }
                }
                if (nsel == 0) break;
                if (++rounds > A.max_rounds) {
                    fail = kFailVisitedFull;
                    break;
                }
                l0.hops += nsel;
                // adjacency rows of the selected entries, new ids gathered
                int n = 0;
                for (int s = 0; s < nsel; s++) {
                    const int v = s == 0 ? sel[0] : s == 1 ? sel[1] : s == 2 ? sel[2] : sel[3];
                    const RowHead rh = load_row_head(L, v, lane);
                    l0.neighbors += rh.deg;
                    for (int b0 = 0; b0 < rh.deg; b0 += kWave) {
                        if (n + kWave > A.list_cap) {
                            flush(df, L, A.vstride, cl, cd, n, W, ef, len, lane);
                            l0.dists += n;
                            n = 0;
                        }
                        const int j = b0 + lane;
                        const bool inrow = j < rh.deg;
                        const int id = b0 == 0 ? rh.first : (inrow ? gbl(L.adj)[rh.base + j] : -1);
                        const bool valid = inrow && id >= 0 && (int64_t)id < L.n;
                        const bool isnew = valid && filter_new(vtab, A.id_mask, A.tag_bits, id);
                        const uint64_t nm = ballot(isnew);
                        if (isnew) cl[n + lanes_below(nm)] = id;
                        n += __builtin_popcountll(nm);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                if (n > 0) {
                    flush(df, L, A.vstride, cl, cd, n, W, ef, len, lane);
                    l0.dists += n;
                }
            }
        }

        // results: the k closest, ascending, external ids
        int32_t* oid = A.out_ids + (size_t)q * A.k_stride;
        float* odist = A.out_dists + (size_t)q * A.k_stride;
        const int kk = fail ? 0 : (k < len ? k : len);
#pragma unroll
        for (int r = 0; r < RR; r++) {
            const int p = 64 * r + lane;
            if (p < kk) {
                oid[p] = A.ext0[key_id(W.key[r])];
                odist[p] = key_dist(W.key[r]);
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

__global__ void k_fill_i32_fast(int* p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class DF, bool INT>
static hipError_t launch_rr(const FastArgs& a, int rr, int grid, int block, size_t lds, hipStream_t s) {
    switch (rr) {
        case 1: hipLaunchKernelGGL((k_search_fast<DF, INT, 1>), dim3(grid), dim3(block), lds, s, a); break;
        case 2: hipLaunchKernelGGL((k_search_fast<DF, INT, 2>), dim3(grid), dim3(block), lds, s, a); break;
        case 4: hipLaunchKernelGGL((k_search_fast<DF, INT, 4>), dim3(grid), dim3(block), lds, s, a); break;
        default: hipLaunchKernelGGL((k_search_fast<DF, INT, 8>), dim3(grid), dim3(block), lds, s, a); break;
    }
    return hipGetLastError();
}

template <class T>
struct TypeTag {
    using type = T;
};

// Calls f(TypeTag<DF>, std::bool_constant<INT>) with the distance functor
// the fast mode uses for this element type, metric and row size class.
template <int DT, int MET, class F>
static hipError_t with_fast_df(int size_class, F&& f) {
    constexpr bool INT = DT == DT_U8 || DT == DT_I8;
    if constexpr (INT) {
        if (size_class == 0) return f(TypeTag<IntDist<DT, MET, 1>>{}, std::true_type{});
        if (size_class == 1) return f(TypeTag<IntDist<DT, MET, 2>>{}, std::true_type{});
        return f(TypeTag<IntDist<DT, MET, 4>>{}, std::true_type{});
    } else {
        if constexpr (DT == DT_F16) {
            if (size_class == kHalfFastClass) return f(TypeTag<HalfDist<MET, 1>>{}, std::true_type{});
            if (size_class == kHalfFastClass + 1) return f(TypeTag<HalfDist<MET, 2>>{}, std::true_type{});
        }
        if (size_class == 0) return f(TypeTag<FloatDist<DT, MET, 8>>{}, std::false_type{});
        return f(TypeTag<FloatDist<DT, MET, 16>>{}, std::false_type{});
    }
}

// with_fast_df for a run-time element type and metric
template <class F>
static hipError_t with_fast_df(int dtype, int metric, int size_class, F&& f) {
    switch (dtype * 4 + metric) {
        case DT_U8 * 4 + MET_L2: return with_fast_df<DT_U8, MET_L2>(size_class, f);
        case DT_U8 * 4 + MET_IP: return with_fast_df<DT_U8, MET_IP>(size_class, f);
        case DT_I8 * 4 + MET_L2: return with_fast_df<DT_I8, MET_L2>(size_class, f);
        case DT_I8 * 4 + MET_IP: return with_fast_df<DT_I8, MET_IP>(size_class, f);
        case DT_F32 * 4 + MET_L2: return with_fast_df<DT_F32, MET_L2>(size_class, f);
        case DT_F32 * 4 + MET_IP: return with_fast_df<DT_F32, MET_IP>(size_class, f);
        case DT_F16 * 4 + MET_L2: return with_fast_df<DT_F16, MET_L2>(size_class, f);
        case DT_F16 * 4 + MET_IP: return with_fast_df<DT_F16, MET_IP>(size_class, f);
        default: return hipErrorInvalidValue;
    }
}

static hipError_t launch_fast(const FastArgs& a, int dtype, int metric, int size_class, int rr, int grid, int block,
                              size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_i32_fast, dim3(1), dim3(64), 0, s, a.qcounter, 0, 1);
    hipLaunchKernelGGL(k_fill_i32_fast, dim3(1), dim3(64), 0, s, a.err, 0, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return with_fast_df(dtype, metric, size_class, [&](auto tag, auto in) {
        return launch_rr<typename decltype(tag)::type, decltype(in)::value>(a, rr, grid, block, lds, s);
    });
}

// ---- query order -----------------------------------------------------------
//
// Queries of one region of the data visit overlapping rows, and rows read by
// one query are still in the L2 or the MI355X's 256 MB Infinity Cache when
// another query of the region asks for them.  For a large batch the
// queries are therefore handed out region by region: the pivots are the
// first nodes of an upper layer (a random sample of the rows, as level
// sampling is), about one per kOrderQueriesPerKey queries and at most
// kOrderMaxKeys; a query's key is its nearest pivot, and a counting sort of
// the keys gives the order.  Results do not depend on the order.
constexpr int64_t kOrderMinQueries = 8192;
constexpr int64_t kOrderQueriesPerKey = 1536;
constexpr int64_t kOrderMaxKeys = 64;
// k_order_keys holds the pivot ids and distances in one wave-wide batch
// (cl / cd, kWave entries each): more pivots need a loop over batches
static_assert(kOrderMaxKeys <= kWave, "the pivot scan takes at most one wave-wide batch");

struct OrderArgs {
    const LayerDev* layers;
    int32_t piv_level;  // the pivots are nodes [0, npiv) of this layer
    int32_t npiv;
    int32_t dim;
    int32_t vstride;
    int32_t qbytes;
    int32_t lds_per_wave;
    const uint8_t* queries;
    int64_t nq;
    int32_t* keys;
    int32_t* ranks;
    int32_t* hist;
};

template <class DF, bool IS_INT>
__global__ __launch_bounds__(256) void k_order_keys(OrderArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    uint8_t* my = smem + (size_t)w * A.lds_per_wave;
    int* cl = reinterpret_cast<int*>(my);
    float* cd = reinterpret_cast<float*>(cl + kWave);
    uint8_t* qst = reinterpret_cast<uint8_t*>(cd + kWave);
    const int g = lane % DF::LPC;
    for (int64_t q = (int64_t)blockIdx.x * 4 + w; q < A.nq; q += (int64_t)gridDim.x * 4) {
        stage_row(qst, A.queries + (size_t)q * (size_t)A.qbytes, A.qbytes, A.vstride, lane);
        DF df;
        df.load_query(qst, g, IS_INT ? (A.vstride >> 4) : A.dim);
        // the nearest pivot (npiv <= 64: one batch of distances)
        if (lane < A.npiv) cl[lane] = lane;
        __builtin_amdgcn_wave_barrier();
        batch_distances(df, A.layers[A.piv_level].vec, A.vstride, cl, cd, A.npiv, lane);
        __builtin_amdgcn_wave_barrier();
        float m;
        const int at = wave_min_lane(lane < A.npiv ? cd[lane] : __builtin_inff(), lane < A.npiv, &m);
        const int key = at >= 0 ? at : 0;
        if (lane == 0) {
            A.keys[q] = key;
            A.ranks[q] = atomicAdd(&A.hist[key], 1);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// hist[0, n) -> exclusive prefix sums, in one workgroup
__global__ __launch_bounds__(1024) void k_order_scan(int32_t* hist, int64_t n) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = t * per < n ? t * per : n;
    const int64_t e = b + per < n ? b + per : n;
    int64_t sum = 0;
    for (int64_t i = b; i < e; i++) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int64_t i = b; i < e; i++) {
        const int32_t c = hist[i];
        hist[i] = (int32_t)run;
        run += c;
    }
}

__global__ void k_order_scatter(const int32_t* __restrict__ keys, const int32_t* __restrict__ ranks,
                                const int32_t* __restrict__ offsets, int32_t* __restrict__ order, int64_t nq) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < nq) order[offsets[keys[q]] + ranks[q]] = (int32_t)q;
}

// rows of the beam for an ef: 64, 128, 256 or 512 keys
static int beam_rows(int ef_max) { return ef_max <= 64 ? 1 : ef_max <= 128 ? 2 : ef_max <= 256 ? 4 : 8; }

static hipError_t fast_occupancy(int dtype, int rr, int block, size_t lds, int* blocks_per_cu) {
    const void* fn = nullptr;
    const bool in = dtype == DT_U8 || dtype == DT_I8;
#define PK_FN(R) \
    (in ? reinterpret_cast<const void*>(&k_search_fast<IntDist<DT_U8, MET_L2, 1>, true, R>) \
        : reinterpret_cast<const void*>(&k_search_fast<FloatDist<DT_F32, MET_L2, 8>, false, R>))
    switch (rr) {
        case 1: fn = PK_FN(1); break;
        case 2: fn = PK_FN(2); break;
        case 4: fn = PK_FN(4); break;
        default: fn = PK_FN(8); break;
    }
#undef PK_FN
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, block, lds);
}

}  // namespace pk

using namespace pk;

namespace pk {

// The order for a batch of nq queries (see the comment above
// kOrderMinQueries), enqueued on c's stream, or null when the batch is too
// small, no upper layer gives enough queries per key, or the stream is
// being captured (the buffers may need to grow).  Used by the fast and the
// exact search; size_class is the fast mode's (search_class_of).
int order_queries(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, int size_class,
                  const int32_t** out) {
    *out = nullptr;
    if (nq < kOrderMinQueries) return PK_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c.stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return PK_OK;
    const int64_t nkeys = std::min<int64_t>(kOrderMaxKeys, std::max<int64_t>(2, nq / kOrderQueriesPerKey));
    // the highest upper layer with that many nodes
    const int enter = h->enter_level;
    int lp = 0;
    for (int l = std::min(enter, PK_MAX_LEVELS - 1); l >= 1; l--)
        if (h->layers[l].n >= nkeys) {
            lp = l;
            break;
        }
    if (lp == 0) return PK_OK;
    const size_t need = (size_t)3 * (size_t)nq + (size_t)nkeys;
    if (c.order_ints < need) {
        if (c.order) {
            // an earlier launch on this stream may still read the old buffer
            if (hipStreamSynchronize(c.stream) != hipSuccess) return set_error(PK_ERR_HIP, "query order: stream sync");
            (void)hipFree(c.order);
            c.order = nullptr;
            c.order_ints = 0;
        }
        if (hipMalloc(reinterpret_cast<void**>(&c.order), need * sizeof(int32_t)) != hipSuccess) {
            c.order = nullptr;
            return set_error(PK_ERR_NOMEM, "query order buffers");
        }
        c.order_ints = need;
    }
    int32_t* keys = c.order;
    int32_t* ranks = keys + nq;
    int32_t* order = ranks + nq;
    int32_t* hist = order + nq;
    OrderArgs a{};
    a.layers = static_cast<const LayerDev*>(h->d_layers);
    a.piv_level = lp;
    a.npiv = (int32_t)nkeys;
    a.dim = h->dim;
    a.vstride = h->vstride;
    a.qbytes = h->qbytes;
    const size_t per_wave = (size_t)kWave * 8 + ((size_t)h->vstride + 15) / 16 * 16;
    a.lds_per_wave = (int32_t)per_wave;
    a.queries = static_cast<const uint8_t*>(d_queries);
    a.nq = nq;
    a.keys = keys;
    a.ranks = ranks;
    a.hist = hist;
    const int grid = (int)std::min<int64_t>((nq + 3) / 4, (int64_t)h->num_cus * 8);
    hipLaunchKernelGGL(k_fill_i32_fast, dim3((unsigned)((nkeys + 255) / 256)), dim3(256), 0, c.stream, hist, 0,
                       (int)nkeys);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = with_fast_df(h->dtype, h->metric, size_class, [&](auto tag, auto in) {
            hipLaunchKernelGGL((k_order_keys<typename decltype(tag)::type, decltype(in)::value>), dim3(grid),
                               dim3(256), per_wave * 4, c.stream, a);
            return hipGetLastError();
        });
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, c.stream, hist, nkeys);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_order_scatter, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, c.stream, keys, ranks,
                           hist, order, nq);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return set_error(PK_ERR_HIP, std::string("query order launch: ") + hipGetErrorString(e));
    *out = order;
    return PK_OK;
}

int fast_search_on(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, const int32_t* d_k,
                   const int32_t* d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride,
                   int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    if (!h->ready) return set_error(PK_ERR_STATE, "index not loaded (call pk_index_set_enter_level)");
    if (nq < 0 || (nq > 0 && (!d_queries || !d_out_ids || !d_out_dists))) return set_error(PK_ERR_ARG, "bad buffers");
    if (k_stride < 1) return set_error(PK_ERR_ARG, "k_stride must be positive");
    if (!d_k && (k_all < 1 || k_all > k_stride)) return set_error(PK_ERR_ARG, "k must be in [1, k_stride]");
    if (!d_ef && (ef_all < 1 || ef_all > ef_max)) return set_error(PK_ERR_ARG, "ef must be in [1, ef_max]");
    if (ef_max < 1 || ef_max > 512) return set_error(PK_ERR_ARG, "fast mode takes ef_max in [1, 512]");
    if (nq > (int64_t)1 << 27) return set_error(PK_ERR_CAPACITY, "at most 2^27 queries per call");
    if (nq == 0) return PK_OK;
    int rc = ensure_ctx_counters(c);
    if (rc) return rc;
    const FastTuning& t = h->fast;
    const int rr = beam_rows(ef_max);
    const int width = t.width > 0 ? t.width : 2;
    const int hb = t.hash_bits > 0 ? t.hash_bits : 11;
    const int wpg = t.waves_per_group > 0 ? t.waves_per_group : 2;
    // the filter takes 4 << hb bytes: 2^(hb + 1) entries of 16 bits, each
    // keeping b - (hb + 1) <= 15 bits of the hashed id (more entries when the
    // capacity needs it); with no more ids than entries every id has its own
    // the hash is a bijection of the ids below 2^b only: indexes past 2^27
    // ids are refused, as the exact search refuses them
    if (h->capacity > ((int64_t)1 << 27))
        return set_error(PK_ERR_CAPACITY, "fast search supports at most 2^27 ids per index");
    int b = 1;
    while (b < 27 && ((int64_t)1 << b) < h->capacity) b++;
    int sb = std::max(hb + 1, b - 15);
    if (b < sb) b = sb;
    // the candidate list holds one round's new neighbours when the rows are
    // within D: width x D slots rounded to whole chunks of 64 (longer rows
    // flush part-way through a round)
    const int D0 = h->layers[0].D > 0 ? h->layers[0].D : 2 * h->M;
    const int list_cap = std::min(kFastCap, std::max(128, width * ((D0 + 63) / 64) * 64));
    size_t per_wave = ((size_t)2 << std::max(sb, 3)) + (size_t)list_cap * 8 + h->vstride;
    per_wave = (per_wave + 15) / 16 * 16;
    const size_t lds = per_wave * wpg;
    if (lds > 160 * 1024) return set_error(PK_ERR_CAPACITY, "fast search needs too much LDS per workgroup");
    int bpc = 0;
    if (fast_occupancy(h->dtype, rr, 64 * wpg, lds, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    if (t.blocks_per_cu > 0) bpc = std::min(bpc, t.blocks_per_cu);
    const int64_t want = (nq + wpg - 1) / wpg;
    const int grid = (int)std::min<int64_t>(want, (int64_t)h->num_cus * bpc);
    FastArgs a{};
    a.layers = static_cast<const LayerDev*>(h->d_layers);
    a.ext0 = h->layers[0].ext;
    a.enter_level = h->enter_level;
    a.dim = h->dim;
    a.vstride = h->vstride;
    a.qbytes = h->qbytes;
    a.queries = static_cast<const uint8_t*>(d_queries);
    a.nq = nq;
    a.k = d_k;
    a.ef = d_ef;
    a.k_all = k_all;
    a.ef_all = ef_all;
    a.k_stride = k_stride;
    a.out_ids = d_out_ids;
    a.out_dists = d_out_dists;
    a.out_count = d_out_count;
    a.ctr = d_counters;
    a.width = width;
    a.slot_bits = sb;
    a.tag_bits = b - sb;
    a.id_mask = b >= 32 ? 0xffffffffu : ((1u << b) - 1u);
    a.lds_per_wave = (int)per_wave;
    a.list_cap = list_cap;
    a.max_rounds = (int)std::min<int64_t>(1 << 30, 4 * h->capacity + 64);
    a.qcounter = c.qcounter;
    a.err = c.err;
    const int size_class = search_class_of(h);
    if ((rc = order_queries(h, c, d_queries, nq, size_class, &a.order))) return rc;
    const hipError_t e = launch_fast(a, h->dtype, h->metric, size_class, rr, grid, 64 * wpg, lds, c.stream);
    if (e != hipSuccess) return set_error(PK_ERR_HIP, std::string("fast search launch: ") + hipGetErrorString(e));
    if (&c == &h->sc) {
        h->last_grid = grid;
        h->last_lds = lds;
    }
    return PK_OK;
}
}  // namespace pk

extern "C" {

int pk_set_fast_tuning(pk_index* h, int width, int hash_bits, int waves_per_group, int blocks_per_cu) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    if (h->servers_alive) return set_error(PK_ERR_STATE, "destroy the serving loops before changing the tuning");
    if (width < 0 || width > 4 || (hash_bits != 0 && (hash_bits < 8 || hash_bits > 14)) ||
        (waves_per_group != 0 && waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4) ||
        blocks_per_cu < 0)
        return set_error(PK_ERR_ARG, "bad fast tuning (width 0..4, hash_bits 0 or 8..14, waves_per_group 0/1/2/4)");
    h->fast.width = width;
    h->fast.hash_bits = hash_bits;
    h->fast.waves_per_group = waves_per_group;
    h->fast.blocks_per_cu = blocks_per_cu;
    return PK_OK;
}

int pk_search_fast_device(pk_index* h, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
                          int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids,
                          float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    return fast_search_on(h, h->sc, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_max, k_stride, d_out_ids,
                          d_out_dists, d_out_count, d_counters);
}

// Coordinate descent over the fast mode's tuning at one (k, ef) on the
// caller's queries: waves per workgroup (queries per workgroup), the LDS
// filter size, the expansion width, then resident workgroups per CU; each
// setting timed over two launches after one warm-up.  Keeps the fastest
// (width may change results: recall at a fixed ef rises with it).
int pk_autotune_fast(pk_index* h, const void* d_queries, int64_t nq, int32_t k, int32_t ef, int32_t* out_tuning,
                     float* out_ms) {
    if (!h || !d_queries || nq < 1 || k < 1 || ef < k || ef > 512) return set_error(PK_ERR_ARG, "bad arguments");
    if (h->servers_alive) return set_error(PK_ERR_STATE, "destroy the serving loops before tuning");
    void* buf = nullptr;
    const size_t out_bytes = (size_t)nq * k * 8 + (size_t)nq * 4;
    if (hipMalloc(&buf, out_bytes) != hipSuccess) return set_error(PK_ERR_NOMEM, "tuner output buffers");
    int32_t* ids = static_cast<int32_t*>(buf);
    float* ds = reinterpret_cast<float*>(static_cast<uint8_t*>(buf) + (size_t)nq * k * 4);
    int32_t* cnt = reinterpret_cast<int32_t*>(static_cast<uint8_t*>(buf) + (size_t)nq * k * 8);
    const FastTuning saved = h->fast;
    auto time_one = [&](const FastTuning& t, float* ms) -> int {
        h->fast = t;
        int r = pk_search_fast_device(h, d_queries, nq, nullptr, nullptr, k, ef, ef, k, ids, ds, cnt, nullptr);
        if (r) return r;
        if (hipEventRecord(h->ev0, h->stream) != hipSuccess) return set_error(PK_ERR_HIP, "event record");
        for (int i = 0; i < 2 && !r; i++)
            r = pk_search_fast_device(h, d_queries, nq, nullptr, nullptr, k, ef, ef, k, ids, ds, cnt, nullptr);
        if (r) return r;
        if (hipEventRecord(h->ev1, h->stream) != hipSuccess || hipEventSynchronize(h->ev1) != hipSuccess ||
            hipEventElapsedTime(ms, h->ev0, h->ev1) != hipSuccess)
            return set_error(PK_ERR_HIP, "event timing");
        *ms /= 2;
        int32_t flags = 0;
        if (hipMemcpy(&flags, h->sc.err, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
            return set_error(PK_ERR_HIP, "flags read");
        if (flags) return set_error(PK_ERR_DEVICE, "fast search failed on the device (flags " + std::to_string(flags) + ")");
        return PK_OK;
    };
    FastTuning best = saved;
    float best_ms = 0;
    int rc = time_one(best, &best_ms);
    if (rc) {
        h->fast = saved;
        (void)hipFree(buf);
        return rc;
    }
    auto try_set = [&](FastTuning t) {
        float ms = 0;
        if (time_one(t, &ms) == PK_OK && ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    };
    for (int w : {1, 2, 4}) {
        FastTuning t = best;
        t.waves_per_group = w;
        try_set(t);
    }
    for (int b : {9, 10, 11, 12, 13}) {
        FastTuning t = best;
        t.hash_bits = b;
        try_set(t);
    }
    for (int w : {1, 2, 3, 4}) {
        FastTuning t = best;
        t.width = w;
        try_set(t);
    }
    {
        // resident workgroups per CU: up to three below what fits
        h->fast = best;
        const int last = h->last_grid;
        const int max_bpc = std::max(1, (last + h->num_cus - 1) / h->num_cus);
        for (int b = max_bpc - 1; b >= std::max(1, max_bpc - 3); b--) {
            FastTuning t = best;
            t.blocks_per_cu = b;
            try_set(t);
        }
    }
    h->fast = best;
    (void)hipFree(buf);
    if (out_tuning) {
        out_tuning[0] = best.width;
        out_tuning[1] = best.hash_bits;
        out_tuning[2] = best.waves_per_group;
        out_tuning[3] = best.blocks_per_cu;
    }
    if (out_ms) *out_ms = best_ms;
    return PK_OK;
}

}  // extern "C"
