// heap.hpp -- pickle's binary heaps, operated by a whole wave.
//
// Reference: EntryHeap (include/queue.hpp:99-289).  The candidate and
// result heaps decide which of several equal-distance entries is expanded
// or evicted first, so exact results need the reference's array layout:
// insert appends and sifts up while the parent strictly ranks below the new
// entry (queue.hpp:119-125, 204-216); pop moves the last entry to the root
// and sifts down, taking the left child first and the right child only when
// it is strictly better (queue.hpp:127-150, 230-236).
//
// The array lives in LDS up to `lcap` entries and continues in a per-wave
// region of device memory (the candidate heap is unbounded in the
// reference, queue.hpp:191-198).  The sequential steps are replaced by
// wave-parallel ones that leave the same array:
//   push_par  lane j reads ancestor j of the new slot; a ballot finds the
//             first ancestor the entry does not pass; the passed ancestors
//             shift down one level and the entry lands below the stop.
//   pop_par   every internal node's preferred child (left unless the right
//             one is strictly better) is found in parallel, then the path
//             from the root is followed with v_readlane until the moved
//             entry no longer ranks below the child; the path shifts up.
//   pop_deep  the sift-down three levels per round: lanes fetch the next
//             three levels below the current node at once.
#pragma once
#include "wave.hpp"

namespace pk {

// entry = float distance bits (low word) | int32 id (high word)
__device__ __forceinline__ uint64_t pack_entry(float d, int id) {
    return (uint64_t)__float_as_uint(d) | ((uint64_t)(uint32_t)id << 32);
}
__device__ __forceinline__ float entry_dist(uint64_t e) { return __uint_as_float((uint32_t)e); }
__device__ __forceinline__ int entry_id(uint64_t e) { return (int)(e >> 32); }

// internal nodes handled per lane by pop_par: 64 * kPopSlots >= len / 2
constexpr int kPopSlots = 5;

template <bool IS_MAX, bool SPILL>
struct WaveHeap {
    uint64_t* lds;   // lcap entries
    uint64_t* glob;  // gcap entries (used only when SPILL)
    int lcap;
    int gcap;
    int len;

    // true when x ranks below y (MaxFirst / MinFirst, queue.hpp:30-40)
    __device__ static __forceinline__ bool below(float x, float y) { return IS_MAX ? (x < y) : (x > y); }

    // per-lane access (any index); device-memory entries written by lane 0
    // are read from L2, not from a possibly stale L1 line.  The device-memory
    // branch waits for its own load before the paths join: otherwise the
    // compiler places a vector-memory wait at the join, which the LDS path
    // then also executes -- and that wait drains every older outstanding
    // store of the wave (the visited-set log), on every heap access.
    __device__ __forceinline__ uint64_t raw_get(int i) const {
        if (!SPILL || i < lcap) return lds[i];
        uint64_t v = __hip_atomic_load(&glob[i - lcap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        return v;
    }
    __device__ __forceinline__ void raw_put(int i, uint64_t e) {
        if (!SPILL || i < lcap)
            lds[i] = e;
        else
            glob[i - lcap] = e;
    }

    // uniform access (same index on every lane)
    __device__ __forceinline__ uint64_t get(int i) const { return uni64(raw_get(uni(i))); }
    __device__ __forceinline__ void put(int i, uint64_t e, bool leader) {
        if (leader) raw_put(i, e);
    }
    __device__ __forceinline__ bool full() const { return len >= lcap + (SPILL ? gcap : 0); }

    // Push in two phases so two heaps can share one LDS round trip:
    // push_read fetches the new slot's ancestors (lane j: ancestor j + 1),
    // push_write decides the stop and moves entries.
    struct PushState {
        int i, t;
        uint64_t v;
    };
    __device__ __forceinline__ PushState push_read(int lane) const {
        PushState p;
        p.i = uni(len);
        p.t = 31 - __builtin_clz((unsigned)(p.i + 1));  // number of ancestors
        const int j = lane + 1;
        const bool act = j <= p.t;
        p.v = act ? raw_get(((p.i + 1) >> j) - 1) : 0;
        return p;
    }
    // returns true when the entry became the root; *last (if given) receives
    // the entry now in the last slot (e, or the parent it displaced)
    __device__ __forceinline__ bool push_write(const PushState& p, uint64_t e, int lane, bool leader,
                                               uint64_t* last = nullptr) {
        const int i = p.i, t = p.t;
        len = i + 1;
        const float d = entry_dist(e);
        const int j = lane + 1;
        const bool act = j <= t;
        const bool up = act && below(entry_dist(p.v), d);
        const uint64_t stopm = ballot(act && !up);
        const int stop = stopm ? (int)__builtin_ctzll(stopm) + 1 : t + 1;
        __builtin_amdgcn_wave_barrier();
        if (act && j < stop) raw_put(((i + 1) >> (j - 1)) - 1, p.v);
        if (leader) raw_put(((i + 1) >> (stop - 1)) - 1, e);
        if (last)
            *last = stop == 1 ? e
                              : ((uint64_t)(uint32_t)readlane((int)(uint32_t)p.v, 0) |
                                 ((uint64_t)(uint32_t)readlane((int)(uint32_t)(p.v >> 32), 0) << 32));
        // device-memory slots written by several lanes: complete the stores
        // before any lane reads them back
        if (SPILL && i >= lcap) __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        return stop == t + 1;
    }
    __device__ bool push_par(uint64_t e, int lane, bool leader) {
        const PushState p = push_read(lane);
        return push_write(p, e, lane, leader);
    }

    // Sift-down reading a three-level subtree per round.  Relative to the
    // current node (position 1), lane t < 14 fetches subtree position t + 2
    // (children 2-3, grandchildren 4-7, great-grandchildren 8-15); up to
    // three levels are then resolved from those values with v_readlane.
    __device__ void pop_deep(int lane, bool leader) {
        const int n = uni(len) - 1;
        const uint64_t x = get(n);
        len = n;
        if (n == 0) return;
        const float xd = entry_dist(x);
        const int p = lane + 2;
        const int L = 31 - __builtin_clz((unsigned)p);
        const int k = p - (1 << L);
        int cur = 0;
        bool going = true;
        while (going) {
            const int idx = ((cur + 1) << L) + k - 1;
            const bool act = lane < 14 && idx < n;
            const uint64_t v = act ? raw_get(idx) : 0;
            const int vlo = (int)(uint32_t)v, vhi = (int)(uint32_t)(v >> 32);
            __builtin_amdgcn_wave_barrier();
            int rp = 1;
            for (int lev = 0; lev < 3; lev++) {
                const int li = 2 * cur + 1;
                if (li >= n) {
                    going = false;
                    break;
                }
                const int lc = 2 * rp - 2;  // lane holding the left child
                uint64_t bv = (uint64_t)(uint32_t)readlane(vlo, lc) | ((uint64_t)(uint32_t)readlane(vhi, lc) << 32);
                int best = li;
                if (li + 1 < n) {
                    uint64_t rv = (uint64_t)(uint32_t)readlane(vlo, lc + 1) |
                                  ((uint64_t)(uint32_t)readlane(vhi, lc + 1) << 32);
                    if (below(entry_dist(bv), entry_dist(rv))) {
                        best = li + 1;
                        bv = rv;
                    }
                }
                if (!below(xd, entry_dist(bv))) {
                    going = false;
                    break;
                }
                put(cur, bv, leader);
                rp = 2 * rp + (best - li);
                cur = best;
            }
        }
        put(cur, x, leader);
        __builtin_amdgcn_wave_barrier();
    }

    // parallel sift-down for heaps of up to 128 * SLOTS entries; returns the
    // new root (undefined when the heap is left empty)
    __device__ uint64_t pop_par(int lane, bool leader) { return pop_with(get(uni(len) - 1), lane, leader); }

    // pop when the caller already holds the entry in the last slot (x)
    __device__ uint64_t pop_with(uint64_t x, int lane, bool leader) {
        if (lcap <= 128) return pop_slots<1>(x, lane, leader);
        if (lcap <= 256) return pop_slots<2>(x, lane, leader);
        return pop_slots<kPopSlots>(x, lane, leader);
    }

    template <int SLOTS>
    __device__ uint64_t pop_slots(const uint64_t x, int lane, bool leader) {
        const int n = uni(len) - 1;
        len = n;
        if (n == 0) return x;
        const float xd = entry_dist(x);
        int pc[SLOTS];
        uint32_t pvl[SLOTS], pvh[SLOTS];
#pragma unroll
        for (int s = 0; s < SLOTS; s++) {
            const int node = lane + kWave * s;
            const int l = 2 * node + 1, r = l + 1;
            int c = -1;
            uint64_t cv = 0;
            if (l < n) {
                uint64_t lv = raw_get(l);
                c = l;
                cv = lv;
                if (r < n) {
                    uint64_t rv = raw_get(r);
                    if (below(entry_dist(lv), entry_dist(rv))) {
                        c = r;
                        cv = rv;
                    }
                }
            }
            pc[s] = c;
            pvl[s] = (uint32_t)cv;
            pvh[s] = (uint32_t)(cv >> 32);
        }
        __builtin_amdgcn_wave_barrier();
        int cur = 0;
        uint64_t root = x;
        for (;;) {
            const int s = cur >> 6, ln = cur & 63;
            int c = -1;
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int k = 0; k < SLOTS; k++)
                if (k == s) {
                    c = readlane(pc[k], ln);
                    lo = (uint32_t)readlane((int)pvl[k], ln);
                    hi = (uint32_t)readlane((int)pvh[k], ln);
                }
            if (c < 0) break;
            const uint64_t cv = (uint64_t)lo | ((uint64_t)hi << 32);
            if (!below(xd, entry_dist(cv))) break;
            if (cur == 0) root = cv;
            put(cur, cv, leader);
            cur = c;
        }
        put(cur, x, leader);
        __builtin_amdgcn_wave_barrier();
        return root;
    }

    __device__ __forceinline__ uint64_t top() const { return get(0); }
};

}  // namespace pk
