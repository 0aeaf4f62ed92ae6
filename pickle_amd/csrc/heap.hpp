// heap.hpp -- pickle's binary heaps, executed by a whole wave.
//
// Reference: EntryHeap (include/queue.hpp:99-289).  The candidate and
// result heaps decide which of several equal-distance entries is expanded
// or evicted first, so exact results need the same array layout: insert
// appends and sifts up while the parent strictly ranks below the new entry
// (queue.hpp:119-125, 204-216); pop moves the last entry to the root and
// sifts down, taking the left child first and the right child only when it
// is strictly better (queue.hpp:127-150, 230-236).
//
// Every lane runs the same (uniform) control flow; values read from memory
// are made scalar with readfirstlane, and only lane 0 stores.  The array
// lives in LDS up to `lcap` entries; deeper entries spill to a per-wave
// region of device memory (the candidate heap is unbounded in the
// reference, queue.hpp:191-198).
#pragma once
#include "wave.hpp"

namespace pk {

// entry = float distance bits (low word) | int32 id (high word)
__device__ __forceinline__ uint64_t pack_entry(float d, int id) {
    return (uint64_t)__float_as_uint(d) | ((uint64_t)(uint32_t)id << 32);
}
__device__ __forceinline__ float entry_dist(uint64_t e) { return __uint_as_float((uint32_t)e); }
__device__ __forceinline__ int entry_id(uint64_t e) { return (int)(e >> 32); }

template <bool IS_MAX, bool SPILL>
struct WaveHeap {
    uint64_t* lds;   // lcap entries
    uint64_t* glob;  // gcap entries (used only when SPILL)
    int lcap;
    int gcap;
    int len;

    // true when x ranks below y (MaxFirst / MinFirst, queue.hpp:30-40)
    __device__ static __forceinline__ bool below(float x, float y) { return IS_MAX ? (x < y) : (x > y); }

    __device__ __forceinline__ uint64_t get(int i) const {
        uint64_t v;
        if (!SPILL || i < lcap)
            v = lds[i];
        else
            v = glob[i - lcap];
        return uni64(v);
    }
    __device__ __forceinline__ void put(int i, uint64_t e, bool leader) {
        if (leader) {
            if (!SPILL || i < lcap)
                lds[i] = e;
            else
                glob[i - lcap] = e;
        }
    }
    __device__ __forceinline__ bool full() const { return len >= lcap + (SPILL ? gcap : 0); }

    __device__ void push(uint64_t e, bool leader) {
        int i = len++;
        float d = entry_dist(e);
        while (i > 0) {
            int p = (i - 1) >> 1;
            uint64_t pe = get(p);
            if (below(entry_dist(pe), d)) {
                put(i, pe, leader);
                i = p;
            } else {
                break;
            }
        }
        put(i, e, leader);
    }

    __device__ void pop(bool leader) {
        uint64_t x = get(len - 1);
        len -= 1;
        if (len == 0) return;
        float xd = entry_dist(x);
        int i = 0;
        for (;;) {
            int l = 2 * i + 1, r = l + 1;
            if (l >= len) break;
            int best = i;
            uint64_t bv = x;
            float bd = xd;
            uint64_t le = get(l);
            if (below(bd, entry_dist(le))) {
                best = l;
                bv = le;
                bd = entry_dist(le);
            }
            if (r < len) {
                uint64_t re = get(r);
                if (below(bd, entry_dist(re))) {
                    best = r;
                    bv = re;
                }
            }
            if (best == i) break;
            put(i, bv, leader);
            i = best;
        }
        put(i, x, leader);
    }

    __device__ __forceinline__ uint64_t top() const { return get(0); }
};

}  // namespace pk
