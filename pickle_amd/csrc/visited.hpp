// visited.hpp -- exact per-search visited set in device memory.
//
// Reference: ByteMap (include/visited.hpp:248-297) is exact with one byte
// per node per thread.  Here each wave owns an open-addressed hash table of
// 2^bits 32-bit slots in device memory plus a log of the slots it filled.
// Test-and-mark is one atomic compare-and-swap per lane against an empty
// slot, so a chunk of 64 adjacency ids costs one memory round trip; after a
// search the wave zeroes the slots in its log, or, when it marked 1/16 of
// the slots or more, the whole table with full-line stores, so the table is
// empty again for the next search.
//
// A slot holds (tag << 27) | (id + 1): the tag is the chunk counter mod 32,
// which tells a lane that finds its own id whether another lane of the same
// chunk put it there (an id repeated within one adjacency row) or an
// earlier chunk did.  Ids must be below 2^27 - 1.
#pragma once
#include "wave.hpp"

namespace pk {

constexpr uint32_t kVisitIdMask = (1u << 27) - 1;
constexpr int64_t kVisitMaxId = (int64_t)kVisitIdMask - 1;

struct WaveVisited {
    uint32_t* tab;  // 2^bits slots, all zero between searches
    int32_t* log;   // 2^bits filled-slot indices
    uint32_t mask;
    int bits;
    int count;      // slots filled by the current search (uniform)
    int peak = 0;        // largest count seen at a clear (diagnostics)
    int full_at = -1;    // count when a probe found the table full (diagnostics)
    uint32_t tag;   // chunk counter

    // marks from which a clear zeroes the whole table instead of the logged
    // slots: at 1/16 of the slots, most of the table's lines hold a mark
    __device__ __forceinline__ int full_clear_at() const { return (1 << bits) / 16; }

    __device__ __forceinline__ uint32_t home(int id) const {
        return ((uint32_t)id * 0x9E3779B1u) >> (32 - bits);
    }

    // Marks `id` (valid lanes only), in two halves so the caller can put
    // other memory work between them: begin() issues the first
    // compare-and-swap, finish() resolves it (and any collision chain).
    // finish() returns a lane mask of ids that were not in the set before;
    // an id repeated inside the call is credited to its first lane.  Sets
    // fail when the table is full.
    struct Pending {
        uint32_t key;
        uint32_t slot;
        uint32_t old;
    };

    __device__ __forceinline__ Pending begin(int id, bool valid) {
        tag = (tag + 1) & 31;
        Pending p;
        p.key = (tag << 27) | (uint32_t)(id + 1);
        p.slot = valid ? home(id) : 0u;
        p.old = valid ? atomicCAS(&tab[p.slot], 0u, p.key) : 0xffffffffu;
        return p;
    }

    __device__ uint64_t finish(const Pending& pd, int id, bool valid, int& fail) {
        int r = 1;  // 0 inserted, 1 present, 2 present from this call, 3 full
        uint32_t at = 0;
        if (valid) {
            uint32_t s = pd.slot, old = pd.old;
            r = 3;
            for (uint32_t p = 0; p <= mask; p++) {
                if (p > 0) {
                    s = (pd.slot + p) & mask;
                    old = atomicCAS(&tab[s], 0u, pd.key);
                }
                if (old == 0u) {
                    r = 0;
                    at = s;
                    break;
                }
                if ((old & kVisitIdMask) == (uint32_t)(id + 1)) {
                    r = (old >> 27) == tag ? 2 : 1;
                    break;
                }
            }
        }
        if (ballot(r == 3)) {
            fail = 1;
            full_at = count;
        }
        uint64_t newm = ballot(r == 0);
        uint64_t dupm = ballot(r == 2);
        if (newm) {
            if (r == 0) log[count + lanes_below(newm)] = (int32_t)at;
            count += __builtin_popcountll(newm);
        }
        // same id twice in one chunk: the first lane is the new one (the
        // reference scans the row in order, include/hnsw.hpp:413-422); a tag
        // match left from 32 chunks ago finds no partner and changes nothing
        while (dupm) {
            int j = __builtin_ctzll(dupm);
            int idj = readlane(id, j);
            uint64_t m = ballot(valid && id == idj) & (newm | dupm);
            uint64_t owner = m & newm;
            if (owner) {
                int first = __builtin_ctzll(m);
                newm = (newm & ~m) | (1ull << first);
            }
            dupm &= ~m;
        }
        return newm;
    }

    __device__ __forceinline__ uint64_t mark(int id, bool valid, int& fail) {
        Pending p = begin(id, valid);
        return finish(p, id, valid, fail);
    }

    // Empties the table (the slots this search filled) and orders the clears
    // before the next search's atomics.
    __device__ void clear(int lane) {
        if (count >= full_clear_at()) {
            // a search that marked this many slots has touched most of the
            // table's lines anyway: zero the whole table with full-line
            // 16-byte stores instead of one scattered store per slot
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            gmem<u32x4>* t = reinterpret_cast<gmem<u32x4>*>(gbl_rw(tab));
            const int n16 = (1 << bits) / 4;
            for (int i = lane; i < n16; i += kWave) t[i] = u32x4{0u, 0u, 0u, 0u};
            peak = count > peak ? count : peak;
            count = 0;
            __builtin_amdgcn_s_waitcnt(0);  // stores reach L2 before the next atomics
            return;
        }
        // log entries were written by other lanes: read them from L2.  Eight
        // reads go out before their eight stores, so the clear costs a few
        // round trips rather than one per 64 slots.
        constexpr int U = 8;
        for (int i0 = 0; i0 < count; i0 += kWave * U) {
            int slot[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * kWave + lane;
                slot[u] = i < count ? __hip_atomic_load(&log[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : -1;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (slot[u] >= 0) tab[slot[u]] = 0u;
        }
        peak = count > peak ? count : peak;
        count = 0;
        __builtin_amdgcn_s_waitcnt(0);  // stores reach L2 before the next atomics
    }
};

}  // namespace pk
