// visited.hpp -- exact per-query visited set: LDS hash table + device spill.
//
// Reference: ByteMap (include/visited.hpp:248-297) gives an exact answer
// with one byte per node per thread.  Here each wave keeps an open-addressed
// table of ids in LDS (linear probing, bounded probe window); an id whose
// window is full goes to a per-wave table in device memory, in the manner
// of the reference's HashBitmap (visited.hpp:17-101: a small direct table
// plus an exact overflow set).  Device entries carry the search's epoch in
// their high word, so a new search needs no clearing there: entries of
// older epochs read as empty.  There are no false positives, so traversal
// is the reference's.
#pragma once
#include "wave.hpp"

namespace pk {

constexpr int kVisitProbe = 16;

struct WaveVisited {
    int* tab;                  // LDS, `mask + 1` slots, -1 = empty
    uint32_t mask;
    int bits;
    unsigned long long* gtab;  // device spill, `gmask + 1` slots
    uint32_t gmask;
    int gbits;
    uint32_t epoch;

    __device__ __forceinline__ uint32_t h1(int id) const {
        return ((uint32_t)id * 0x9E3779B1u) >> (32 - bits);
    }
    __device__ __forceinline__ uint32_t h2(int id) const {
        return ((uint32_t)id * 0x85EBCA6Bu) >> (32 - gbits);
    }

    // all lanes call; empties the LDS table
    __device__ void reset(uint32_t new_epoch, int lane) {
        epoch = new_epoch;
        int4* t4 = reinterpret_cast<int4*>(tab);
        int n4 = (int)((mask + 1) >> 2);
        for (int i = lane; i < n4; i += kWave) t4[i] = make_int4(-1, -1, -1, -1);
    }

    __device__ bool test(int id) const {
        uint32_t h = h1(id);
#pragma unroll 4
        for (int p = 0; p < kVisitProbe; p++) {
            int v = tab[(h + p) & mask];
            if (v == id) return true;
            if (v == -1) return false;
        }
        unsigned long long key = ((unsigned long long)epoch << 32) | (uint32_t)id;
        uint32_t g = h2(id);
        for (uint32_t p = 0; p <= gmask; p++) {
            unsigned long long v = gtab[(g + p) & gmask];
            if ((uint32_t)(v >> 32) != epoch) return false;
            if (v == key) return true;
        }
        return false;
    }

    // 0: inserted by this lane, 1: already present (possibly inserted by a
    // lane of the same call), 2: device table full
    __device__ int insert(int id) {
        uint32_t h = h1(id);
        for (int p = 0; p < kVisitProbe; p++) {
            int old = atomicCAS(&tab[(h + p) & mask], -1, id);
            if (old == -1) return 0;
            if (old == id) return 1;
        }
        unsigned long long key = ((unsigned long long)epoch << 32) | (uint32_t)id;
        uint32_t g = h2(id);
        uint32_t p = 0;
        for (uint32_t it = 0; it < 4 * (gmask + 1) && p <= gmask; it++) {
            unsigned long long* slot = &gtab[(g + p) & gmask];
            unsigned long long v = *slot;
            if ((uint32_t)(v >> 32) == epoch) {
                if (v == key) return 1;
                p++;
                continue;
            }
            unsigned long long prev = atomicCAS(slot, v, key);
            if (prev == v) return 0;
            if (prev == key) return 1;
            // another lane took the slot: look at it again
        }
        return 2;
    }
};

}  // namespace pk
