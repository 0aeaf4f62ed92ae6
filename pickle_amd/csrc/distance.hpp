// distance.hpp -- L2 / negative-IP distances on wave lane groups.
//
// Reference: include/distance/avx512.hpp:14-243 (and generic.hpp:28-65).
// A wave computes 64 / LPC distances at a time, one aligned group of LPC
// lanes per candidate row; every lane keeps its slice of the query in
// registers.
//
// Integer rows (u8, i8): exact int32 sums from packed dot instructions
// (v_dot4_u32_u8 / v_dot4c_i32_i8).  L2 is |q|^2 + |b|^2 - 2 q.b, which is
// the exact sum of squared differences, converted to float with one
// round-to-nearest like the reference's int -> float return.
//
// Float rows (f32, f16 widened to f32): the order the reference build uses
// (-O3 -funsafe-math-optimizations): lane j of the 16-lane group accumulates
// elements j, j+16, ... with fused multiply-adds, a halving tree joins the 16
// lanes, and the dim % 16 tail runs in the group's first lane in the order
// of the compiled remainder loop (8-wide FMA block + tree, a 4-wide block
// seeded with that sum + tree, scalar FMAs), added last.
#pragma once
#include "wave.hpp"

namespace pk {

enum { DT_U8 = 0, DT_I8 = 1, DT_F32 = 2, DT_F16 = 3 };
enum { MET_L2 = 1, MET_IP = 2 };

__device__ __forceinline__ uint32_t dot4(bool is_signed, uint32_t a, uint32_t b, uint32_t acc) {
    if (is_signed) return (uint32_t)__builtin_amdgcn_sdot4((int)a, (int)b, (int)acc, false);
    return __builtin_amdgcn_udot4(a, b, acc, false);
}

// Integer distance: groups of 8 lanes, lane g owns 16-byte chunks g, g+8, ...
template <int DT, int MET, int NCH>
struct IntDist {
    static constexpr int LPC = 8;
    static constexpr bool SIGNED = DT == DT_I8;
    struct Row {
        uint4 c[NCH];
    };
    uint4 q[NCH];
    int qq;  // |q|^2 (whole query)
    int nchunks;

    __device__ void load_query(const uint8_t* qst, int g, int nchunks_) {
        nchunks = nchunks_;
        uint32_t part = 0;
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            int c = g + 8 * i;
            q[i] = c < nchunks ? ld16_lds(qst + 16 * c) : make_uint4(0, 0, 0, 0);
            part = dot4(SIGNED, q[i].x, q[i].x, part);
            part = dot4(SIGNED, q[i].y, q[i].y, part);
            part = dot4(SIGNED, q[i].z, q[i].z, part);
            part = dot4(SIGNED, q[i].w, q[i].w, part);
        }
        qq = sum8((int)part);
    }

    // Loads are issued unconditionally (chunks past the row are read from
    // its first chunk and ignored in finish): no branch separates them, so
    // all of a batch's gathers are in flight together.
    __device__ void load_row(Row& r, const uint8_t* row, int g, bool /*active*/) const {
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int c = g + 8 * i;
            r.c[i] = ld16_global(row + 16 * (c < nchunks ? c : 0));
        }
    }

    // the same chunks from a row staged in LDS
    __device__ void load_row_lds(Row& r, const uint8_t* row, int g) const {
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int c = g + 8 * i;
            r.c[i] = ld16_lds(row + 16 * (c < nchunks ? c : 0));
        }
    }

    // every lane of the group returns the distance
    template <bool LDS = false>
    __device__ float finish(const Row& r, int g, const uint8_t* row, bool active) const {
        uint32_t ab = 0, bb = 0;
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            if (g + 8 * i >= nchunks) continue;
            ab = dot4(SIGNED, q[i].x, r.c[i].x, ab);
            ab = dot4(SIGNED, q[i].y, r.c[i].y, ab);
            ab = dot4(SIGNED, q[i].z, r.c[i].z, ab);
            ab = dot4(SIGNED, q[i].w, r.c[i].w, ab);
            if (MET == MET_L2) {
                bb = dot4(SIGNED, r.c[i].x, r.c[i].x, bb);
                bb = dot4(SIGNED, r.c[i].y, r.c[i].y, bb);
                bb = dot4(SIGNED, r.c[i].z, r.c[i].z, bb);
                bb = dot4(SIGNED, r.c[i].w, r.c[i].w, bb);
            }
        }
        int sab = sum8((int)ab);
        if (MET == MET_L2) {
            int sbb = sum8((int)bb);
            return (float)(qq + sbb - 2 * sab);
        }
        return -(float)sab;
    }
};

// element i of a row in device memory / of the query staged in LDS
template <bool HALF>
__device__ __forceinline__ float felem(const uint8_t* p, int i) {
    if (HALF) return (float)gbl(reinterpret_cast<const _Float16*>(p))[i];
    return gbl(reinterpret_cast<const float*>(p))[i];
}
template <bool HALF>
__device__ __forceinline__ float qelem(const uint8_t* p, int i) {
    if (HALF) return (float)lds_ptr(reinterpret_cast<const _Float16*>(p))[i];
    return lds_ptr(reinterpret_cast<const float*>(p))[i];
}

template <int MET>
__device__ __forceinline__ float fterm(float x, float y, float acc) {
    if (MET == MET_L2) {
        float t = x - y;
        return __builtin_fmaf(t, t, acc);
    }
    return __builtin_fmaf(x, y, acc);
}

// Float distance: groups of 16 lanes; lane g owns elements g, g+16, ...
template <int DT, int MET, int NEL>
struct FloatDist {
    static constexpr int LPC = 16;
    static constexpr bool HALF = DT == DT_F16;
    struct Row {
        uint32_t raw[NEL];  // element bits (binary16 in the low half for fp16)
    };
    float q[NEL];
    int nmain;  // dim / 16
    int rem;    // dim % 16
    const uint8_t* qst;

    __device__ void load_query(const uint8_t* qst_, int g, int dim) {
        qst = qst_;
        nmain = dim >> 4;
        rem = dim & 15;
#pragma unroll
        for (int i = 0; i < NEL; i++) q[i] = i < nmain ? qelem<HALF>(qst, 16 * i + g) : 0.0f;
    }

    // unconditional loads of element bits (elements past the row re-read
    // element g); conversion waits until finish, so all loads are in flight
    __device__ void load_row(Row& r, const uint8_t* row, int g, bool /*active*/) const {
#pragma unroll
        for (int i = 0; i < NEL; i++) {
            const int e = i < nmain ? 16 * i + g : g;
            if (HALF)
                r.raw[i] = gbl(reinterpret_cast<const uint16_t*>(row))[e];
            else
                r.raw[i] = gbl(reinterpret_cast<const uint32_t*>(row))[e];
        }
    }
    // the same element bits from a row staged in LDS
    __device__ void load_row_lds(Row& r, const uint8_t* row, int g) const {
#pragma unroll
        for (int i = 0; i < NEL; i++) {
            const int e = i < nmain ? 16 * i + g : g;
            if (HALF)
                r.raw[i] = lds_ptr(reinterpret_cast<const uint16_t*>(row))[e];
            else
                r.raw[i] = lds_ptr(reinterpret_cast<const uint32_t*>(row))[e];
        }
    }
    __device__ static __forceinline__ float elem(uint32_t bits) {
        if (HALF) return (float)__builtin_bit_cast(_Float16, (uint16_t)bits);
        return __uint_as_float(bits);
    }
    template <bool LDS>
    __device__ static __forceinline__ float row_elem(const uint8_t* row, int i) {
        if (LDS) return qelem<HALF>(row, i);
        return felem<HALF>(row, i);
    }

    // sequential tail in the compiled reference's order (first lane only);
    // the row is in device memory, or in LDS when LDS is set
    template <bool LDS = false>
    __device__ float tail(const uint8_t* row) const {
        int base = nmain * 16, j = 0;
        float r = 0.0f;
        if (rem >= 8) {
            float a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = fterm<MET>(qelem<HALF>(qst, base + u), row_elem<LDS>(row, base + u), 0.0f);
#pragma unroll
            for (int h = 4; h >= 1; h >>= 1)
#pragma unroll
                for (int u = 0; u < h; u++) a[u] = a[u] + a[u + h];
            r = a[0];
            j = 8;
        }
        if (rem - j >= 4) {
            float a[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                a[u] = fterm<MET>(qelem<HALF>(qst, base + j + u), row_elem<LDS>(row, base + j + u), u == 0 ? r : 0.0f);
            a[0] = a[0] + a[2];
            a[1] = a[1] + a[3];
            r = a[0] + a[1];
            j += 4;
        }
        for (; j < rem; j++) r = fterm<MET>(qelem<HALF>(qst, base + j), row_elem<LDS>(row, base + j), r);
        return r;
    }

    template <bool LDS = false>
    __device__ float finish(const Row& r, int g, const uint8_t* row, bool active) const {
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < NEL; i++)
            if (i < nmain) acc = fterm<MET>(q[i], elem(r.raw[i]), acc);
        float s = tree_sum16(acc);
        if (rem != 0 && g == 0 && active) s = s + tail<LDS>(row);
        return MET == MET_IP ? -s : s;
    }
};

// Order-free FP16 distance, for the traversal copy of the precision form
// only (an FP16 index with its FP32 rows attached, whose results the FP32
// re-rank recomputes in the reference's order).  Groups of 16 lanes; lane g
// owns 16-byte chunks g, g + 16, ...: one load per lane per row where the
// reference-order FloatDist needs eight 2-byte loads, so a batch keeps
// 32 rows in flight.  The eight halves of a chunk widen to FP32 and
// accumulate per lane; the 16 lanes then sum by the halving tree.  Rows and
// the staged query are zero past the vector.
template <int MET, int NCH>
struct HalfDist {
    static constexpr int LPC = 16;
    struct Row {
        uint4 c[NCH];
    };
    float q[NCH * 8];
    int nchunks;

    __device__ static __forceinline__ float half_at(uint32_t w, int hi) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> (16 * hi)));
    }

    __device__ void load_query(const uint8_t* qst, int g, int nchunks_) {
        nchunks = nchunks_;
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int c = g + 16 * i;
            const uint4 u = c < nchunks ? ld16_lds(qst + 16 * c) : make_uint4(0, 0, 0, 0);
            const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int e = 0; e < 8; e++) q[8 * i + e] = half_at(w[e >> 1], e & 1);
        }
    }

    __device__ void load_row(Row& r, const uint8_t* row, int g, bool /*active*/) const {
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int c = g + 16 * i;
            r.c[i] = ld16_global(row + 16 * (c < nchunks ? c : 0));
        }
    }

    __device__ void load_row_lds(Row& r, const uint8_t* row, int g) const {
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            const int c = g + 16 * i;
            r.c[i] = ld16_lds(row + 16 * (c < nchunks ? c : 0));
        }
    }

    // every lane of the group returns the distance
    template <bool LDS = false>
    __device__ float finish(const Row& r, int g, const uint8_t* /*row*/, bool /*active*/) const {
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < NCH; i++) {
            if (g + 16 * i >= nchunks) continue;
            const uint32_t w[4] = {r.c[i].x, r.c[i].y, r.c[i].z, r.c[i].w};
#pragma unroll
            for (int e = 0; e < 8; e++) acc = fterm<MET>(q[8 * i + e], half_at(w[e >> 1], e & 1), acc);
        }
        const float s = tree_sum16(acc);
        return MET == MET_IP ? -s : s;
    }
};

}  // namespace pk
