// brute.hip -- exhaustive k-nearest-neighbour search on the matrix cores.
//
// No reference counterpart on the hot path: the reference's benchmark reads
// ground-truth files computed elsewhere (benchmark/bench_util.hpp:196-202,
// benchmark/hnsw_pickle.cpp:94-112).  This produces them.
//
// The distance matrix is a dense contraction, so the dot products run as
// MFMA tile products and only the top-k selection is vector work:
//   u8 / i8: v_mfma_i32_32x32x32_i8 over 32 base rows x 32 queries.  u8 is
//     moved to the signed range by x' = x - 128 (an XOR of the top bit);
//     L2 is |q'|^2 + |x'|^2 - 2 q'.x' (the shift cancels) and -IP adds
//     back 128 (sum q' + sum x') + 128^2 dim.  Every sum is exact in int32.
//   f16 / f32: v_mfma_f64_16x16x4f64 over 16 base rows x 16 queries, on
//     the values widened to FP64 (exactly); L2 is |x|^2 - 2 q.x + |q|^2 in
//     FP64.
// The order of K inside a tile does not change a dot product, so every lane
// feeds one contiguous 16-byte chunk of its row to each operand; the A and
// B operands pair the same chunk bytes lane group by lane group.  Staged
// query rows are zero past the vector (after the u8 shift), so bytes past
// `dim` in a base row contribute nothing.
//
// Selection: each lane owns one query and a fixed half (integer tiles) or
// quarter (FP64 tiles) of every tile's rows, and keeps a sorted list of its
// 32 best (key, row) pairs in registers.  Rows reach a lane in increasing
// order, so a strict compare against the list's last key keeps the lower
// row first among equal distances.  The grid splits the base rows as well
// as the queries; k_brute_merge takes the 32 best entries of every query's
// lists, ascending by (distance, row id).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pickle_gpu.h"
#include "brute.hpp"
#include "index_impl.hpp"

namespace pk {

__global__ void k_brute_fill(int32_t* __restrict__ p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- per-row constants -------------------------------------------------

// integer rows: |x'|^2 (L2) or -128 sum x' (u8 -IP); rows past nbase get 0
template <int DT, int MET>
__global__ __launch_bounds__(256) void k_brute_rowconst(const uint8_t* __restrict__ base, int64_t nbase, int vstride,
                                                        int dim, int64_t npad, int32_t* __restrict__ crow) {
    constexpr uint32_t FLIP = DT == DT_U8 ? 0x80808080u : 0u;
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= npad) return;
    int acc = 0;
    if (r < nbase) {
        const uint8_t* rp = base + r * (int64_t)vstride;
        for (int c = 0; c * 16 < dim; c++) {
            const uint4 u = ld16_global(rp + 16 * c);
            const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int nb = dim - (16 * c + 4 * i);  // bytes of this word inside the vector
                const uint32_t m = nb >= 4 ? ~0u : nb <= 0 ? 0u : ((1u << (8 * nb)) - 1u);
                const int x = (int)((w[i] ^ FLIP) & m);
                acc = MET == MET_L2 ? __builtin_amdgcn_sdot4(x, x, acc, false)
                                    : __builtin_amdgcn_sdot4(x, 0x01010101, acc, false);
            }
        }
    }
    crow[r] = MET == MET_L2 ? acc : -128 * acc;
}

// ---- integer tiles -------------------------------------------------------

// Queries are staged NKB * 32 bytes apart (u8 padded with 0x80, i8 with 0).
// One-dimensional grid of qblocks x splits workgroups: query blocks of kIntQ
// queries times row splits of tiles_per_split tiles (brute_block_coords).
// Lists: pd / pi [q][split * 2 + lane / 32][kBruteK].
template <int DT, int MET, int NKB>
__global__ __launch_bounds__(256) void k_brute_imma(const uint8_t* __restrict__ base, int64_t nbase, int vstride,
                                                    const int32_t* __restrict__ crow, const uint8_t* __restrict__ qs,
                                                    int64_t nq, int dim, int64_t qblocks, int xcd_map,
                                                    int64_t tiles_per_split, int nlists, double* __restrict__ pd,
                                                    int32_t* __restrict__ pi, int32_t* __restrict__ qbound) {
    constexpr int KB = kBruteK;
    constexpr int QST = NKB * 32;
    constexpr uint32_t FLIP = DT == DT_U8 ? 0x80808080u : 0u;
    constexpr int COEF = MET == MET_L2 ? 2 : 1;
    constexpr bool CROW = !(DT == DT_I8 && MET == MET_IP);
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5, col = lane & 31;
    int64_t qb, split;
    brute_block_coords(qblocks, xcd_map, qb, split);
    const int64_t q = qb * kIntQ + (threadIdx.x >> 6) * 32 + col;
    const int64_t qc = q < nq ? q : nq - 1;

    v4i bq[NKB];
    int part = 0;
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) {
        bq[kb] = as_v4i(ld16_global(qs + qc * QST + kb * 32 + h * 16), FLIP);
        const int w[4] = {bq[kb].x, bq[kb].y, bq[kb].z, bq[kb].w};
#pragma unroll
        for (int i = 0; i < 4; i++)
            part = MET == MET_L2 ? __builtin_amdgcn_sdot4(w[i], w[i], part, false)
                                 : __builtin_amdgcn_sdot4(w[i], 0x01010101, part, false);
    }
    part += __shfl_xor(part, 32, kWave);
    const int cq = MET == MET_L2 ? part : (DT == DT_U8 ? -128 * part - 16384 * dim : 0);

    int coff[NKB];  // the lane's chunk offsets, kept inside the row
#pragma unroll
    for (int kb = 0; kb < NKB; kb++) coff[kb] = kb * 32 + h * 16 < vstride ? kb * 32 + h * 16 : 0;

    __shared__ int skey_all[4][16 * kWave];
    int* skey = skey_all[threadIdx.x >> 6];

    int kd[KB], ki[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) {
        kd[j] = INT_MAX;
        ki[j] = INT_MAX;
    }
    int thr = INT_MAX;
    int gbound = INT_MAX;  // smallest 32nd-best key any list of this query has published

    const int64_t ntiles = (nbase + 31) / 32;
    const int64_t t0 = split * tiles_per_split;
    const int64_t t1 = t0 + tiles_per_split < ntiles ? t0 + tiles_per_split : ntiles;

    // Base-row tiles arrive two ahead of their use (three register
    // buffers in rotation): one tile of MFMA and selection work is shorter
    // than a gather's latency at two waves per SIMD.
    struct Tile {
        v4i a[NKB];
        int4 c[4];
    };
    auto load_tile = [&](Tile& T, int64_t t) {
        int64_t row = t * 32 + col;
        row = row < nbase ? row : nbase - 1;
        const uint8_t* rp = base + row * (int64_t)vstride;
#pragma unroll
        for (int kb = 0; kb < NKB; kb++) T.a[kb] = as_v4i(ld16_global(rp + coff[kb]), FLIP);
        if (CROW) {
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const uint4 u = ld16_global(crow + t * 32 + 8 * g + 4 * h);
                T.c[g] = make_int4((int)u.x, (int)u.y, (int)u.z, (int)u.w);
            }
        }
    };
    auto process = [&](const Tile& T, int64_t t) {
        v16i acc = {};
#pragma unroll
        for (int kb = 0; kb < NKB; kb++) acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(T.a[kb], bq[kb], acc, 0, 0, 0);

        // Lanes l and l + 32 hold the same query.  A key above the other
        // half's 32nd-best key loses to all 32 of that half's entries, so it
        // cannot be among the query's 32 best: the filter takes the tighter
        // bound (equal keys pass; the merge orders ties by row id).
        // Every 32 tiles the lane also publishes its 32nd-best key to the
        // query's bound in device memory (shared by all row splits) and
        // takes back the smallest published so far, for the same reason.
        if (((t - t0) & 31) == 31 && q < nq) {
            const int got = atomicMin(&qbound[q], kd[KB - 1]);
            gbound = min(gbound, min(got, kd[KB - 1]));
        }
        const int other = min(__shfl_xor(kd[KB - 1], 32, kWave), gbound);
        const int obound = other == INT_MAX ? INT_MAX : other + 1;
        thr = min(kd[KB - 1], obound);

        // acc[r] is row t*32 + 8 (r / 4) + 4 h + r % 4 against query col
        int key[16];
        uint32_t pend = 0;  // rows of this tile that pass the lane's filter
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int4 c4 = T.c[r >> 2];
            const int c = (r & 3) == 0 ? c4.x : (r & 3) == 1 ? c4.y : (r & 3) == 2 ? c4.z : c4.w;
            key[r] = (CROW ? c : 0) - COEF * acc[r];
            pend |= (uint32_t)(key[r] < thr) << r;
        }
        // every lane inserts its next passing row in the same step, in row
        // order, re-checking against its tightened bound: a step per row
        // of the busiest lane, not per row that any lane passes.  The keys
        // wait in LDS ([row][lane], no bank conflicts) so that registers
        // need not hold all sixteen through the loop.
        if (ballot(pend != 0)) {
#pragma unroll
            for (int r = 0; r < 16; r++) skey[r * kWave + lane] = key[r];
            __builtin_amdgcn_wave_barrier();
#pragma unroll 1
            for (int step = 0; step < 16; step++) {
                if (!ballot(pend != 0)) break;
                if (pend) {
                    const int r = __builtin_ctz(pend);
                    pend &= pend - 1;
                    const int kr = skey[r * kWave + lane];
                    const int64_t m = t * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                    if (kr < thr && m < nbase) {
                        sorted_insert(kd, ki, kr, (int)m);
                        thr = min(kd[KB - 1], obound);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    };
    Tile B0, B1, B2;
    if (t0 < t1) load_tile(B0, t0);
    if (t0 + 1 < t1) load_tile(B1, t0 + 1);
    for (int64_t t = t0; t < t1; t += 3) {
        if (t + 2 < t1) load_tile(B2, t + 2);
        process(B0, t);
        if (t + 1 >= t1) break;
        if (t + 3 < t1) load_tile(B0, t + 3);
        process(B1, t + 1);
        if (t + 2 >= t1) break;
        if (t + 4 < t1) load_tile(B1, t + 4);
        process(B2, t + 2);
    }

    if (q < nq) {
        const size_t o = ((size_t)q * nlists + (size_t)split * 2 + h) * KB;
#pragma unroll
        for (int j = 0; j < KB; j++) {
            pd[o + j] = ki[j] == INT_MAX ? __builtin_inf() : (double)(kd[j] + cq);
            pi[o + j] = ki[j];
        }
    }
}

// ---- merge -----------------------------------------------------------------

// One wave per query: lane i holds the head of list i (nlists <= 64); each
// round the wave takes the least (distance, row id) head and its lane moves
// to the next entry.
__global__ __launch_bounds__(256) void k_brute_merge(const double* __restrict__ pd, const int32_t* __restrict__ pi,
                                                     int nlists, int64_t nq, int32_t* __restrict__ out_ids,
                                                     float* __restrict__ out_d) {
    constexpr int KB = kBruteK;
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nq) return;
    const bool has = lane < nlists;
    const size_t o = ((size_t)q * nlists + (has ? lane : 0)) * KB;
    int p = 0;
    double hd = has ? pd[o] : __builtin_inf();
    int hi = has ? pi[o] : INT_MAX;
    for (int j = 0; j < KB; j++) {
        double bd = hd;
        int bi = hi, bl = lane;
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
            const double od = __shfl_xor(bd, s, kWave);
            const int oi = __shfl_xor(bi, s, kWave);
            const int ol = __shfl_xor(bl, s, kWave);
            const bool take = od < bd || (od == bd && (oi < bi || (oi == bi && ol < bl)));
            bd = take ? od : bd;
            bi = take ? oi : bi;
            bl = take ? ol : bl;
        }
        if (lane == 0) {
            out_ids[q * KB + j] = bi == INT_MAX ? -1 : bi;
            out_d[q * KB + j] = bi == INT_MAX ? __builtin_inff() : (float)bd;
        }
        if (lane == bl) {
            p++;
            hd = p < KB ? pd[o + p] : __builtin_inf();
            hi = p < KB ? pi[o + p] : INT_MAX;
        }
    }
}

template <int DT, int MET>
static hipError_t launch_int(int nkb, const BrutePlan& p, hipStream_t s, const uint8_t* base, int64_t nbase,
                             int vstride, const int32_t* crow, const uint8_t* dq, int64_t nq, int dim, double* pd,
                             int32_t* pi, int32_t* qbound) {
    const dim3 grid((unsigned)(p.qblocks * p.splits));
    const int xcd_map = p.splits % kXcds == 0;
#define PK_IMMA_CASE(N)                                                                                       \
    if (nkb <= N) {                                                                                           \
        hipLaunchKernelGGL((k_brute_imma<DT, MET, N>), grid, dim3(256), 0, s, base, nbase, vstride, crow, dq, \
                           nq, dim, p.qblocks, xcd_map, p.tiles_per_split, p.nlists, pd, pi, qbound);        \
        return hipGetLastError();                                                                             \
    }
    PK_IMMA_CASE(1)
    PK_IMMA_CASE(2)
    PK_IMMA_CASE(3)
    PK_IMMA_CASE(4)
    PK_IMMA_CASE(6)
    PK_IMMA_CASE(8)
#undef PK_IMMA_CASE
    return hipErrorInvalidValue;
}

}  // namespace pk

using namespace pk;

extern "C" int pk_brute_force(pk_index* h, const void* queries, int64_t nq, int32_t* out_ids, float* out_dists) {
    if (!h || (nq > 0 && (!queries || !out_ids || !out_dists))) return set_error(PK_ERR_ARG, "null argument");
    if (!h->base || h->base_rows < 1) return set_error(PK_ERR_STATE, "no base rows");
    if (nq == 0) return PK_OK;
    const bool is_int = h->dtype == PK_UINT8 || h->dtype == PK_INT8;
    const int dim = h->dim;
    const int64_t nbase = h->base_rows;
    // staged query stride: 32-byte K blocks (integer) or 64-byte super-chunks (float)
    const int unit = is_int ? 32 : 64;
    const int nunits = (h->qbytes + unit - 1) / unit;
    if (nunits > 8) return set_error(PK_ERR_ARG, "row too wide for the brute-force kernel (at most 256 bytes of u8/i8, 512 of f16/f32)");
    static const int kRound[9] = {1, 1, 2, 3, 4, 6, 6, 8, 8};
    const int nu = kRound[nunits];
    const int qst = nu * unit;
    const BrutePlan p = is_int ? plan_brute_xcd(nq, nbase, kIntQ, 32, 2, PK_BRUTE_INT_WGS)
                               : plan_brute_xcd(nq, nbase, kFltQ, 16, 4, PK_BRUTE_WGS);
    const int64_t npad = (nbase + 31) / 32 * 32;

    std::vector<uint8_t> staged((size_t)nq * qst, h->dtype == PK_UINT8 ? 0x80 : 0x00);
    for (int64_t i = 0; i < nq; i++)
        std::memcpy(staged.data() + (size_t)i * qst, (const uint8_t*)queries + (size_t)i * h->qbytes, h->qbytes);

    uint8_t* dq = nullptr;
    void* rc = nullptr;  // crow (int32) or row norms (FP64)
    double* pd = nullptr;
    int32_t* pi = nullptr;
    int32_t* di = nullptr;
    float* dd = nullptr;
    int32_t* qb = nullptr;  // per-query bound shared by the integer tiles' row splits
    auto cleanup = [&]() {
        for (void* x : {(void*)dq, rc, (void*)pd, (void*)pi, (void*)di, (void*)dd, (void*)qb})
            if (x) (void)hipFree(x);
    };
    const size_t nl = (size_t)nq * p.nlists * kBruteK;
    if (hipMalloc(&dq, staged.size()) != hipSuccess || hipMalloc(&rc, (size_t)npad * 8) != hipSuccess ||
        hipMalloc(&pd, nl * 8) != hipSuccess || hipMalloc(&pi, nl * 4) != hipSuccess ||
        hipMalloc(&di, (size_t)nq * kBruteK * 4) != hipSuccess || hipMalloc(&dd, (size_t)nq * kBruteK * 4) != hipSuccess ||
        hipMalloc(&qb, (size_t)nq * 4) != hipSuccess) {
        cleanup();
        return set_error(PK_ERR_NOMEM, "brute force: device allocation failed");
    }
    hipStream_t s = h->stream;
    hipError_t e = hipMemcpyAsync(dq, staged.data(), staged.size(), hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_brute_fill, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, qb, INT_MAX, nq);
    const dim3 rgrid((unsigned)(npad / 256 + 1));
    const uint8_t* b = h->base;
    const int vs = h->vstride;
    if (e == hipSuccess) {
        switch (h->dtype * 4 + h->metric) {
            case DT_U8 * 4 + MET_L2:
                hipLaunchKernelGGL((k_brute_rowconst<DT_U8, MET_L2>), rgrid, dim3(256), 0, s, b, nbase, vs, dim, npad, (int32_t*)rc);
                e = launch_int<DT_U8, MET_L2>(nu, p, s, b, nbase, vs, (int32_t*)rc, dq, nq, dim, pd, pi, qb);
                break;
            case DT_U8 * 4 + MET_IP:
                hipLaunchKernelGGL((k_brute_rowconst<DT_U8, MET_IP>), rgrid, dim3(256), 0, s, b, nbase, vs, dim, npad, (int32_t*)rc);
                e = launch_int<DT_U8, MET_IP>(nu, p, s, b, nbase, vs, (int32_t*)rc, dq, nq, dim, pd, pi, qb);
                break;
            case DT_I8 * 4 + MET_L2:
                hipLaunchKernelGGL((k_brute_rowconst<DT_I8, MET_L2>), rgrid, dim3(256), 0, s, b, nbase, vs, dim, npad, (int32_t*)rc);
                e = launch_int<DT_I8, MET_L2>(nu, p, s, b, nbase, vs, (int32_t*)rc, dq, nq, dim, pd, pi, qb);
                break;
            case DT_I8 * 4 + MET_IP:
                e = launch_int<DT_I8, MET_IP>(nu, p, s, b, nbase, vs, nullptr, dq, nq, dim, pd, pi, qb);
                break;
            case DT_F32 * 4 + MET_L2:
            case DT_F32 * 4 + MET_IP:
            case DT_F16 * 4 + MET_L2:
            case DT_F16 * 4 + MET_IP:
                e = launch_brute_fp64(h->dtype, h->metric, nu, p, s, b, nbase, vs, dim, npad, (double*)rc, dq, nq, pd, pi);
                break;
            default: e = hipErrorInvalidValue;
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_brute_merge, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0, s, pd, pi, p.nlists, nq, di, dd);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_ids, di, (size_t)nq * kBruteK * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out_dists, dd, (size_t)nq * kBruteK * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    cleanup();
    if (e != hipSuccess) return set_error(PK_ERR_HIP, std::string("brute force: ") + hipGetErrorString(e));
    return PK_OK;
}
