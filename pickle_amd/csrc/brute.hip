// brute.hip -- exhaustive k-nearest-neighbour search (ground truth).
//
// No reference counterpart on the hot path: the reference's benchmark reads
// ground-truth files computed elsewhere (benchmark/bench_util.hpp:196-202,
// benchmark/hnsw_pickle.cpp:94-112).  This is the brute-force side of the
// recall measurement.
//
// A workgroup owns 64 queries; its four waves stream disjoint quarters of
// the base rows.  Each lane keeps one query in registers and its own sorted
// top-KB list; a row address is the same for every lane of a wave, so each
// row is fetched once per wave and broadcast.  The four partial lists of a
// query are merged through LDS at the end.
//   u8 / i8: exact int32 distances (v_dot4 on 16-byte chunks)
//   f16 / f32: FP32 fused multiply-adds in element order; callers re-rank
//   the KB candidates in FP64.
// Ties rank by the lower row id.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pickle_gpu.h"
#include "distance.hpp"
#include "index_impl.hpp"

namespace pk {

constexpr int kBruteQ = 64;
constexpr int kBruteK = 32;  // candidates kept per query

template <int KB>
__device__ __forceinline__ void topk_insert(float (&bd)[KB], int (&bi)[KB], float d, int id) {
    // lexicographic (distance, id) order; list sorted ascending
    bool better = d < bd[KB - 1] || (d == bd[KB - 1] && id < bi[KB - 1]);
    if (!better) return;
    bd[KB - 1] = d;
    bi[KB - 1] = id;
#pragma unroll
    for (int j = KB - 1; j > 0; j--) {
        bool sw = bd[j] < bd[j - 1] || (bd[j] == bd[j - 1] && bi[j] < bi[j - 1]);
        float td = sw ? bd[j - 1] : bd[j];
        int ti = sw ? bi[j - 1] : bi[j];
        bd[j - 1] = sw ? bd[j] : bd[j - 1];
        bi[j - 1] = sw ? bi[j] : bi[j - 1];
        bd[j] = td;
        bi[j] = ti;
    }
}

// CH = 16-byte chunks per row
template <int DT, int MET, int CH>
__global__ __launch_bounds__(256) void k_brute(const uint8_t* __restrict__ base, int64_t nbase, int vstride,
                                               const uint8_t* __restrict__ queries, int64_t nq, int qbytes, int nch,
                                               int32_t* __restrict__ out_ids, float* __restrict__ out_d) {
    __shared__ float sd[4][kBruteK][kBruteQ];
    __shared__ int si[4][kBruteK][kBruteQ];
    const int lane = threadIdx.x & 63;
    const int sub = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * kBruteQ + lane;
    const bool live = q < nq;
    constexpr bool INT = DT == DT_U8 || DT == DT_I8;

    uint4 qv[CH];
    const uint8_t* qp = queries + (live ? q : 0) * (int64_t)qbytes;
#pragma unroll
    for (int c = 0; c < CH; c++) {
        uint32_t w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            uint32_t x = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                int off = c * 16 + u * 4 + b;
                uint32_t by = off < qbytes ? qp[off] : 0;
                x |= by << (8 * b);
            }
            w[u] = x;
        }
        qv[c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    int qq = 0;
    if (INT) {
        uint32_t s = 0;
#pragma unroll
        for (int c = 0; c < CH; c++) {
            s = dot4(DT == DT_I8, qv[c].x, qv[c].x, s);
            s = dot4(DT == DT_I8, qv[c].y, qv[c].y, s);
            s = dot4(DT == DT_I8, qv[c].z, qv[c].z, s);
            s = dot4(DT == DT_I8, qv[c].w, qv[c].w, s);
        }
        qq = (int)s;
    }

    float bd[kBruteK];
    int bi[kBruteK];
#pragma unroll
    for (int j = 0; j < kBruteK; j++) {
        bd[j] = __builtin_inff();
        bi[j] = 0x7fffffff;
    }

    for (int64_t r = sub; r < nbase; r += 4) {
        const uint4* row = reinterpret_cast<const uint4*>(base + r * (int64_t)vstride);
        float d;
        if (INT) {
            uint32_t ab = 0, bb = 0;
#pragma unroll
            for (int c = 0; c < CH; c++) {
                uint4 b = c < nch ? row[c] : make_uint4(0, 0, 0, 0);
                ab = dot4(DT == DT_I8, qv[c].x, b.x, ab);
                ab = dot4(DT == DT_I8, qv[c].y, b.y, ab);
                ab = dot4(DT == DT_I8, qv[c].z, b.z, ab);
                ab = dot4(DT == DT_I8, qv[c].w, b.w, ab);
                if (MET == MET_L2) {
                    bb = dot4(DT == DT_I8, b.x, b.x, bb);
                    bb = dot4(DT == DT_I8, b.y, b.y, bb);
                    bb = dot4(DT == DT_I8, b.z, b.z, bb);
                    bb = dot4(DT == DT_I8, b.w, b.w, bb);
                }
            }
            d = MET == MET_L2 ? (float)(qq + (int)bb - 2 * (int)ab) : -(float)(int)ab;
        } else {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < CH; c++) {
                uint4 b = c < nch ? row[c] : make_uint4(0, 0, 0, 0);
                uint32_t bw[4] = {b.x, b.y, b.z, b.w};
                uint32_t aw[4] = {qv[c].x, qv[c].y, qv[c].z, qv[c].w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (DT == DT_F32) {
                        float x = __uint_as_float(aw[u]), y = __uint_as_float(bw[u]);
                        acc = MET == MET_L2 ? __builtin_fmaf(x - y, x - y, acc) : __builtin_fmaf(x, y, acc);
                    } else {
#pragma unroll
                        for (int hh = 0; hh < 2; hh++) {
                            _Float16 xa, ya;
                            uint16_t xs = (uint16_t)(aw[u] >> (16 * hh)), ys = (uint16_t)(bw[u] >> (16 * hh));
                            __builtin_memcpy(&xa, &xs, 2);
                            __builtin_memcpy(&ya, &ys, 2);
                            float x = (float)xa, y = (float)ya;
                            acc = MET == MET_L2 ? __builtin_fmaf(x - y, x - y, acc) : __builtin_fmaf(x, y, acc);
                        }
                    }
                }
            }
            d = MET == MET_L2 ? acc : -acc;
        }
        topk_insert(bd, bi, d, (int)r);
    }

#pragma unroll
    for (int j = 0; j < kBruteK; j++) {
        sd[sub][j][lane] = bd[j];
        si[sub][j][lane] = bi[j];
    }
    __syncthreads();
    if (sub == 0 && live) {
        int pos[4] = {0, 0, 0, 0};
        for (int j = 0; j < kBruteK; j++) {
            int best = 0;
            float bdv = __builtin_inff();
            int biv = 0x7fffffff;
            for (int s = 0; s < 4; s++) {
                if (pos[s] >= kBruteK) continue;
                float dv = sd[s][pos[s]][lane];
                int iv = si[s][pos[s]][lane];
                if (dv < bdv || (dv == bdv && iv < biv)) {
                    bdv = dv;
                    biv = iv;
                    best = s;
                }
            }
            pos[best]++;
            out_d[q * kBruteK + j] = bdv;
            out_ids[q * kBruteK + j] = biv == 0x7fffffff ? -1 : biv;
        }
    }
}

template <int DT, int MET>
static hipError_t brute_dt(int chunks, dim3 grid, hipStream_t s, const uint8_t* base, int64_t nbase, int vstride,
                           const uint8_t* q, int64_t nq, int qbytes, int nch, int32_t* ids, float* d) {
#define PK_BRUTE_CASE(CHV)                                                                                       \
    if (chunks <= CHV) {                                                                                        \
        hipLaunchKernelGGL((k_brute<DT, MET, CHV>), grid, dim3(256), 0, s, base, nbase, vstride, q, nq, qbytes, \
                           nch, ids, d);                                                                        \
        return hipGetLastError();                                                                               \
    }
    PK_BRUTE_CASE(2)
    PK_BRUTE_CASE(4)
    PK_BRUTE_CASE(6)
    PK_BRUTE_CASE(8)
    if (DT == DT_F32 || DT == DT_F16) {
        PK_BRUTE_CASE(16)
        PK_BRUTE_CASE(32)
    }
#undef PK_BRUTE_CASE
    return hipErrorInvalidValue;
}

}  // namespace pk

using namespace pk;

extern "C" int pk_brute_force(pk_index* h, const void* queries, int64_t nq, int32_t* out_ids, float* out_dists) {
    if (!h || (nq > 0 && (!queries || !out_ids || !out_dists))) return set_error(PK_ERR_ARG, "null argument");
    if (!h->base || h->base_rows < 1) return set_error(PK_ERR_STATE, "no base rows");
    if (nq == 0) return PK_OK;
    int chunks = h->vstride / 16;
    bool is_int = h->dtype == PK_UINT8 || h->dtype == PK_INT8;
    if ((is_int && chunks > 8) || chunks > 32) return set_error(PK_ERR_ARG, "row too wide for the brute-force kernel");
    uint8_t* dq = nullptr;
    int32_t* di = nullptr;
    float* dd = nullptr;
    auto cleanup = [&]() {
        if (dq) (void)hipFree(dq);
        if (di) (void)hipFree(di);
        if (dd) (void)hipFree(dd);
    };
    if (hipMalloc(&dq, (size_t)nq * h->qbytes) != hipSuccess || hipMalloc(&di, (size_t)nq * kBruteK * 4) != hipSuccess ||
        hipMalloc(&dd, (size_t)nq * kBruteK * 4) != hipSuccess) {
        cleanup();
        return set_error(PK_ERR_NOMEM, "brute force: device allocation failed");
    }
    hipError_t e = hipMemcpyAsync(dq, queries, (size_t)nq * h->qbytes, hipMemcpyHostToDevice, h->stream);
    dim3 grid((unsigned)((nq + kBruteQ - 1) / kBruteQ));
    if (e == hipSuccess) {
        switch (h->dtype * 4 + h->metric) {
            case DT_U8 * 4 + MET_L2: e = brute_dt<DT_U8, MET_L2>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_U8 * 4 + MET_IP: e = brute_dt<DT_U8, MET_IP>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_I8 * 4 + MET_L2: e = brute_dt<DT_I8, MET_L2>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_I8 * 4 + MET_IP: e = brute_dt<DT_I8, MET_IP>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_F32 * 4 + MET_L2: e = brute_dt<DT_F32, MET_L2>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_F32 * 4 + MET_IP: e = brute_dt<DT_F32, MET_IP>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_F16 * 4 + MET_L2: e = brute_dt<DT_F16, MET_L2>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            case DT_F16 * 4 + MET_IP: e = brute_dt<DT_F16, MET_IP>(chunks, grid, h->stream, h->base, h->base_rows, h->vstride, dq, nq, h->qbytes, chunks, di, dd); break;
            default: e = hipErrorInvalidValue;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out_ids, di, (size_t)nq * kBruteK * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out_dists, dd, (size_t)nq * kBruteK * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    cleanup();
    if (e != hipSuccess) return set_error(PK_ERR_HIP, std::string("brute force: ") + hipGetErrorString(e));
    return PK_OK;
}
