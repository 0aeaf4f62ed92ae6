// precision.hip -- low-precision traversal copies and the exact FP32 re-rank.
//
// The precision form keeps two copies of the base vectors in HBM:
//   * the index's own rows (its dtype), a traversal copy in FP16 or in
//     per-dimension-scaled int8, which the graph search reads;
//   * FP32 rows by external id, read only to re-rank the final candidates.
// A search converts the FP32 queries to the traversal type on the device,
// runs the ordinary search (include/hnsw.hpp:509-541) with k = ef, and
// re-ranks the ef candidates with FP32 distances computed in the
// reference's own order (include/distance/avx512.hpp, FloatDist), so every
// returned distance is the one the reference's FP32 search reports for that
// id.  The re-rank is a gathered per-query product (ef rows x d): memory
// bound at about 2 flop per byte, computed with vector FMA in the
// reference's summation order rather than as a matrix-core tile.
//
// int8 copy: q[d] = clamp(rint((x[d] - offset[d]) * inv_scale[d]), -127, 127)
// in FP32 arithmetic; traversal then uses the int8 L2 / IP of the scaled
// values.  FP16 copy: round to nearest even.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "index_impl.hpp"
#include "traverse.hpp"

namespace pk {

namespace {

#define PCHK(expr)                                                                                        \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return set_error(PK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int palloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(PK_ERR_NOMEM, "device allocation failed (" + std::to_string(count * sizeof(T)) + " bytes)");
    }
    return PK_OK;
}

int scratch(IoBuf& b, size_t bytes, void** p) {
    if (b.bytes < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        int rc = palloc(reinterpret_cast<uint8_t**>(&b.p), bytes);
        if (rc) return rc;
        b.bytes = bytes;
    }
    *p = b.p;
    return PK_OK;
}

// FP32 rows (in_stride floats apart) -> traversal rows (out_stride bytes,
// zero padded past dim).  OUT: DT_F16 or DT_I8.
template <int OUT>
__global__ void __launch_bounds__(256) k_quantize(const float* __restrict__ in, int64_t in_stride, int64_t n,
                                                  int dim, uint8_t* __restrict__ out, int64_t out_stride,
                                                  const float* __restrict__ off, const float* __restrict__ inv) {
    constexpr int ES = OUT == DT_F16 ? 2 : 1;
    const int per_row = (int)(out_stride / ES);
    const int64_t total = n * per_row;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / per_row;
        const int d = (int)(t - r * per_row);
        const float x = d < dim ? in[r * in_stride + d] : 0.0f;
        if (OUT == DT_F16) {
            reinterpret_cast<_Float16*>(out + r * out_stride)[d] = d < dim ? (_Float16)x : (_Float16)0.0f;
        } else {
            float v = 0.0f;
            if (d < dim) {
                float u = x - off[d];
                u = u * inv[d];
                v = fminf(fmaxf(__builtin_rintf(u), -127.0f), 127.0f);
            }
            reinterpret_cast<int8_t*>(out + r * out_stride)[d] = (int8_t)(int)v;
        }
    }
}

struct RerankArgs {
    const float* queries;  // nq x dim
    int64_t nq;
    int dim;
    const uint8_t* full;  // FP32 rows by external id, full_vstride bytes apart
    int full_vstride;
    int64_t full_rows;
    const int32_t* cand;  // nq x cand_stride external ids in traversal order
    int cand_stride;
    const int32_t* cand_cnt;
    int k, k_stride;
    int32_t* out_ids;
    float* out_dists;
    int32_t* out_count;
    int32_t* qcounter;
    int lds_per_wave;
    int cap;  // candidate slots staged per wave (multiple of 64)
};

constexpr int kRerankBlock = 128;

// order-preserving map of a float to an unsigned integer
__device__ __forceinline__ uint32_t ord_f32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, s, kWave);
        uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), s, kWave);
        uint64_t o = (uint64_t)lo | ((uint64_t)hi << 32);
        v = o < v ? o : v;
    }
    return v;
}

template <int MET, int NEL>
__global__ __launch_bounds__(kRerankBlock) void k_rerank(RerankArgs A) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    uint8_t* my = smem + (size_t)w * A.lds_per_wave;
    int* cl = reinterpret_cast<int*>(my);
    float* cd = reinterpret_cast<float*>(cl + A.cap);
    uint8_t* qst = reinterpret_cast<uint8_t*>(cd + A.cap);
    const int g = lane % 16;
    constexpr int kSlots = 10;  // cap <= 640
    for (;;) {
        int qi = 0;
        if (lane == 0) qi = atomicAdd(A.qcounter, 1);
        const int64_t q = readlane(qi, 0);
        if (q >= A.nq) break;
        int32_t* oid = A.out_ids + (size_t)q * A.k_stride;
        float* odist = A.out_dists + (size_t)q * A.k_stride;
        const int cnt = uni(A.cand_cnt[q]);
        int kk = 0;
        if (cnt >= 0) {
            stage_row(qst, reinterpret_cast<const uint8_t*>(A.queries + (size_t)q * A.dim), A.dim * 4,
                      A.full_vstride, lane);
            FloatDist<DT_F32, MET, NEL> df;
            df.load_query(qst, g, A.dim);
            // candidates whose id is not a stored row are dropped
            int n = 0;
            for (int b = 0; b < cnt; b += kWave) {
                const int i = b + lane;
                const int id = i < cnt ? A.cand[(size_t)q * A.cand_stride + i] : -1;
                const bool ok = i < cnt && id >= 0 && (int64_t)id < A.full_rows;
                const uint64_t m = ballot(ok);
                if (ok) cl[n + lanes_below(m)] = id;
                n += __builtin_popcountll(m);
            }
            __builtin_amdgcn_wave_barrier();
            batch_distances(df, A.full, A.full_vstride, cl, cd, n, lane);
            __builtin_amdgcn_wave_barrier();
            // ascending (distance, traversal rank)
            uint64_t key[kSlots];
#pragma unroll
            for (int s = 0; s < kSlots; s++) {
                const int i = lane + kWave * s;
                key[s] = i < n ? ((uint64_t)ord_f32(cd[i]) << 32) | (uint32_t)i : ~0ull;
            }
            kk = A.k < n ? A.k : n;
            for (int r = 0; r < kk; r++) {
                uint64_t m = ~0ull;
#pragma unroll
                for (int s = 0; s < kSlots; s++) m = key[s] < m ? key[s] : m;
                m = wave_min_u64(m);
                const int i = (int)(uint32_t)m;
                if (lane == 0) {
                    oid[r] = cl[i];
                    odist[r] = cd[i];
                }
#pragma unroll
                for (int s = 0; s < kSlots; s++)
                    if (lane + kWave * s == i) key[s] = ~0ull;
            }
        }
        for (int i = kk + lane; i < A.k_stride; i += kWave) {
            oid[i] = -1;
            odist[i] = __builtin_inff();
        }
        if (lane == 0 && A.out_count) A.out_count[q] = cnt < 0 ? -1 : kk;
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ void k_zero_i32(int32_t* p) {
    if (threadIdx.x == 0) *p = 0;
}

hipError_t launch_rerank(const RerankArgs& a, int metric, int nel, int grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, a.qcounter);
    if (metric == MET_L2) {
        if (nel <= 8)
            hipLaunchKernelGGL((k_rerank<MET_L2, 8>), dim3(grid), dim3(kRerankBlock), lds, s, a);
        else
            hipLaunchKernelGGL((k_rerank<MET_L2, 16>), dim3(grid), dim3(kRerankBlock), lds, s, a);
    } else {
        if (nel <= 8)
            hipLaunchKernelGGL((k_rerank<MET_IP, 8>), dim3(grid), dim3(kRerankBlock), lds, s, a);
        else
            hipLaunchKernelGGL((k_rerank<MET_IP, 16>), dim3(grid), dim3(kRerankBlock), lds, s, a);
    }
    return hipGetLastError();
}

int quantize_rows(pk_index* h, const float* d_in, int64_t in_stride, int64_t n, uint8_t* d_out, int64_t out_stride) {
    if (n <= 0) return PK_OK;
    const int64_t per_row = out_stride / (h->dtype == PK_FLOAT16 ? 2 : 1);
    const int64_t total = n * per_row;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, (int64_t)h->num_cus * 16);
    if (h->dtype == PK_FLOAT16)
        hipLaunchKernelGGL(k_quantize<DT_F16>, dim3(grid), dim3(256), 0, h->stream, d_in, in_stride, n, h->dim, d_out,
                           out_stride, h->q_offset, h->q_inv_scale);
    else
        hipLaunchKernelGGL(k_quantize<DT_I8>, dim3(grid), dim3(256), 0, h->stream, d_in, in_stride, n, h->dim, d_out,
                           out_stride, h->q_offset, h->q_inv_scale);
    PCHK(hipGetLastError());
    return PK_OK;
}

int check_precision_form(pk_index* h) {
    if (h->dtype != PK_FLOAT16 && h->dtype != PK_INT8)
        return set_error(PK_ERR_ARG, "the precision form needs an FP16 or int8 traversal index");
    if (h->dtype == PK_INT8 && !h->has_quantizer)
        return set_error(PK_ERR_STATE, "set the int8 quantizer (pk_index_set_quantizer) first");
    if (h->dim > 256) return set_error(PK_ERR_ARG, "the FP32 re-rank supports dim <= 256");
    return PK_OK;
}

}  // namespace
}  // namespace pk

using namespace pk;

extern "C" {

int pk_index_set_quantizer(pk_index* h, const float* offset, const float* scale) {
    if (!h || !offset || !scale) return set_error(PK_ERR_ARG, "null argument");
    if (h->dtype != PK_INT8) return set_error(PK_ERR_ARG, "a quantizer applies to an int8 traversal index");
    std::vector<float> inv(h->dim);
    for (int d = 0; d < h->dim; d++) {
        if (!(scale[d] > 0.0f) || !std::isfinite(scale[d]) || !std::isfinite(offset[d]))
            return set_error(PK_ERR_ARG, "scale[" + std::to_string(d) + "] must be positive and finite");
        inv[d] = 1.0f / scale[d];
    }
    int rc;
    if (!h->q_offset && ((rc = palloc(&h->q_offset, h->dim)) || (rc = palloc(&h->q_inv_scale, h->dim)))) return rc;
    PCHK(hipMemcpy(h->q_offset, offset, h->dim * sizeof(float), hipMemcpyHostToDevice));
    PCHK(hipMemcpy(h->q_inv_scale, inv.data(), h->dim * sizeof(float), hipMemcpyHostToDevice));
    h->has_quantizer = true;
    return PK_OK;
}

int pk_index_set_full_rows(pk_index* h, const float* rows, int64_t first, int64_t n) {
    if (!h || (!rows && n > 0)) return set_error(PK_ERR_ARG, "null argument");
    int rc = check_precision_form(h);
    if (rc) return rc;
    if (first < 0 || n < 0 || first + n > h->capacity) return set_error(PK_ERR_ARG, "rows outside the capacity");
    h->full_vstride = (h->dim * 4 + 15) / 16 * 16;
    if (!h->full) {
        if ((rc = palloc(&h->full, (size_t)h->capacity * h->full_vstride))) return rc;
        PCHK(hipMemsetAsync(h->full, 0, (size_t)h->capacity * h->full_vstride, h->stream));
    }
    if (!h->base) {
        if ((rc = palloc(&h->base, (size_t)h->capacity * h->vstride))) return rc;
        PCHK(hipMemsetAsync(h->base, 0, (size_t)h->capacity * h->vstride, h->stream));
    }
    if (n > 0) {
        uint8_t* dst = h->full + (size_t)first * h->full_vstride;
        PCHK(upload_rows(dst, h->full_vstride, rows, (size_t)h->dim * 4, n, h->stream));
        if ((rc = quantize_rows(h, reinterpret_cast<const float*>(dst), h->full_vstride / 4, n,
                                h->base + (size_t)first * h->vstride, h->vstride)))
            return rc;
    }
    PCHK(hipStreamSynchronize(h->stream));
    h->full_rows = std::max(h->full_rows, first + n);
    h->base_rows = std::max(h->base_rows, first + n);
    return PK_OK;
}

int pk_index_copy_graph(pk_index* dst, pk_index* src) {
    if (!dst || !src || dst == src) return set_error(PK_ERR_ARG, "two distinct handles needed");
    if (!src->ready) return set_error(PK_ERR_STATE, "the source index holds no complete graph");
    if (dst->capacity != src->capacity || dst->dim != src->dim || dst->M != src->M)
        return set_error(PK_ERR_ARG, "indexes differ in capacity, dim or M");
    if (!dst->base) return set_error(PK_ERR_STATE, "set the destination's rows first");
    if (dst->graphs_alive || dst->servers_alive)
        return set_error(PK_ERR_STATE, "destroy the destination's search graphs and serving loops first");
    PCHK(hipStreamSynchronize(src->stream));
    PCHK(hipStreamSynchronize(dst->stream));
    // the destination's layers are freed below: no search may use them until
    // pk_index_set_enter_level publishes the copy, whatever fails meanwhile
    dst->ready = false;
    int rc;
    for (int l = 0; l < PK_MAX_LEVELS; l++) {
        Layer& S = src->layers[l];
        Layer& D = dst->layers[l];
        free_layer_public(D);
        if (!S.loaded) continue;
        const int64_t pad = 64;
        D.n = S.n;
        D.cap = S.n;
        D.D = S.D;
        D.adj_len = S.adj_len;  // rows may be read past D, up to the source's allocation
        if ((rc = palloc(&D.adj, (size_t)(D.adj_len + pad))) || (rc = palloc(&D.deg, (size_t)D.n)) ||
            (rc = palloc(&D.heur, (size_t)D.n)) || (rc = palloc(&D.ext, (size_t)D.n)) ||
            (rc = palloc(&D.vec, (size_t)D.n * dst->vstride)) || (rc = palloc(&D.has, (size_t)D.n)) ||
            (rc = palloc(&D.down, (size_t)D.n)))
            return rc;
        PCHK(hipMemcpyAsync(D.adj, S.adj, D.adj_len * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
        PCHK(hipMemsetAsync(D.adj + D.adj_len, 0xff, pad * sizeof(int32_t), dst->stream));
        PCHK(hipMemcpyAsync(D.deg, S.deg, D.n * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
        PCHK(hipMemcpyAsync(D.heur, S.heur, D.n * sizeof(int16_t), hipMemcpyDeviceToDevice, dst->stream));
        PCHK(hipMemcpyAsync(D.ext, S.ext, D.n * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
        PCHK(hipMemcpyAsync(D.has, S.has, D.n, hipMemcpyDeviceToDevice, dst->stream));
        PCHK(hipMemcpyAsync(D.down, S.down, D.n * sizeof(int32_t), hipMemcpyDeviceToDevice, dst->stream));
        D.loaded = true;
        D.built = true;  // down maps copied, not recomputed
        if ((rc = gather_layer_vectors(dst, l))) return rc;
    }
    PCHK(hipStreamSynchronize(dst->stream));
    dst->building = false;
    dst->ready = false;
    return pk_index_set_enter_level(dst, src->enter_level);
}

}  // extern "C"

// pk_search_rerank_device with the traversal's visited-table size set for
// this call only (vbits_override > 0: the rerun of queries that filled
// their tables, on no more waves than the allocated scratch holds)
static int rerank_device(pk_index* h, const float* d_queries, int64_t nq, int32_t k, int32_t ef, int32_t k_stride,
                         int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters,
                         int vbits_override, bool fast = false) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    int rc = check_precision_form(h);
    if (rc) return rc;
    if (!h->full || h->full_rows == 0) return set_error(PK_ERR_STATE, "no FP32 rows (pk_index_set_full_rows)");
    if (nq < 0 || (nq > 0 && (!d_queries || !d_out_ids || !d_out_dists))) return set_error(PK_ERR_ARG, "bad buffers");
    if (k < 1 || k > k_stride || ef < k) return set_error(PK_ERR_ARG, "need 1 <= k <= k_stride and ef >= k");
    if (ef > 630) return set_error(PK_ERR_ARG, "ef above 630");
    if (nq == 0) return PK_OK;
    void *qc, *cids, *cd, *cc, *ctr;
    if ((rc = scratch(h->rr[0], (size_t)nq * h->vstride, &qc)) ||
        (rc = scratch(h->rr[1], (size_t)nq * ef * 4, &cids)) || (rc = scratch(h->rr[2], (size_t)nq * ef * 4, &cd)) ||
        (rc = scratch(h->rr[3], (size_t)nq * 4, &cc)) || (rc = scratch(h->rr[4], 64, &ctr)))
        return rc;
    // queries -> traversal type (rows of qbytes, as the search reads them)
    if ((rc = quantize_rows(h, d_queries, h->dim, nq, static_cast<uint8_t*>(qc), h->qbytes))) return rc;
    if (fast)
        rc = pk_search_fast_device(h, qc, nq, nullptr, nullptr, ef, ef, ef, ef, static_cast<int32_t*>(cids),
                                   static_cast<float*>(cd), static_cast<int32_t*>(cc), d_counters);
    else
        rc = search_on(h, h->sc, qc, nq, nullptr, nullptr, ef, ef, ef, ef, static_cast<int32_t*>(cids),
                       static_cast<float*>(cd), static_cast<int32_t*>(cc), d_counters, vbits_override);
    if (rc) return rc;
    RerankArgs a{};
    a.queries = d_queries;
    a.nq = nq;
    a.dim = h->dim;
    a.full = h->full;
    a.full_vstride = h->full_vstride;
    a.full_rows = h->full_rows;
    a.cand = static_cast<const int32_t*>(cids);
    a.cand_stride = ef;
    a.cand_cnt = static_cast<const int32_t*>(cc);
    a.k = k;
    a.k_stride = k_stride;
    a.out_ids = d_out_ids;
    a.out_dists = d_out_dists;
    a.out_count = d_out_count;
    a.qcounter = static_cast<int32_t*>(ctr);
    a.cap = (ef + 63) / 64 * 64;
    a.lds_per_wave = (a.cap * 8 + h->full_vstride + 15) / 16 * 16;
    const size_t lds = (size_t)a.lds_per_wave * (kRerankBlock / kWave);
    const int nel = (h->dim + 15) / 16;
    const int grid = (int)std::min<int64_t>((nq + 1) / 2, (int64_t)h->num_cus * 8);
    PCHK(launch_rerank(a, h->metric, nel, grid, lds, h->stream));
    return PK_OK;
}

extern "C" {

int pk_search_rerank_device(pk_index* h, const float* d_queries, int64_t nq, int32_t k, int32_t ef,
                            int32_t k_stride, int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count,
                            pk_query_counters* d_counters) {
    return rerank_device(h, d_queries, nq, k, ef, k_stride, d_out_ids, d_out_dists, d_out_count, d_counters, 0);
}

int pk_search_rerank_fast_device(pk_index* h, const float* d_queries, int64_t nq, int32_t k, int32_t ef,
                                 int32_t k_stride, int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count,
                                 pk_query_counters* d_counters) {
    if (ef > 512) return set_error(PK_ERR_ARG, "fast mode takes ef in [k, 512]");
    return rerank_device(h, d_queries, nq, k, ef, k_stride, d_out_ids, d_out_dists, d_out_count, d_counters, 0,
                         true);
}

int pk_search_rerank(pk_index* h, const float* queries, int64_t nq, int32_t k, int32_t ef, int32_t k_stride,
                     int32_t* out_ids, float* out_dists, int32_t* out_count, pk_query_counters* counters) {
    if (!h) return set_error(PK_ERR_ARG, "null handle");
    if (nq < 0 || (nq > 0 && (!queries || !out_ids || !out_dists))) return set_error(PK_ERR_ARG, "bad buffers");
    if (k < 1 || k > k_stride || ef < k) return set_error(PK_ERR_ARG, "need 1 <= k <= k_stride and ef >= k");
    if (nq == 0) return PK_OK;
    void *dq = nullptr, *di = nullptr, *dd = nullptr, *dc = nullptr, *dctr = nullptr;
    int rc;
    if ((rc = palloc(reinterpret_cast<uint8_t**>(&dq), (size_t)nq * h->dim * 4)) ||
        (rc = palloc(reinterpret_cast<uint8_t**>(&di), (size_t)nq * k_stride * 4)) ||
        (rc = palloc(reinterpret_cast<uint8_t**>(&dd), (size_t)nq * k_stride * 4)) ||
        (rc = palloc(reinterpret_cast<uint8_t**>(&dc), (size_t)nq * 4)) ||
        (counters && (rc = palloc(reinterpret_cast<uint8_t**>(&dctr), (size_t)nq * sizeof(pk_query_counters))))) {
        for (void* p : {dq, di, dd, dc, dctr})
            if (p) (void)hipFree(p);
        return rc;
    }
    auto done = [&](int code) {
        for (void* p : {dq, di, dd, dc, dctr})
            if (p) (void)hipFree(p);
        return code;
    };
    if (hipMemcpy(dq, queries, (size_t)nq * h->dim * 4, hipMemcpyHostToDevice) != hipSuccess)
        return done(set_error(PK_ERR_HIP, "query upload failed"));
    rc = pk_search_rerank_device(h, static_cast<const float*>(dq), nq, k, ef, k_stride, static_cast<int32_t*>(di),
                                 static_cast<float*>(dd), static_cast<int32_t*>(dc),
                                 static_cast<pk_query_counters*>(dctr));
    if (rc) return done(rc);
    std::vector<int32_t> cnt((size_t)nq);
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(out_ids, di, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(out_dists, dd, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(cnt.data(), dc, (size_t)nq * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        (counters && hipMemcpy(counters, dctr, (size_t)nq * sizeof(pk_query_counters), hipMemcpyDeviceToHost) != hipSuccess))
        return done(set_error(PK_ERR_HIP, "result download failed"));
    int32_t flags = 0;
    if (hipMemcpy(&flags, h->sc.err, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
        return done(set_error(PK_ERR_HIP, "failure flags download failed"));
    // Queries whose traversal filled its visited table (flag 1 alone, count
    // -1) run again with tables twice the size, up to 2^20 slots, as in
    // pk_search; the size does not change results.  Not while captured
    // search graphs pin the scratch.
    const int dim = h->dim;
    int vb = h->sc.last_vbits;
    while (flags == 1 && vb < 20 && h->graphs_alive == 0) {
        vb++;
        std::vector<int64_t> bad;
        for (int64_t i = 0; i < nq; i++)
            if (cnt[i] < 0) bad.push_back(i);
        const int64_t nb = (int64_t)bad.size();
        if (nb == 0) break;
        std::vector<float> q2((size_t)nb * dim);
        for (int64_t j = 0; j < nb; j++) std::memcpy(&q2[(size_t)j * dim], queries + bad[j] * dim, (size_t)dim * 4);
        if (hipMemcpy(dq, q2.data(), q2.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return done(set_error(PK_ERR_HIP, "query upload failed"));
        rc = rerank_device(h, static_cast<const float*>(dq), nb, k, ef, k_stride, static_cast<int32_t*>(di),
                           static_cast<float*>(dd), static_cast<int32_t*>(dc), static_cast<pk_query_counters*>(dctr),
                           vb);
        if (rc) return done(rc);
        std::vector<int32_t> ids2((size_t)nb * k_stride), cnt2((size_t)nb);
        std::vector<float> d2((size_t)nb * k_stride);
        std::vector<pk_query_counters> ctr2(counters ? nb : 0);
        if (hipStreamSynchronize(h->stream) != hipSuccess ||
            hipMemcpy(ids2.data(), di, ids2.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(d2.data(), dd, d2.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(cnt2.data(), dc, cnt2.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (counters && hipMemcpy(ctr2.data(), dctr, (size_t)nb * sizeof(pk_query_counters), hipMemcpyDeviceToHost) !=
                             hipSuccess) ||
            hipMemcpy(&flags, h->sc.err, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess)
            return done(set_error(PK_ERR_HIP, "result download failed"));
        for (int64_t j = 0; j < nb; j++) {
            std::memcpy(out_ids + bad[j] * k_stride, &ids2[(size_t)j * k_stride], (size_t)k_stride * 4);
            std::memcpy(out_dists + bad[j] * k_stride, &d2[(size_t)j * k_stride], (size_t)k_stride * 4);
            cnt[bad[j]] = cnt2[j];
            if (counters) counters[bad[j]] = ctr2[j];
        }
    }
    if (out_count) std::memcpy(out_count, cnt.data(), (size_t)nq * 4);
    if (flags)
        return done(set_error(PK_ERR_DEVICE, "traversal failed on the device (flags " + std::to_string(flags) +
                                                 ": 1 = scratch capacity, 2 = bad k/ef)"));
    return done(PK_OK);
}

}  // extern "C"
