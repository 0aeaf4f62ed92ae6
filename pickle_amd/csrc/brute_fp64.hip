// brute_fp64.hip -- exhaustive nearest neighbours of f16 / f32 rows on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64); the entry point, the integer
// tiles and the merge are in brute.hip.
#include "brute.hpp"

namespace pk {

// float rows: |x|^2 in FP64; rows past nbase get 0
template <int DT>
__global__ __launch_bounds__(256) void k_brute_rownorm(const uint8_t* __restrict__ base, int64_t nbase, int vstride,
                                                       int dim, int64_t npad, double* __restrict__ nrm) {
    constexpr int E = DT == DT_F32 ? 4 : 8;
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= npad) return;
    double s = 0.0;
    if (r < nbase) {
        const uint8_t* rp = base + r * (int64_t)vstride;
        for (int c = 0; c * E < dim; c++) {
            const uint4 u = ld16_global(rp + 16 * c);
#pragma unroll
            for (int j = 0; j < E; j++) {
                const double v = c * E + j < dim ? chunk_elem<DT>(u, j) : 0.0;
                s = __builtin_fma(v, v, s);
            }
        }
    }
    nrm[r] = s;
}

// ---- FP64 tiles ----------------------------------------------------------

// Queries are staged NSC * 64 bytes apart, zero padded.
// One-dimensional grid of qblocks x splits workgroups (brute_block_coords):
// query blocks of kFltQ queries times row splits of tiles_per_split tiles.
// Lists: [q][split * 4 + lane / 16][kBruteK].
template <int DT, int MET, int NSC>
__global__ __launch_bounds__(256) void k_brute_dmma(const uint8_t* __restrict__ base, int64_t nbase, int vstride,
                                                    const double* __restrict__ nrm, const uint8_t* __restrict__ qs,
                                                    int64_t nq, int64_t qblocks, int xcd_map, int64_t tiles_per_split,
                                                    int nlists, double* __restrict__ pd, int32_t* __restrict__ pi) {
    constexpr int KB = kBruteK;
    constexpr int QST = NSC * 64;
    constexpr int E = DT == DT_F32 ? 4 : 8;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, col = lane & 15;
    int64_t qb, split;
    brute_block_coords(qblocks, xcd_map, qb, split);
    const int64_t q = qb * kFltQ + (threadIdx.x >> 6) * 16 + col;
    const int64_t qc = q < nq ? q : nq - 1;

    uint4 bq[NSC];
    double qq = 0.0;
#pragma unroll
    for (int sc = 0; sc < NSC; sc++) {
        bq[sc] = ld16_global(qs + qc * QST + sc * 64 + g * 16);
#pragma unroll
        for (int j = 0; j < E; j++) {
            const double v = chunk_elem<DT>(bq[sc], j);
            qq = __builtin_fma(v, v, qq);
        }
    }
    qq += __shfl_xor(qq, 16, kWave);
    qq += __shfl_xor(qq, 32, kWave);

    int coff[NSC];
#pragma unroll
    for (int sc = 0; sc < NSC; sc++) coff[sc] = sc * 64 + g * 16 < vstride ? sc * 64 + g * 16 : 0;

    double kd[KB];
    int ki[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) {
        kd[j] = __builtin_inf();
        ki[j] = INT_MAX;
    }
    double thr = __builtin_inf();

    const int64_t ntiles = (nbase + 15) / 16;
    const int64_t t0 = split * tiles_per_split;
    const int64_t t1 = t0 + tiles_per_split < ntiles ? t0 + tiles_per_split : ntiles;

    uint4 an[NSC];
    double nn[4];
    auto load_tile = [&](int64_t t) {
        int64_t row = t * 16 + col;
        row = row < nbase ? row : nbase - 1;
        const uint8_t* rp = base + row * (int64_t)vstride;
#pragma unroll
        for (int sc = 0; sc < NSC; sc++) an[sc] = ld16_global(rp + coff[sc]);
        if (MET == MET_L2) {
            const gmem<double>* np = gbl(nrm + t * 16 + g);
#pragma unroll
            for (int r = 0; r < 4; r++) nn[r] = np[4 * r];
        }
    };
    if (t0 < t1) load_tile(t0);
    for (int64_t t = t0; t < t1; t++) {
        uint4 a[NSC];
#pragma unroll
        for (int sc = 0; sc < NSC; sc++) a[sc] = an[sc];
        double nr[4];
#pragma unroll
        for (int r = 0; r < 4; r++) nr[r] = nn[r];
        if (t + 1 < t1) load_tile(t + 1);

        v4d acc = {};
#pragma unroll
        for (int sc = 0; sc < NSC; sc++)
#pragma unroll
            for (int j = 0; j < E; j++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(chunk_elem<DT>(a[sc], j), chunk_elem<DT>(bq[sc], j), acc,
                                                           0, 0, 0);

        // acc[r] is row t*16 + g + 4 r against query col (the FP64 16x16
        // output layout; the FP32 16x16 one differs)
        double key[4];
        bool any = false;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            key[r] = MET == MET_L2 ? __builtin_fma(-2.0, acc[r], nr[r]) : -acc[r];
            any |= key[r] < thr;
        }
        if (ballot(any)) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t m = t * 16 + g + 4 * r;
                if (key[r] < thr && m < nbase) {
                    sorted_insert(kd, ki, key[r], (int)m);
                    thr = kd[KB - 1];
                }
            }
        }
    }

    if (q < nq) {
        const size_t o = ((size_t)q * nlists + (size_t)split * 4 + g) * KB;
#pragma unroll
        for (int j = 0; j < KB; j++) {
            pd[o + j] = ki[j] == INT_MAX ? __builtin_inf() : (MET == MET_L2 ? kd[j] + qq : kd[j]);
            pi[o + j] = ki[j];
        }
    }
}

template <int DT, int MET>
static hipError_t launch_flt(int nsc, const BrutePlan& p, hipStream_t s, const uint8_t* base, int64_t nbase,
                             int vstride, const double* nrm, const uint8_t* dq, int64_t nq, double* pd, int32_t* pi) {
    const dim3 grid((unsigned)(p.qblocks * p.splits));
    const int xcd_map = p.splits % kXcds == 0;
#define PK_DMMA_CASE(N)                                                                                         \
    if (nsc <= N) {                                                                                             \
        hipLaunchKernelGGL((k_brute_dmma<DT, MET, N>), grid, dim3(256), 0, s, base, nbase, vstride, nrm, dq, nq, \
                           p.qblocks, xcd_map, p.tiles_per_split, p.nlists, pd, pi);                           \
        return hipGetLastError();                                                                               \
    }
    PK_DMMA_CASE(1)
    PK_DMMA_CASE(2)
    PK_DMMA_CASE(3)
    PK_DMMA_CASE(4)
    PK_DMMA_CASE(6)
    PK_DMMA_CASE(8)
#undef PK_DMMA_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_brute_fp64(int dtype, int metric, int nsc, const BrutePlan& p, hipStream_t s,
                             const uint8_t* base, int64_t nbase, int vstride, int dim, int64_t npad, double* nrm,
                             const uint8_t* dq, int64_t nq, double* pd, int32_t* pi) {
    const dim3 rgrid((unsigned)(npad / 256 + 1));
    if (metric == MET_L2) {
        if (dtype == DT_F32)
            hipLaunchKernelGGL((k_brute_rownorm<DT_F32>), rgrid, dim3(256), 0, s, base, nbase, vstride, dim, npad, nrm);
        else
            hipLaunchKernelGGL((k_brute_rownorm<DT_F16>), rgrid, dim3(256), 0, s, base, nbase, vstride, dim, npad, nrm);
    }
    switch (dtype * 4 + metric) {
        case DT_F32 * 4 + MET_L2: return launch_flt<DT_F32, MET_L2>(nsc, p, s, base, nbase, vstride, nrm, dq, nq, pd, pi);
        case DT_F32 * 4 + MET_IP: return launch_flt<DT_F32, MET_IP>(nsc, p, s, base, nbase, vstride, nullptr, dq, nq, pd, pi);
        case DT_F16 * 4 + MET_L2: return launch_flt<DT_F16, MET_L2>(nsc, p, s, base, nbase, vstride, nrm, dq, nq, pd, pi);
        case DT_F16 * 4 + MET_IP: return launch_flt<DT_F16, MET_IP>(nsc, p, s, base, nbase, vstride, nullptr, dq, nq, pd, pi);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace pk
