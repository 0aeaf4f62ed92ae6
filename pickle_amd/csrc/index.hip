// index.hip -- the C-ABI (include/pickle_gpu.h): HBM-resident index state,
// graph loading / saving, and the search entry points.
//
// Reference: HNSWGraph (include/hnsw.hpp:23-220) and HNSWLayer
// (include/hnsw_layer.hpp:57-465) own the same state on the host; here it
// lives in device memory laid out for the kernels:
//   base        capacity rows x vstride bytes, by external id (m_data)
//   layer l     adj   n_l * D_l int32 (+ slots read past the last row)
//               deg   n_l int32, heur n_l int16, dist n_l * D_l float
//               ext   n_l int32 (internal -> external id)
//               vec   n_l rows x vstride (the layer's own copy, m_data)
//               down  n_l int32: internal id of the node one layer below
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pickle_gpu.h"
#include "index_impl.hpp"
#include "search.hpp"
#include "visited.hpp"

using namespace pk;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(PK_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

size_t elem_size(int dtype) {
    switch (dtype) {
        case PK_UINT8:
        case PK_INT8: return 1;
        case PK_FLOAT16: return 2;
        case PK_FLOAT32: return 4;
        default: return 0;
    }
}

__global__ void k_gather_rows(uint8_t* dst, const uint8_t* base, const int32_t* ext, const uint8_t* has, int64_t n,
                              int vstride, int64_t base_rows) {
    int64_t row = (int64_t)blockIdx.x * (blockDim.x / 16) + threadIdx.x / 16;
    int c = threadIdx.x % 16;
    if (row >= n) return;
    int chunks = vstride / 16;
    int32_t e = ext[row];
    bool ok = has[row] != 0 && e >= 0 && e < base_rows;
    uint4* d = reinterpret_cast<uint4*>(dst + row * (int64_t)vstride);
    const uint4* s = reinterpret_cast<const uint4*>(base + (int64_t)(ok ? e : 0) * vstride);
    for (int i = c; i < chunks; i += 16) d[i] = ok ? s[i] : make_uint4(0, 0, 0, 0);
}

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) return fail(PK_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    return PK_OK;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_layer(Layer& L) {
    dfree(L.lock);
    dfree(L.adj);
    dfree(L.deg);
    dfree(L.heur);
    dfree(L.dist);
    dfree(L.ext);
    dfree(L.vec);
    dfree(L.down);
    dfree(L.has);
    L = Layer{};
}

}  // namespace

namespace pk {
int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

void free_layer_public(Layer& L) { free_layer(L); }

// Uploads the layer descriptors the kernels read (sizes, pointers).  Waits
// for the index stream first, so no running kernel sees a half-updated table.
int publish_layers(pk_index* h) {
    struct LayerBuildDesc {
        int32_t* adj;
        int32_t* deg;
        int16_t* heur;
        float* dist;
        int32_t* ext;
        uint8_t* vec;
        uint8_t* has;
        int32_t* down;
        int32_t* lock;
        int64_t adj_len;
        int32_t D;
        int32_t pad_;
    };
    LayerDev desc[PK_MAX_LEVELS] = {};
    LayerBuildDesc bdesc[PK_MAX_LEVELS] = {};
    for (int l = 0; l < PK_MAX_LEVELS; l++) {
        const Layer& L = h->layers[l];
        if (!L.loaded) continue;
        desc[l] = LayerDev{L.adj, L.deg, L.vec, L.down, L.n, L.adj_len, L.D, 0};
        bdesc[l] = LayerBuildDesc{L.adj, L.deg, L.heur, L.dist, L.ext, L.vec, L.has, L.down, L.lock, L.adj_len, L.D, 0};
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!h->d_layers) HIPCHK(hipMalloc(&h->d_layers, sizeof(desc)));
    if (!h->d_build) HIPCHK(hipMalloc(&h->d_build, sizeof(bdesc)));
    HIPCHK(hipMemcpy(h->d_layers, desc, sizeof(desc), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_build, bdesc, sizeof(bdesc), hipMemcpyHostToDevice));
    return PK_OK;
}

// Per-node lists of a batch (R and Pruned per level).
static int ensure_insert_lists(pk_index* h, int64_t nodes, int maxl, int topk) {
    InsertScratch& s = h->ins;
    if (nodes > s.cap_nodes || maxl > s.cap_maxl || topk > s.topk) {
        dfree(s.d_ext);
        dfree(s.d_lvl);
        dfree(s.d_in);
        dfree(s.r_ids);
        dfree(s.r_d);
        dfree(s.r_cnt);
        s.cap_nodes = 0;
        int rc;
        size_t lists = (size_t)nodes * maxl * 2;
        if ((rc = dalloc(&s.d_ext, nodes)) || (rc = dalloc(&s.d_lvl, nodes)) ||
            (rc = dalloc(&s.d_in, (size_t)nodes * maxl)) || (rc = dalloc(&s.r_ids, lists * topk)) ||
            (rc = dalloc(&s.r_d, lists * topk)) || (rc = dalloc(&s.r_cnt, lists)))
            return rc;
        s.cap_nodes = nodes;
        s.cap_maxl = maxl;
        s.topk = topk;
    }
    return PK_OK;
}

// Visited tables, their logs and the candidate-heap spill areas of the
// construction searches.  The first call sizes them for `waves` waves at
// 2^vbits slots; a later, larger table (searches that filled theirs run
// again) runs on fewer waves in the same memory (insert_scratch_waves), so
// only a table too large for even one wave grows the allocation.  The
// candidate heap has room for 2^vbits entries per wave: it only holds
// first-visited ids, so it cannot fill before the visited table does.
static int ensure_insert_tables(pk_index* h, int waves, int vbits) {
    InsertScratch& s = h->ins;
    const size_t cg_stride = std::max<size_t>((size_t)h->cg_cap, (size_t)1 << vbits);
    if (s.v_slots > 0 && ((size_t)1 << vbits) <= s.v_slots && cg_stride <= s.cg_entries) return PK_OK;
    const size_t need_v = std::max(s.v_slots, (size_t)std::max(waves, 1) << vbits);
    const size_t need_cg = std::max(s.cg_entries, (size_t)std::max(waves, 1) * cg_stride);
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(s.cglob);
    dfree(s.vtab);
    dfree(s.vlog);
    s.v_slots = s.cg_entries = 0;
    int rc;
    if ((rc = dalloc(&s.cglob, need_cg)) || (rc = dalloc(&s.vtab, need_v)) || (rc = dalloc(&s.vlog, need_v)))
        return rc;
    HIPCHK(hipMemsetAsync(s.vtab, 0, need_v * sizeof(uint32_t), h->stream));
    s.v_slots = need_v;
    s.cg_entries = need_cg;
    return PK_OK;
}

int insert_scratch_waves(const pk_index* h, int vbits) {
    const InsertScratch& s = h->ins;
    const size_t cg_stride = std::max<size_t>((size_t)h->cg_cap, (size_t)1 << vbits);
    return (int)std::min<size_t>(s.v_slots >> vbits, s.cg_entries / cg_stride);
}

int ensure_insert_scratch(pk_index* h, int64_t nodes, int maxl, int topk, int waves, int vbits) {
    InsertScratch& s = h->ins;
    int rc;
    if ((rc = ensure_insert_lists(h, nodes, maxl, topk))) return rc;
    if ((rc = ensure_insert_tables(h, waves, vbits))) return rc;
    if (!s.counter) {
        // err[0]: flags; err[1..4]: the first failing node (ext id, level,
        // ids in its visited table, hops) for the error message
        if ((rc = dalloc(&s.counter, 1)) || (rc = dalloc(&s.err, 8))) return rc;
        HIPCHK(hipMemsetAsync(s.err, 0, 32, h->stream));
    }
    return PK_OK;
}

int upload_insert_batch(pk_index* h, const int32_t* ext, const int32_t* lvl, const int32_t* in, int64_t nb,
                        int maxl) {
    InsertScratch& s = h->ins;
    if (nb > s.cap_nodes || maxl > s.cap_maxl) {
        int rc = ensure_insert_lists(h, std::max(nb, s.cap_nodes), std::max(maxl, s.cap_maxl), s.topk);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpy(s.d_ext, ext, nb * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.d_lvl, lvl, nb * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.d_in, in, (size_t)nb * maxl * 4, hipMemcpyHostToDevice));
    return PK_OK;
}
}  // namespace pk

extern "C" {

const char* pk_last_error(void) { return g_err.c_str(); }

int pk_version(void) { return 100; }

int pk_index_create(int dtype, int metric, int dim, int M, int efc, int64_t capacity, pk_index** out) {
    if (!out) return fail(PK_ERR_ARG, "out is null");
    *out = nullptr;
    if (elem_size(dtype) == 0) return fail(PK_ERR_ARG, "unsupported dtype %d", dtype);
    if (metric != PK_L2 && metric != PK_IP) return fail(PK_ERR_ARG, "unsupported metric %d (L2=1, IP=2)", metric);
    if (dim < 1) return fail(PK_ERR_ARG, "dim must be positive");
    if (M < 2 || M > 16383) return fail(PK_ERR_ARG, "M out of range");
    if (capacity < 1 || capacity > INT32_MAX) return fail(PK_ERR_ARG, "capacity out of range");
    auto* h = new pk_index();
    h->dtype = dtype;
    h->metric = metric;
    h->dim = dim;
    h->M = M;
    h->efc = efc;
    h->capacity = capacity;
    h->qbytes = (int)(dim * elem_size(dtype));
    h->vstride = (h->qbytes + 15) / 16 * 16;
    h->size_class = size_class_for(dtype, dim, h->vstride);
    if (h->size_class < 0) {
        delete h;
        return fail(PK_ERR_ARG, "dim %d too large for this build (max 512 bytes per int row, 271 floats)", dim);
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        delete h;
        return fail(PK_ERR_HIP, "stream / event creation failed");
    }
    h->sc.stream = h->stream;
    int dev = 0;
    (void)hipGetDevice(&dev);
    h->device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) h->num_cus = prop.multiProcessorCount;
    if (h->num_cus <= 0) h->num_cus = 256;
    *out = h;
    return PK_OK;
}

int pk_index_destroy(pk_index* h) {
    if (h && (h->graphs_alive || h->servers_alive))
        return fail(PK_ERR_STATE, "destroy the index's search graphs and serving loops first");
    if (!h) return PK_OK;
    (void)hipStreamSynchronize(h->stream);
    for (auto& L : h->layers) free_layer(L);
    dfree(h->base);
    free_search_ctx(h->sc);
    if (h->d_layers) (void)hipFree(h->d_layers);
    if (h->d_build) (void)hipFree(h->d_build);
    dfree(h->ins.d_ext);
    dfree(h->ins.d_lvl);
    dfree(h->ins.d_in);
    dfree(h->ins.r_ids);
    dfree(h->ins.r_d);
    dfree(h->ins.r_cnt);
    dfree(h->ins.cglob);
    dfree(h->ins.vtab);
    dfree(h->ins.vlog);
    dfree(h->ins.counter);
    dfree(h->ins.err);
    for (auto& ex : h->ins.graph)
        if (ex) (void)hipGraphExecDestroy(ex);
    for (auto& b : h->io) dfree(b.p);
    for (auto& b : h->rr) dfree(b.p);
    dfree(h->full);
    dfree(h->q_offset);
    dfree(h->q_inv_scale);
    (void)hipEventDestroy(h->ev0);
    (void)hipEventDestroy(h->ev1);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return PK_OK;
}

int pk_index_set_base(pk_index* h, const void* rows, int64_t first, int64_t n) {
    if (!h || (!rows && n > 0)) return fail(PK_ERR_ARG, "null argument");
    if (first < 0 || n < 0 || first + n > h->capacity) return fail(PK_ERR_ARG, "rows [%lld, %lld) outside capacity %lld",
                                                                   (long long)first, (long long)(first + n),
                                                                   (long long)h->capacity);
    if (!h->base) {
        int rc = dalloc(&h->base, (size_t)h->capacity * h->vstride);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(h->base, 0, (size_t)h->capacity * h->vstride, h->stream));
    }
    if (n > 0) HIPCHK(upload_rows(h->base + (size_t)first * h->vstride, h->vstride, rows, h->qbytes, n, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->base_rows = std::max(h->base_rows, first + n);
    return PK_OK;
}

}  // extern "C"

namespace pk {
// A pitched copy from pageable host memory goes row by row (microseconds per
// row: over a minute for 10M rows), so rows are copied flat when the pitches
// match and otherwise padded on the host in 64 MB pieces.
hipError_t upload_rows(uint8_t* dst, size_t dst_pitch, const void* src, size_t row_bytes, int64_t n, hipStream_t s) {
    const uint8_t* p = static_cast<const uint8_t*>(src);
    if (row_bytes == dst_pitch) {
        hipError_t e = hipMemcpyAsync(dst, p, (size_t)n * row_bytes, hipMemcpyHostToDevice, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    const int64_t chunk = std::max<int64_t>(1, (int64_t)((64u << 20) / dst_pitch));
    std::vector<uint8_t> tmp;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t m = std::min(chunk, n - r0);
        tmp.assign((size_t)m * dst_pitch, 0);  // padding bytes are zero
        for (int64_t i = 0; i < m; i++)
            std::memcpy(tmp.data() + (size_t)i * dst_pitch, p + (size_t)(r0 + i) * row_bytes, row_bytes);
        hipError_t e = hipMemcpyAsync(dst + (size_t)r0 * dst_pitch, tmp.data(), tmp.size(), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// the layer's own vector copy: base row ext[i] where has[i], zeros elsewhere
int gather_layer_vectors(pk_index* h, int level) {
    Layer& L = h->layers[level];
    if (L.n > 0) {
        int rows_per_block = 16;
        int grid = (int)((L.n + rows_per_block - 1) / rows_per_block);
        hipLaunchKernelGGL(k_gather_rows, dim3(grid), dim3(256), 0, h->stream, L.vec, h->base, L.ext, L.has, L.n,
                           h->vstride, h->capacity);
        HIPCHK(hipGetLastError());
    }
    return PK_OK;
}
}  // namespace pk

extern "C" {

int pk_index_load_layer(pk_index* h, int level, int64_t n, int D, int64_t adj_len, const int32_t* ext,
                        const int32_t* degree, const int16_t* heur, const int32_t* adj, const float* dist,
                        const uint8_t* has_vector) {
    if (!h || !ext || !degree || !adj || !has_vector) return fail(PK_ERR_ARG, "null argument");
    if (h->graphs_alive || h->servers_alive)
        return fail(PK_ERR_STATE, "destroy the search graphs and serving loops before replacing layers");
    if (level < 0 || level >= PK_MAX_LEVELS) return fail(PK_ERR_ARG, "level %d out of range", level);
    if (n < 0 || n > h->capacity || D < 1 || adj_len < n * D) return fail(PK_ERR_ARG, "bad layer shape");
    if (!h->base) return fail(PK_ERR_STATE, "set the base rows before loading layers");
    Layer& L = h->layers[level];
    // no search may run on the layer from here until pk_index_set_enter_level
    // republishes a complete graph, whatever fails below
    h->ready = false;
    free_layer(L);
    L.n = n;
    L.cap = n;
    L.D = D;
    L.adj_len = adj_len;
    L.built = false;
    h->building = false;
    const int64_t pad = 64;  // a chunk may read up to one wave past the degree
    int rc;
    if ((rc = dalloc(&L.adj, (size_t)(adj_len + pad))) || (rc = dalloc(&L.deg, (size_t)n)) ||
        (rc = dalloc(&L.heur, (size_t)n)) || (rc = dalloc(&L.dist, (size_t)n * D)) ||
        (rc = dalloc(&L.ext, (size_t)n)) || (rc = dalloc(&L.vec, (size_t)n * h->vstride)) ||
        (rc = dalloc(&L.has, (size_t)n)) || (rc = dalloc(&L.down, (size_t)n)))
        return rc;
    HIPCHK(hipMemsetAsync(L.adj + adj_len, 0xff, pad * sizeof(int32_t), h->stream));
    HIPCHK(hipMemcpyAsync(L.adj, adj, adj_len * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(L.deg, degree, n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    if (heur)
        HIPCHK(hipMemcpyAsync(L.heur, heur, n * sizeof(int16_t), hipMemcpyHostToDevice, h->stream));
    else
        HIPCHK(hipMemsetAsync(L.heur, 0, n * sizeof(int16_t), h->stream));
    if (dist)
        HIPCHK(hipMemcpyAsync(L.dist, dist, (size_t)n * D * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(L.ext, ext, n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(L.has, has_vector, n, hipMemcpyHostToDevice, h->stream));
    if ((rc = gather_layer_vectors(h, level))) return rc;
    L.h_ext.assign(ext, ext + n);
    HIPCHK(hipStreamSynchronize(h->stream));
    L.loaded = true;
    h->ready = false;
    return PK_OK;
}

int pk_index_set_enter_level(pk_index* h, int enter_level) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (enter_level < 0 || enter_level >= PK_MAX_LEVELS) return fail(PK_ERR_ARG, "enter level out of range");
    for (int l = 0; l <= enter_level; l++)
        if (!h->layers[l].loaded || h->layers[l].n == 0)
            return fail(PK_ERR_STATE, "layer %d is not loaded (or empty) below the enter level", l);
    // down maps: GetEntInID(GetExtID(v, l), l - 1) (include/hnsw.hpp:515-521)
    std::vector<int32_t> ext2in((size_t)h->capacity, -1), down;
    for (int l = 1; l <= enter_level; l++) {
        Layer& lo = h->layers[l - 1];
        Layer& up = h->layers[l];
        if (up.built) continue;  // the insertion kernels write these
        std::fill(ext2in.begin(), ext2in.end(), -1);
        for (int64_t i = 0; i < lo.n; i++) {
            int32_t e = lo.h_ext[i];
            if (e >= 0 && e < h->capacity) ext2in[e] = (int32_t)i;
        }
        down.assign(up.n, 0);
        for (int64_t i = 0; i < up.n; i++) {
            int32_t e = up.h_ext[i];
            int32_t d = (e >= 0 && e < h->capacity) ? ext2in[e] : -1;
            if (d < 0) return fail(PK_ERR_ARG, "node %lld of layer %d (ext %d) is missing from layer %d", (long long)i, l,
                                   e, l - 1);
            down[i] = d;
        }
        HIPCHK(hipMemcpy(up.down, down.data(), up.n * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    int rc = publish_layers(h);
    if (rc) return rc;
    h->enter_level = enter_level;
    h->ready = true;
    return PK_OK;
}

int pk_index_info(pk_index* h, int* nlayers, int* enter_level, int64_t* sizes, int* degrees) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    int nl = 0;
    for (int l = 0; l < PK_MAX_LEVELS; l++)
        if (h->layers[l].loaded) nl = l + 1;
    if (nlayers) *nlayers = nl;
    if (enter_level) *enter_level = h->enter_level;
    for (int l = 0; l < nl; l++) {
        if (sizes) sizes[l] = h->layers[l].n;
        if (degrees) degrees[l] = h->layers[l].D;
    }
    return PK_OK;
}

int pk_index_read_layer(pk_index* h, int level, int32_t* ext, int32_t* degree, int16_t* heur, int32_t* adj,
                        float* dist) {
    if (!h || level < 0 || level >= PK_MAX_LEVELS || !h->layers[level].loaded)
        return fail(PK_ERR_ARG, "no such layer");
    Layer& L = h->layers[level];
    HIPCHK(hipStreamSynchronize(h->stream));
    if (ext) HIPCHK(hipMemcpy(ext, L.ext, L.n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (degree) HIPCHK(hipMemcpy(degree, L.deg, L.n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (heur) HIPCHK(hipMemcpy(heur, L.heur, L.n * sizeof(int16_t), hipMemcpyDeviceToHost));
    if (adj) HIPCHK(hipMemcpy(adj, L.adj, (size_t)L.n * L.D * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (dist) HIPCHK(hipMemcpy(dist, L.dist, (size_t)L.n * L.D * sizeof(float), hipMemcpyDeviceToHost));
    return PK_OK;
}

// ---------------------------------------------------------------------------
// search
// ---------------------------------------------------------------------------

}  // extern "C"

namespace pk {
void free_search_ctx(SearchCtx& c) {
    dfree(c.cglob);
    dfree(c.vtab);
    dfree(c.vlog);
    dfree(c.qcounter);
    dfree(c.err);
    dfree(c.order);
    c.v_slots = c.cg_entries = c.order_ints = 0;
}

// Scratch for a launch of `waves` waves with 2^vbits-slot visited tables.
// Each buffer is kept as large as the largest launch so far needed, counted
// in slots / entries whatever the layout: the tables are all zero between
// searches, so any later layout fits inside them, and a launch of a few
// waves with large tables (the queries searched again after filling theirs)
// needs no more memory than a full launch with small ones.
static int ensure_scratch(pk_index* h, SearchCtx& c, int waves, int vbits) {
    // the candidate heap only holds first-visited ids, so with room for
    // 2^vbits entries per wave it cannot fill before the visited table does
    const int cg_stride = std::max(h->cg_cap, 1 << vbits);
    const size_t need_v = (size_t)waves << vbits, need_cg = (size_t)waves * cg_stride;
    if (need_v <= c.v_slots && need_cg <= c.cg_entries && c.qcounter && c.err) {
        c.cg_cap = cg_stride;
        return PK_OK;
    }
    if (&c == &h->sc && h->graphs_alive)
        return fail(PK_ERR_STATE, "this search needs more scratch than the captured search graphs use; destroy them "
                                  "first");
    HIPCHK(hipStreamSynchronize(c.stream));
    int rc;
    if (need_cg > c.cg_entries) {
        dfree(c.cglob);
        c.cg_entries = 0;
        if ((rc = dalloc(&c.cglob, need_cg))) return rc;
        c.cg_entries = need_cg;
    }
    if (need_v > c.v_slots) {
        dfree(c.vtab);
        dfree(c.vlog);
        c.v_slots = 0;
        if ((rc = dalloc(&c.vtab, need_v)) || (rc = dalloc(&c.vlog, need_v))) return rc;
        HIPCHK(hipMemsetAsync(c.vtab, 0, need_v * sizeof(uint32_t), c.stream));
        c.v_slots = need_v;
    }
    if (!c.qcounter && (rc = dalloc(&c.qcounter, 1))) return rc;
    if (!c.err && (rc = dalloc(&c.err, 1))) return rc;
    c.cg_cap = cg_stride;
    return PK_OK;
}
}  // namespace pk

extern "C" {

}  // extern "C"

namespace pk {
struct SearchPlan {
    SearchArgs a;
    int grid = 0;
    size_t lds = 0;
    int vbits = 0;  // visited-table size of the launch (log2 slots per wave)
};

// validation, LDS plan and scratch (the part that may allocate)
// Visited-table size per wave (2^bits slots): room for the ids an ef-wide
// beam visits on large graphs (hard queries on 10M rows visit over 2^15 at
// ef 128), and never more than twice the capacity, which cannot fill.
static int default_vbits(int ef_max, int64_t capacity) {
    int v = ef_max <= 32 ? 14 : (ef_max <= 64 ? 15 : (ef_max <= 256 ? 16 : (ef_max <= 512 ? 17 : 18)));
    int cap_bits = 10;
    while (cap_bits < 20 && ((int64_t)1 << cap_bits) < 2 * capacity) cap_bits++;
    return std::min(v, cap_bits);
}

// An FP16 index with FP32 rows attached is a traversal copy: its results
// are re-ranked in exact FP32, so it searches with the order-free FP16
// distance (16-byte gathers) instead of the reference's summation order.
static int search_class(const pk_index* h) {
    if (h->dtype == PK_FLOAT16 && h->full && h->vstride <= 512)
        return kHalfFastClass + (h->vstride <= 256 ? 0 : 1);
    return h->size_class;
}

// vbits_override > 0 sets the visited-table size for this launch only (the
// rerun of queries that filled their tables); such a launch takes no more
// waves than the scratch already allocated holds at that size, since the
// persistent queue serves any number of queries with fewer waves.
// the size class a search on h runs with (fast.hip)
int search_class_of(const pk_index* h) { return search_class(h); }

// a search context's work counter and failure flags
int ensure_ctx_counters(SearchCtx& c) {
    int rc;
    if (!c.qcounter && (rc = dalloc(&c.qcounter, 1))) return rc;
    if (!c.err && (rc = dalloc(&c.err, 1))) return rc;
    return PK_OK;
}

static int plan_search(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, const int32_t* d_k,
                       const int32_t* d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride,
                       int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters,
                       SearchPlan* P, int vbits_override = 0, bool bind_scratch = true) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (!h->ready) return fail(PK_ERR_STATE, "index not loaded (call pk_index_set_enter_level)");
    if (nq < 0 || (nq > 0 && (!d_queries || !d_out_ids || !d_out_dists))) return fail(PK_ERR_ARG, "bad buffers");
    if (k_stride < 1) return fail(PK_ERR_ARG, "k_stride must be positive");
    if (!d_k && (k_all < 1 || k_all > k_stride)) return fail(PK_ERR_ARG, "k must be in [1, k_stride]");
    if (!d_ef && ef_all < 1) return fail(PK_ERR_ARG, "ef must be positive");
    if (ef_max < 1) return fail(PK_ERR_ARG, "ef_max must be positive");
    if (nq > (int64_t)1 << 27) return fail(PK_ERR_CAPACITY, "at most 2^27 queries per call");

    SearchArgs& a = P->a;
    a = SearchArgs{};
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

    // LDS plan per wave: result heap, candidate heap (deeper levels spill to
    // device memory), candidate id / distance lists, query row.  The visited
    // table lives in device memory, sized for the visits ef implies (a full
    // table fails the query).
    const SearchTuning& t = h->tuning;
    int wcap = result_capacity_for(ef_max);
    if (wcap < 0) return fail(PK_ERR_ARG, "ef_max %d above the largest supported ef (630)", ef_max);
    int ccap = t.ccap > 0 ? t.ccap : (ef_max <= 64 ? 512 : 1024);
    int vbits = vbits_override > 0 ? vbits_override : t.vbits > 0 ? t.vbits : default_vbits(ef_max, h->capacity);
    if (h->capacity > kVisitMaxId) return fail(PK_ERR_CAPACITY, "search supports at most %lld ids", (long long)kVisitMaxId);
    size_t per_wave = (size_t)wcap * 8 + (size_t)ccap * 8 + 64 * 4 * 2 + h->vstride;
    per_wave = (per_wave + 15) / 16 * 16;
    size_t lds = per_wave * (kSearchBlock / 64);
    if (lds > 160 * 1024) return fail(PK_ERR_CAPACITY, "ef_max %d needs %zu bytes of LDS per workgroup", ef_max, lds);
    int bpc = 0;
    if (search_occupancy(h->dtype, h->metric, h->size_class, lds, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    if (t.blocks_per_cu > 0) bpc = std::min(bpc, t.blocks_per_cu);
    int64_t want = std::max<int64_t>(1, (nq + (kSearchBlock / 64) - 1) / (kSearchBlock / 64));
    int grid = (int)std::min<int64_t>(want, (int64_t)h->num_cus * bpc);
    if (vbits_override > 0) {
        const int64_t cg_stride = std::max<int64_t>(h->cg_cap, (int64_t)1 << vbits);
        const int64_t fit = std::min<int64_t>((int64_t)(c.v_slots >> vbits), (int64_t)c.cg_entries / cg_stride);
        grid = (int)std::max<int64_t>(1, std::min<int64_t>(grid, fit / (kSearchBlock / 64)));
    }
    a.vg_bits = vbits;
    a.wcap = wcap;
    a.ccap = ccap;
    a.lds_per_wave = (int)per_wave;
    P->grid = grid;
    P->lds = lds;
    P->vbits = vbits;
    if (!bind_scratch) return PK_OK;  // the caller places the scratch
    int waves = grid * (kSearchBlock / 64);
    int rc = ensure_scratch(h, c, waves, vbits);
    if (rc) return rc;
    c.last_vbits = vbits;
    a.cglob = c.cglob;
    a.vtab = c.vtab;
    a.vlog = c.vlog;
    a.cg_cap = c.cg_cap;
    a.qcounter = c.qcounter;
    a.err = c.err;
    return PK_OK;
}

// the kernels only (capturable)
static int enqueue_search(pk_index* h, SearchCtx& c, const SearchPlan& P) {
    if (&c == &h->sc) {
        h->last_grid = P.grid;
        h->last_lds = P.lds;
    }
    HIPCHK(launch_search(P.a, h->dtype, h->metric, search_class(h), P.grid, P.lds, c.stream));
    return PK_OK;
}

int search_on(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
              int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids, float* d_out_dists,
              int32_t* d_out_count, pk_query_counters* d_counters, int vbits_override) {
    if (h && nq == 0 && d_out_ids) return PK_OK;
    SearchPlan P;
    int rc = plan_search(h, c, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_max, k_stride, d_out_ids, d_out_dists,
                         d_out_count, d_counters, &P, vbits_override);
    if (rc) return rc;
    if (nq == 0) return PK_OK;
    // a large batch is handed out region by region (results do not depend
    // on the order); graph captures and the static entry point keep the
    // caller's order
    if ((rc = order_queries(h, c, d_queries, nq, search_class_of(h), &P.a.order))) return rc;
    return enqueue_search(h, c, P);
}
}  // namespace pk

extern "C" {

int pk_search_device_ex(pk_index* h, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
                        int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids,
                        float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    return search_on(h, h->sc, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_max, k_stride, d_out_ids, d_out_dists,
                     d_out_count, d_counters);
}

// ---- HIP graphs of one search launch ----

int pk_search_graph_create(pk_index* h, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
                           int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids,
                           float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters,
                           pk_search_graph** out) {
    if (!out) return fail(PK_ERR_ARG, "null argument");
    *out = nullptr;
    if (nq < 1) return fail(PK_ERR_ARG, "a search graph needs at least one query");
    if (!h) return fail(PK_ERR_ARG, "null handle");
    SearchPlan P;
    int rc = plan_search(h, h->sc, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_max, k_stride, d_out_ids, d_out_dists,
                         d_out_count, d_counters, &P);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_search(h, h->sc, P);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(PK_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(PK_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    auto* sg = new pk_search_graph{h, g, ex, P.grid};
    h->graphs_alive++;
    *out = sg;
    return PK_OK;
}

int pk_search_graph_launch(pk_search_graph* g) {
    if (!g) return fail(PK_ERR_ARG, "null graph");
    HIPCHK(hipGraphLaunch(g->exec, g->h->stream));
    g->h->last_grid = g->grid;
    return PK_OK;
}

int pk_search_graph_destroy(pk_search_graph* g) {
    if (!g) return PK_OK;
    (void)hipStreamSynchronize(g->h->stream);
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    g->h->graphs_alive--;
    delete g;
    return PK_OK;
}

// ---- autotuner ----

int pk_autotune_search(pk_index* h, const void* d_queries, int64_t nq, int32_t k, int32_t ef, int32_t* out_tuning,
                       float* out_ms) {
    if (!h || !d_queries || nq < 1 || k < 1 || ef < k) return fail(PK_ERR_ARG, "bad arguments");
    if (h->graphs_alive) return fail(PK_ERR_STATE, "destroy the search graphs before tuning");
    // serving loops plan their launches from the handle's tuning, which the
    // tuner rewrites while it runs
    if (h->servers_alive) return fail(PK_ERR_STATE, "destroy the serving loops before tuning");
    int rc;
    uint8_t* buf = nullptr;
    const size_t out_bytes = (size_t)nq * k * 8 + (size_t)nq * 4;
    if ((rc = dalloc(&buf, out_bytes))) return rc;
    int32_t* ids = reinterpret_cast<int32_t*>(buf);
    float* ds = reinterpret_cast<float*>(buf + (size_t)nq * k * 4);
    int32_t* cnt = reinterpret_cast<int32_t*>(buf + (size_t)nq * k * 8);
    const SearchTuning saved = h->tuning;
    auto time_one = [&](const SearchTuning& t, float* ms) -> int {
        h->tuning = t;
        int r = pk_search_device_ex(h, d_queries, nq, nullptr, nullptr, k, ef, ef, k, ids, ds, cnt, nullptr);
        if (r) return r;
        HIPCHK(hipEventRecord(h->ev0, h->stream));
        for (int i = 0; i < 2 && !r; i++)
            r = pk_search_device_ex(h, d_queries, nq, nullptr, nullptr, k, ef, ef, k, ids, ds, cnt, nullptr);
        if (r) return r;
        HIPCHK(hipEventRecord(h->ev1, h->stream));
        HIPCHK(hipEventSynchronize(h->ev1));
        HIPCHK(hipEventElapsedTime(ms, h->ev0, h->ev1));
        *ms /= 2;
        // a configuration whose searches fail (full visited table or
        // candidate heap) ends those queries early: never a candidate
        int32_t flags = 0;
        HIPCHK(hipMemcpy(&flags, h->sc.err, sizeof(flags), hipMemcpyDeviceToHost));
        if (flags) return fail(PK_ERR_DEVICE, "search failed on the device (flags %d)", flags);
        return PK_OK;
    };
    // coordinate descent: resident workgroups per CU, then the candidate
    // heap's LDS depth, then the visited-table size (configurations that do
    // not fit are skipped)
    SearchTuning best = saved;
    float best_ms = 0;
    // the starting point must complete every query: grow its visited table
    // until it does
    for (;;) {
        rc = time_one(best, &best_ms);
        if (rc != PK_ERR_DEVICE) break;
        const int v = best.vbits > 0 ? best.vbits : default_vbits(ef, h->capacity);
        if (v >= 20) break;
        best.vbits = v + 1;
    }
    if (rc) {
        h->tuning = saved;
        dfree(buf);
        return rc;
    }
    auto try_set = [&](SearchTuning t) {
        float ms = 0;
        if (time_one(t, &ms) == PK_OK && ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    };
    int max_bpc = 0;
    {
        SearchPlan P;
        h->tuning = best;
        if ((rc = plan_search(h, h->sc, d_queries, nq, nullptr, nullptr, k, ef, ef, k, ids, ds, cnt, nullptr, &P)) ==
            PK_OK)
            max_bpc = (P.grid + h->num_cus - 1) / h->num_cus;
    }
    for (int b = max_bpc - 1; b >= std::max(1, max_bpc - 3); b--) {
        SearchTuning t = best;
        t.blocks_per_cu = b;
        try_set(t);
    }
    for (int c : {256, 512, 1024, 2048}) {
        SearchTuning t = best;
        t.ccap = c;
        try_set(t);
    }
    const int vb0 = best.vbits > 0 ? best.vbits : default_vbits(ef, h->capacity);
    for (int v : {vb0 - 1, vb0 + 1}) {
        if (v < 8 || v > 20) continue;
        SearchTuning t = best;
        t.vbits = v;
        try_set(t);
    }
    h->tuning = best;
    dfree(buf);
    if (out_tuning) {
        out_tuning[0] = best.ccap;
        out_tuning[1] = best.vbits;
        out_tuning[2] = best.blocks_per_cu;
    }
    if (out_ms) *out_ms = best_ms;
    return PK_OK;
}

}  // extern "C"

// ---- the search without allocation: scratch in a static device array ----
//
// pk_search_device_static runs the kernels of pk_search_device_ex with
// every piece of device scratch -- the work counter, the failure flags, the
// visited tables with their logs and the candidate-heap spill areas -- in
// the file-scope array `scratch` (kStaticScratchBytes per device, held until
// the process ends).  Host code never names the array as an address: the
// device address comes from hipGetSymbolAddress, once per device with that
// device current, and is kept per device.  The visited tables are zero
// between searches (each search clears what it marked); the first call on a
// device zeroes the whole array with a fill kernel.  Calls that overlap on
// one device share the array and race.
namespace {
// fixed regions, so the visited tables occupy the same bytes whatever a
// call's table size (they are zero between searches; the logs and the heap
// spill are not): head (work counter, failure flags), tables, logs, spill
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
constexpr size_t kStaticHead = 256;
constexpr size_t kStaticTables = (size_t)32 << 20;
constexpr size_t kStaticLogs = (size_t)32 << 20;
constexpr size_t kStaticSpill = kStaticScratchBytes - kStaticHead - kStaticTables - kStaticLogs;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];
std::mutex g_static_mu;
uint8_t* g_static_ptr[kMaxDevices] = {};
bool g_static_zeroed[kMaxDevices] = {};

__global__ void k_zero16(uint4* p, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

// the array's device address on h's device (looked up on first use there)
int static_scratch(pk_index* h, uint8_t** out) {
    if (h->device < 0 || h->device >= kMaxDevices) return fail(PK_ERR_STATE, "device %d out of range", h->device);
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_ptr[h->device]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != h->device) HIPCHK(hipSetDevice(h->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != h->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(PK_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_static_ptr[h->device] = static_cast<uint8_t*>(p);
    }
    if (!g_static_zeroed[h->device]) {
        hipLaunchKernelGGL(k_zero16, dim3(1024), dim3(256), 0, h->stream,
                           reinterpret_cast<uint4*>(g_static_ptr[h->device]), kStaticScratchBytes / 16);
        HIPCHK(hipGetLastError());
        g_static_zeroed[h->device] = true;
    }
    *out = g_static_ptr[h->device];
    return PK_OK;
}
}  // namespace

extern "C" {

int pk_search_device_static(pk_index* h, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
                            int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids,
                            float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (nq == 0 && d_out_ids) return PK_OK;
    SearchPlan P;
    int rc = plan_search(h, h->sc, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_max, k_stride, d_out_ids, d_out_dists,
                         d_out_count, d_counters, &P, 0, /*bind_scratch=*/false);
    if (rc) return rc;
    if (nq == 0) return PK_OK;
    // per wave: visited table and log (4 B per slot each), candidate-heap
    // spill (8 B per entry, room for 2^vbits); the persistent queue serves
    // any number of queries with the waves that fit
    const size_t slots = (size_t)1 << P.vbits;
    const size_t cg_stride = std::max<size_t>((size_t)h->cg_cap, slots);
    const size_t waves_fit = std::min(kStaticTables / (slots * 4), kStaticSpill / (cg_stride * 8));
    constexpr int kWavesPerGroup = kSearchBlock / 64;
    if (waves_fit < (size_t)kWavesPerGroup)
        return fail(PK_ERR_CAPACITY,
                    "a visited table of 2^%d slots needs %zu bytes of tables and %zu of heap spill per workgroup; the "
                    "static scratch holds %zu and %zu",
                    P.vbits, slots * 4 * kWavesPerGroup, cg_stride * 8 * kWavesPerGroup, kStaticTables, kStaticSpill);
    P.grid = (int)std::min<size_t>((size_t)P.grid, waves_fit / kWavesPerGroup);
    uint8_t* base = nullptr;
    if ((rc = static_scratch(h, &base))) return rc;
    SearchArgs& a = P.a;
    a.qcounter = reinterpret_cast<int32_t*>(base);
    a.err = reinterpret_cast<int32_t*>(base) + 1;
    a.vtab = reinterpret_cast<uint32_t*>(base + kStaticHead);
    a.vlog = reinterpret_cast<int32_t*>(base + kStaticHead + kStaticTables);
    a.cglob = reinterpret_cast<uint64_t*>(base + kStaticHead + kStaticTables + kStaticLogs);
    a.cg_cap = (int)cg_stride;
    HIPCHK(launch_search(a, h->dtype, h->metric, search_class(h), P.grid, P.lds, h->stream));
    h->last_grid = P.grid;
    h->last_lds = P.lds;
    return PK_OK;
}

int pk_search_device(pk_index* h, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
                     int32_t k_all, int32_t ef_all, int32_t k_stride, int32_t* d_out_ids, float* d_out_dists,
                     int32_t* d_out_count, pk_query_counters* d_counters) {
    if (d_ef) return fail(PK_ERR_ARG, "per-query ef on device needs pk_search_device_ex (ef_max)");
    return pk_search_device_ex(h, d_queries, nq, d_k, d_ef, k_all, ef_all, ef_all, k_stride, d_out_ids, d_out_dists,
                               d_out_count, d_counters);
}

int pk_device_error(pk_index* h, int* flags) {
    if (!h || !h->sc.err) return fail(PK_ERR_STATE, "no search has run");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(flags, h->sc.err, sizeof(int32_t), hipMemcpyDeviceToHost));
    return PK_OK;
}

static int io_buf(pk_index* h, int slot, size_t bytes, void** p) {
    auto& b = h->io[slot];
    if (b.bytes < bytes) {
        dfree(b.p);
        b.bytes = 0;
        int rc = dalloc(reinterpret_cast<uint8_t**>(&b.p), bytes);
        if (rc) return rc;
        b.bytes = bytes;
    }
    *p = b.p;
    return PK_OK;
}

int pk_search_fast(pk_index* h, const void* queries, int64_t nq, const int32_t* k, const int32_t* ef, int32_t k_all,
                   int32_t ef_all, int32_t k_stride, int32_t* out_ids, float* out_dists, int32_t* out_count,
                   pk_query_counters* counters) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (nq < 0 || (nq > 0 && (!queries || !out_ids || !out_dists))) return fail(PK_ERR_ARG, "bad buffers");
    if (nq == 0) return PK_OK;
    int ef_max = ef_all;
    if (ef) {
        ef_max = 1;
        for (int64_t i = 0; i < nq; i++) {
            if (ef[i] < 1) return fail(PK_ERR_ARG, "ef[%lld] < 1", (long long)i);
            ef_max = std::max(ef_max, ef[i]);
        }
    }
    if (k)
        for (int64_t i = 0; i < nq; i++)
            if (k[i] < 1 || k[i] > k_stride) return fail(PK_ERR_ARG, "k[%lld] outside [1, k_stride]", (long long)i);
    void *dq, *dk = nullptr, *def = nullptr, *dids, *dd, *dc, *dctr = nullptr;
    int rc;
    if ((rc = io_buf(h, 0, (size_t)nq * h->qbytes, &dq)) || (rc = io_buf(h, 3, (size_t)nq * k_stride * 4, &dids)) ||
        (rc = io_buf(h, 4, (size_t)nq * k_stride * 4, &dd)) || (rc = io_buf(h, 5, (size_t)nq * 4, &dc)))
        return rc;
    if (k && (rc = io_buf(h, 1, (size_t)nq * 4, &dk))) return rc;
    if (ef && (rc = io_buf(h, 2, (size_t)nq * 4, &def))) return rc;
    if (counters && (rc = io_buf(h, 6, (size_t)nq * sizeof(pk_query_counters), &dctr))) return rc;
    HIPCHK(hipMemcpyAsync(dq, queries, (size_t)nq * h->qbytes, hipMemcpyHostToDevice, h->stream));
    if (k) HIPCHK(hipMemcpyAsync(dk, k, (size_t)nq * 4, hipMemcpyHostToDevice, h->stream));
    if (ef) HIPCHK(hipMemcpyAsync(def, ef, (size_t)nq * 4, hipMemcpyHostToDevice, h->stream));
    rc = fast_search_on(h, h->sc, dq, nq, (const int32_t*)dk, (const int32_t*)def, k_all, ef_all, ef_max, k_stride,
                        (int32_t*)dids, (float*)dd, (int32_t*)dc, (pk_query_counters*)dctr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out_ids, dids, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(out_dists, dd, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost, h->stream));
    if (out_count) HIPCHK(hipMemcpyAsync(out_count, dc, (size_t)nq * 4, hipMemcpyDeviceToHost, h->stream));
    if (counters)
        HIPCHK(hipMemcpyAsync(counters, dctr, (size_t)nq * sizeof(pk_query_counters), hipMemcpyDeviceToHost, h->stream));
    int32_t errflag = 0;
    HIPCHK(hipMemcpyAsync(&errflag, h->sc.err, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // the fast mode keeps no visited table that could fill: a flag here
    // means bad per-query arguments or a query past its round limit
    if (errflag) return fail(PK_ERR_DEVICE, "fast search failed on the device (flags %d)", errflag);
    return PK_OK;
}

int pk_search(pk_index* h, const void* queries, int64_t nq, const int32_t* k, const int32_t* ef, int32_t k_all,
              int32_t ef_all, int32_t k_stride, int32_t* out_ids, float* out_dists, int32_t* out_count,
              pk_query_counters* counters) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (nq < 0 || (nq > 0 && (!queries || !out_ids || !out_dists))) return fail(PK_ERR_ARG, "bad buffers");
    if (nq == 0) return PK_OK;
    int ef_max = ef_all;
    if (ef) {
        ef_max = 1;
        for (int64_t i = 0; i < nq; i++) {
            if (ef[i] < 1) return fail(PK_ERR_ARG, "ef[%lld] < 1", (long long)i);
            ef_max = std::max(ef_max, ef[i]);
        }
    }
    if (k)
        for (int64_t i = 0; i < nq; i++)
            if (k[i] < 1 || k[i] > k_stride) return fail(PK_ERR_ARG, "k[%lld] outside [1, k_stride]", (long long)i);
    void *dq, *dk = nullptr, *def = nullptr, *dids, *dd, *dc, *dctr = nullptr;
    int rc;
    if ((rc = io_buf(h, 0, (size_t)nq * h->qbytes, &dq)) || (rc = io_buf(h, 3, (size_t)nq * k_stride * 4, &dids)) ||
        (rc = io_buf(h, 4, (size_t)nq * k_stride * 4, &dd)) || (rc = io_buf(h, 5, (size_t)nq * 4, &dc)))
        return rc;
    if (k && (rc = io_buf(h, 1, (size_t)nq * 4, &dk))) return rc;
    if (ef && (rc = io_buf(h, 2, (size_t)nq * 4, &def))) return rc;
    if (counters && (rc = io_buf(h, 6, (size_t)nq * sizeof(pk_query_counters), &dctr))) return rc;
    HIPCHK(hipMemcpyAsync(dq, queries, (size_t)nq * h->qbytes, hipMemcpyHostToDevice, h->stream));
    if (k) HIPCHK(hipMemcpyAsync(dk, k, (size_t)nq * 4, hipMemcpyHostToDevice, h->stream));
    if (ef) HIPCHK(hipMemcpyAsync(def, ef, (size_t)nq * 4, hipMemcpyHostToDevice, h->stream));
    rc = pk_search_device_ex(h, dq, nq, (const int32_t*)dk, (const int32_t*)def, k_all, ef_all, ef_max, k_stride,
                             (int32_t*)dids, (float*)dd, (int32_t*)dc, (pk_query_counters*)dctr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out_ids, dids, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(out_dists, dd, (size_t)nq * k_stride * 4, hipMemcpyDeviceToHost, h->stream));
    std::vector<int32_t> cnt;
    int32_t* cdst = out_count;
    if (!cdst) {
        cnt.resize(nq);
        cdst = cnt.data();
    }
    HIPCHK(hipMemcpyAsync(cdst, dc, (size_t)nq * 4, hipMemcpyDeviceToHost, h->stream));
    if (counters)
        HIPCHK(hipMemcpyAsync(counters, dctr, (size_t)nq * sizeof(pk_query_counters), hipMemcpyDeviceToHost, h->stream));
    int32_t errflag = 0;
    HIPCHK(hipMemcpyAsync(&errflag, h->sc.err, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // Queries that filled their visited table (flag 1 alone) are searched
    // again with a table twice the size, up to 2^20 slots; the table size
    // does not change results.  Not while captured search graphs pin the
    // scratch.
    constexpr int kVisitedFull = 1;
    int vb = h->sc.last_vbits;
    while (errflag == kVisitedFull && vb < 20 && h->graphs_alive == 0) {
        vb++;
        std::vector<int64_t> bad;
        for (int64_t i = 0; i < nq; i++)
            if (cdst[i] < 0) bad.push_back(i);
        const int64_t nb = (int64_t)bad.size();
        if (nb == 0) break;
        const size_t qb = (size_t)h->qbytes;
        std::vector<uint8_t> q2((size_t)nb * qb);
        std::vector<int32_t> k2(k ? nb : 0), e2(ef ? nb : 0);
        for (int64_t j = 0; j < nb; j++) {
            std::memcpy(q2.data() + j * qb, static_cast<const uint8_t*>(queries) + bad[j] * qb, qb);
            if (k) k2[j] = k[bad[j]];
            if (ef) e2[j] = ef[bad[j]];
        }
        HIPCHK(hipMemcpyAsync(dq, q2.data(), q2.size(), hipMemcpyHostToDevice, h->stream));
        if (k) HIPCHK(hipMemcpyAsync(dk, k2.data(), (size_t)nb * 4, hipMemcpyHostToDevice, h->stream));
        if (ef) HIPCHK(hipMemcpyAsync(def, e2.data(), (size_t)nb * 4, hipMemcpyHostToDevice, h->stream));
        rc = search_on(h, h->sc, dq, nb, (const int32_t*)dk, (const int32_t*)def, k_all, ef_all, ef_max, k_stride,
                       (int32_t*)dids, (float*)dd, (int32_t*)dc, (pk_query_counters*)dctr, vb);
        if (rc) return rc;
        std::vector<int32_t> ids2((size_t)nb * k_stride), cnt2(nb);
        std::vector<float> d2((size_t)nb * k_stride);
        std::vector<pk_query_counters> ctr2(counters ? nb : 0);
        HIPCHK(hipMemcpyAsync(ids2.data(), dids, ids2.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(d2.data(), dd, d2.size() * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(cnt2.data(), dc, (size_t)nb * 4, hipMemcpyDeviceToHost, h->stream));
        if (counters)
            HIPCHK(hipMemcpyAsync(ctr2.data(), dctr, (size_t)nb * sizeof(pk_query_counters), hipMemcpyDeviceToHost,
                                  h->stream));
        HIPCHK(hipMemcpyAsync(&errflag, h->sc.err, 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int64_t j = 0; j < nb; j++) {
            std::memcpy(out_ids + bad[j] * k_stride, ids2.data() + j * k_stride, (size_t)k_stride * 4);
            std::memcpy(out_dists + bad[j] * k_stride, d2.data() + j * k_stride, (size_t)k_stride * 4);
            cdst[bad[j]] = cnt2[j];
            if (counters) counters[bad[j]] = ctr2[j];
        }
    }
    if (errflag) {
        int64_t bad = 0;
        for (int64_t i = 0; i < nq; i++) bad += cdst[i] < 0;
        return fail(PK_ERR_DEVICE, "%lld queries failed on the device (flags %d: 1 = scratch capacity, 2 = bad k/ef)",
                    (long long)bad, errflag);
    }
    return PK_OK;
}

int pk_wall_clock_khz(int* khz) {
    if (!khz) return fail(PK_ERR_ARG, "null argument");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev));
    return PK_OK;
}

int pk_sync(pk_index* h) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return PK_OK;
}

int pk_device_alloc(void** p, size_t bytes) {
    if (!p) return fail(PK_ERR_ARG, "null argument");
    return dalloc(reinterpret_cast<uint8_t**>(p), bytes);
}
int pk_device_free(void* p) {
    if (p) HIPCHK(hipFree(p));
    return PK_OK;
}
int pk_copy_to_device(void* d, const void* s, size_t bytes) {
    HIPCHK(hipMemcpy(d, s, bytes, hipMemcpyHostToDevice));
    return PK_OK;
}
int pk_copy_to_host(void* d, const void* s, size_t bytes) {
    HIPCHK(hipMemcpy(d, s, bytes, hipMemcpyDeviceToHost));
    return PK_OK;
}

}  // extern "C"

namespace {
// plain streaming copy of 16-byte units (the calibration kernel for HBM
// byte counters: it reads and writes exactly `units` * 16 bytes)
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                 int64_t units) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < units; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" {

int pk_device_copy(pk_index* h, void* d_dst, const void* d_src, size_t bytes) {
    if (!h || !d_dst || !d_src) return fail(PK_ERR_ARG, "null argument");
    if (bytes % 16 || (uintptr_t)d_dst % 16 || (uintptr_t)d_src % 16)
        return fail(PK_ERR_ARG, "pk_device_copy needs 16-byte aligned pointers and sizes");
    if (bytes == 0) return PK_OK;
    int64_t units = (int64_t)(bytes / 16);
    int grid = (int)std::min<int64_t>((units + 255) / 256, (int64_t)h->num_cus * 8);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, h->stream, static_cast<const uint4*>(d_src),
                       static_cast<uint4*>(d_dst), units);
    HIPCHK(hipGetLastError());
    return PK_OK;
}

int pk_timer_begin(pk_index* h) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    return PK_OK;
}
int pk_timer_end(pk_index* h, float* ms) {
    if (!h || !ms) return fail(PK_ERR_ARG, "null argument");
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipEventSynchronize(h->ev1));
    HIPCHK(hipEventElapsedTime(ms, h->ev0, h->ev1));
    return PK_OK;
}

int pk_set_tuning(pk_index* h, int ccap, int vbits, int blocks_per_cu) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (h->servers_alive) return fail(PK_ERR_STATE, "destroy the serving loops before changing the tuning");
    if (ccap < 0 || ccap > 16384 || (vbits != 0 && (vbits < 8 || vbits > 20)) || blocks_per_cu < 0)
        return fail(PK_ERR_ARG, "bad tuning (ccap <= 16384, vbits 0 or 8..20)");
    h->tuning.ccap = ccap;
    h->tuning.vbits = vbits;
    h->tuning.blocks_per_cu = blocks_per_cu;
    return PK_OK;
}

int pk_last_launch(pk_index* h, int* grid, int* lds_bytes) {
    if (!h) return fail(PK_ERR_ARG, "null handle");
    if (grid) *grid = h->last_grid;
    if (lds_bytes) *lds_bytes = (int)h->last_lds;
    return PK_OK;
}

}  // extern "C"
