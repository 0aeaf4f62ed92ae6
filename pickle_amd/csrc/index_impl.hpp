// index_impl.hpp -- the handle behind pk_index (host-side bookkeeping).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/pickle_gpu.h"

namespace pk {

struct Layer {
    bool loaded = false;
    bool built = false;  // created by pk_index_init_build (rows written by the GPU)
    int64_t n = 0;       // nodes present (NewInID's next id)
    int64_t cap = 0;     // rows allocated
    int32_t D = 0;
    int64_t adj_len = 0;
    int32_t* adj = nullptr;
    int32_t* deg = nullptr;
    int16_t* heur = nullptr;
    float* dist = nullptr;
    int32_t* ext = nullptr;
    uint8_t* vec = nullptr;
    int32_t* down = nullptr;
    uint8_t* has = nullptr;
    int32_t* lock = nullptr;     // per-node AddEdge locks (built layers)
    std::vector<int32_t> h_ext;  // host copy, for the down maps of loaded layers
};

// device scratch of the insertion kernels
struct InsertScratch {
    int32_t* d_ext = nullptr;
    int32_t* d_lvl = nullptr;
    int32_t* d_in = nullptr;
    int64_t cap_nodes = 0;
    int cap_maxl = 0;
    int32_t* r_ids = nullptr;
    float* r_d = nullptr;
    int32_t* r_cnt = nullptr;
    int topk = 0;
    uint64_t* cglob = nullptr;
    uint32_t* vtab = nullptr;
    int32_t* vlog = nullptr;
    size_t v_slots = 0;     // visited-table / log slots (any waves x 2^vbits layout that fits)
    size_t cg_entries = 0;  // candidate-heap spill entries
    int vbits = 0;          // visited-table size the last batch needed
    int32_t* counter = nullptr;
    int32_t* err = nullptr;
    // PK_BUILD_GRAPH: the executable graph of each insert phase (0: the
    // construction searches, 1: AddNode + AddEdge), updated batch by batch
    hipGraphExec_t graph[2] = {nullptr, nullptr};
};

struct SearchTuning {
    int ccap = 0;           // candidate heap entries in LDS (0 = by ef)
    int vbits = 0;          // log2 slots of each wave's visited table (0 = by ef)
    int blocks_per_cu = 0;  // cap on resident workgroups per CU (0 = occupancy)
};

// fast-mode search (fast.hip): 0 = automatic for every field
struct FastTuning {
    int width = 0;            // entries expanded per round (1..4)
    int hash_bits = 0;        // log2 slots of each wave's LDS visited filter (8..14)
    int waves_per_group = 0;  // queries per workgroup (1, 2, 4)
    int blocks_per_cu = 0;    // cap on resident workgroups per CU
};

// search scratch bound to one stream: per resident wave a candidate-heap
// spill area, a visited table and its log; the work counter and failure flags
struct SearchCtx {
    hipStream_t stream = nullptr;
    uint64_t* cglob = nullptr;
    uint32_t* vtab = nullptr;
    int32_t* vlog = nullptr;
    size_t v_slots = 0;     // visited-table / log slots allocated (any waves x 2^vbits layout that fits)
    size_t cg_entries = 0;  // candidate-heap spill entries allocated
    int cg_cap = 0;  // candidate-heap entries per wave of the last planned launch (>= 2^vbits)
    int last_vbits = 0;  // visited-table size of the last search planned on this context
    int32_t* qcounter = nullptr;
    int32_t* err = nullptr;
    int32_t* order = nullptr;  // fast mode's query order: keys, ranks, order, key counts
    size_t order_ints = 0;
};

struct IoBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// sets the thread's last-error message; returns code
int set_error(int code, const std::string& msg);

}  // namespace pk

struct pk_index;

namespace pk {
void free_layer_public(Layer& L);
int publish_layers(pk_index* h);
int ensure_insert_scratch(pk_index* h, int64_t nodes, int maxl, int topk, int waves, int vbits);
// waves whose visited tables of 2^vbits slots (and candidate heaps) fit the
// insertion scratch
int insert_scratch_waves(const pk_index* h, int vbits);
int upload_insert_batch(pk_index* h, const int32_t* ext, const int32_t* lvl, const int32_t* in, int64_t nb,
                        int maxl);
int gather_layer_vectors(pk_index* h, int level);
// n host rows of row_bytes each, to device rows dst_pitch apart (padding
// bytes zero), in stream order; returns when the host memory is free
hipError_t upload_rows(uint8_t* dst, size_t dst_pitch, const void* src, size_t row_bytes, int64_t n,
                       hipStream_t s);
// one batched search on a given context (its stream and scratch); the
// handle's own calls use h->sc.  vbits_override > 0 sets the visited-table
// size instead of the handle's tuning.
int search_on(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, const int32_t* d_k, const int32_t* d_ef,
              int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride, int32_t* d_out_ids, float* d_out_dists,
              int32_t* d_out_count, pk_query_counters* d_counters, int vbits_override = 0);
// one fast-mode search (fast.hip) on a given context: its stream and its
// work counter and failure flags; the handle's own calls use h->sc
int fast_search_on(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, const int32_t* d_k,
                   const int32_t* d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max, int32_t k_stride,
                   int32_t* d_out_ids, float* d_out_dists, int32_t* d_out_count, pk_query_counters* d_counters);
// the context's work counter and failure flags, allocated once
int ensure_ctx_counters(SearchCtx& c);
// a processing order that groups a large batch's queries by region (fast.hip)
int order_queries(pk_index* h, SearchCtx& c, const void* d_queries, int64_t nq, int size_class,
                  const int32_t** out);
int search_class_of(const pk_index* h);
void free_search_ctx(SearchCtx& c);
}  // namespace pk

struct pk_index {
    int dtype = 0, metric = 0, dim = 0, M = 0, efc = 0;
    int64_t capacity = 0;
    int qbytes = 0;   // bytes per caller row
    int vstride = 0;  // bytes per stored row (16-byte multiple, zero padded)
    int size_class = 0;
    int num_cus = 0;
    int device = 0;  // the device current when the handle was created
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint8_t* base = nullptr;
    int64_t base_rows = 0;
    pk::Layer layers[PK_MAX_LEVELS];
    void* d_layers = nullptr;  // device copy of the kernels' layer descriptors
    void* d_build = nullptr;   // device copy of the insertion kernels' layer descriptors
    std::vector<uint8_t> levels;  // insert level per external id (InitBuffer)
    bool building = false;
    pk::InsertScratch ins;
    std::atomic<int> enter_level{0};  // raised by inserts while a serving loop's searches read it
    bool ready = false;
    // search scratch of the handle's own stream
    pk::SearchCtx sc;
    int cg_cap = 16384;
    int row_limit = 0;  // reference-rows builds: longest row AddEdge writes (0 = 511)
    pk::SearchTuning tuning;
    pk::FastTuning fast;
    pk::IoBuf io[8];
    int last_grid = 0;
    size_t last_lds = 0;
    // precision form: full-precision rows by external id (FP32) and the
    // per-dimension int8 quantizer of the traversal copy
    uint8_t* full = nullptr;
    int full_vstride = 0;
    int64_t full_rows = 0;
    float* q_offset = nullptr;
    float* q_inv_scale = nullptr;
    bool has_quantizer = false;
    pk::IoBuf rr[5];  // re-rank scratch: converted queries, candidate ids / dists / counts, counter
    int graphs_alive = 0;   // captured search graphs hold pointers into scratch and layers
    int servers_alive = 0;  // serving loops read the layers (their scratch is their own)
};

struct pk_search_graph {
    pk_index* h;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int grid;
};
