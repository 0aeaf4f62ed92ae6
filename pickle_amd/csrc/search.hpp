// search.hpp -- kernel arguments shared by the search launcher and the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pickle_gpu.h"

namespace pk {

constexpr int kSearchBlock = 128;  // two waves per workgroup

// One HNSW layer in HBM (internal-id indexed, as HNSWLayer keeps it,
// include/hnsw_layer.hpp:75-83).
struct LayerDev {
    const int32_t* adj;   // n * D ids (stride D) + slots read past the last row
    const int32_t* deg;   // n degrees
    const uint8_t* vec;   // n rows of vstride bytes (zeros where AddNode never ran)
    const int32_t* down;  // n: internal id of the same node one layer down (levels >= 1)
    int64_t n;
    int64_t adj_len;
    int32_t D;
    int32_t pad_;
};

struct SearchArgs {
    const LayerDev* layers;  // device array, PK_MAX_LEVELS entries
    const int32_t* ext0;  // layer-0 internal -> external id
    int32_t enter_level;
    int32_t dim;
    int32_t vstride;  // bytes per stored row (multiple of 16)
    int32_t qbytes;   // bytes per query row (dim * element size)
    const uint8_t* queries;
    int64_t nq;
    const int32_t* k;
    const int32_t* ef;
    int32_t k_all, ef_all, k_stride;
    int32_t* out_ids;
    float* out_dists;
    int32_t* out_count;
    pk_query_counters* ctr;
    uint64_t* cglob;  // per-wave candidate-heap spill, cg_cap entries each
    uint32_t* vtab;   // per-wave visited tables, 2^vg_bits slots each (zero between searches)
    int32_t* vlog;    // per-wave filled-slot logs, 2^vg_bits entries each
    int32_t cg_cap;
    int32_t vg_bits;
    int32_t wcap;  // result heap entries in LDS (>= ef + 1)
    int32_t ccap;  // candidate heap entries in LDS
    int32_t lds_per_wave;
    int32_t* qcounter;
    int32_t* err;
    const int32_t* order;  // queue position -> query (null: in order)
};

int size_class_for(int dtype, int dim, int vstride);
// size classes of the order-free FP16 traversal distance (rows of at most
// 256 and 512 bytes), used for FP16 indexes that carry FP32 rows
constexpr int kHalfFastClass = 8;
int result_capacity_for(int ef_max);
hipError_t launch_search(const SearchArgs& a, int dtype, int metric, int size_class, int grid, size_t lds,
                         hipStream_t s);
hipError_t search_occupancy(int dtype, int metric, int size_class, size_t lds, int* blocks_per_cu);

}  // namespace pk
