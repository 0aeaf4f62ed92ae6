// index_impl.hpp -- the handle behind pk_index (host-side bookkeeping).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pickle_gpu.h"

namespace pk {

struct Layer {
    bool loaded = false;
    int64_t n = 0;
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
    std::vector<int32_t> h_ext;  // host copy, for the down maps
};

struct SearchTuning {
    int ccap = 0;           // candidate heap entries in LDS (0 = by ef)
    int vbits = 0;          // log2 slots of each wave's visited table (0 = by ef)
    int blocks_per_cu = 0;  // cap on resident workgroups per CU (0 = occupancy)
};

struct IoBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// sets the thread's last-error message; returns code
int set_error(int code, const std::string& msg);

}  // namespace pk

struct pk_index {
    int dtype = 0, metric = 0, dim = 0, M = 0, efc = 0;
    int64_t capacity = 0;
    int qbytes = 0;   // bytes per caller row
    int vstride = 0;  // bytes per stored row (16-byte multiple, zero padded)
    int size_class = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint8_t* base = nullptr;
    int64_t base_rows = 0;
    pk::Layer layers[PK_MAX_LEVELS];
    void* d_layers = nullptr;  // device copy of the kernels' layer descriptors
    int enter_level = 0;
    bool ready = false;
    // search scratch (per resident wave)
    uint64_t* cglob = nullptr;
    uint32_t* vtab = nullptr;
    int32_t* vlog = nullptr;
    int scratch_waves = 0;
    int scratch_vbits = 0;
    int cg_cap = 16384;
    int32_t* qcounter = nullptr;
    int32_t* err = nullptr;
    pk::SearchTuning tuning;
    pk::IoBuf io[8];
    int last_grid = 0;
    size_t last_lds = 0;
};
