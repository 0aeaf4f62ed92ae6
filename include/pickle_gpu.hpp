/*
 * pickle_gpu.hpp -- header-only C++ facade with the shape of
 * pickle::v2::HNSWGraph (include/hnsw.hpp:23-220), backed by the C-ABI in
 * pickle_gpu.h.  A caller of the reference swaps
 *     #include "hnsw.hpp"          pickle::v2::HNSWGraph
 * for
 *     #include "pickle_gpu.hpp"    pickle_gpu::HNSWGraph
 * and keeps its calls: constructor / Init, InitBuffer, Build<T, Dim>,
 * Insert<T, Dim>, AnnSearch<T, Dim, collect_metric> over a span or an
 * NDArray, with results as Entry {m_dist, m_vid} carrying external ids.
 *
 * Differences a caller can observe:
 *   - errors throw pickle_gpu::Error (code + message) instead of aborting
 *     or exiting (include/marco.hpp:20-27, 64-66);
 *   - Build inserts in batches on the GPU (set_build_batch); batch size 1
 *     with reference rows reproduces the reference's single-threaded Build;
 *   - the level sampler is seeded (set_seed; default std::random_device as
 *     the reference, include/util.hpp:34-64);
 *   - Save / Load use the reference serializer layout
 *     (include/serializer.hpp:27-81) plus a "<path>.v2meta" sidecar.
 * AnnSearch(span) may be called from many host threads at once, as the
 * reference's callers do inside OpenMP loops: such calls share the graph and
 * go through a serving loop on the index (pk_server_*), which packs the
 * queries waiting when a launch slot frees into one launch, so concurrent
 * single-query calls run as batches.  Their results equal a batch search's.
 * Calls that change the graph (InitBuffer, Build, Insert, InsertBatch, Load)
 * wait for the searches in flight and retire that loop first.
 */
#ifndef PICKLE_GPU_HPP
#define PICKLE_GPU_HPP

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <random>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "pickle_gpu.h"

namespace pickle_gpu {

typedef int32_t external_id_t;
typedef int32_t internal_id_t;
typedef float distance_t;

/* include/common.hpp:21-33 (same values) */
enum class DataType { Uint8 = 0, Int8 = 1, Float16 = 3, Float32 = 2 };
enum class DistFunc { L1, L2, IP, RUNTIME };

/* include/queue.hpp:17-28 */
struct Entry {
    distance_t m_dist{std::numeric_limits<distance_t>::max()};
    internal_id_t m_vid{-1};
    Entry(distance_t d, internal_id_t v) : m_dist{d}, m_vid{v} {}
    Entry() = default;
    bool operator==(const Entry& o) const { return m_dist == o.m_dist && m_vid == o.m_vid; }
};

/* IEEE binary16 storage (the reference uses _Float16, which g++ 11 lacks) */
struct half {
    uint16_t bits;
};

template <class T>
struct dtype_of;
template <>
struct dtype_of<uint8_t> {
    static constexpr DataType value = DataType::Uint8;
};
template <>
struct dtype_of<int8_t> {
    static constexpr DataType value = DataType::Int8;
};
template <>
struct dtype_of<float> {
    static constexpr DataType value = DataType::Float32;
};
template <>
struct dtype_of<half> {
    static constexpr DataType value = DataType::Float16;
};
#ifdef __FLT16_MAX__
template <>
struct dtype_of<_Float16> {
    static constexpr DataType value = DataType::Float16;
};
#endif

inline size_t dtype_bytes(DataType t) {
    switch (t) {
        case DataType::Float32: return 4;
        case DataType::Float16: return 2;
        default: return 1;
    }
}

/* include/ndarray.hpp:17-101: shared-owned row-major host bytes */
struct NDArray {
    std::shared_ptr<std::vector<char>> m_data{nullptr};
    std::vector<size_t> m_shape;
    size_t m_byte{0};
    size_t m_stride{0};
    DataType m_dtype{DataType::Float32};

    NDArray() = default;
    NDArray(DataType dtype, const std::vector<size_t>& shape)
        : NDArray(nullptr, dtype, shape) {}
    NDArray(const std::shared_ptr<std::vector<char>>& data, DataType dtype, const std::vector<size_t>& shape)
        : m_shape(shape), m_dtype(dtype) {
        if (shape.size() != 2) throw std::invalid_argument("NDArray is two-dimensional");
        m_byte = dtype_bytes(dtype);
        m_stride = m_byte * shape[1];
        m_data = data ? data : std::make_shared<std::vector<char>>(m_stride * shape[0]);
        if (m_data->size() < m_stride * shape[0]) throw std::invalid_argument("NDArray buffer too small");
    }
    void* get_raw(size_t row) { return m_data->data() + row * m_stride; }
    const void* get_raw(size_t row) const { return m_data->data() + row * m_stride; }
    template <class T, size_t Dim = std::dynamic_extent>
    std::span<const T, Dim> get_span(size_t row) const {
        return std::span<const T, Dim>(static_cast<const T*>(get_raw(row)), m_shape[1]);
    }
};

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), m_code(code) {}
    int code() const { return m_code; }

   private:
    int m_code;
};

inline void check(int rc) {
    if (rc != PK_OK) throw Error(rc, pk_last_error());
}

class HNSWGraph {
   public:
    HNSWGraph() = default;
    HNSWGraph(size_t max_degree, size_t build_ef, DistFunc df) { Init(max_degree, build_ef, df); }

    /* include/hnsw.hpp:31-36 */
    void Init(size_t max_degree, size_t build_ef, DistFunc df) {
        if (df != DistFunc::L2 && df != DistFunc::IP) throw Error(PK_ERR_ARG, "only L2 and IP are supported");
        m_max_degree = max_degree;
        m_build_ef = build_ef;
        m_df = df;
    }

    /* include/hnsw.hpp:222-245: fixes the capacity; the layers are sized
     * when the element type and width are known (first Insert). */
    void InitBuffer(size_t max_nodes) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        retire_server();
        m_capacity = max_nodes;
        m_h.reset();
    }

    /* include/hnsw.hpp:497-507: rows of `data` are indexed by external id */
    template <class T, size_t Dim = std::dynamic_extent>
    void Build(const std::vector<external_id_t>& ext_ids, const NDArray& data) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        retire_server();
        check_type<T>(data.m_dtype);
        if (Dim != std::dynamic_extent && data.m_shape[1] != Dim) throw Error(PK_ERR_ARG, "row width != Dim");
        m_capacity = ext_ids.size();
        create(dtype_of<T>::value, data.m_shape[1]);
        size_t rows = std::min(data.m_shape[0], m_capacity);
        check(pk_index_set_base(m_h.get(), data.get_raw(0), 0, (int64_t)rows));
        init_levels();
        check(pk_index_insert(m_h.get(), ext_ids.data(), (int64_t)ext_ids.size(), m_batch, m_flags));
    }

    /* include/hnsw.hpp:247-302 (one node; InsertBatch is the fast path) */
    template <class T, size_t Dim>
    void Insert(external_id_t ext_id, std::span<const T, Dim> query) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        retire_server();
        if (!m_h) {
            create(dtype_of<T>::value, query.size());
            init_levels();
        }
        check_type<T>(m_dtype);
        if (query.size() != m_dim) throw Error(PK_ERR_ARG, "query width != index dim");
        check(pk_index_set_base(m_h.get(), query.data(), ext_id, 1));
        check(pk_index_insert(m_h.get(), &ext_id, 1, 1, m_flags));
    }

    /* many Insert calls at once: rows[i] is the vector of ext_ids[i] */
    template <class T>
    void InsertBatch(const std::vector<external_id_t>& ext_ids, const NDArray& rows) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        retire_server();
        if (!m_h) {
            create(dtype_of<T>::value, rows.m_shape[1]);
            init_levels();
        }
        check_type<T>(rows.m_dtype);
        if (rows.m_shape[0] != ext_ids.size()) throw Error(PK_ERR_ARG, "one row per id");
        for (size_t i = 0; i < ext_ids.size(); i++)
            check(pk_index_set_base(m_h.get(), rows.get_raw(i), ext_ids[i], 1));
        check(pk_index_insert(m_h.get(), ext_ids.data(), (int64_t)ext_ids.size(), m_batch, m_flags));
    }

    /* include/hnsw.hpp:509-525.  Without metrics (and with k <= ef <= 630,
     * 512 in fast mode) the query goes through the graph's serving loop, so
     * calls from many threads at once run as batched launches; with metrics
     * it runs alone on the handle's stream, whose counters LastCounters()
     * reports. */
    template <class T, size_t Dim, bool collect_metric = false>
    std::vector<Entry> AnnSearch(int top_k, int search_ef, std::span<const T, Dim> query) const {
        if (top_k < 1 || search_ef < 1) throw Error(PK_ERR_ARG, "k and ef must be positive");
        for (;;) {
            {
                std::shared_lock<std::shared_mutex> g(m_mu);
                ready();
                check_type<T>(m_dtype);
                if (query.size() != m_dim) throw Error(PK_ERR_ARG, "query width != index dim");
                if (collect_metric || search_ef < top_k || search_ef > serve_max_ef()) {
                    std::lock_guard<std::mutex> d(m_direct);
                    std::vector<std::vector<Entry>> r = search(top_k, search_ef, query.data(), 1, collect_metric);
                    return std::move(r[0]);
                }
                if (m_server && top_k <= m_serve_k) return serve_one(top_k, search_ef, query.data());
            }
            // no loop yet, or one whose result rows are narrower than k
            std::unique_lock<std::shared_mutex> x(m_mu);
            ready();
            if (!m_server || top_k > m_serve_k) start_server(top_k);
        }
    }

    /* Fast mode (no counterpart in the reference): AnnSearch runs the fast
     * search (pk_search_fast, and a PK_SERVE_FAST loop for concurrent
     * single-query calls), whose results are recall-equivalent to the
     * reference's, not bit-exact; ef above 512 keeps the exact search.  Off
     * by default. */
    void set_fast_search(bool on) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        if (on == m_fast) return;
        retire_server();
        m_fast = on;
    }

    /* include/hnsw.hpp:528-541 */
    template <class T, size_t Dim, bool collect_metric = false>
    std::vector<std::vector<Entry>> AnnSearch(int top_k, int search_ef, NDArray query) const {
        std::shared_lock<std::shared_mutex> g(m_mu);
        std::lock_guard<std::mutex> d(m_direct);
        ready();
        check_type<T>(query.m_dtype);
        if (query.m_shape[1] != m_dim) throw Error(PK_ERR_ARG, "query width != index dim");
        if (query.m_shape[0] == 0) return {};
        return search(top_k, search_ef, query.get_raw(0), query.m_shape[0], collect_metric);
    }

    /* serializer layout (include/serializer.hpp:27-81) + sidecar */
    void Save(const std::string& path) const {
        std::shared_lock<std::shared_mutex> g(m_mu);
        std::lock_guard<std::mutex> d(m_direct);
        ready();
        check(pk_index_save_file(m_h.get(), path.c_str()));
    }
    /* the file holds the graph; `data` holds the vectors by external id */
    template <class T>
    void Load(const std::string& path, const NDArray& data) {
        std::unique_lock<std::shared_mutex> g(m_mu);
        retire_server();
        check_type<T>(data.m_dtype);
        m_capacity = data.m_shape[0];
        create(dtype_of<T>::value, data.m_shape[1]);
        check(pk_index_set_base(m_h.get(), data.get_raw(0), 0, (int64_t)data.m_shape[0]));
        check(pk_index_load_file(m_h.get(), path.c_str()));
    }

    int GetEnterLevel() const {
        std::shared_lock<std::shared_mutex> g(m_mu);
        ready();
        int nl = 0, el = 0;
        check(pk_index_info(m_h.get(), &nl, &el, nullptr, nullptr));
        return el;
    }
    /* nodes per layer */
    std::vector<int64_t> LayerSizes() const {
        std::shared_lock<std::shared_mutex> g(m_mu);
        ready();
        int nl = 0, el = 0;
        std::vector<int64_t> sizes(PK_MAX_LEVELS, 0);
        check(pk_index_info(m_h.get(), &nl, &el, sizes.data(), nullptr));
        sizes.resize(nl);
        return sizes;
    }
    /* per-query work counters of the last AnnSearch<..., true>
     * (include/profiler.hpp:13-69 units) */
    const std::vector<pk_query_counters>& LastCounters() const { return m_counters; }

    /* build knobs: insert batch size and whether rows may grow past D as
     * the reference's AddEdge does (include/hnsw_layer.hpp:441-447) */
    void set_build_batch(int max_batch, bool reference_rows = true) {
        m_batch = std::max(1, max_batch);
        m_flags = reference_rows ? PK_BUILD_EXACT_ROWS : 0;
    }
    void set_seed(uint64_t seed) { m_seed = seed; }

   private:
    template <class T>
    void check_type(DataType have) const {
        if (have != dtype_of<T>::value) throw Error(PK_ERR_ARG, "element type does not match the array / index");
    }
    void ready() const {
        if (!m_h) throw Error(PK_ERR_STATE, "graph is empty (Build, Insert or Load first)");
    }
    void create(DataType dt, size_t dim) {
        if (m_max_degree == 0) throw Error(PK_ERR_STATE, "Init was not called");
        if (m_capacity == 0) throw Error(PK_ERR_STATE, "capacity is zero (InitBuffer / Build)");
        pk_index* raw = nullptr;
        int metric = m_df == DistFunc::L2 ? PK_L2 : PK_IP;
        check(pk_index_create((int)dt, metric, (int)dim, (int)m_max_degree, (int)m_build_ef, (int64_t)m_capacity,
                              &raw));
        m_h = std::shared_ptr<pk_index>(raw, [](pk_index* p) { pk_index_destroy(p); });
        m_dtype = dt;
        m_dim = dim;
    }
    void init_levels() {
        uint64_t seed = m_seed ? m_seed : ((uint64_t)std::random_device{}() << 32 | std::random_device{}());
        std::vector<uint8_t> levels(m_capacity);
        check(pk_sample_levels((int64_t)m_capacity, (int)m_max_degree, seed, levels.data()));
        check(pk_index_init_build(m_h.get(), levels.data()));
    }
    // the serving loop behind concurrent AnnSearch(span) calls: result rows
    // of `k` entries (at least 100), any ef in [k, 630]
    int serve_max_ef() const { return m_fast ? kFastMaxEf : kServeMaxEf; }
    void start_server(int k) const {
        retire_server();
        const int kk = std::min(serve_max_ef(), std::max({k, 100, m_serve_k}));
        pk_server* s = nullptr;
        check(pk_server_create_ex(m_h.get(), kServeSlots, kServeQueries, kk, serve_max_ef(),
                                  m_fast ? PK_SERVE_FAST : 0, &s));
        m_server.reset(s);
        m_serve_k = kk;
    }
    void retire_server() const {
        if (m_server) {
            pk_server* s = m_server.release();
            check(pk_server_destroy(s));
        }
    }
    std::vector<Entry> serve_one(int k, int ef, const void* q) const {
        std::vector<int32_t> ids(k);
        std::vector<float> ds(k);
        int32_t cnt = 0;
        int64_t ticket = 0;
        check(pk_server_submit(m_server.get(), q, 1, k, ef, ids.data(), ds.data(), &cnt, &ticket));
        check(pk_server_wait(m_server.get(), ticket, -1, nullptr, nullptr));
        if (cnt < 0) throw Error(PK_ERR_DEVICE, "query failed on the device");
        std::vector<Entry> out;
        out.reserve(cnt);
        for (int j = 0; j < cnt; j++) out.emplace_back(ds[j], ids[j]);
        return out;
    }
    struct ServerFree {
        void operator()(pk_server* s) const { (void)pk_server_destroy(s); }
    };
    static constexpr int kServeMaxEf = 630;  // the largest ef a search supports
    static constexpr int kFastMaxEf = 512;   // the largest ef of the fast mode
    static constexpr int kServeSlots = 4;
    static constexpr int kServeQueries = 1024;

    std::vector<std::vector<Entry>> search(int k, int ef, const void* q, size_t nq, bool collect) const {
        if (k < 1 || ef < 1) throw Error(PK_ERR_ARG, "k and ef must be positive");
        std::vector<int32_t> ids(nq * k), cnt(nq);
        std::vector<float> ds(nq * k);
        if (collect) m_counters.assign(nq, pk_query_counters{});
        auto run = m_fast && ef <= kFastMaxEf ? pk_search_fast : pk_search;
        check(run(m_h.get(), q, (int64_t)nq, nullptr, nullptr, k, ef, k, ids.data(), ds.data(), cnt.data(),
                  collect ? m_counters.data() : nullptr));
        std::vector<std::vector<Entry>> out(nq);
        for (size_t i = 0; i < nq; i++) {
            if (cnt[i] < 0) throw Error(PK_ERR_DEVICE, "query " + std::to_string(i) + " failed on the device");
            out[i].reserve(cnt[i]);
            for (int j = 0; j < cnt[i]; j++) out[i].emplace_back(ds[i * k + j], ids[i * k + j]);
        }
        return out;
    }

    size_t m_max_degree{0}, m_build_ef{0}, m_capacity{0}, m_dim{0};
    DistFunc m_df{DistFunc::L2};
    DataType m_dtype{DataType::Float32};
    int m_batch{4096};
    int m_flags{PK_BUILD_EXACT_ROWS};
    uint64_t m_seed{0};
    std::shared_ptr<pk_index> m_h;
    mutable std::vector<pk_query_counters> m_counters;
    // shared by searches; exclusive for calls that change the graph
    mutable std::shared_mutex m_mu;
    // calls that use the handle's own stream and scratch run one at a time
    mutable std::mutex m_direct;
    // declared after m_h: retired before the index it searches
    mutable std::unique_ptr<pk_server, ServerFree> m_server;
    mutable int m_serve_k{0};
    bool m_fast{false};
};

}  // namespace pickle_gpu

#endif /* PICKLE_GPU_HPP */
