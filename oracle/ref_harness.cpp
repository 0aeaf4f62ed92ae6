// ref_harness.cpp -- drives the reference's own pickle::v2::HNSWGraph.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference's include/ directory (the reference is not copied); the binary
// lands in oracle/_ref/.  It is used to
//   * write golden fixtures (tests/golden/) in the container that has the
//     reference sources, and
//   * time the reference's CPU build and search for the CPU baseline.
//
// The reference keeps HNSWGraph's state private and has no v2 loader or
// dumper, so this file makes the class members visible before including it.
// All standard headers the reference uses are included first so that only
// the reference's own classes are affected.
#include <algorithm>
#include <array>
#include <atomic>
#include <cassert>
#include <cerrno>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <immintrin.h>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <limits>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <omp.h>
#include <optional>
#include <queue>
#include <random>
#include <set>
#include <shared_mutex>
#include <span>
#include <sstream>
#include <stack>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <variant>
#include <vector>
#include <xmmintrin.h>

#define private public
#include "hnsw.hpp"
#undef private

using namespace pickle;
using namespace pickle::v2;

namespace {

struct Args {
    std::map<std::string, std::string> kv;
    std::string get(const std::string &k, const std::string &d = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    long long geti(const std::string &k, long long d = 0) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::stoll(it->second);
    }
    bool has(const std::string &k) const { return kv.count(k) > 0; }
};

Args parse(int argc, char **argv, int start) {
    Args a;
    for (int i = start; i < argc; i++) {
        std::string s = argv[i];
        if (s.rfind("--", 0) == 0) {
            std::string key = s.substr(2);
            std::string val = "1";
            if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) val = argv[++i];
            a.kv[key] = val;
        }
    }
    return a;
}

std::vector<char> read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::cerr << "cannot open " << path << "\n";
        std::exit(2);
    }
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

DataType parse_dtype(const std::string &s) {
    if (s == "u8" || s == "uint8") return DataType::Uint8;
    if (s == "i8" || s == "int8") return DataType::Int8;
    if (s == "f16" || s == "float16") return DataType::Float16;
    if (s == "f32" || s == "float32") return DataType::Float32;
    std::cerr << "bad dtype " << s << "\n";
    std::exit(2);
}

DistFunc parse_metric(const std::string &s) {
    if (s == "l2") return DistFunc::L2;
    if (s == "ip") return DistFunc::IP;
    std::cerr << "bad metric " << s << "\n";
    std::exit(2);
}

size_t dtype_size(DataType t) {
    switch (t) {
    case DataType::Uint8:
    case DataType::Int8: return 1;
    case DataType::Float16: return 2;
    default: return 4;
    }
}

template <class T> void wr(std::ofstream &f, const T &v) { f.write(reinterpret_cast<const char *>(&v), sizeof(T)); }
template <class T> void wrn(std::ofstream &f, const T *p, size_t n) {
    f.write(reinterpret_cast<const char *>(p), sizeof(T) * n);
}
template <class T> T rd(const char *&p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    return v;
}

// Graph file: the serializer's per-layer byte layout
// (include/serializer.hpp:27-55) followed by nothing else; the v2 state the
// layout has no room for goes to PREFIX.v2meta.
void dump_graph(const HNSWGraph &g, const std::string &prefix, DataType dtype, DistFunc df,
                size_t dim, size_t M, size_t efc) {
    std::ofstream f(prefix, std::ios::binary | std::ios::trunc);
    uint64_t nl = g.m_hnsw_layers.size();
    wr(f, nl);
    for (const auto &L : g.m_hnsw_layers) {
        uint64_t n = L->m_capacity, D = L->m_max_degree;
        uint64_t size = 24 + 4 * n + 4 * n + 4 * n * D + 4 * n * D;
        wr(f, size);
        wr(f, n);
        wr(f, D);
        for (uint64_t i = 0; i < n; i++) wr(f, (int32_t)L->m_node_list[i].m_ext_id);
        for (uint64_t i = 0; i < n; i++) wr(f, (int32_t)L->m_node_list[i].m_degree);
        wrn(f, L->m_adj_list.data(), n * D);
        wrn(f, L->m_dist_list.data(), n * D);
    }
    f.close();

    std::ofstream m(prefix + ".v2meta", std::ios::binary | std::ios::trunc);
    m.write("PKLV2SC", 8);
    wr(m, (uint32_t)1);
    wr(m, (uint32_t)dtype);
    wr(m, (uint32_t)(df == DistFunc::IP ? 2 : 1));
    wr(m, (uint32_t)dim);
    wr(m, (uint32_t)M);
    wr(m, (uint32_t)efc);
    wr(m, (uint64_t)g.m_capacity);
    wr(m, (int32_t)g.m_enter_level);
    wr(m, (uint32_t)nl);
    for (const auto &L : g.m_hnsw_layers) {
        wr(m, (int32_t)L->m_ent_id);
        wr(m, (int32_t)L->m_next_id);
        uint64_t n = L->m_capacity;
        for (uint64_t i = 0; i < n; i++) wr(m, (int16_t)L->m_node_list[i].m_heuristic);
        // a row holds the vector only if AddNode ran for it; detect by
        // comparing with the base row (rows never written stay zero)
        for (uint64_t i = 0; i < n; i++) {
            int32_t ext = L->m_node_list[i].m_ext_id;
            const char *row = static_cast<const char *>(L->m_data.get_raw(i));
            const char *src = static_cast<const char *>(g.m_data.get_raw(ext));
            size_t rb = L->m_data.m_stride;
            bool same = ext >= 0 && std::memcmp(row, src, rb) == 0;
            bool zero = std::all_of(row, row + rb, [](char c) { return c == 0; });
            uint8_t flag = same && !zero ? 1 : (zero ? 0 : 1);
            wr(m, flag);
        }
    }
    wrn(m, g.m_insert_level.data(), g.m_insert_level.size());
    m.close();
}

// Rebuilds HNSWGraph state from the files dump_graph writes.
void load_graph(HNSWGraph &g, const std::string &prefix, NDArray base) {
    auto buf = read_file(prefix);
    auto meta = read_file(prefix + ".v2meta");
    const char *p = buf.data();
    const char *q = meta.data() + 8;
    (void)rd<uint32_t>(q);
    auto dtype = (DataType)rd<uint32_t>(q);
    auto metric = rd<uint32_t>(q);
    (void)rd<uint32_t>(q);
    auto M = rd<uint32_t>(q);
    auto efc = rd<uint32_t>(q);
    auto ntot = rd<uint64_t>(q);
    auto enter = rd<int32_t>(q);
    auto nlm = rd<uint32_t>(q);
    g.Init(M, efc, metric == 2 ? DistFunc::IP : DistFunc::L2);
    g.m_data = base;
    g.m_capacity = ntot;
    g.m_enter_level = enter;
    uint64_t nl = rd<uint64_t>(p);
    if (nl != nlm) {
        std::cerr << "layer count mismatch\n";
        std::exit(2);
    }
    g.m_hnsw_layers.clear();
    g.m_layer_capacity.clear();
    for (uint64_t l = 0; l < nl; l++) {
        (void)rd<uint64_t>(p);
        uint64_t n = rd<uint64_t>(p), D = rd<uint64_t>(p);
        auto L = std::make_shared<HNSWLayer>();
        std::vector<size_t> shape = {n, base.m_shape[1]};
        L->Init(D, n, (external_id_t)ntot, dtype, shape, g.m_df);
        for (uint64_t i = 0; i < n; i++) L->m_node_list[i].m_ext_id = rd<int32_t>(p);
        for (uint64_t i = 0; i < n; i++) L->m_node_list[i].m_degree = (int16_t)rd<int32_t>(p);
        std::memcpy(L->m_adj_list.data(), p, 4 * n * D);
        p += 4 * n * D;
        std::memcpy(L->m_dist_list.data(), p, 4 * n * D);
        p += 4 * n * D;
        L->m_ent_id = rd<int32_t>(q);
        L->m_next_id = rd<int32_t>(q);
        L->m_is_empty = L->m_next_id == 0;
        for (uint64_t i = 0; i < n; i++) L->m_node_list[i].m_heuristic = rd<int16_t>(q);
        for (uint64_t i = 0; i < n; i++) {
            uint8_t has = rd<uint8_t>(q);
            int32_t ext = L->m_node_list[i].m_ext_id;
            if (has && ext >= 0)
                std::memcpy(L->m_data.get_raw(i), base.get_raw(ext), base.m_stride);
            if (ext >= 0) L->m_ext2in_table[ext] = (internal_id_t)i;
        }
        g.m_hnsw_layers.push_back(L);
        g.m_layer_capacity.push_back(n);
    }
    g.m_insert_level.assign(reinterpret_cast<const uint8_t *>(q),
                            reinterpret_cast<const uint8_t *>(q) + ntot);
}

template <class T, size_t Dim> void do_build(HNSWGraph &g, const std::vector<external_id_t> &ids, const NDArray &d) {
    g.Build<T, Dim>(ids, d);
}

// Dim switch mirroring benchmark/hnsw_pickle.cpp:9-27 (static extents for the
// dimensions the benchmark specialises, dynamic otherwise).
#define DIM_SWITCH(val, DIM, ...)                                   \
    do {                                                            \
        if ((val) == 128) {                                         \
            constexpr size_t DIM = 128;                             \
            { __VA_ARGS__ }                                         \
        } else if ((val) == 100) {                                  \
            constexpr size_t DIM = 100;                             \
            { __VA_ARGS__ }                                         \
        } else if ((val) == 96) {                                   \
            constexpr size_t DIM = 96;                              \
            { __VA_ARGS__ }                                         \
        } else if ((val) == 200) {                                  \
            constexpr size_t DIM = 200;                             \
            { __VA_ARGS__ }                                         \
        } else {                                                    \
            constexpr size_t DIM = std::dynamic_extent;             \
            { __VA_ARGS__ }                                         \
        }                                                           \
    } while (0)

std::vector<int> parse_list(const std::string &s) {
    std::vector<int> out;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (!tok.empty()) out.push_back(std::stoi(tok));
    return out;
}

int cmd_build(const Args &a) {
    DataType dtype = parse_dtype(a.get("dtype"));
    DistFunc df = parse_metric(a.get("metric", "l2"));
    size_t n = a.geti("n"), dim = a.geti("dim"), M = a.geti("M", 16), efc = a.geti("efc", 100);
    int threads = (int)a.geti("threads", 1);
    omp_set_num_threads(threads);
    auto raw = std::make_shared<std::vector<char>>(read_file(a.get("base")));
    if (raw->size() < n * dim * dtype_size(dtype)) {
        std::cerr << "base file too small\n";
        return 2;
    }
    NDArray data(raw, dtype, {n, dim});
    std::vector<external_id_t> ids(n);
    if (a.has("order")) {
        auto ob = read_file(a.get("order"));
        std::memcpy(ids.data(), ob.data(), 4 * n);
    } else if (a.get("shuffle", "0") == "1") {
        ids = GetRandIndices<external_id_t>(n); // util.hpp:66-80, as the benchmark does
    } else {
        std::iota(ids.begin(), ids.end(), 0);
    }
    HNSWGraph g(M, efc, df);
    auto t0 = std::chrono::steady_clock::now();
    ATEN_DTYPE_SWITCH(dtype, T, { DIM_SWITCH(dim, Dim, { do_build<T, Dim>(g, ids, data); }); });
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    if (a.has("out")) dump_graph(g, a.get("out"), dtype, df, dim, M, efc);
    if (a.has("order-out")) {
        std::ofstream f(a.get("order-out"), std::ios::binary | std::ios::trunc);
        wrn(f, ids.data(), ids.size());
    }
    std::printf("{\"build_seconds\": %.6f, \"threads\": %d, \"n\": %zu, \"layers\": %zu, \"enter_level\": %d}\n", secs,
                threads, n, g.m_hnsw_layers.size(), (int)g.m_enter_level);
    return 0;
}

template <class T, size_t Dim>
std::vector<std::vector<Entry>> run_search(const HNSWGraph &g, int k, int ef, const NDArray &qs, bool collect) {
    if (collect) return g.AnnSearch<T, Dim, true>(k, ef, qs);
    return g.AnnSearch<T, Dim, false>(k, ef, qs);
}

int cmd_search(const Args &a) {
    DataType dtype = parse_dtype(a.get("dtype"));
    size_t n = a.geti("n"), dim = a.geti("dim"), nq = a.geti("nq");
    int k = (int)a.geti("k", 10);
    int threads = (int)a.geti("threads", 1);
    omp_set_num_threads(threads);
    auto raw = std::make_shared<std::vector<char>>(read_file(a.get("base")));
    NDArray base(raw, dtype, {n, dim});
    auto qraw = std::make_shared<std::vector<char>>(read_file(a.get("queries")));
    if (qraw->size() < nq * dim * dtype_size(dtype)) {
        std::cerr << "query file too small\n";
        return 2;
    }
    NDArray qs(qraw, dtype, {nq, dim});
    HNSWGraph g;
    load_graph(g, a.get("graph"), base);
    bool counters = a.get("counters", "0") == "1";
    int repeat = (int)a.geti("repeat", 1);
    auto efs = parse_list(a.get("ef", "10"));
    std::string out = a.get("out", "");
    std::printf("[");
    for (size_t ei = 0; ei < efs.size(); ei++) {
        int ef = efs[ei];
        std::vector<std::vector<Entry>> res;
        double best = 1e30;
        for (int r = 0; r < repeat; r++) {
            auto t0 = std::chrono::steady_clock::now();
            ATEN_DTYPE_SWITCH(dtype, T, { DIM_SWITCH(dim, Dim, { res = run_search<T, Dim>(g, k, ef, qs, false); }); });
            auto t1 = std::chrono::steady_clock::now();
            best = std::min(best, std::chrono::duration<double>(t1 - t0).count());
        }
        long long hop0 = 0, hop = 0, dist0 = 0, dist = 0, nb = 0;
        if (counters) {
            // second, untimed pass with the profiler on (benchmark/hnsw_pickle.cpp:88-92)
            Profiler::Global()->Reset();
            ATEN_DTYPE_SWITCH(dtype, T, { res = run_search<T, std::dynamic_extent>(g, k, ef, qs, true); });
            hop = Profiler::Global()->GetHop();
            hop0 = Profiler::Global()->GetHop(0);
            dist = Profiler::Global()->GetDist();
            dist0 = Profiler::Global()->GetDist(0);
            nb = Profiler::Global()->GetNeighbor();
        }
        if (!out.empty()) {
            std::ofstream f(out + ".ef" + std::to_string(ef), std::ios::binary | std::ios::trunc);
            for (size_t i = 0; i < nq; i++) {
                int32_t cnt = (int32_t)res[i].size();
                wr(f, cnt);
                for (int j = 0; j < k; j++) {
                    int32_t id = j < cnt ? res[i][j].m_vid : -1;
                    float d = j < cnt ? res[i][j].m_dist : std::numeric_limits<float>::infinity();
                    wr(f, id);
                    wr(f, d);
                }
            }
        }
        std::printf("%s{\"ef\": %d, \"k\": %d, \"seconds\": %.6f, \"qps\": %.3f, \"threads\": %d, \"hops\": %lld, "
                    "\"hops0\": %lld, \"dists\": %lld, \"dists0\": %lld, \"neighbors\": %lld}",
                    ei ? ", " : "", ef, k, best, nq / best, threads, hop, hop0, dist, dist0, nb);
    }
    std::printf("]\n");
    return 0;
}

// Heap script: int32 count, then per op: int8 op (1 push, 0 pop, 2 = extractMinTopK with k),
// float dist, int32 id.  Output: per emitted entry "dist id" lines.
int cmd_heap(const Args &a) {
    auto buf = read_file(a.get("script"));
    const char *p = buf.data();
    int is_max = (int)rd<int32_t>(p);
    int n = (int)rd<int32_t>(p);
    std::FILE *out = std::fopen(a.get("out").c_str(), "wb");
    auto emit = [&](const Entry &e) {
        std::fwrite(&e.m_dist, 4, 1, out);
        std::fwrite(&e.m_vid, 4, 1, out);
    };
    auto run = [&](auto &h) {
        for (int i = 0; i < n; i++) {
            int8_t op = rd<int8_t>(p);
            float d = rd<float>(p);
            int32_t id = rd<int32_t>(p);
            if (op == 1) {
                // alternate the two insert overloads the reference has
                // (queue.hpp:191-202 heapifyUp, :204-216 std::push_heap)
                if (i % 2) h.insert(Entry(d, id));
                else h.insert(d, id);
            } else if (op == 0) {
                if (!h.empty()) {
                    emit(h.top());
                    h.pop();
                }
            } else if (op == 2) {
                auto v = h.extractMinTopK(id);
                for (auto &e : v) emit(e);
            }
        }
    };
    if (is_max) {
        MaxHeap h;
        run(h);
    } else {
        MinHeap h;
        run(h);
    }
    std::fclose(out);
    return 0;
}

// Distance known answers: pairs of rows from two files.
int cmd_dist(const Args &a) {
    DataType dtype = parse_dtype(a.get("dtype"));
    DistFunc df = parse_metric(a.get("metric", "l2"));
    size_t n = a.geti("n"), dim = a.geti("dim");
    auto ra = std::make_shared<std::vector<char>>(read_file(a.get("a")));
    auto rb = std::make_shared<std::vector<char>>(read_file(a.get("b")));
    NDArray A(ra, dtype, {n, dim}), B(rb, dtype, {n, dim});
    std::vector<float> out(n);
    ATEN_DTYPE_SWITCH(dtype, T, {
        DIM_SWITCH(dim, Dim, {
            for (size_t i = 0; i < n; i++) out[i] = Distance<T, Dim>(A.get_span<T, Dim>(i), B.get_span<T, Dim>(i), df);
        });
    });
    std::ofstream f(a.get("out"), std::ios::binary | std::ios::trunc);
    wrn(f, out.data(), n);
    return 0;
}

// AddNode / AddEdge script on one standalone layer (hnsw_layer.hpp:297-465).
// File: int32 D, int32 cap, int32 dim, int32 nops, then cap*dim float32 vectors,
// then ops: int32 kind (0 = new id, 1 = add node, 2 = add edge), int32 a, int32 b,
// float d, int32 nr, int32 np, then nr+np (float, int32) entries.
// Output: the layer's adj / dist / degree / heuristic after every op (cap+1 rows).
int cmd_layer(const Args &a) {
    auto buf = read_file(a.get("script"));
    const char *p = buf.data();
    int D = rd<int32_t>(p), cap = rd<int32_t>(p), dim = rd<int32_t>(p), nops = rd<int32_t>(p);
    std::vector<float> vecs(size_t(cap) * dim);
    std::memcpy(vecs.data(), p, 4 * vecs.size());
    p += 4 * vecs.size();
    // one spare row so a write one past a full last row stays inside the buffer
    HNSWLayer L;
    std::vector<size_t> shape = {size_t(cap) + 1, size_t(dim)};
    L.Init(D, cap + 1, cap + 1, DataType::Float32, shape, DistFunc::L2);
    std::ofstream f(a.get("out"), std::ios::binary | std::ios::trunc);
    for (int o = 0; o < nops; o++) {
        int kind = rd<int32_t>(p), x = rd<int32_t>(p), y = rd<int32_t>(p);
        float d = rd<float>(p);
        int nr = rd<int32_t>(p), np = rd<int32_t>(p);
        std::vector<Entry> R, P;
        for (int i = 0; i < nr; i++) {
            float dd = rd<float>(p);
            int32_t id = rd<int32_t>(p);
            R.emplace_back(dd, id);
        }
        for (int i = 0; i < np; i++) {
            float dd = rd<float>(p);
            int32_t id = rd<int32_t>(p);
            P.emplace_back(dd, id);
        }
        if (kind == 0) {
            (void)L.NewInID(x);
        } else if (kind == 1) {
            std::span<const float, std::dynamic_extent> v{vecs.data() + size_t(x) * dim, size_t(dim)};
            L.AddNode<float, std::dynamic_extent>(x, R, P, v);
        } else {
            L.AddEdge<float, std::dynamic_extent>(x, y, d);
        }
        wrn(f, L.m_adj_list.data(), L.m_adj_list.size());
        wrn(f, L.m_dist_list.data(), L.m_dist_list.size());
        for (int i = 0; i <= cap; i++) wr(f, (int16_t)L.m_node_list[i].m_degree);
        for (int i = 0; i <= cap; i++) wr(f, (int16_t)L.m_node_list[i].m_heuristic);
    }
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: pickle_ref build|search|heap|dist|layer --key value ...\n";
        return 2;
    }
    std::string cmd = argv[1];
    Args a = parse(argc, argv, 2);
    if (cmd == "build") return cmd_build(a);
    if (cmd == "search") return cmd_search(a);
    if (cmd == "heap") return cmd_heap(a);
    if (cmd == "dist") return cmd_dist(a);
    if (cmd == "layer") return cmd_layer(a);
    std::cerr << "unknown command " << cmd << "\n";
    return 2;
}
