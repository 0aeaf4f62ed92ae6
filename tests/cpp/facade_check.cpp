// facade_check.cpp -- drives pickle_gpu::HNSWGraph (include/pickle_gpu.hpp)
// the way a caller of pickle::v2::HNSWGraph would.  Built by
// __graft_entry__.build(); run by tests/test_facade.py (errors) and
// tests/test_gpu_facade.py (search, build).
//
//   facade_check errors
//   facade_check search <graph> <base> <queries> <n> <nq> <dim> <dtype> <metric> <M> <efc> <k> <ef> <out>
//   facade_check build  <base> <queries> <n> <nq> <dim> <dtype> <metric> <M> <efc> <k> <ef> <seed> <batch>
//                       <reference_rows> <out> <graph_out>
//
// <out> receives nq*k int32 ids (-1 padded) then nq*k float32 distances
// (+inf padded).  dtype: u8 i8 f16 f32; metric: l2 ip.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <thread>

#include "pickle_gpu.hpp"

using namespace pickle_gpu;

static NDArray read_rows(const std::string& path, DataType dt, size_t n, size_t dim) {
    NDArray a(dt, {n, dim});
    std::ifstream f(path, std::ios::binary);
    if (!f.read(a.m_data->data(), (std::streamsize)(a.m_stride * n))) throw std::runtime_error("short read " + path);
    return a;
}

static void write_results(const std::string& path, const std::vector<std::vector<Entry>>& res, size_t k) {
    std::vector<int32_t> ids(res.size() * k, -1);
    std::vector<float> ds(res.size() * k, INFINITY);
    for (size_t i = 0; i < res.size(); i++)
        for (size_t j = 0; j < res[i].size() && j < k; j++) {
            ids[i * k + j] = res[i][j].m_vid;
            ds[i * k + j] = res[i][j].m_dist;
        }
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(ids.data()), (std::streamsize)(ids.size() * 4));
    f.write(reinterpret_cast<const char*>(ds.data()), (std::streamsize)(ds.size() * 4));
}

static DataType parse_dtype(const std::string& s) {
    if (s == "u8") return DataType::Uint8;
    if (s == "i8") return DataType::Int8;
    if (s == "f16") return DataType::Float16;
    if (s == "f32") return DataType::Float32;
    throw std::runtime_error("dtype " + s);
}

// single-query AnnSearch must give what the batch gave
template <class T>
static int compare_single(const HNSWGraph& g, const NDArray& q, size_t nq, int k, int ef,
                          const std::vector<std::vector<Entry>>& batch) {
    size_t m = std::min<size_t>(nq, 32);
    for (size_t i = 0; i < m; i++) {
        auto one = g.AnnSearch<T, std::dynamic_extent>(k, ef, q.get_span<T>(i));
        if (one != batch[i]) {
            std::cerr << "single-query result differs from batch at query " << i << "\n";
            return 1;
        }
    }
    return 0;
}

// single-query AnnSearch from 8 host threads at once (a caller's OpenMP loop
// over queries): every result equals the batch's
template <class T>
static int compare_threads(const HNSWGraph& g, const NDArray& q, size_t nq, int k, int ef,
                           const std::vector<std::vector<Entry>>& batch) {
    const size_t m = std::min<size_t>(nq, 512);
    const int nt = 8;
    std::vector<int> bad(nt, -1);
    std::vector<std::string> err(nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            try {
                for (size_t i = t; i < m; i += nt)
                    if (g.AnnSearch<T, std::dynamic_extent>(k, ef, q.get_span<T>(i)) != batch[i]) {
                        bad[t] = (int)i;
                        return;
                    }
            } catch (const std::exception& e) {
                err[t] = e.what();
            }
        });
    for (auto& x : th) x.join();
    for (int t = 0; t < nt; t++) {
        if (!err[t].empty()) {
            std::cerr << "thread " << t << ": " << err[t] << "\n";
            return 1;
        }
        if (bad[t] >= 0) {
            std::cerr << "concurrent single-query result differs from batch at query " << bad[t] << "\n";
            return 1;
        }
    }
    return 0;
}

// fast mode: concurrent single-query calls equal its batch, and the batch
// finds nearly the same neighbours as the exact search
template <class T>
static int check_fast(HNSWGraph& g, const NDArray& q, size_t nq, int k, int ef,
                      const std::vector<std::vector<Entry>>& exact) {
    if (ef > 512) return 0;
    g.set_fast_search(true);
    auto fres = g.AnnSearch<T, std::dynamic_extent>(k, ef, q);
    if (compare_threads<T>(g, q, nq, k, ef, fres)) return 1;
    size_t same = 0, total = 0;
    for (size_t i = 0; i < nq; i++) {
        for (const Entry& e : fres[i]) {
            for (const Entry& x : exact[i])
                if (x.m_vid == e.m_vid) {
                    same++;
                    break;
                }
        }
        total += exact[i].size();
    }
    g.set_fast_search(false);
    if (total && same < 0.9 * total) {
        std::cerr << "fast mode shares only " << same << " of " << total << " ids with the exact search\n";
        return 1;
    }
    return 0;
}

template <class T>
static int run_search(char** a) {
    const std::string graph = a[2], base = a[3], qpath = a[4];
    size_t n = std::stoul(a[5]), nq = std::stoul(a[6]), dim = std::stoul(a[7]);
    DataType dt = parse_dtype(a[8]);
    DistFunc df = std::string(a[9]) == "ip" ? DistFunc::IP : DistFunc::L2;
    size_t M = std::stoul(a[10]), efc = std::stoul(a[11]);
    int k = std::atoi(a[12]), ef = std::atoi(a[13]);
    HNSWGraph g(M, efc, df);
    g.Load<T>(graph, read_rows(base, dt, n, dim));
    NDArray q = read_rows(qpath, dt, nq, dim);
    auto res = g.AnnSearch<T, std::dynamic_extent, true>(k, ef, q);
    if (g.LastCounters().size() != nq) {
        std::cerr << "counters missing\n";
        return 1;
    }
    write_results(a[14], res, k);
    if (compare_single<T>(g, q, nq, k, ef, res) || compare_threads<T>(g, q, nq, k, ef, res)) return 1;
    return check_fast<T>(g, q, nq, k, ef, res);
}

template <class T>
static int run_build(char** a) {
    const std::string base = a[2], qpath = a[3];
    size_t n = std::stoul(a[4]), nq = std::stoul(a[5]), dim = std::stoul(a[6]);
    DataType dt = parse_dtype(a[7]);
    DistFunc df = std::string(a[8]) == "ip" ? DistFunc::IP : DistFunc::L2;
    size_t M = std::stoul(a[9]), efc = std::stoul(a[10]);
    int k = std::atoi(a[11]), ef = std::atoi(a[12]);
    uint64_t seed = std::stoull(a[13]);
    int batch = std::atoi(a[14]);
    bool refrows = std::atoi(a[15]) != 0;
    HNSWGraph g(M, efc, df);
    g.set_seed(seed);
    g.set_build_batch(batch, refrows);
    std::vector<external_id_t> ids(n);
    for (size_t i = 0; i < n; i++) ids[i] = (external_id_t)i;
    g.Build<T>(ids, read_rows(base, dt, n, dim));
    NDArray q = read_rows(qpath, dt, nq, dim);
    auto res = g.AnnSearch<T, std::dynamic_extent>(k, ef, q);
    write_results(a[16], res, k);
    g.Save(a[17]);
    return compare_single<T>(g, q, nq, k, ef, res) || compare_threads<T>(g, q, nq, k, ef, res);
}

// Behaviour that needs no device: argument and state errors are thrown as
// pickle_gpu::Error with the C-ABI's codes.
static int run_errors() {
    int bad = 0;
    try {
        HNSWGraph g(16, 64, DistFunc::L1);
        bad |= 1;
    } catch (const Error& e) {
        if (e.code() != PK_ERR_ARG) bad |= 2;
    }
    HNSWGraph g(16, 64, DistFunc::L2);
    NDArray q(DataType::Uint8, {2, 8});
    try {
        (void)g.AnnSearch<uint8_t, std::dynamic_extent>(10, 20, q);
        bad |= 4;
    } catch (const Error& e) {
        if (e.code() != PK_ERR_STATE) bad |= 8;
    }
    try {
        g.Save("/nonexistent/x");
        bad |= 16;
    } catch (const Error& e) {
        if (e.code() != PK_ERR_STATE) bad |= 32;
    }
    try {
        HNSWGraph h;
        std::vector<external_id_t> ids{0, 1};
        h.Build<uint8_t>(ids, q);
        bad |= 64;
    } catch (const Error& e) {
        if (e.code() != PK_ERR_STATE) bad |= 128;
    }
    try {
        std::vector<external_id_t> ids{0, 1};
        g.Build<float>(ids, q);  // u8 rows passed as float
        bad |= 256;
    } catch (const Error& e) {
        if (e.code() != PK_ERR_ARG) bad |= 512;
    }
    if (dtype_bytes(DataType::Float16) != 2 || q.m_stride != 8) bad |= 1024;
    std::cout << (bad ? "FAIL " : "ok ") << bad << std::endl;
    return bad ? 1 : 0;
}

int main(int argc, char** argv) {
    try {
        std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "errors") return run_errors();
        if (mode == "search" && argc == 15) {
            switch (parse_dtype(argv[8])) {
                case DataType::Uint8: return run_search<uint8_t>(argv);
                case DataType::Int8: return run_search<int8_t>(argv);
                case DataType::Float16: return run_search<half>(argv);
                case DataType::Float32: return run_search<float>(argv);
            }
        }
        if (mode == "build" && argc == 18) {
            switch (parse_dtype(argv[7])) {
                case DataType::Uint8: return run_build<uint8_t>(argv);
                case DataType::Int8: return run_build<int8_t>(argv);
                case DataType::Float16: return run_build<half>(argv);
                case DataType::Float32: return run_build<float>(argv);
            }
        }
        std::cerr << "usage: facade_check errors | search ... | build ...\n";
        return 2;
    } catch (const Error& e) {
        std::cerr << "pickle_gpu::Error " << e.code() << ": " << e.what() << "\n";
        return 3;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 4;
    }
}
