// graph_file.hip -- the on-disk graph format (host code).
//
// Main file: the per-layer byte layout of the reference serializer
// (include/serializer.hpp:27-55): u64 num_layers, then per layer u64
// graph_size, u64 num_nodes, u64 max_degree, i32 ext[N], i32 deg[N],
// i32 adj[N*D], f32 dist[N*D].  Sidecar "<path>.v2meta": magic "PKLV2SC\0",
// u32 version 1, u32 dtype, metric, dim, M, efc, u64 n_total, i32
// enter_level, u32 num_layers, per layer i32 ent_id, i32 next_id, i16
// heuristic[N], u8 has_vector[N], then u8 insert_level[n_total].
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/pickle_gpu.h"
#include "index_impl.hpp"

using pk::set_error;

namespace {

struct Reader {
    std::vector<char> buf;
    size_t off = 0;
    bool ok = true;
    bool open(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        return true;
    }
    template <class T>
    T get() {
        T v{};
        if (off + sizeof(T) > buf.size()) {
            ok = false;
            return v;
        }
        std::memcpy(&v, buf.data() + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
    const char* take(size_t bytes) {
        if (off + bytes > buf.size()) {
            ok = false;
            return nullptr;
        }
        const char* p = buf.data() + off;
        off += bytes;
        return p;
    }
};

template <class T>
void put(std::ofstream& f, const T& v) {
    f.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

}  // namespace

extern "C" int pk_index_load_file(pk_index* h, const char* path) {
    if (!h || !path) return set_error(PK_ERR_ARG, "null argument");
    Reader g, m;
    if (!g.open(path)) return set_error(PK_ERR_IO, std::string("cannot read ") + path);
    if (!m.open(std::string(path) + ".v2meta")) return set_error(PK_ERR_IO, std::string("cannot read ") + path + ".v2meta");
    const char* magic = m.take(8);
    if (!magic || std::memcmp(magic, "PKLV2SC", 8) != 0) return set_error(PK_ERR_IO, "bad sidecar magic");
    uint32_t ver = m.get<uint32_t>(), dtype = m.get<uint32_t>(), metric = m.get<uint32_t>(), dim = m.get<uint32_t>();
    (void)m.get<uint32_t>();
    (void)m.get<uint32_t>();
    uint64_t ntot = m.get<uint64_t>();
    int32_t enter = m.get<int32_t>();
    uint32_t nlm = m.get<uint32_t>();
    if (!m.ok || ver != 1) return set_error(PK_ERR_IO, "unsupported sidecar");
    if ((int)dtype != h->dtype || (int)metric != h->metric || (int)dim != h->dim)
        return set_error(PK_ERR_ARG, "graph file dtype / metric / dim differ from the index");
    if ((int64_t)ntot > h->capacity) return set_error(PK_ERR_CAPACITY, "graph holds more ids than the index capacity");
    uint64_t nl = g.get<uint64_t>();
    if (!g.ok || nl != nlm || nl < 1 || nl > PK_MAX_LEVELS) return set_error(PK_ERR_IO, "bad layer count");
    for (uint64_t l = 0; l < nl; l++) {
        uint64_t size = g.get<uint64_t>(), n = g.get<uint64_t>(), D = g.get<uint64_t>();
        // bounded before any size arithmetic, so a corrupt header cannot wrap it
        if (!g.ok || D == 0 || D > 65536 || n > (uint64_t)h->capacity || size != 24 + 8 * n + 8 * n * D)
            return set_error(PK_ERR_IO, "bad layer header");
        const int32_t* ext = reinterpret_cast<const int32_t*>(g.take(4 * n));
        const int32_t* deg = reinterpret_cast<const int32_t*>(g.take(4 * n));
        const int32_t* adj = reinterpret_cast<const int32_t*>(g.take(4 * n * D));
        const float* dist = reinterpret_cast<const float*>(g.take(4 * n * D));
        (void)m.get<int32_t>();
        (void)m.get<int32_t>();
        const int16_t* heur = reinterpret_cast<const int16_t*>(m.take(2 * n));
        const uint8_t* has = reinterpret_cast<const uint8_t*>(m.take(n));
        if (!g.ok || !m.ok) return set_error(PK_ERR_IO, "truncated graph file");
        // the file layout is byte packed; copy into aligned storage
        std::vector<int32_t> e(ext, ext + n), d(deg, deg + n), a(adj, adj + n * D);
        std::vector<float> ds(dist, dist + n * D);
        std::vector<int16_t> hh(heur, heur + n);
        int rc = pk_index_load_layer(h, (int)l, (int64_t)n, (int)D, (int64_t)(n * D), e.data(), d.data(), hh.data(),
                                     a.data(), ds.data(), has);
        if (rc) return rc;
    }
    return pk_index_set_enter_level(h, enter);
}

extern "C" int pk_index_save_file(pk_index* h, const char* path) {
    if (!h || !path) return set_error(PK_ERR_ARG, "null argument");
    if (!h->ready) return set_error(PK_ERR_STATE, "index has no graph");
    int nl = 0;
    for (int l = 0; l < PK_MAX_LEVELS; l++)
        if (h->layers[l].loaded) nl = l + 1;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    std::ofstream m(std::string(path) + ".v2meta", std::ios::binary | std::ios::trunc);
    if (!f || !m) return set_error(PK_ERR_IO, std::string("cannot write ") + path);
    put(f, (uint64_t)nl);
    m.write("PKLV2SC", 8);
    put(m, (uint32_t)1);
    put(m, (uint32_t)h->dtype);
    put(m, (uint32_t)h->metric);
    put(m, (uint32_t)h->dim);
    put(m, (uint32_t)h->M);
    put(m, (uint32_t)h->efc);
    put(m, (uint64_t)h->capacity);
    put(m, (int32_t)h->enter_level);
    put(m, (uint32_t)nl);
    std::vector<uint8_t> levels((size_t)h->capacity, 0);
    for (int l = 0; l < nl; l++) {
        pk::Layer& L = h->layers[l];
        uint64_t n = L.n, D = L.D;
        std::vector<int32_t> ext(n), deg(n), adj(n * D);
        std::vector<float> dist(n * D);
        std::vector<int16_t> heur(n);
        std::vector<uint8_t> has(n);
        int rc = pk_index_read_layer(h, l, ext.data(), deg.data(), heur.data(), adj.data(), dist.data());
        if (rc) return rc;
        if (hipMemcpy(has.data(), L.has, n, hipMemcpyDeviceToHost) != hipSuccess)
            return set_error(PK_ERR_HIP, "copy of vector flags failed");
        put(f, (uint64_t)(24 + 8 * n + 8 * n * D));
        put(f, n);
        put(f, D);
        f.write(reinterpret_cast<const char*>(ext.data()), 4 * n);
        f.write(reinterpret_cast<const char*>(deg.data()), 4 * n);
        f.write(reinterpret_cast<const char*>(adj.data()), 4 * n * D);
        f.write(reinterpret_cast<const char*>(dist.data()), 4 * n * D);
        put(m, (int32_t)0);
        put(m, (int32_t)n);
        m.write(reinterpret_cast<const char*>(heur.data()), 2 * n);
        m.write(reinterpret_cast<const char*>(has.data()), n);
        for (uint64_t i = 0; i < n; i++)
            if (ext[i] >= 0 && ext[i] < h->capacity) levels[ext[i]] = (uint8_t)l;
    }
    m.write(reinterpret_cast<const char*>(levels.data()), levels.size());
    if (!f || !m) return set_error(PK_ERR_IO, "write failed");
    return PK_OK;
}
