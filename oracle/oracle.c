/*
 * oracle.c -- CPU restatement of pickle v2's HNSW query and insertion path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP product; see oracle.h).
 * Compiled with -ffp-contract=off so float distances keep the separate
 * multiply and add roundings of the reference's AVX-512 loop.
 *
 * Citations are file:line in the reference repository.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* half precision                                                      */
/* ------------------------------------------------------------------ */

/* IEEE-754 binary16 -> binary32, exact (every half is a float). */
static float half_to_float(uint16_t h)
{
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t exp = (h >> 10) & 0x1f;
    uint32_t man = h & 0x3ff;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else { /* subnormal: renormalise */
            int e = -1;
            do {
                e++;
                man <<= 1;
            } while ((man & 0x400) == 0);
            man &= 0x3ff;
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp - 15 + 127) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

double orc_half_to_float(uint16_t h) { return half_to_float(h); }

/* ------------------------------------------------------------------ */
/* distances                                                           */
/* ------------------------------------------------------------------ */

/* Integer types: include/distance/avx512.hpp:14-129 widens to int16,
 * squares or multiplies (every product fits the 16-bit lane with the
 * zero/sign extension chosen there), and accumulates in int32; the
 * generic path (generic.hpp:28-51) accumulates in int.  Both are the exact
 * integer sum, returned through an int -> float conversion. */
static float dist_int(int is_signed, int metric, const void *pa, const void *pb, int dim)
{
    int32_t acc = 0;
    if (is_signed) {
        const int8_t *a = (const int8_t *)pa, *b = (const int8_t *)pb;
        if (metric == ORC_L2)
            for (int i = 0; i < dim; i++) {
                int t = (int)a[i] - (int)b[i];
                acc += t * t;
            }
        else
            for (int i = 0; i < dim; i++) acc += (int)a[i] * (int)b[i];
    } else {
        const uint8_t *a = (const uint8_t *)pa, *b = (const uint8_t *)pb;
        if (metric == ORC_L2)
            for (int i = 0; i < dim; i++) {
                int t = (int)a[i] - (int)b[i];
                acc += t * t;
            }
        else
            for (int i = 0; i < dim; i++) acc += (int)a[i] * (int)b[i];
    }
    float r = (float)acc;
    return metric == ORC_IP ? -r : r; /* avx512.hpp:239 returns -IP */
}

/* Float types: avx512.hpp:131-231 as the reference's build compiles it
 * (-O3 -funsafe-math-optimizations, CMakeLists.txt:15-17).  Sixteen lane
 * accumulators; lane j sums elements j, j+16, ... in order with a fused
 * multiply-add (for L2 of the rounded difference x - y).  The fast-math
 * build reassociates Sum16 (marco.hpp:128-131) into a halving tree: lanes
 * j and j+8, then j and j+4, j+2, j+1.  The dim % 16 remainder loop is
 * vectorised the same way: blocks of 8 (FMA, halving tree), then a 4-wide
 * block whose lane 0 starts from that partial sum, then scalar FMAs; the
 * remainder is added to the tree sum last.
 *
 * This order reproduces the compiled reference bit for bit whenever
 * dim % 16 < 8 (every dimension of the benchmark workloads: 96, 100, 128)
 * and for a dim of 200.  For other remainders of 8..15 it is within a few
 * ulp of the compiled code. */
static inline float load_elem(int is_half, const void *p, int i)
{
    return is_half ? half_to_float(((const uint16_t *)p)[i]) : ((const float *)p)[i];
}

static inline float fma_term(int metric, float x, float y, float acc)
{
    if (metric == ORC_L2) {
        float t = x - y;
        return fmaf(t, t, acc);
    }
    return fmaf(x, y, acc);
}

static float tree_sum(float *t, int n)
{
    for (int h = n / 2; h >= 1; h /= 2)
        for (int j = 0; j < h; j++) t[j] = t[j] + t[j + h];
    return t[0];
}

static float dist_float(int is_half, int metric, const void *pa, const void *pb, int dim)
{
    float tmp[16];
    for (int j = 0; j < 16; j++) tmp[j] = 0.0f;
    int main_len = dim - dim % 16;
    for (int i = 0; i < main_len; i += 16)
        for (int j = 0; j < 16; j++)
            tmp[j] = fma_term(metric, load_elem(is_half, pa, i + j), load_elem(is_half, pb, i + j),
                              tmp[j]);
    float s = tree_sum(tmp, 16);

    int R = dim - main_len, j = 0;
    float r = 0.0f;
    if (R >= 8) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (; R - j >= 8; j += 8)
            for (int u = 0; u < 8; u++)
                acc[u] = fma_term(metric, load_elem(is_half, pa, main_len + j + u),
                                  load_elem(is_half, pb, main_len + j + u), acc[u]);
        r = tree_sum(acc, 8);
    }
    if (R - j >= 4) {
        float acc[4] = {r, 0, 0, 0};
        for (; R - j >= 4; j += 4)
            for (int u = 0; u < 4; u++)
                acc[u] = fma_term(metric, load_elem(is_half, pa, main_len + j + u),
                                  load_elem(is_half, pb, main_len + j + u), acc[u]);
        r = tree_sum(acc, 4);
    }
    for (; j < R; j++)
        r = fma_term(metric, load_elem(is_half, pa, main_len + j), load_elem(is_half, pb, main_len + j), r);
    s = s + r;
    return metric == ORC_IP ? -s : s;
}

float orc_distance(int dtype, int metric, const void *a, const void *b, int dim)
{
    switch (dtype) {
    case ORC_U8: return dist_int(0, metric, a, b, dim);
    case ORC_I8: return dist_int(1, metric, a, b, dim);
    case ORC_F32: return dist_float(0, metric, a, b, dim);
    case ORC_F16: return dist_float(1, metric, a, b, dim);
    default: abort();
    }
}

/* ------------------------------------------------------------------ */
/* EntryHeap (include/queue.hpp:99-289)                                */
/* ------------------------------------------------------------------ */

typedef struct {
    orc_entry *a;
    int len;
    int cap;
    int is_max; /* MaxFirst (root = largest) vs MinFirst (root = smallest) */
} heap_t;

/* comp(x, y): true when x ranks below y, so y must sit above x.
 * MaxFirst: x.dist < y.dist (queue.hpp:30-34); MinFirst: x.dist > y.dist (:36-40). */
static inline int h_comp(const heap_t *h, orc_entry x, orc_entry y)
{
    return h->is_max ? (x.dist < y.dist) : (x.dist > y.dist);
}

static void h_init(heap_t *h, int is_max)
{
    h->cap = 512; /* queue.hpp:152 */
    h->len = 0;
    h->is_max = is_max;
    h->a = (orc_entry *)malloc(sizeof(orc_entry) * (size_t)h->cap);
}

static void h_free(heap_t *h)
{
    free(h->a);
    h->a = NULL;
    h->len = h->cap = 0;
}

/* insert(): append then sift up while the parent strictly ranks below the
 * new entry (queue.hpp:119-125, 191-216; std::push_heap is the same rule). */
static void h_push(heap_t *h, float dist, int32_t id)
{
    if (h->len >= h->cap) {
        h->cap *= 8; /* queue.hpp:194 */
        h->a = (orc_entry *)realloc(h->a, sizeof(orc_entry) * (size_t)h->cap);
    }
    orc_entry e = {dist, id};
    int i = h->len++;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (h_comp(h, h->a[p], e)) {
            h->a[i] = h->a[p];
            i = p;
        } else {
            break;
        }
    }
    h->a[i] = e;
}

/* pop(): move the last entry to the root and sift down; the left child is
 * taken first and the right child replaces it only if strictly better
 * (queue.hpp:127-150, 230-236). */
static void h_pop(heap_t *h)
{
    h->a[0] = h->a[h->len - 1];
    h->len -= 1;
    if (h->len == 0) return;
    orc_entry e = h->a[0];
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = 2 * i + 2, best = i;
        orc_entry bv = e;
        if (l < h->len && h_comp(h, bv, h->a[l])) {
            best = l;
            bv = h->a[l];
        }
        if (r < h->len && h_comp(h, bv, h->a[r])) {
            best = r;
            bv = h->a[r];
        }
        if (best == i) break;
        h->a[i] = bv;
        i = best;
    }
    h->a[i] = e;
}

static inline orc_entry h_top(const heap_t *h) { return h->a[0]; }

static orc_entry h_extract_top(heap_t *h)
{
    orc_entry t = h->a[0];
    h_pop(h);
    return t;
}

/* extractMinTopK (queue.hpp:265-286): ascending output of length min(k, len). */
static int h_extract_min_topk(heap_t *h, int k, orc_entry *out)
{
    if (k > h->len) k = h->len;
    if (h->is_max) {
        while (h->len > k) h_pop(h);
        for (int i = 0; i < k; i++) out[k - i - 1] = h_extract_top(h);
    } else {
        for (int i = 0; i < k; i++) out[i] = h_extract_top(h);
    }
    return k;
}

int orc_heap_script(int is_max, const int8_t *ops, int nops, const float *dists,
                    const int32_t *ids, int final_k, orc_entry *out, int out_cap)
{
    heap_t h;
    h_init(&h, is_max);
    int n_out = 0, n_push = 0;
    for (int i = 0; i < nops; i++) {
        if (ops[i] > 0) {
            h_push(&h, dists[n_push], ids[n_push]);
            n_push++;
        } else if (h.len > 0) {
            if (n_out < out_cap) out[n_out] = h_top(&h);
            n_out++;
            h_pop(&h);
        }
    }
    if (final_k >= 0) {
        orc_entry *tmp = (orc_entry *)malloc(sizeof(orc_entry) * (size_t)(h.len + 1));
        int k = h_extract_min_topk(&h, final_k, tmp);
        for (int i = 0; i < k; i++) {
            if (n_out < out_cap) out[n_out] = tmp[i];
            n_out++;
        }
        free(tmp);
    }
    h_free(&h);
    return n_out;
}

/* ------------------------------------------------------------------ */
/* level sampling (include/util.hpp:13-64)                             */
/* ------------------------------------------------------------------ */

int orc_cumulative_probability(int M, double *out, int cap)
{
    double r_logm = 1.0 / log((double)M);
    double c_prob = 0.0;
    int n = 0;
    for (;;) {
        int c_level = n;
        double delta = exp(-1.0 * c_level / r_logm) * (1 - exp(-1.0 / r_logm));
        if (delta < 1e-9) break;
        c_prob += delta;
        if (n < cap) out[n] = c_prob;
        n++;
    }
    return n;
}

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

/* A fresh uniform draw per level, stopping at the first level whose
 * cumulative probability exceeds the draw (util.hpp:48-61).  The reference
 * seeds mt19937 from random_device; only the distribution is reproducible,
 * so this uses a seeded splitmix64 stream instead. */
void orc_sample_levels(int64_t n, int M, uint64_t seed, uint8_t *out)
{
    double cum[64];
    int nc = orc_cumulative_probability(M, cum, 64);
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        int set = 0;
        for (int level = 0; level < nc; level++) {
            double u = (double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
            if (u < cum[level]) {
                out[i] = (uint8_t)level;
                set = 1;
                break;
            }
        }
        if (!set) out[i] = (uint8_t)(nc - 1);
    }
}

/* ------------------------------------------------------------------ */
/* graph storage (include/hnsw_layer.hpp:57-465)                       */
/* ------------------------------------------------------------------ */

typedef struct {
    int64_t cap;     /* m_capacity */
    int D;           /* m_max_degree: 2M at layer 0, M above (hnsw.hpp:235) */
    int32_t next_id; /* m_next_id */
    int32_t ent_id;  /* m_ent_id, always 0 in v2 */
    int is_empty;
    int32_t *adj;    /* (cap + 1) * D, padded -1 (hnsw_layer.hpp:257) */
    float *dist;     /* (cap + 1) * D, padded FLT_MAX (:258) */
    int16_t *deg;
    int16_t *heur;
    int32_t *ext;
    uint8_t *data;   /* cap rows; zero until AddNode copies the vector */
    uint8_t *has_data;
    int32_t *ext2in; /* over the global N */
    uint32_t *vis;   /* ByteMap equivalent: epoch per node (visited.hpp:248-297) */
    uint32_t vis_epoch;
} layer_t;

struct orc_graph {
    int dtype, metric, dim, M, efc;
    int64_t n_total;
    size_t row_bytes;
    uint8_t *base;
    uint8_t *levels;
    int nlayers;
    int enter_level; /* m_enter_level */
    layer_t *layers;
};

static size_t dtype_bytes(int dtype)
{
    switch (dtype) {
    case ORC_U8:
    case ORC_I8: return 1;
    case ORC_F16: return 2;
    default: return 4;
    }
}

static void layer_alloc(layer_t *L, int64_t cap, int D, int64_t n_total, size_t row_bytes)
{
    memset(L, 0, sizeof(*L));
    L->cap = cap;
    L->D = D;
    L->is_empty = 1;
    size_t nslot = (size_t)(cap + 1) * (size_t)D;
    L->adj = (int32_t *)malloc(sizeof(int32_t) * nslot);
    L->dist = (float *)malloc(sizeof(float) * nslot);
    for (size_t i = 0; i < nslot; i++) {
        L->adj[i] = -1;
        L->dist[i] = 3.402823466e+38f;
    }
    L->deg = (int16_t *)calloc((size_t)cap + 1, sizeof(int16_t));
    L->heur = (int16_t *)calloc((size_t)cap + 1, sizeof(int16_t));
    L->ext = (int32_t *)malloc(sizeof(int32_t) * ((size_t)cap + 1));
    for (int64_t i = 0; i <= cap; i++) L->ext[i] = -1;
    L->data = (uint8_t *)calloc((size_t)cap + 1, row_bytes);
    L->has_data = (uint8_t *)calloc((size_t)cap + 1, 1);
    L->ext2in = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_total);
    for (int64_t i = 0; i < n_total; i++) L->ext2in[i] = -1;
    L->vis = (uint32_t *)calloc((size_t)cap + 1, sizeof(uint32_t));
    L->vis_epoch = 1;
}

static void layer_free(layer_t *L)
{
    free(L->adj);
    free(L->dist);
    free(L->deg);
    free(L->heur);
    free(L->ext);
    free(L->data);
    free(L->has_data);
    free(L->ext2in);
    free(L->vis);
}

/* HNSWGraph::InitBuffer (hnsw.hpp:222-245) with the levels given. */
orc_graph *orc_create(int dtype, int metric, int dim, int M, int efc, int64_t n_total,
                      const void *base, const uint8_t *levels)
{
    orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
    g->dtype = dtype;
    g->metric = metric;
    g->dim = dim;
    g->M = M;
    g->efc = efc;
    g->n_total = n_total;
    g->row_bytes = dtype_bytes(dtype) * (size_t)dim;
    g->base = (uint8_t *)malloc(g->row_bytes * (size_t)(n_total > 0 ? n_total : 1));
    if (base) memcpy(g->base, base, g->row_bytes * (size_t)n_total);
    g->levels = (uint8_t *)malloc((size_t)(n_total > 0 ? n_total : 1));
    int max_level = 0;
    for (int64_t i = 0; i < n_total; i++) {
        g->levels[i] = levels[i];
        if (levels[i] > max_level) max_level = levels[i];
    }
    g->nlayers = max_level + 1;
    g->enter_level = 0;
    g->layers = (layer_t *)calloc((size_t)g->nlayers, sizeof(layer_t));
    for (int l = 0; l < g->nlayers; l++) {
        int64_t cap = 0;
        for (int64_t i = 0; i < n_total; i++)
            if (levels[i] >= l) cap++;
        layer_alloc(&g->layers[l], cap, l == 0 ? 2 * M : M, n_total, g->row_bytes);
    }
    return g;
}

void orc_destroy(orc_graph *g)
{
    if (!g) return;
    for (int l = 0; l < g->nlayers; l++) layer_free(&g->layers[l]);
    free(g->layers);
    free(g->base);
    free(g->levels);
    free(g);
}

static inline const uint8_t *layer_vec(const orc_graph *g, const layer_t *L, int32_t vid)
{
    return L->data + (size_t)vid * g->row_bytes;
}

static inline float gdist(const orc_graph *g, const void *a, const void *b)
{
    return orc_distance(g->dtype, g->metric, a, b, g->dim);
}

static inline int32_t *row_adj(layer_t *L, int32_t vid) { return L->adj + (size_t)vid * L->D; }
static inline float *row_dist(layer_t *L, int32_t vid) { return L->dist + (size_t)vid * L->D; }

/* GetEntInID (hnsw_layer.hpp:186-190) */
static inline int32_t ent_in_id(const layer_t *L, int32_t ext_id)
{
    return ext_id == -1 ? L->ent_id : L->ext2in[ext_id];
}

/* NewInID (hnsw_layer.hpp:168-177) */
int32_t orc_new_in_id(orc_graph *g, int level, int32_t ext_id)
{
    layer_t *L = &g->layers[level];
    int32_t v = L->next_id++;
    L->ext[v] = ext_id;
    L->ext2in[ext_id] = v;
    L->is_empty = 0;
    return v;
}

/* ByteMap::Advance (visited.hpp:273-280): a new epoch empties the set. */
static void vis_advance(layer_t *L)
{
    L->vis_epoch++;
    if (L->vis_epoch == 0) {
        memset(L->vis, 0, sizeof(uint32_t) * (size_t)(L->cap + 1));
        L->vis_epoch = 1;
    }
}

static inline int vis_test(const layer_t *L, int32_t id) { return L->vis[id] == L->vis_epoch; }
static inline void vis_mark(layer_t *L, int32_t id) { L->vis[id] = L->vis_epoch; }

/* SearchLayer / SearchLayerFast (hnsw.hpp:304-371, 374-444).  Both compute
 * the same thing: best-first search bounded by ef.  W (nearest_neighbors,
 * MaxHeap) must be initialised by the caller and receives the result. */
static void search_layer(orc_graph *g, int32_t enter_id, int level, int ef, const void *q,
                         heap_t *W, orc_counters *ctr)
{
    layer_t *L = &g->layers[level];
    heap_t C;
    h_init(&C, 0);
    int64_t nhop = 0, ndist = 0, nnb = 0;
    float ent_dist = gdist(g, q, layer_vec(g, L, enter_id));
    h_push(W, ent_dist, enter_id);
    h_push(&C, ent_dist, enter_id);
    vis_mark(L, enter_id);
    while (C.len > 0) {
        orc_entry top = h_extract_top(&C);
        float f_dist = h_top(W).dist;
        if (top.dist > f_dist) break; /* hnsw.hpp:403 */
        int deg = L->deg[top.id];
        const int32_t *adj = row_adj(L, top.id);
        nhop++;
        nnb += deg;
        for (int i = 0; i < deg; i++) {
            int32_t vid = adj[i];
            if (vis_test(L, vid)) continue;
            ndist++;
            vis_mark(L, vid); /* marked before the distance (hnsw.hpp:422) */
            float d = gdist(g, q, layer_vec(g, L, vid));
            if (W->len < ef || h_top(W).dist > d) { /* hnsw.hpp:425-431 */
                h_push(W, d, vid);
                h_push(&C, d, vid);
                if (W->len > ef) h_pop(W);
            }
        }
    }
    vis_advance(L);
    h_free(&C);
    if (ctr) {
        ctr->hops[level] += nhop;
        ctr->dists[level] += ndist;
        ctr->neighbors[level] += nnb;
    }
}

/* SlideLayer (hnsw.hpp:446-494): greedy descent.  The adjacency scanned in
 * one pass is that of the node current at the start of the pass; the best
 * strictly-closer neighbour (first in row order on ties) wins. */
static int32_t slide_layer(orc_graph *g, int32_t enter_id, int level, const void *q,
                           orc_counters *ctr)
{
    layer_t *L = &g->layers[level];
    int32_t c_id = enter_id;
    float c_dist = gdist(g, q, layer_vec(g, L, enter_id));
    vis_mark(L, enter_id);
    int64_t nhop = 0, ndist = 0, nnb = 0;
    int updated;
    do {
        updated = 0;
        int32_t scan = c_id;
        int deg = L->deg[scan];
        const int32_t *adj = row_adj(L, scan);
        nhop++;
        nnb += deg;
        for (int i = 0; i < deg; i++) {
            int32_t vid = adj[i];
            if (vis_test(L, vid)) continue;
            ndist++;
            float d = gdist(g, q, layer_vec(g, L, vid));
            if (d < c_dist) {
                c_id = vid;
                c_dist = d;
                updated = 1;
            }
            vis_mark(L, vid);
        }
    } while (updated);
    vis_advance(L);
    if (ctr) {
        ctr->hops[level] += nhop;
        ctr->dists[level] += ndist;
        ctr->neighbors[level] += nnb;
    }
    return c_id;
}

typedef struct {
    orc_entry *a;
    int n, cap;
} vec_t;

static void v_push(vec_t *v, orc_entry e)
{
    if (v->n >= v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->a = (orc_entry *)realloc(v->a, sizeof(orc_entry) * (size_t)v->cap);
    }
    v->a[v->n++] = e;
}

/* SelectHeuristic (hnsw.hpp:70-116) with keep_pruned = true.  W is filled
 * from the candidate MaxHeap's array in storage order. */
static void select_heuristic(orc_graph *g, int top_k, int level, const heap_t *cand, vec_t *R,
                             vec_t *P)
{
    layer_t *L = &g->layers[level];
    heap_t W, Wd;
    h_init(&W, 0);
    h_init(&Wd, 0);
    R->n = 0;
    P->n = 0;
    for (int i = 0; i < cand->len; i++) h_push(&W, cand->a[i].dist, cand->a[i].id);
    while (W.len > 0 && R->n < top_k) {
        orc_entry v1 = h_top(&W);
        const uint8_t *v1d = layer_vec(g, L, v1.id);
        int ins = 1;
        for (int j = 0; j < R->n; j++) {
            float dd = gdist(g, v1d, layer_vec(g, L, R->a[j].id));
            if (v1.dist > dd) {
                ins = 0;
                break;
            }
        }
        if (ins)
            v_push(R, v1);
        else
            h_push(&Wd, v1.dist, v1.id);
        h_pop(&W);
    }
    while (Wd.len > 0 && R->n + P->n < top_k) v_push(P, h_extract_top(&Wd));
    h_free(&W);
    h_free(&Wd);
}

/* AddNode (hnsw_layer.hpp:297-332) */
int orc_add_node(orc_graph *g, int level, int32_t vid, const orc_entry *R, int nr,
                 const orc_entry *P, int np, const void *vdata)
{
    layer_t *L = &g->layers[level];
    memcpy(L->data + (size_t)vid * g->row_bytes, vdata, g->row_bytes);
    L->has_data[vid] = 1;
    int32_t *adj = row_adj(L, vid);
    float *dist = row_dist(L, vid);
    int h = nr < L->D ? nr : L->D;
    for (int i = 0; i < h; i++) {
        adj[i] = R[i].id;
        dist[i] = R[i].dist;
    }
    int degree = h + np < L->D ? h + np : L->D;
    for (int i = 0; i < degree - h; i++) {
        adj[i + h] = P[i].id;
        dist[i + h] = P[i].dist;
    }
    L->deg[vid] = (int16_t)degree;
    L->heur[vid] = (int16_t)h;
    return 0;
}

/* first index in [0, n) whose value is not less than x (std::lower_bound) */
static int lower_bound_f(const float *a, int n, float x)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

/* AddEdge (hnsw_layer.hpp:360-465).  Row accesses are flat offsets from
 * the row start, exactly as the reference's spans: when a full row keeps
 * every entry in the heuristic re-run, the write at slot D lands in the next
 * row's slot 0 and the degree becomes D + 1 (:441-447, not capped). */
int orc_add_edge(orc_graph *g, int level, int32_t vid, int32_t nid, float distance)
{
    layer_t *L = &g->layers[level];
    int32_t *adj = row_adj(L, vid);
    float *dist = row_dist(L, vid);
    int heuristic = L->heur[vid];
    int degree = L->deg[vid];
    int offset = lower_bound_f(dist, heuristic, distance);
    const uint8_t *nd = layer_vec(g, L, nid);
    int insert_n = 1;
    for (int i = 0; i < offset; i++) {
        float dd = gdist(g, nd, layer_vec(g, L, adj[i]));
        if (distance > dd) {
            insert_n = 0;
            break;
        }
    }
    if (!insert_n) {
        int32_t *adj_p = adj + heuristic;
        float *dist_p = dist + heuristic;
        int pruned = degree - heuristic;
        int off = lower_bound_f(dist_p, pruned, distance);
        if (off == pruned && degree < L->D) {
            adj_p[off] = nid;
            dist_p[off] = distance;
            L->deg[vid] = (int16_t)(degree + 1);
        } else if (off < pruned) {
            int end = (degree < L->D) ? pruned : pruned - 1;
            for (int i = end; i > off; i--) {
                adj_p[i] = adj_p[i - 1];
                dist_p[i] = dist_p[i - 1];
            }
            adj_p[off] = nid;
            dist_p[off] = distance;
            degree += (degree < L->D);
            L->deg[vid] = (int16_t)degree;
        }
    } else {
        heap_t W, Wd;
        h_init(&W, 0);
        h_init(&Wd, 0);
        for (int i = offset; i < degree; i++) h_push(&W, dist[i], adj[i]);
        adj[offset] = nid;
        dist[offset] = distance;
        int cur_h = offset + 1;
        while (W.len > 0) {
            orc_entry v1 = h_extract_top(&W);
            const uint8_t *v1d = layer_vec(g, L, v1.id);
            int ins = 1;
            for (int i = 0; i < cur_h; i++) {
                float dd = gdist(g, v1d, layer_vec(g, L, adj[i]));
                if (dd < v1.dist) {
                    ins = 0;
                    break;
                }
            }
            if (ins) {
                adj[cur_h] = v1.id;
                dist[cur_h] = v1.dist;
                cur_h++;
            } else {
                h_push(&Wd, v1.dist, v1.id);
            }
        }
        int cur_deg = cur_h;
        while (cur_deg < L->D && Wd.len > 0) {
            orc_entry e = h_extract_top(&Wd);
            adj[cur_deg] = e.id;
            dist[cur_deg] = e.dist;
            cur_deg++;
        }
        L->heur[vid] = (int16_t)cur_h;
        L->deg[vid] = (int16_t)cur_deg;
        h_free(&W);
        h_free(&Wd);
    }
    return 0;
}

/* Insert (hnsw.hpp:247-302) */
int orc_insert(orc_graph *g, int32_t ext_id)
{
    if (ext_id < 0 || ext_id >= g->n_total) return -1;
    const uint8_t *q = g->base + (size_t)ext_id * g->row_bytes;
    int l = g->levels[ext_id];
    int Lv = g->enter_level;
    int32_t enter_ext = -1, enter_in = -1;
    for (int level = Lv; level >= l + 1; level--) {
        if (!g->layers[level].is_empty) {
            enter_in = ent_in_id(&g->layers[level], enter_ext);
            enter_in = slide_layer(g, enter_in, level, q, NULL);
            enter_ext = g->layers[level].ext[enter_in];
        }
    }
    int top_search = l < Lv ? l : Lv;
    int top_inserted = 0;
    vec_t R[ORC_MAX_LEVELS], P[ORC_MAX_LEVELS];
    memset(R, 0, sizeof(R));
    memset(P, 0, sizeof(P));
    for (int level = top_search; level >= 0; level--) {
        layer_t *Ly = &g->layers[level];
        if (Ly->is_empty) continue;
        enter_in = ent_in_id(Ly, enter_ext);
        int top_k = level == 0 ? 2 * g->M : g->M;
        heap_t Wr;
        h_init(&Wr, 1);
        search_layer(g, enter_in, level, g->efc, q, &Wr, NULL);
        select_heuristic(g, top_k, level, &Wr, &R[level], &P[level]);
        h_free(&Wr);
        enter_in = R[level].a[0].id;
        enter_ext = Ly->ext[enter_in];
        if (level > top_inserted) top_inserted = level;
    }
    for (int level = 0; level <= l; level++) {
        int32_t vid = orc_new_in_id(g, level, ext_id);
        if (level <= top_inserted) {
            orc_add_node(g, level, vid, R[level].a, R[level].n, P[level].a, P[level].n, q);
            for (int i = 0; i < R[level].n; i++)
                orc_add_edge(g, level, R[level].a[i].id, vid, R[level].a[i].dist);
            for (int i = 0; i < P[level].n; i++)
                orc_add_edge(g, level, P[level].a[i].id, vid, P[level].a[i].dist);
        }
    }
    for (int i = 0; i < ORC_MAX_LEVELS; i++) {
        free(R[i].a);
        free(P[i].a);
    }
    if (l > g->enter_level) g->enter_level = l; /* SetEntLevel (hnsw.hpp:146-151) */
    return 0;
}

int orc_build(orc_graph *g, const int32_t *order, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (orc_insert(g, order[i]) != 0) return -1;
    return 0;
}

/* AnnSearch (hnsw.hpp:509-525) + SelectSimpleExt (:118-125) */
int orc_search(orc_graph *g, int k, int ef, const void *query, int32_t *out_ids,
               float *out_dists, orc_counters *ctr)
{
    if (g->layers[0].is_empty) return 0;
    int32_t enter_ext = -1, enter_in = -1;
    for (int level = g->enter_level; level >= 1; level--) {
        layer_t *L = &g->layers[level];
        enter_in = ent_in_id(L, enter_ext);
        enter_in = slide_layer(g, enter_in, level, query, ctr);
        enter_ext = L->ext[enter_in];
    }
    layer_t *L0 = &g->layers[0];
    enter_in = ent_in_id(L0, enter_ext);
    heap_t W;
    h_init(&W, 1);
    search_layer(g, enter_in, 0, ef, query, &W, ctr);
    orc_entry *tmp = (orc_entry *)malloc(sizeof(orc_entry) * (size_t)(W.len + 1));
    int n = h_extract_min_topk(&W, k, tmp);
    for (int i = 0; i < n; i++) {
        out_ids[i] = L0->ext[tmp[i].id];
        out_dists[i] = tmp[i].dist;
    }
    free(tmp);
    h_free(&W);
    return n;
}

int orc_search_batch(orc_graph *g, int k, int ef, const void *queries, int64_t nq,
                     int32_t *out_ids, float *out_dists, int32_t *out_n, orc_counters *ctr)
{
    for (int64_t i = 0; i < nq; i++) {
        int32_t *ids = out_ids + (size_t)i * k;
        float *ds = out_dists + (size_t)i * k;
        for (int j = 0; j < k; j++) {
            ids[j] = -1;
            ds[j] = INFINITY;
        }
        int n = orc_search(g, k, ef, (const uint8_t *)queries + (size_t)i * g->row_bytes, ids, ds,
                           ctr);
        if (out_n) out_n[i] = n;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* accessors                                                           */
/* ------------------------------------------------------------------ */

int orc_nlayers(const orc_graph *g) { return g->nlayers; }
int orc_enter_level(const orc_graph *g) { return g->enter_level; }
int64_t orc_layer_capacity(const orc_graph *g, int l) { return g->layers[l].cap; }
int orc_layer_max_degree(const orc_graph *g, int l) { return g->layers[l].D; }
int32_t orc_layer_next_id(const orc_graph *g, int l) { return g->layers[l].next_id; }
int32_t *orc_layer_adj(orc_graph *g, int l) { return g->layers[l].adj; }
float *orc_layer_dist(orc_graph *g, int l) { return g->layers[l].dist; }
int16_t *orc_layer_degree(orc_graph *g, int l) { return g->layers[l].deg; }
int16_t *orc_layer_heuristic(orc_graph *g, int l) { return g->layers[l].heur; }
int32_t *orc_layer_ext(orc_graph *g, int l) { return g->layers[l].ext; }
uint8_t *orc_layer_data(orc_graph *g, int l) { return g->layers[l].data; }
uint8_t *orc_layer_has_data(orc_graph *g, int l) { return g->layers[l].has_data; }

int orc_set_layer(orc_graph *g, int level, int64_t cap, int D, int32_t next_id,
                  const int32_t *adj, const float *dist, const int16_t *deg,
                  const int16_t *heur, const int32_t *ext, const uint8_t *data,
                  const uint8_t *has_data)
{
    if (level < 0 || level >= g->nlayers) return -1;
    layer_t *L = &g->layers[level];
    layer_free(L);
    layer_alloc(L, cap, D, g->n_total, g->row_bytes);
    size_t nslot = (size_t)cap * (size_t)D;
    memcpy(L->adj, adj, sizeof(int32_t) * nslot);
    memcpy(L->dist, dist, sizeof(float) * nslot);
    memcpy(L->deg, deg, sizeof(int16_t) * (size_t)cap);
    memcpy(L->heur, heur, sizeof(int16_t) * (size_t)cap);
    memcpy(L->ext, ext, sizeof(int32_t) * (size_t)cap);
    memcpy(L->data, data, g->row_bytes * (size_t)cap);
    memcpy(L->has_data, has_data, (size_t)cap);
    L->next_id = next_id;
    L->is_empty = next_id == 0;
    for (int32_t i = 0; i < next_id; i++)
        if (L->ext[i] >= 0 && L->ext[i] < g->n_total) L->ext2in[L->ext[i]] = i;
    return 0;
}

int orc_set_enter_level(orc_graph *g, int level)
{
    if (level < 0 || level >= g->nlayers) return -1;
    g->enter_level = level;
    return 0;
}

/* IsValid (hnsw_layer.hpp:278-295): heuristic and pruned segments each
 * sorted by distance with no empty ids and no adjacent repeats. */
int64_t orc_layer_invalid_count(orc_graph *g, int level)
{
    layer_t *L = &g->layers[level];
    int64_t bad = 0;
    for (int32_t v = 0; v < L->next_id; v++) {
        const int32_t *adj = row_adj(L, v);
        const float *dist = row_dist(L, v);
        int h = L->heur[v], d = L->deg[v];
        int ok = h <= d;
        for (int i = 1; i < h && ok; i++)
            if (adj[i] == -1 || adj[i] == adj[i - 1] || dist[i] < dist[i - 1]) ok = 0;
        for (int i = h + 1; i < d && ok; i++)
            if (adj[i] == -1 || adj[i] == adj[i - 1] || dist[i] < dist[i - 1]) ok = 0;
        if (!ok) bad++;
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* FP64 brute force ground truth                                       */
/* ------------------------------------------------------------------ */

static double elem(int dtype, const void *p, int64_t i)
{
    switch (dtype) {
    case ORC_U8: return ((const uint8_t *)p)[i];
    case ORC_I8: return ((const int8_t *)p)[i];
    case ORC_F16: return half_to_float(((const uint16_t *)p)[i]);
    default: return ((const float *)p)[i];
    }
}

void orc_brute_force(int dtype, int metric, int dim, const void *base, int64_t n,
                     const void *queries, int64_t nq, int k, int32_t *out_ids, double *out_dists)
{
    double *bd = (double *)malloc(sizeof(double) * (size_t)k);
    int32_t *bi = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    for (int64_t qi = 0; qi < nq; qi++) {
        int cnt = 0;
        for (int64_t v = 0; v < n; v++) {
            double s = 0.0;
            for (int j = 0; j < dim; j++) {
                double a = elem(dtype, queries, qi * dim + j);
                double b = elem(dtype, base, v * dim + j);
                if (metric == ORC_L2)
                    s += (a - b) * (a - b);
                else
                    s += a * b;
            }
            if (metric == ORC_IP) s = -s;
            /* insertion into a sorted top-k; strict compare keeps lower ids first */
            if (cnt < k || s < bd[cnt - 1]) {
                int pos = cnt < k ? cnt : k - 1;
                while (pos > 0 && bd[pos - 1] > s) {
                    bd[pos] = bd[pos - 1];
                    bi[pos] = bi[pos - 1];
                    pos--;
                }
                bd[pos] = s;
                bi[pos] = (int32_t)v;
                if (cnt < k) cnt++;
            }
        }
        for (int j = 0; j < k; j++) {
            out_ids[qi * k + j] = j < cnt ? bi[j] : -1;
            out_dists[qi * k + j] = j < cnt ? bd[j] : INFINITY;
        }
    }
    free(bd);
    free(bi);
}
