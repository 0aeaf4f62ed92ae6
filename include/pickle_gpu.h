/*
 * pickle_gpu.h -- C-ABI of the MI355X HNSW query and insertion path.
 *
 * libpickle_amd.so implements pickle v2's HNSWGraph search and insertion
 * (reference include/hnsw.hpp, include/hnsw_layer.hpp) with HIP kernels for
 * gfx950.  Every entry point takes plain pointers and sizes, returns a status
 * (PK_OK = 0, negative on error, never aborts) and leaves the message for
 * pk_last_error().  The library owns the HBM-resident index and its scratch
 * memory; callers own every buffer they pass in.  Calls on one handle are
 * serialised on the handle's HIP stream.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * the reference repository).
 */
#ifndef PICKLE_GPU_H
#define PICKLE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PK_OK 0
#define PK_ERR_ARG -1      /* bad argument (shape, dtype, metric, range) */
#define PK_ERR_HIP -2      /* a HIP runtime call failed */
#define PK_ERR_NOMEM -3    /* device allocation failed */
#define PK_ERR_STATE -4    /* call not valid in the handle's current state */
#define PK_ERR_IO -5       /* file read / write failed or malformed file */
#define PK_ERR_CAPACITY -6 /* a fixed capacity would be exceeded */
#define PK_ERR_DEVICE -7   /* a kernel reported a failure (see pk_last_error) */

/* element types: DataType, include/common.hpp:21-26 (same codes) */
#define PK_UINT8 0
#define PK_INT8 1
#define PK_FLOAT32 2
#define PK_FLOAT16 3

/* metrics: DistFunc, include/common.hpp:28-33 (same codes; L1 unsupported,
 * as in include/distance/avx512.hpp:233-243) */
#define PK_L2 1
#define PK_IP 2

#define PK_MAX_LEVELS 16

typedef struct pk_index pk_index;
typedef struct pk_search_graph pk_search_graph;
typedef struct pk_server pk_server;

/* Per-query work counters with the reference Profiler's definitions
 * (include/profiler.hpp:13-69, updated at include/hnsw.hpp:438-442 and
 * :488-492): hops = expanded nodes, neighbors = adjacency entries scanned,
 * dists = distance evaluations.  Index [0] is layer 0, [1] all upper layers.
 * latency_ticks: device wall-clock ticks (pk_wall_clock_khz) from the moment
 * a wave takes the query to the moment its results are written. */
typedef struct {
    int32_t hops[2];
    int32_t neighbors[2];
    int32_t dists[2];
    int32_t latency_ticks;
} pk_query_counters;

/* Message of the last failed call on this thread ("" if none). */
const char *pk_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int pk_version(void);

/* HNSWGraph(max_degree, build_ef, df) + InitBuffer(max_nodes)
 * (include/hnsw.hpp:27-38, 222-245).  M is the reference's max_degree
 * argument (layer-0 rows hold 2M ids, upper rows M); efc is build_ef;
 * capacity is the number of external ids the index can ever hold. */
int pk_index_create(int dtype, int metric, int dim, int M, int efc, int64_t capacity,
                    pk_index **out);

/* Frees the index and all device memory it owns. */
int pk_index_destroy(pk_index *h);

/* Base vectors by external id: the NDArray the reference keeps as m_data
 * (include/hnsw.hpp:140, 497-501).  Host rows, n x dim of the index dtype.
 * Rows [first, first + n) are (over)written. */
int pk_index_set_base(pk_index *h, const void *host_rows, int64_t first, int64_t n);

/* One layer of a graph in the serializer's layout (include/serializer.hpp:
 * 27-55) plus v2's per-node heuristic count and vector flag
 * (include/hnsw_layer.hpp:60-65, 83).  adj/dist hold adj_len >= n * D items:
 * rows of stride D, plus any slots a degree above D reads past the last row
 * (include/hnsw_layer.hpp:441-447 lets a row grow past D).  has_vector[i]
 * != 0 means the layer's copy of node i is base row ext[i]; 0 means zeros.
 * dist and heur may be NULL for a search-only index. */
int pk_index_load_layer(pk_index *h, int level, int64_t n, int D, int64_t adj_len,
                        const int32_t *ext, const int32_t *degree, const int16_t *heur,
                        const int32_t *adj, const float *dist, const uint8_t *has_vector);

/* m_enter_level (include/hnsw.hpp:133, 142-151).  Loads are complete after
 * this call: the search kernels read the layers from here on. */
int pk_index_set_enter_level(pk_index *h, int enter_level);

/* Reads a graph file in the serializer layout plus its ".v2meta" sidecar
 * (the format is described in pickle_amd/serializer.py); base rows must
 * already be set.  Replaces Serializer::from_disk (include/serializer.hpp:57-81). */
int pk_index_load_file(pk_index *h, const char *path);

/* Writes the index's graph in the same format.  Replaces
 * Serializer::to_disk (include/serializer.hpp:27-55). */
int pk_index_save_file(pk_index *h, const char *path);

/* Graph shape. */
int pk_index_info(pk_index *h, int *nlayers, int *enter_level, int64_t *layer_sizes,
                  int *layer_degrees);

/* Copies layer `level` back to host arrays (any pointer may be NULL).
 * adj/dist receive n * D items. */
int pk_index_read_layer(pk_index *h, int level, int32_t *ext, int32_t *degree, int16_t *heur,
                        int32_t *adj, float *dist);

/* Batched AnnSearch (include/hnsw.hpp:509-541 and SelectSimpleExt :118-125):
 * for each query i, greedy descent through the upper layers, the ef[i]-bounded
 * layer-0 beam search and the ascending top-k[i] with external ids.  Results
 * are bit-identical to the reference for every dtype: f16/f32 distances sum
 * in the order of the reference's compiled AVX-512 loop, dim % 16 tail
 * included.
 *   queries   nq x dim rows of the index dtype
 *   k, ef     per-query arrays (or NULL with k_all / ef_all used for all)
 *   out_ids   nq x k_stride external ids, -1 past the result count
 *   out_dists nq x k_stride distances (float32; -IP for IP), +inf past it
 *   out_count nq result counts (min(k, |W|)), may be NULL
 *   counters  nq per-query counters, may be NULL
 * Host pointers; the call copies in, runs and copies out synchronously.
 * ef is at most 630 (the result heap lives in LDS; the reference has no such
 * bound): a larger ef, like a build ef above 630, returns PK_ERR_ARG.
 * Queries that fill their visited table are searched again with larger
 * tables (up to 2^20 slots; results do not depend on the size), so this call
 * returns PK_ERR_DEVICE only if that still fails or captured search graphs
 * keep the scratch from growing.  The pk_search_device* calls report such
 * queries through their counts and pk_device_error instead.
 * Batches of 8,192 queries or more (here, in pk_search_device and
 * pk_search_device_ex, and in pk_search_fast*) are processed region by
 * region: each query is keyed by its nearest of up to 64 pivots (the first
 * nodes of an upper layer), and the keys are sorted on the device, so that
 * queries sharing rows run together.  Results, counts and counters are per
 * query and do not depend on the order.  The first such call on a handle
 * allocates about 12 bytes per query for it (kept and grown with the
 * handle); calls during stream capture keep the caller's order. */
int pk_search(pk_index *h, const void *queries, int64_t nq, const int32_t *k, const int32_t *ef,
              int32_t k_all, int32_t ef_all, int32_t k_stride, int32_t *out_ids, float *out_dists,
              int32_t *out_count, pk_query_counters *counters);

/* Same as pk_search with every pointer a device pointer (HBM-resident
 * inputs and outputs); enqueued on the handle's stream, returns without
 * waiting.  k/ef may be NULL (k_all/ef_all then apply). */
int pk_search_device(pk_index *h, const void *d_queries, int64_t nq, const int32_t *d_k,
                     const int32_t *d_ef, int32_t k_all, int32_t ef_all, int32_t k_stride,
                     int32_t *d_out_ids, float *d_out_dists, int32_t *d_out_count,
                     pk_query_counters *d_counters);

/* pk_search_device with per-query ef on the device: ef_max bounds every
 * d_ef[i] (it sizes the LDS result heap).  Same reference interface. */
int pk_search_device_ex(pk_index *h, const void *d_queries, int64_t nq, const int32_t *d_k,
                        const int32_t *d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max,
                        int32_t k_stride, int32_t *d_out_ids, float *d_out_dists,
                        int32_t *d_out_count, pk_query_counters *d_counters);

/* Fast mode of the batched search (no counterpart in the reference; its
 * caller is the same AnnSearch(k, ef, NDArray), include/hnsw.hpp:528-541).
 * The reference's accept and stopping rules (hnsw.hpp:398-434) run on a
 * sorted beam kept in registers: each round expands the `width` closest
 * unexpanded entries together, new neighbours are merged with wave bitonic
 * networks, and the visited filter lives in LDS (forgotten ids are only
 * evaluated again, never returned twice).  Results are recall-equivalent,
 * not bit-exact: ties are ordered by id, and width > 1 expands a few more
 * nodes.  Device buffers, enqueued on the handle's stream; ef_max <= 512;
 * no visited state in device memory.  Failures: count -1 and
 * pk_device_error flags as pk_search_device_ex. */
int pk_search_fast_device(pk_index *h, const void *d_queries, int64_t nq, const int32_t *d_k,
                          const int32_t *d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max,
                          int32_t k_stride, int32_t *d_out_ids, float *d_out_dists,
                          int32_t *d_out_count, pk_query_counters *d_counters);

/* pk_search_fast_device with host buffers, as pk_search (copies in, runs on
 * the handle's stream, copies out; synchronous); ef <= 512.  The caller
 * AnnSearch(k, ef, NDArray) (include/hnsw.hpp:528-541) in fast mode.
 * PK_ERR_DEVICE if a query failed on the device (flags in the message). */
int pk_search_fast(pk_index *h, const void *queries, int64_t nq, const int32_t *k, const int32_t *ef,
                   int32_t k_all, int32_t ef_all, int32_t k_stride, int32_t *out_ids, float *out_dists,
                   int32_t *out_count, pk_query_counters *counters);

/* Fast-mode tuning (0 = automatic for each): width 1..4 entries expanded
 * per round (default 2), hash_bits 8..14 (each query's LDS visited filter
 * takes 4 << hash_bits bytes: 2^(hash_bits + 1) 16-bit entries; default 11), waves_per_group 1, 2 or 4 (queries per
 * workgroup, default 2), blocks_per_cu (cap on resident workgroups).  No
 * tuning changes a result except width. */
int pk_set_fast_tuning(pk_index *h, int width, int hash_bits, int waves_per_group, int blocks_per_cu);

/* Autotuner of the fast mode (no counterpart in the reference): coordinate
 * descent at one (k, ef) on the caller's device queries over queries per
 * workgroup, the LDS filter size, the expansion width and resident
 * workgroups per CU; keeps the fastest setting on the handle.  out_tuning
 * (4 ints: width, hash_bits, waves_per_group, blocks_per_cu) and out_ms may
 * be NULL.  Width changes results (recall rises with it); the rest do not. */
int pk_autotune_fast(pk_index *h, const void *d_queries, int64_t nq, int32_t k, int32_t ef,
                     int32_t *out_tuning, float *out_ms);

/* pk_search_device_ex without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels run with all of their device scratch (work counter, failure flags,
 * visited tables and logs, candidate-heap spill) in one static device array
 * of 128 MiB per device, held until the process ends; the call makes no
 * hipMalloc, hipFree or other allocating call.  Launches take as many waves
 * as their visited tables fit in the array (one workgroup needs 2 x 2^vbits
 * x 16 B plus the heap spill; every supported size fits), so a large batch
 * runs on fewer waves than pk_search_device_ex uses.  A query that fills its
 * table ends with count -1, as in pk_search_device_ex; its flags are not
 * reported through pk_device_error.  Calls on one device share the array:
 * calls that overlap in time on one device race. */
int pk_search_device_static(pk_index *h, const void *d_queries, int64_t nq, const int32_t *d_k,
                            const int32_t *d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max,
                            int32_t k_stride, int32_t *d_out_ids, float *d_out_dists,
                            int32_t *d_out_count, pk_query_counters *d_counters);

/* One pk_search_device_ex launch captured into a HIP graph (no counterpart
 * in the reference): the same buffers, shapes and per-query k / ef arrays
 * are replayed by pk_search_graph_launch on the index's stream, so a serving
 * loop or a repeated batch pays one graph launch.  Buffer contents may change
 * between replays; pointers may not.  While a graph exists, calls that would
 * reallocate what it points to (a search needing more scratch, loading
 * layers, init_build, copy_graph) fail with PK_ERR_STATE. */
int pk_search_graph_create(pk_index *h, const void *d_queries, int64_t nq, const int32_t *d_k,
                           const int32_t *d_ef, int32_t k_all, int32_t ef_all, int32_t ef_max,
                           int32_t k_stride, int32_t *d_out_ids, float *d_out_dists,
                           int32_t *d_out_count, pk_query_counters *d_counters,
                           pk_search_graph **out);
int pk_search_graph_launch(pk_search_graph *g);
int pk_search_graph_destroy(pk_search_graph *g);

/* Autotuner (no counterpart in the reference): times the search on the
 * caller's sample queries (device buffer) at one (k, ef) over the resident
 * workgroups per CU, the candidate heap's LDS depth and the visited-table
 * size, keeps the fastest as the index's tuning, and returns it as
 * out_tuning[3] = {ccap, vbits, blocks_per_cu} (0 = automatic) with its time
 * per launch in *out_ms.  Results do not depend on the tuning. */
int pk_autotune_search(pk_index *h, const void *d_queries, int64_t nq, int32_t k, int32_t ef,
                       int32_t *out_tuning, float *out_ms);

/* ---- serving loop (no counterpart in the reference; its callers run
 * AnnSearch per request on host threads, include/hnsw.hpp:509-541) ----
 * A request loop over one resident index with continuous batching: `slots`
 * launch slots (1..8), each with its own stream, search scratch and pinned
 * staging for `slot_queries` queries; a free slot takes waiting requests in
 * arrival order (requests are not split), so requests that arrive while
 * others run start as slots free up, and slots overlap on the device.  Each
 * request has its own k <= max_k and k <= ef <= max_ef.  Inserts keep their
 * own arrival order and run on a second thread beside the searches (the
 * index must come from pk_index_init_build): as with the reference's
 * concurrent Insert and AnnSearch, a search running meanwhile may see part
 * of an insert, and every search submitted after the insert's
 * pk_server_wait returns sees all of it.  All calls are thread-safe.  While a server exists
 * the index's layers cannot be replaced; the handle's own searches still run. */
int pk_server_create(pk_index *h, int slots, int slot_queries, int max_k, int max_ef, pk_server **out);
/* pk_server_create with flags.  PK_SERVE_FAST: launches run the fast mode
 * (pk_search_fast_device and the handle's fast tuning; max_ef <= 512).  A
 * fast-mode query's result depends only on the graph, the query, its k and
 * ef and the handle's fast tuning (which launches do not change), so it is
 * still the same alone and in the mix; it is recall-equivalent to the exact search, not bit-exact. */
#define PK_SERVE_FAST 1
int pk_server_create_ex(pk_index *h, int slots, int slot_queries, int max_k, int max_ef, int flags,
                        pk_server **out);
/* Queues nq queries (host rows, copied at once); results go to the caller's
 * out_ids / out_dists (nq x k) and out_count (nq, may be NULL) when the
 * request completes, so those buffers must outlive the request.  Queries that
 * fill their visited table run again with tables twice the size, up to 2^20
 * slots, before their request completes (results do not depend on the size);
 * past that the request fails with PK_ERR_DEVICE and count -1 for those
 * queries.  Other requests are not held back. */
int pk_server_submit(pk_server *s, const void *queries, int64_t nq, int32_t k, int32_t ef,
                     int32_t *out_ids, float *out_dists, int32_t *out_count, int64_t *ticket);
/* Queues Insert of n rows (host, by external id; copied at once). */
int pk_server_insert(pk_server *s, const int32_t *ext_ids, const void *rows, int64_t n, int flags,
                     int64_t *ticket);
/* Waits for a request (timeout_ms < 0: no limit) and releases its ticket;
 * latency_us = submit to completion, queue_us = submit to admission.  One
 * caller at a time may wait on a ticket; after a timeout it stays valid. */
int pk_server_wait(pk_server *s, int64_t ticket, int32_t timeout_ms, double *latency_us, double *queue_us);
/* out[5] = {launches, queries done, inserts done, most slots busy at once, requests waiting}. */
int pk_server_stats(pk_server *s, int64_t *out);
/* Finishes queued work, then frees the slots. */
int pk_server_destroy(pk_server *s);

/* Device failure flags of the last search, or'ed over its queries (0 = none;
 * 1 = a visited table filled up, 4 = a candidate heap filled up -- both
 * scratch capacities; 2 = a query had k or ef out of range).  Waits for the
 * stream. */
int pk_device_error(pk_index *h, int *flags);

/* Search launch shape: LDS candidate-heap entries, log2 of the slots of each
 * wave's visited table and a cap on resident workgroups per CU (0 = automatic).  No reference
 * counterpart (the reference has no device launch); set by the autotuner. */
int pk_set_tuning(pk_index *h, int ccap, int vbits, int blocks_per_cu);

/* Grid size and LDS bytes per workgroup of the last search launch. */
int pk_last_launch(pk_index *h, int *grid, int *lds_bytes);

/* Exhaustive nearest neighbours of nq host queries over the base rows set
 * with pk_index_set_base (rows [0, highest row set)): the 32 closest per
 * query, ascending by (distance, row id), to out_ids / out_dists (nq x 32);
 * -1 / +inf past the row count.  Dot products run on the matrix cores:
 * u8 / i8 as int8 MFMA with exact int32 distances; f16 / f32 as FP64 MFMA
 * (distances in FP64, returned rounded to FP32).  Rows up to 256 bytes
 * (u8 / i8) or 512 bytes (f16 / f32).  No reference counterpart on the hot
 * path: the reference reads precomputed ground-truth files
 * (benchmark/bench_util.hpp:196-202); this produces them. */
int pk_brute_force(pk_index *h, const void *queries, int64_t nq, int32_t *out_ids,
                   float *out_dists);

/* Frequency of the device wall clock behind latency_ticks, in kHz. */
int pk_wall_clock_khz(int *khz);

/* ---- insertion ---- */

/* pk_index_insert flag: reproduce the reference's AddEdge exactly, including
 * its write one slot past a full row (include/hnsw_layer.hpp:441-447); with
 * max_batch = 1 the build equals the reference's single-threaded Build.
 * Without it rows never exceed D entries. */
#define PK_BUILD_EXACT_ROWS 1
/* pk_index_insert flag (no counterpart in the reference): each batch's two
 * kernel phases (the construction searches; AddNode + AddEdge) run as HIP
 * graphs, captured from the launch sequence, instantiated once per index and
 * updated in place for every later batch.  Results are those of the stream
 * launches.  While it runs, nothing else may be enqueued on the index's own
 * stream (its own searches), as for any insertion beside searches. */
#define PK_BUILD_GRAPH 2

/* Level sampler of include/util.hpp:13-64: a fresh uniform per level,
 * P(level >= l) = M^(-l(l+1)/2); seeded (the reference seeds from
 * random_device). */
int pk_sample_levels(int64_t n, int M, uint64_t seed, uint8_t *out);

/* InitBuffer (include/hnsw.hpp:222-245): levels[capacity] is every external
 * id's insert level; sizes every layer for them and empties the graph.  Base
 * rows must be set. */
int pk_index_init_build(pk_index *h, const uint8_t *levels);

/* Insert (include/hnsw.hpp:247-302) for ext_ids in order, on the GPU in
 * batches of up to max_batch nodes (batches grow with the graph; a node that
 * opens a new top level goes alone).  Build (:497-507) is init_build + one
 * insert of all ids.  flags: PK_BUILD_EXACT_ROWS or 0.  Synchronous.
 * A batch whose construction searches fill their visited tables reruns
 * those searches with larger tables, up to 2^20 slots.  A device failure
 * past that returns PK_ERR_DEVICE with its flags in the message (1 a visited
 * table filled up, 4 a candidate heap filled up, 8 a node lock stayed taken
 * for seconds, 16 a row already held 511 ids, the longest row AddEdge
 * handles). */
int pk_index_insert(pk_index *h, const int32_t *ext_ids, int64_t n, int max_batch, int flags);

/* Longest row a PK_BUILD_EXACT_ROWS insertion writes: the reference's rows
 * grow without bound when its heuristic re-run keeps more than D ids
 * (hnsw_layer.hpp:441-447); here at most `limit` ids (0: 511, else 2M..511)
 * are kept and a neighbour that would land past them is dropped, as a full
 * row's occluded neighbour is.  No counterpart in the reference. */
int pk_set_build_row_limit(pk_index *h, int limit);

/* ---- precision form: low-precision traversal copy + exact FP32 re-rank ----
 * No counterpart in the reference (its search reads one copy of the data).
 * An index of dtype PK_FLOAT16 or PK_INT8 holds the traversal copy; these
 * calls attach the FP32 rows and search with FP32 queries.  The search runs
 * the ordinary batched AnnSearch (include/hnsw.hpp:509-541) on the copy with
 * k = ef, then re-ranks those ef candidates by FP32 distances computed in the
 * reference's own order (include/distance/avx512.hpp), ascending, ties by
 * traversal rank; returned distances are the reference's FP32 values. */

/* Per-dimension int8 quantizer of an int8 traversal copy:
 * q[d] = clamp(rint((x[d] - offset[d]) * (1 / scale[d])), -127, 127). */
int pk_index_set_quantizer(pk_index *h, const float *offset, const float *scale);

/* FP32 rows [first, first + n) by external id (host, n x dim); the traversal
 * copy of the same rows is converted on the device and becomes the index's
 * base rows (replacing pk_index_set_base for this index). */
int pk_index_set_full_rows(pk_index *h, const float *host_rows, int64_t first, int64_t n);

/* Makes dst search the graph of src (same capacity, dim and M): layers are
 * copied on the device and dst's layer vectors gathered from dst's rows.
 * The copy is search-only (no insertion, no save). */
int pk_index_copy_graph(pk_index *dst, pk_index *src);

/* Precision-form search with FP32 queries (device buffers, nq x dim);
 * ef >= k, one ef for all queries.  Outputs as pk_search_device. */
int pk_search_rerank_device(pk_index *h, const float *d_queries, int64_t nq, int32_t k, int32_t ef,
                            int32_t k_stride, int32_t *d_out_ids, float *d_out_dists, int32_t *d_out_count,
                            pk_query_counters *d_counters);

/* pk_search_rerank_device with the fast-mode traversal (pk_search_fast_device
 * and its tuning) on the low-precision copy; ef <= 512.  The re-rank is the
 * same exact FP32 one. */
int pk_search_rerank_fast_device(pk_index *h, const float *d_queries, int64_t nq, int32_t k, int32_t ef,
                                 int32_t k_stride, int32_t *d_out_ids, float *d_out_dists,
                                 int32_t *d_out_count, pk_query_counters *d_counters);
/* The same with host buffers (copies in, runs, copies out).  As in
 * pk_search, queries whose traversal fills its visited table run again with
 * larger tables (up to 2^20 slots); PK_ERR_DEVICE only if that still fails,
 * with count -1 for those queries. */
int pk_search_rerank(pk_index *h, const float *queries, int64_t nq, int32_t k, int32_t ef, int32_t k_stride,
                     int32_t *out_ids, float *out_dists, int32_t *out_count, pk_query_counters *counters);

/* Waits for all work enqueued on the handle's stream. */
int pk_sync(pk_index *h);

/* Device memory helpers for callers without their own allocator. */
int pk_device_alloc(void **d_ptr, size_t bytes);
int pk_device_free(void *d_ptr);
int pk_copy_to_device(void *d_dst, const void *h_src, size_t bytes);
int pk_copy_to_host(void *h_dst, const void *d_src, size_t bytes);

/* Device-to-device copy on the handle's stream by a plain streaming kernel
 * (k_copy16) that reads and writes exactly `bytes`; 16-byte aligned
 * pointers and size.  No reference counterpart: it is the known-size kernel
 * against which hardware byte counters are calibrated. */
int pk_device_copy(pk_index *h, void *d_dst, const void *d_src, size_t bytes);

/* HIP-event timing of the work enqueued between begin and end on the
 * handle's stream; pk_timer_end waits and returns milliseconds. */
int pk_timer_begin(pk_index *h);
int pk_timer_end(pk_index *h, float *ms);

#ifdef __cplusplus
}
#endif
#endif
