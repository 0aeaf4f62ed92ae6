// server.hip -- continuous-batching request loop over one resident index.
//
// Callers submit query batches of any size (up to a slot's capacity) with
// their own k and ef from any host thread; the loop keeps the index in HBM
// and runs several launch slots at once.  Each slot owns a HIP stream, the
// search scratch for its resident waves (visited tables, candidate-heap
// spill) and pinned staging for up to `slot_queries` queries: admission is
// by that per-query scratch.  Whenever a slot is free, the dispatcher packs
// waiting requests into it in arrival order (a request is never split), so
// requests that arrive while others are on the GPU start as soon as a slot
// frees, and slots overlap on the device.  One launch carries per-query k
// and ef arrays, so requests with different k / ef share it.
//
// Inserts queue in their own arrival order and run on a second thread, on
// the index's own stream, beside the search launches (see insert_worker for
// why that is memory-safe).  As with the reference's concurrent Insert and
// AnnSearch, a search running while an insert is applied may see some of its
// edges; every search submitted after the insert's wait returns sees all of
// it.  A query's result depends only on the graph it ran on, so with no
// insert in progress it is the same alone and in any mix.
//
// Search launches are issued by the dispatcher thread, inserts by the insert
// thread.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "index_impl.hpp"

using namespace pk;

namespace {

#define SCHK(expr)                                                                                        \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return set_error(PK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t ticket = 0;
    std::vector<uint8_t> queries;  // nq rows (copied at submit)
    int64_t nq = 0;
    int k = 0, ef = 0;
    int32_t* out_ids = nullptr;  // caller-owned, nq x k
    float* out_dists = nullptr;
    int32_t* out_count = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    bool done = false;
    bool waited = false;  // a caller is in pk_server_wait on it
    int vbits = 0;        // visited-table size of its next run (0 = the handle's tuning)
    // an internal rerun of some rows of `parent`: their indices there, and
    // the results until they are copied into the parent's buffers
    Request* parent = nullptr;
    std::vector<int64_t> rows;
    std::vector<int32_t> res_ids, res_count;
    std::vector<float> res_dists;
    int status = PK_OK;
    std::string err;  // the failure's message (set on the thread that saw it)
    // insert request
    bool is_insert = false;
    std::vector<int32_t> ext_ids;
    int flags = 0;
};

struct Slot {
    SearchCtx ctx;
    uint8_t* d_q = nullptr;
    int32_t *d_k = nullptr, *d_ef = nullptr, *d_ids = nullptr, *d_cnt = nullptr;
    float* d_d = nullptr;
    uint8_t* h_q = nullptr;  // pinned
    int32_t *h_k = nullptr, *h_ef = nullptr, *h_ids = nullptr, *h_cnt = nullptr;
    float* h_d = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Request*> reqs;
    std::vector<int64_t> first;  // first query row of each request in the launch
    bool busy = false;
    int64_t launched = 0;
    int vbits = 0;  // visited-table size of the launch (0 = the handle's tuning)
};

}  // namespace

struct pk_server {
    pk_index* h = nullptr;
    int slot_queries = 0, max_k = 0, max_ef = 0;
    bool fast = false;  // launches run the fast mode (k_search_fast)
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;         // searches, in arrival order
    std::deque<Request*> pending_insert;  // inserts, in arrival order
    std::unordered_map<int64_t, Request*> by_ticket;
    int64_t next_ticket = 1;
    bool stop = false;
    int device = 0;
    std::thread th, th_insert;
    // statistics
    int64_t launches = 0, queries_done = 0, inserts_done = 0, max_concurrent = 0;
};

namespace {

void free_slot(Slot& s) {
    free_search_ctx(s.ctx);
    for (void* p : {(void*)s.d_q, (void*)s.d_k, (void*)s.d_ef, (void*)s.d_ids, (void*)s.d_cnt, (void*)s.d_d})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)s.h_q, (void*)s.h_k, (void*)s.h_ef, (void*)s.h_ids, (void*)s.h_cnt, (void*)s.h_d})
        if (p) (void)hipHostFree(p);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.ctx.stream) (void)hipStreamDestroy(s.ctx.stream);
}

int alloc_slot(pk_server* S, Slot& s) {
    const int64_t q = S->slot_queries;
    const size_t rb = (size_t)S->h->qbytes;
    SCHK(hipStreamCreateWithFlags(&s.ctx.stream, hipStreamNonBlocking));
    SCHK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_q), q * rb));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_k), q * 4));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_ef), q * 4));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_ids), q * S->max_k * 4));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_d), q * S->max_k * 4));
    SCHK(hipMalloc(reinterpret_cast<void**>(&s.d_cnt), q * 4));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_q), q * rb, hipHostMallocDefault));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_k), q * 4, hipHostMallocDefault));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_ef), q * 4, hipHostMallocDefault));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_ids), q * S->max_k * 4, hipHostMallocDefault));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_d), q * S->max_k * 4, hipHostMallocDefault));
    SCHK(hipHostMalloc(reinterpret_cast<void**>(&s.h_cnt), q * 4, hipHostMallocDefault));
    return PK_OK;
}

// launch: copy in, search with per-query k / ef, copy out, record the event
int launch_slot(pk_server* S, Slot& s) {
    const size_t rb = (size_t)S->h->qbytes;
    int64_t nq = 0;
    int ef_launch = 1;  // the launch's LDS and visited tables follow its largest ef, not the server's limit
    s.first.clear();
    for (Request* r : s.reqs) {
        s.first.push_back(nq);
        ef_launch = std::max(ef_launch, (int)r->ef);
        std::memcpy(s.h_q + nq * rb, r->queries.data(), r->nq * rb);
        for (int64_t i = 0; i < r->nq; i++) {
            s.h_k[nq + i] = r->k;
            s.h_ef[nq + i] = r->ef;
        }
        nq += r->nq;
    }
    hipStream_t st = s.ctx.stream;
    SCHK(hipMemcpyAsync(s.d_q, s.h_q, nq * rb, hipMemcpyHostToDevice, st));
    SCHK(hipMemcpyAsync(s.d_k, s.h_k, nq * 4, hipMemcpyHostToDevice, st));
    SCHK(hipMemcpyAsync(s.d_ef, s.h_ef, nq * 4, hipMemcpyHostToDevice, st));
    int rc = S->fast ? fast_search_on(S->h, s.ctx, s.d_q, nq, s.d_k, s.d_ef, 0, 0, ef_launch, S->max_k, s.d_ids,
                                      s.d_d, s.d_cnt, nullptr)
                     : search_on(S->h, s.ctx, s.d_q, nq, s.d_k, s.d_ef, 0, 0, ef_launch, S->max_k, s.d_ids, s.d_d,
                                 s.d_cnt, nullptr, s.vbits);
    if (rc) return rc;
    SCHK(hipMemcpyAsync(s.h_ids, s.d_ids, nq * S->max_k * 4, hipMemcpyDeviceToHost, st));
    SCHK(hipMemcpyAsync(s.h_d, s.d_d, nq * S->max_k * 4, hipMemcpyDeviceToHost, st));
    SCHK(hipMemcpyAsync(s.h_cnt, s.d_cnt, nq * 4, hipMemcpyDeviceToHost, st));
    SCHK(hipEventRecord(s.done, st));
    s.launched = nq;
    return PK_OK;
}

// a launch that failed part-way may still have copies queued that read or
// write the slot's pinned staging: drain its stream before the slot is reused
int fail_launch(Slot& s, int rc) {
    const std::string msg = pk_last_error();
    (void)hipStreamSynchronize(s.ctx.stream);
    return set_error(rc, msg);
}

// A finished top-level request, or one whose rerun of its failed rows ended.
void finish_request(pk_server* S, Request* r, int status, const std::string& err, double t) {
    r->status = status;
    r->err = err;
    r->t_done = t;
    r->done = true;
    S->queries_done += r->nq;
}

// The end of a rerun (child request) of some rows of its parent: its
// results go into the parent's rows, and the parent completes.
void finish_child(pk_server* S, Request* c, int status, const std::string& err, double t) {
    Request* p = c->parent;
    for (size_t j = 0; j < c->rows.size(); j++) {
        const int64_t i = c->rows[j];
        std::memcpy(p->out_ids + i * p->k, c->out_ids + (int64_t)j * c->k, c->k * 4);
        std::memcpy(p->out_dists + i * p->k, c->out_dists + (int64_t)j * c->k, c->k * 4);
        if (p->out_count) p->out_count[i] = c->out_count[j];
    }
    finish_request(S, p, status, err, t);
    delete c;
}

// Results of a finished slot into the requests' own buffers.  A query that
// filled its visited table ends early with count -1 (k and ef were checked
// at submit, and the candidate heap cannot fill first).  Its rows run again
// as one internal request, ahead of the waiting ones, with a table twice the
// size, up to 2^20 slots (the size does not change results); the request
// completes when they do, and fails with PK_ERR_DEVICE past that size.  The
// other requests of the launch are unaffected.
void complete_slot(pk_server* S, Slot& s, int status) {
    const double t = now_us();
    const std::string launch_err = status == PK_OK ? std::string() : std::string(pk_last_error());
    const size_t qb = (size_t)S->h->qbytes;
    std::vector<Request*> rerun;
    for (size_t j = 0; j < s.reqs.size(); j++) {
        Request* r = s.reqs[j];
        const int64_t f = s.first[j];
        int rs = status;
        std::string err = launch_err;
        if (status == PK_OK) {
            for (int64_t i = 0; i < r->nq; i++) {
                std::memcpy(r->out_ids + i * r->k, s.h_ids + (f + i) * S->max_k, r->k * 4);
                std::memcpy(r->out_dists + i * r->k, s.h_d + (f + i) * S->max_k, r->k * 4);
            }
            if (r->out_count) std::memcpy(r->out_count, s.h_cnt + f, r->nq * 4);
            std::vector<int64_t> bad;
            for (int64_t i = 0; i < r->nq; i++)
                if (s.h_cnt[f + i] < 0) bad.push_back(i);
            // (the fast mode keeps no visited table: its queries fail only
            // on arguments, which submit checked, so they are not rerun)
            if (!bad.empty() && !S->fast && s.ctx.last_vbits < 20) {
                Request* c = r;  // a rerun that failed again runs again, whole
                if (!r->parent) {
                    c = new Request;
                    c->parent = r;
                    c->rows = bad;
                    c->nq = (int64_t)bad.size();
                    c->k = r->k;
                    c->ef = r->ef;
                    c->queries.resize((size_t)c->nq * qb);
                    for (int64_t b = 0; b < c->nq; b++)
                        std::memcpy(c->queries.data() + b * qb, r->queries.data() + bad[b] * qb, qb);
                    c->res_ids.resize((size_t)c->nq * c->k);
                    c->res_dists.resize((size_t)c->nq * c->k);
                    c->res_count.resize((size_t)c->nq);
                    c->out_ids = c->res_ids.data();
                    c->out_dists = c->res_dists.data();
                    c->out_count = c->res_count.data();
                }
                c->vbits = s.ctx.last_vbits + 1;
                rerun.push_back(c);
                continue;
            }
            if (!bad.empty()) {
                rs = PK_ERR_DEVICE;
                err = std::to_string(bad.size()) +
                      (S->fast ? " queries failed on the device (fast mode)"
                               : " queries failed on the device (a visited table of 2^20 slots filled up)");
            }
        }
        if (r->parent)
            finish_child(S, r, rs, err, t);
        else
            finish_request(S, r, rs, err, t);
    }
    for (auto it = rerun.rbegin(); it != rerun.rend(); ++it) S->pending.push_front(*it);
    s.reqs.clear();
    s.busy = false;
}

void dispatcher(pk_server* S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        bool progress = false;
        // finished launches
        for (Slot& s : S->slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            int status = PK_OK;
            if (q != hipSuccess) status = set_error(PK_ERR_HIP, std::string("search launch: ") + hipGetErrorString(q));
            complete_slot(S, s, status);
            progress = true;
        }
        int busy = 0;
        for (Slot& s : S->slots) busy += s.busy;
        if (progress) S->cv_done.notify_all();
        if (S->stop && S->pending.empty() && busy == 0) break;
        {
            // admission: free slots take waiting searches in arrival order
            // (reruns first), one visited-table size per launch; searches are
            // not held behind inserts
            for (Slot& s : S->slots) {
                if (s.busy || S->pending.empty()) continue;
                int64_t nq = 0;
                const int vb = S->pending.front()->vbits;
                while (!S->pending.empty() && S->pending.front()->vbits == vb &&
                       nq + S->pending.front()->nq <= S->slot_queries) {
                    Request* r = S->pending.front();
                    S->pending.pop_front();
                    if (r->t_start == 0) r->t_start = now_us();
                    nq += r->nq;
                    s.reqs.push_back(r);
                }
                if (s.reqs.empty()) continue;
                s.vbits = vb;
                s.busy = true;
                S->launches++;
                int rc = launch_slot(S, s);
                if (rc) complete_slot(S, s, fail_launch(s, rc));
                progress = true;
            }
            busy = 0;
            for (Slot& s : S->slots) busy += s.busy;
            S->max_concurrent = std::max<int64_t>(S->max_concurrent, busy);
            if (progress) S->cv_done.notify_all();
        }
        if (progress) continue;
        if (busy > 0) {
            // launches in flight: poll their events without holding the lock
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        } else {
            S->cv_work.wait_for(lk, std::chrono::milliseconds(50));
        }
    }
}

// Inserts run on their own thread and the index's own stream, beside the
// search launches: memory-safe because the layers never move (capacity is
// fixed by init_build), searches take a layer's node count from the
// published table -- new ids stay out of reach until their batch is
// published, after its kernels finish -- and AddEdge only writes valid ids
// (or -1) into rows, whose degrees searches clamp to the allocation.
void insert_worker(pk_server* S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        S->cv_work.wait_for(lk, std::chrono::milliseconds(50),
                            [&] { return S->stop || !S->pending_insert.empty(); });
        if (S->pending_insert.empty()) {
            if (S->stop) break;
            continue;
        }
        Request* r = S->pending_insert.front();
        S->pending_insert.pop_front();
        r->t_start = now_us();
        lk.unlock();
        // rows by external id (contiguous runs in one copy each), then Insert
        int rc = PK_OK;
        const int64_t n = (int64_t)r->ext_ids.size();
        for (int64_t i = 0; i < n && rc == PK_OK;) {
            int64_t j = i + 1;
            while (j < n && r->ext_ids[j] == r->ext_ids[j - 1] + 1) j++;
            rc = pk_index_set_base(S->h, r->queries.data() + i * S->h->qbytes, r->ext_ids[i], j - i);
            i = j;
        }
        if (rc == PK_OK) rc = pk_index_insert(S->h, r->ext_ids.data(), n, 4096, r->flags);
        const std::string err = rc == PK_OK ? std::string() : std::string(pk_last_error());
        lk.lock();
        r->status = rc;
        r->err = err;
        r->t_done = now_us();
        r->done = true;
        S->inserts_done++;
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int pk_server_create(pk_index* h, int slots, int slot_queries, int max_k, int max_ef, pk_server** out) {
    return pk_server_create_ex(h, slots, slot_queries, max_k, max_ef, 0, out);
}

int pk_server_create_ex(pk_index* h, int slots, int slot_queries, int max_k, int max_ef, int flags, pk_server** out) {
    if (!h || !out) return set_error(PK_ERR_ARG, "null argument");
    *out = nullptr;
    if (!h->ready) return set_error(PK_ERR_STATE, "index not loaded");
    if (slots < 1 || slots > 8 || slot_queries < 1 || max_k < 1 || max_ef < max_k || max_ef > 630)
        return set_error(PK_ERR_ARG, "need 1..8 slots, slot_queries >= 1, 1 <= max_k <= max_ef <= 630");
    if (flags & ~PK_SERVE_FAST) return set_error(PK_ERR_ARG, "unknown serving flags");
    if ((flags & PK_SERVE_FAST) && max_ef > 512) return set_error(PK_ERR_ARG, "the fast mode takes max_ef <= 512");
    auto* S = new pk_server;
    S->h = h;
    S->fast = (flags & PK_SERVE_FAST) != 0;
    S->slot_queries = slot_queries;
    S->max_k = max_k;
    S->max_ef = max_ef;
    S->slots.resize(slots);
    for (Slot& s : S->slots) {
        int rc = alloc_slot(S, s);
        if (rc) {
            for (Slot& t : S->slots) free_slot(t);
            delete S;
            return rc;
        }
    }
    (void)hipGetDevice(&S->device);
    h->servers_alive++;  // the slots' launches read the layers: no reallocation meanwhile
    S->th = std::thread(dispatcher, S);
    S->th_insert = std::thread(insert_worker, S);
    *out = S;
    return PK_OK;
}

int pk_server_submit(pk_server* S, const void* queries, int64_t nq, int32_t k, int32_t ef, int32_t* out_ids,
                     float* out_dists, int32_t* out_count, int64_t* ticket) {
    if (!S || !queries || !out_ids || !out_dists || !ticket) return set_error(PK_ERR_ARG, "null argument");
    if (nq < 1 || nq > S->slot_queries) return set_error(PK_ERR_CAPACITY, "request larger than a slot");
    if (k < 1 || k > S->max_k || ef < k || ef > S->max_ef) return set_error(PK_ERR_ARG, "k / ef outside the server's limits");
    auto* r = new Request;
    r->queries.assign(static_cast<const uint8_t*>(queries), static_cast<const uint8_t*>(queries) + nq * S->h->qbytes);
    r->nq = nq;
    r->k = k;
    r->ef = ef;
    r->out_ids = out_ids;
    r->out_dists = out_dists;
    r->out_count = out_count;
    r->t_submit = now_us();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) {
            delete r;
            return set_error(PK_ERR_STATE, "server is stopping");
        }
        r->ticket = S->next_ticket++;
        S->by_ticket[r->ticket] = r;
        S->pending.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_one();
    return PK_OK;
}

int pk_server_insert(pk_server* S, const int32_t* ext_ids, const void* rows, int64_t n, int flags, int64_t* ticket) {
    if (!S || !ext_ids || !rows || !ticket || n < 1) return set_error(PK_ERR_ARG, "bad arguments");
    auto* r = new Request;
    r->is_insert = true;
    r->ext_ids.assign(ext_ids, ext_ids + n);
    r->queries.assign(static_cast<const uint8_t*>(rows), static_cast<const uint8_t*>(rows) + n * S->h->qbytes);
    r->flags = flags;
    r->nq = n;
    r->t_submit = now_us();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) {
            delete r;
            return set_error(PK_ERR_STATE, "server is stopping");
        }
        r->ticket = S->next_ticket++;
        S->by_ticket[r->ticket] = r;
        S->pending_insert.push_back(r);
        *ticket = r->ticket;
    }
    S->cv_work.notify_all();
    return PK_OK;
}

int pk_server_wait(pk_server* S, int64_t ticket, int32_t timeout_ms, double* latency_us, double* queue_us) {
    if (!S) return set_error(PK_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->by_ticket.find(ticket);
    if (it == S->by_ticket.end()) return set_error(PK_ERR_ARG, "unknown ticket");
    Request* r = it->second;
    if (r->waited) return set_error(PK_ERR_STATE, "another caller is waiting on this ticket");
    r->waited = true;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0) {
        S->cv_done.wait(lk, pred);
    } else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
        r->waited = false;  // the ticket stays valid: wait again later
        return set_error(PK_ERR_STATE, "request not finished within the timeout");
    }
    if (latency_us) *latency_us = r->t_done - r->t_submit;
    if (queue_us) *queue_us = r->t_start - r->t_submit;
    const int status = r->status;
    const std::string err = r->err;
    S->by_ticket.erase(ticket);  // by key: submits during the wait may have rehashed the map
    delete r;
    return status == PK_OK ? PK_OK : set_error(status, "request failed: " + err);
}

int pk_server_stats(pk_server* S, int64_t* out) {
    if (!S || !out) return set_error(PK_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(S->mu);
    out[0] = S->launches;
    out[1] = S->queries_done;
    out[2] = S->inserts_done;
    out[3] = S->max_concurrent;
    out[4] = (int64_t)(S->pending.size() + S->pending_insert.size());
    return PK_OK;
}

int pk_server_destroy(pk_server* S) {
    if (!S) return PK_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->th.joinable()) S->th.join();
    if (S->th_insert.joinable()) S->th_insert.join();
    for (Slot& s : S->slots) {
        if (s.ctx.stream) (void)hipStreamSynchronize(s.ctx.stream);
        free_slot(s);
    }
    for (auto& kv : S->by_ticket) delete kv.second;
    S->h->servers_alive--;
    delete S;
    return PK_OK;
}

}  // extern "C"
