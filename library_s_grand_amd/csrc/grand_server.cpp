// grand_server.cpp -- continuous-batching decode server over device sessions.
//
// Every lane has a worker thread of its own:
//
//   lane = one GrandSession (its own workspace, reserved for a full batch)
//        + one non-blocking stream + device buffers for inputs and results
//        + pinned host staging for the same.
//
// A free lane takes every pending request of its code (whole requests, in
// arrival order, up to max_batch_shots shots), copies their rows into its pinned
// staging, queues H2D copy -> decode -> D2H copy -> event on its stream, waits
// for the event and scatters the results to the requests' buffers.  Lanes of
// different codes, and the lanes of one code, run concurrently (their host
// copies too), so new arrivals never wait for a batch that is already running.
//
// Nothing on the request path allocates device memory: every workspace and
// buffer is sized when the code is admitted, and admission is refused when the
// code's footprint does not fit the server's HBM budget.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/grand_decoder.h"
#include "../../include/grand_server.h"
#include "../../include/grand_session.h"

namespace {

using Clock = std::chrono::steady_clock;

bool ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "grand server: %s failed: %s\n", what, hipGetErrorString(e));
    return false;
}

struct PinnedBuf {
    void *p = nullptr;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    bool alloc(size_t b) { return ok(hipHostMalloc(&p, std::max<size_t>(b, 64), hipHostMallocDefault), "hipHostMalloc"); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t b) { return ok(hipMalloc(&p, std::max<size_t>(b, 256)), "hipMalloc"); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Request {
    long long ticket = 0;
    int code = 0;
    const int *input = nullptr;
    int input_errors = 0;
    const int *logicals = nullptr;
    int shots = 0;
    GrandDecoderResult *results = nullptr;
    int *decoded = nullptr;
    Clock::time_point t_submit;
    double latency_ms = 0.0;
    std::atomic<int> status{GRAND_REQ_PENDING};
};
using ReqPtr = std::shared_ptr<Request>;

struct Lane {
    GrandSession *sess = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf d_in, d_lg, d_res, d_dec;
    PinnedBuf h_in, h_lg, h_res, h_dec, h_err;
    std::vector<ReqPtr> batch;
    int batch_shots = 0;
    bool want_dec = false;
    bool busy = false;
    ~Lane() {
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
        if (sess) grand_session_destroy(sess);
    }
};

struct Code {
    int kind = 0, n = 0, m = 0, NL = 0, M32 = 0;
    bool css = true;
    long long footprint = 0;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::deque<ReqPtr> queue;     // guarded by the server mutex
    std::condition_variable cv;   // submit / stop -> this code's lanes
};

}  // namespace

struct GrandServer {
    GrandServerConfig cfg{};
    int device = 0;
    std::mutex mu;
    std::condition_variable done_cv;   // lanes -> waiters
    std::vector<std::unique_ptr<Code>> codes;
    std::unordered_map<long long, ReqPtr> reqs;
    long long next_ticket = 0;
    long long queued_shots = 0;
    long long used_bytes = 0;
    bool stop = false;
    std::vector<std::thread> workers;  // one per lane
    std::atomic<long long> n_requests{0}, n_shots{0}, n_batches{0};

    void lane_loop(Code *c, Lane *L);
    bool launch(Code &c, Lane &L);
    void finish(Code &c, Lane &L, bool failed);
};

// Stages the lane's batch and queues its decode; returns false on a HIP error.
bool GrandServer::launch(Code &c, Lane &L) {
    const int S = L.batch_shots;
    const int ie = L.batch.front()->input_errors;
    const int cols = (c.css && ie) ? c.m : c.n;
    size_t off = 0, loff = 0;
    for (const ReqPtr &r : L.batch) {
        std::memcpy(L.h_in.as<int>() + off, r->input, sizeof(int) * (size_t)r->shots * cols);
        off += (size_t)r->shots * cols;
        if (!c.css && c.NL > 0) {
            std::memcpy(L.h_lg.as<int>() + loff, r->logicals, sizeof(int) * (size_t)r->shots * c.NL);
            loff += (size_t)r->shots * c.NL;
        }
    }
    if (!ok(hipMemcpyAsync(L.d_in.p, L.h_in.p, sizeof(int) * off, hipMemcpyHostToDevice, L.stream), "H2D")) return false;
    if (loff && !ok(hipMemcpyAsync(L.d_lg.p, L.h_lg.p, sizeof(int) * loff, hipMemcpyHostToDevice, L.stream), "H2D"))
        return false;
    if (grand_session_decode_device(L.sess, L.d_in.as<int>(), ie, L.d_lg.as<int>(), S,
                                    L.d_res.as<GrandDecoderResult>(), L.want_dec ? L.d_dec.as<unsigned>() : nullptr,
                                    nullptr, nullptr, L.stream) != 0)
        return false;
    if (!ok(hipMemcpyAsync(L.h_res.p, L.d_res.p, sizeof(GrandDecoderResult) * (size_t)S, hipMemcpyDeviceToHost,
                           L.stream), "D2H"))
        return false;
    if (L.want_dec && !ok(hipMemcpyAsync(L.h_dec.p, L.d_dec.p, sizeof(unsigned) * (size_t)S * c.M32,
                                         hipMemcpyDeviceToHost, L.stream), "D2H"))
        return false;
    if (grand_session_error_bits_async(L.sess, L.h_err.as<unsigned>(), L.stream) != 0) return false;
    return ok(hipEventRecord(L.done, L.stream), "hipEventRecord");
}

// Scatters a finished batch to its requests and wakes their waiters.
void GrandServer::finish(Code &c, Lane &L, bool failed) {
    if (!failed && *L.h_err.as<unsigned>() != 0u) {
        const unsigned bits = *L.h_err.as<unsigned>();
        if (bits & 1u) fprintf(stderr, "grand server: input rows must hold only 0 and 1\n");
        if (bits & 2u) fprintf(stderr, "grand server: an active set exceeds the supported size\n");
        failed = true;
    }
    const auto now = Clock::now();
    size_t off = 0;
    for (const ReqPtr &r : L.batch) {
        if (!failed) {
            std::memcpy(r->results, L.h_res.as<GrandDecoderResult>() + off, sizeof(GrandDecoderResult) * r->shots);
            if (r->decoded) {
                const unsigned *w = L.h_dec.as<unsigned>() + off * c.M32;
                for (int i = 0; i < r->shots; ++i)
                    for (int j = 0; j < c.m; ++j)
                        r->decoded[(size_t)i * c.m + j] = (int)((w[(size_t)i * c.M32 + (j >> 5)] >> (j & 31)) & 1u);
            }
        }
        off += r->shots;
        r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
    }
    {
        std::lock_guard<std::mutex> g(mu);
        for (const ReqPtr &r : L.batch) r->status.store(failed ? GRAND_REQ_FAILED : GRAND_REQ_DONE);
        if (!failed) {
            n_requests += (long long)L.batch.size();
            n_shots += L.batch_shots;
        }
    }
    done_cv.notify_all();
    L.batch.clear();
    L.batch_shots = 0;
    L.busy = false;
}

// A lane's worker: takes every pending request of its code, decodes them as
// one batch and scatters the results; sleeps while its code has nothing queued.
void GrandServer::lane_loop(Code *c, Lane *L) {
    (void)hipSetDevice(device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            c->cv.wait(lk, [&] { return stop || !c->queue.empty(); });
            if (c->queue.empty()) break;                 // stopping, nothing left for this code
            const int ie = c->queue.front()->input_errors;
            while (!c->queue.empty() && c->queue.front()->input_errors == ie &&
                   L->batch_shots + c->queue.front()->shots <= cfg.max_batch_shots) {
                ReqPtr r = c->queue.front();
                c->queue.pop_front();
                L->batch_shots += r->shots;
                L->batch.push_back(std::move(r));
            }
            queued_shots -= L->batch_shots;
        }
        L->want_dec = false;
        for (const ReqPtr &r : L->batch) L->want_dec = L->want_dec || r->decoded;
        L->busy = true;
        ++n_batches;
        bool failed = !launch(*c, *L);
        if (failed) (void)hipStreamSynchronize(L->stream);
        else failed = !ok(hipEventSynchronize(L->done), "hipEventSynchronize");
        finish(*c, *L, failed);
    }
}

extern "C" {

GrandServer *grand_server_create(const GrandServerConfig *cfg) {
    auto *s = new GrandServer();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_batch_shots <= 0) s->cfg.max_batch_shots = 16384;
    if (s->cfg.lanes_per_code <= 0) s->cfg.lanes_per_code = 2;
    if (s->cfg.max_queued_shots <= 0) s->cfg.max_queued_shots = 1 << 20;
    if (!ok(hipGetDevice(&s->device), "hipGetDevice")) { delete s; return nullptr; }
    if (s->cfg.hbm_budget_bytes <= 0) {
        size_t free_b = 0, total_b = 0;
        if (!ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) { delete s; return nullptr; }
        s->cfg.hbm_budget_bytes = (long long)(free_b / 2);
    }
    return s;
}

int grand_server_add_code(GrandServer *srv, int kind, const int *H, const int *H2, int n, int m, int n_logicos,
                          char type_of_error, const GrandDecoderConfig *config, const char *file) {
    if (!srv) return -1;
    const int B = srv->cfg.max_batch_shots;
    const int lanes = srv->cfg.lanes_per_code;
    auto c = std::make_unique<Code>();
    c->kind = kind; c->n = n; c->m = m; c->NL = (kind >= GRAND_KIND_CLN) ? n_logicos : 0;
    c->css = kind <= GRAND_KIND_PROB_CC;
    for (int i = 0; i < lanes; ++i) {
        auto L = std::make_unique<Lane>();
        L->sess = grand_session_create(kind, H, H2, n, m, n_logicos, type_of_error, config, file, nullptr);
        if (!L->sess) return -1;
        c->lanes.push_back(std::move(L));
    }
    GrandSession *s0 = c->lanes[0]->sess;
    c->M32 = grand_session_pattern_words(s0);
    // code once, each lane's batch scratch, each lane's input / result buffers
    const long long code_b = (long long)grand_session_footprint(s0, 1);
    const long long scratch_b = (long long)grand_session_footprint(s0, B) - code_b;
    const int cols = std::max(m, n);
    const long long lane_b = (long long)B * (4LL * cols + 4LL * c->NL + (long long)sizeof(GrandDecoderResult) +
                                             4LL * c->M32);
    c->footprint = code_b + lanes * (scratch_b + lane_b);
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (srv->used_bytes + c->footprint > srv->cfg.hbm_budget_bytes) {
            fprintf(stderr, "grand server: code refused: footprint %lld bytes, %lld of %lld in use\n", c->footprint,
                    srv->used_bytes, srv->cfg.hbm_budget_bytes);
            return -2;
        }
        srv->used_bytes += c->footprint;   // reserved before allocating; released on failure below
    }
    bool good = true;
    for (auto &L : c->lanes) {
        good = good && grand_session_reserve(L->sess, B) == 0;
        good = good && L->d_in.alloc(sizeof(int) * (size_t)B * cols) && L->d_res.alloc(sizeof(GrandDecoderResult) * B) &&
               L->d_dec.alloc(sizeof(unsigned) * (size_t)B * c->M32) &&
               L->d_lg.alloc(sizeof(int) * (size_t)B * std::max(c->NL, 1));
        good = good && L->h_in.alloc(sizeof(int) * (size_t)B * cols) && L->h_res.alloc(sizeof(GrandDecoderResult) * B) &&
               L->h_dec.alloc(sizeof(unsigned) * (size_t)B * c->M32) &&
               L->h_lg.alloc(sizeof(int) * (size_t)B * std::max(c->NL, 1)) && L->h_err.alloc(sizeof(unsigned));
        good = good && ok(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking), "hipStreamCreate");
        good = good && ok(hipEventCreateWithFlags(&L->done, hipEventDisableTiming), "hipEventCreate");
        if (!good) break;
    }
    std::lock_guard<std::mutex> g(srv->mu);
    if (!good) {
        srv->used_bytes -= c->footprint;
        return -1;
    }
    Code *cp = c.get();
    srv->codes.push_back(std::move(c));
    for (auto &L : cp->lanes) {
        Lane *lp = L.get();
        srv->workers.emplace_back([srv, cp, lp] { srv->lane_loop(cp, lp); });
    }
    return (int)srv->codes.size() - 1;
}

long long grand_server_footprint(GrandServer *srv, int code_id) {
    if (!srv) return 0;
    std::lock_guard<std::mutex> g(srv->mu);
    if (code_id < 0) return srv->used_bytes;
    if (code_id >= (int)srv->codes.size()) return 0;
    return srv->codes[code_id]->footprint;
}

long long grand_server_submit(GrandServer *srv, int code_id, const int *input, int input_is_errors,
                              const int *logicals, int num_shots, GrandDecoderResult *results, int *decoded) {
    if (!srv || !input || !results || num_shots <= 0) return -1;
    auto r = std::make_shared<Request>();
    Code *cp = nullptr;
    r->code = code_id; r->input = input; r->input_errors = input_is_errors ? 1 : 0; r->logicals = logicals;
    r->shots = num_shots; r->results = results; r->decoded = decoded;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(srv->mu);
        if (code_id < 0 || code_id >= (int)srv->codes.size()) return -1;
        Code &c = *srv->codes[code_id];
        if (!c.css) r->input_errors = 0;
        if (!c.css && c.NL > 0 && !logicals) return -1;
        if (num_shots > srv->cfg.max_batch_shots) {
            fprintf(stderr, "grand server: request of %d shots exceeds max_batch_shots = %d\n", num_shots,
                    srv->cfg.max_batch_shots);
            return -1;
        }
        if (srv->queued_shots + num_shots > srv->cfg.max_queued_shots) return -2;
        r->ticket = srv->next_ticket++;
        srv->queued_shots += num_shots;
        srv->reqs[r->ticket] = r;
        c.queue.push_back(r);
        cp = &c;
    }
    cp->cv.notify_one();
    return r->ticket;
}

int grand_server_poll(GrandServer *srv, long long ticket, double *latency_ms) {
    if (!srv) return GRAND_REQ_FAILED;
    std::lock_guard<std::mutex> g(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return GRAND_REQ_FAILED;
    const int st = it->second->status.load();
    if (st != GRAND_REQ_PENDING && latency_ms) *latency_ms = it->second->latency_ms;
    return st;
}

int grand_server_wait(GrandServer *srv, long long ticket, double *latency_ms) {
    if (!srv) return GRAND_REQ_FAILED;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return GRAND_REQ_FAILED;
    ReqPtr r = it->second;
    srv->done_cv.wait(lk, [&] { return r->status.load() != GRAND_REQ_PENDING; });
    if (latency_ms) *latency_ms = r->latency_ms;
    srv->reqs.erase(ticket);
    return r->status.load();
}

void grand_server_counters(GrandServer *srv, long long *requests, long long *shots, long long *batches) {
    if (!srv) return;
    if (requests) *requests = srv->n_requests.load();
    if (shots) *shots = srv->n_shots.load();
    if (batches) *batches = srv->n_batches.load();
}

void grand_server_destroy(GrandServer *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> g(srv->mu);
        srv->stop = true;
        for (auto &c : srv->codes) c->cv.notify_all();
    }
    for (auto &t : srv->workers)
        if (t.joinable()) t.join();
    delete srv;
}

}  // extern "C"
