// cgm_engine.hip -- continuous-batching serving engine for SW + WFA pairs.
//
// Requests (a few pairs each, with their packed sequences) are copied at
// submit time into host memory and queued.  One dispatcher thread owns the
// GPU side: it keeps `n_slots` slots, each with a stream, device buffers,
// pinned staging and a mixed-kernel plan (cgm_mixed.hip), all allocated at
// creation and charged against a fixed HBM budget.  Whenever a slot is free
// and requests are waiting, every waiting request that fits the slot forms
// the next batch (first come, first served), so requests that arrive while
// other batches run on the GPU start as soon as a slot frees and their
// kernels overlap with the running ones.  Admission is bounded: a request
// is refused with CGM_ERR_BUSY while the queued pairs would exceed the
// engine's pending limit (back-pressure instead of unbounded queueing).
// Every pair's result depends only on the pair, never on the batch it ran
// in, so a request gets the same result alone and in any mix.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cgm_internal.h"

namespace cgm {

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    int32_t n = 0;
    std::vector<uint64_t> qp, rp;
    std::vector<int32_t> ql, rl, qo, ro, kind, band, me;
    std::vector<int32_t> score, qend, rend;
    double t_arrival = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 on the GPU, 2 done
};

struct Slot {
    hipStream_t stream = nullptr;
    MixPlan *plan = nullptr;
    uint64_t *d_qp = nullptr, *d_rp = nullptr, *h_qp = nullptr, *h_rp = nullptr;
    int32_t *d_arr = nullptr, *h_arr = nullptr;  // 7 arrays of cap_pairs
    int32_t *d_out = nullptr, *h_out = nullptr;  // 3 arrays of cap_pairs
    hipEvent_t done = nullptr;
    std::vector<Request *> reqs;
    bool busy = false;
};

}  // namespace

struct Engine {
    int device = 0;
    int n_slots = 0;
    int64_t cap_pairs = 0, cap_words = 0, max_pending_pairs = 0, device_bytes = 0;
    SwScoring sc{};
    WfaPen pen{};
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    int64_t queued_pairs = 0;
    std::unordered_map<int64_t, Request *> live;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::atomic<int64_t> batches{0}, batch_pairs{0}, done_requests{0}, done_pairs{0}, refused{0};
    int worker_error = 0;
    std::string worker_msg;

    ~Engine() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        (void)hipSetDevice(device);
        for (Slot &s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.plan) mix_plan_destroy(s.plan);
            void *dev_ptrs[] = {s.d_qp, s.d_rp, s.d_arr, s.d_out};
            for (void *p : dev_ptrs)
                if (p) (void)hipFree(p);
            void *host_ptrs[] = {s.h_qp, s.h_rp, s.h_arr, s.h_out};
            for (void *p : host_ptrs)
                if (p) (void)hipHostFree(p);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        for (auto &kv : live) delete kv.second;
    }

    // Copies finished slots' results into their requests.
    bool retire() {
        bool any = false;
        for (Slot &s : slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            any = true;
            const double t = now_s();
            int64_t off = 0;
            const int64_t n_batch = [&] {
                int64_t k = 0;
                for (Request *r : s.reqs) k += r->n;
                return k;
            }();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : s.reqs) {
                    r->score.assign(s.h_out + off, s.h_out + off + r->n);
                    r->qend.assign(s.h_out + cap_pairs + off, s.h_out + cap_pairs + off + r->n);
                    r->rend.assign(s.h_out + 2 * cap_pairs + off,
                                   s.h_out + 2 * cap_pairs + off + r->n);
                    if (q != hipSuccess) std::fill(r->score.begin(), r->score.end(), CGM_BAD_PAIR);
                    r->t_done = t;
                    r->state = 2;
                    off += r->n;
                }
                done_requests += (int64_t)s.reqs.size();
                done_pairs += n_batch;
            }
            s.reqs.clear();
            s.busy = false;
            cv_done.notify_all();
        }
        return any;
    }

    // Forms and launches one batch on a free slot; false when nothing waits.
    bool launch(Slot &s) {
        std::vector<Request *> take;
        int64_t np = 0, qw = 0, rw = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!queue.empty()) {
                Request *r = queue.front();
                if (np + r->n > cap_pairs || qw + (int64_t)r->qp.size() > cap_words ||
                    rw + (int64_t)r->rp.size() > cap_words)
                    break;
                queue.pop_front();
                queued_pairs -= r->n;
                np += r->n;
                qw += (int64_t)r->qp.size();
                rw += (int64_t)r->rp.size();
                r->state = 1;
                take.push_back(r);
            }
        }
        if (take.empty()) return false;
        // pack into pinned staging: words, then 7 arrays of cap_pairs
        int64_t p = 0, wq = 0, wr = 0;
        int32_t *A = s.h_arr;
        for (Request *r : take) {
            memcpy(s.h_qp + wq, r->qp.data(), r->qp.size() * 8);
            memcpy(s.h_rp + wr, r->rp.data(), r->rp.size() * 8);
            for (int32_t i = 0; i < r->n; ++i) {
                A[0 * cap_pairs + p + i] = r->ql[i];
                A[1 * cap_pairs + p + i] = r->rl[i];
                A[2 * cap_pairs + p + i] = r->qo[i] + (int32_t)wq;
                A[3 * cap_pairs + p + i] = r->ro[i] + (int32_t)wr;
                A[4 * cap_pairs + p + i] = r->kind[i];
                A[5 * cap_pairs + p + i] = r->band[i];
                A[6 * cap_pairs + p + i] = r->me[i];
            }
            p += r->n;
            wq += (int64_t)r->qp.size();
            wr += (int64_t)r->rp.size();
        }
        hipStream_t st = s.stream;
        int rc = CGM_OK;
        auto h2d = [&](void *d, const void *h, int64_t bytes) {
            if (rc == CGM_OK && bytes > 0 &&
                hipMemcpyAsync(d, h, (size_t)bytes, hipMemcpyHostToDevice, st) != hipSuccess)
                rc = CGM_ERR_HIP;
        };
        h2d(s.d_qp, s.h_qp, wq * 8);
        h2d(s.d_rp, s.h_rp, wr * 8);
        for (int k = 0; k < 7; ++k) h2d(s.d_arr + k * cap_pairs, s.h_arr + k * cap_pairs, p * 4);
        int32_t *D = s.d_arr;
        Batch b = {s.d_qp, s.d_rp, D + 0 * cap_pairs, D + 1 * cap_pairs, D + 2 * cap_pairs,
                   D + 3 * cap_pairs, p, wq, wr};
        SwOut out = {s.d_out, s.d_out + cap_pairs, s.d_out + 2 * cap_pairs};
        if (rc == CGM_OK)
            rc = mix_enqueue(s.plan, b, D + 4 * cap_pairs, 0, D + 5 * cap_pairs, 0,
                             D + 6 * cap_pairs, 0, sc, pen, out, st);
        for (int k = 0; k < 3 && rc == CGM_OK; ++k)
            if (hipMemcpyAsync(s.h_out + k * cap_pairs, s.d_out + k * cap_pairs, (size_t)p * 4,
                               hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = CGM_ERR_HIP;
        if (rc == CGM_OK && hipEventRecord(s.done, st) != hipSuccess) rc = CGM_ERR_HIP;
        s.reqs = take;
        s.busy = true;
        batches += 1;
        batch_pairs += p;
        if (rc != CGM_OK) {
            // results of this batch are reported as bad pairs
            (void)hipStreamSynchronize(st);
            std::lock_guard<std::mutex> lk(mu);
            worker_error = rc;
            for (int64_t i = 0; i < p; ++i) s.h_out[i] = CGM_BAD_PAIR;
            (void)hipEventRecord(s.done, st);
        }
        return true;
    }

    void run() {
        (void)hipSetDevice(device);
        for (;;) {
            bool did = retire();
            for (Slot &s : slots)
                if (!s.busy && launch(s)) did = true;
            std::unique_lock<std::mutex> lk(mu);
            if (stop) break;
            bool busy_any = false;
            for (Slot &s : slots) busy_any |= s.busy;
            if (!did) {
                if (busy_any || !queue.empty())
                    cv_work.wait_for(lk, std::chrono::microseconds(20));
                else
                    cv_work.wait_for(lk, std::chrono::milliseconds(2));
            }
        }
    }
};

}  // namespace cgm

using namespace cgm;

extern "C" {

int cgm_engine_create(int64_t hbm_budget_bytes, int32_t n_slots, int64_t max_batch_pairs,
                      int64_t max_batch_words, int32_t max_error_cap, int32_t match,
                      int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t x, int32_t o,
                      int32_t e, void **engine) {
    if (!engine || n_slots < 1 || n_slots > 16 || max_batch_pairs < 1 || max_batch_words < 1 ||
        max_batch_pairs > (int64_t(1) << 28) || max_batch_words > (int64_t(1) << 30))
        return set_error(CGM_ERR_INVALID, "cgm_engine_create: bad arguments");
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    WfaPen pen = {x, o, e};
    if (int rc = check_sw_scoring(sc)) return rc;
    if (int rc = check_wfa_pen(pen)) return rc;
    Engine *E = new Engine();
    auto fail = [&](int rc) {
        delete E;
        return rc;
    };
    if (hipGetDevice(&E->device) != hipSuccess) return fail(set_error(CGM_ERR_NODEV, "no device"));
    E->n_slots = n_slots;
    E->cap_pairs = max_batch_pairs;
    E->cap_words = max_batch_words;
    E->max_pending_pairs = (int64_t)n_slots * max_batch_pairs;
    E->sc = sc;
    E->pen = pen;
    E->slots.resize((size_t)n_slots);
    for (Slot &s : E->slots) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
            return fail(set_error(CGM_ERR_HIP, "engine: stream/event"));
        int rc = mix_plan_create(max_batch_pairs, max_error_cap, pen, 0, 0, &s.plan);
        if (rc) return fail(rc);
        E->device_bytes += mix_plan_bytes(s.plan);
        const int64_t wb = max_batch_words * 8, ab = 7 * max_batch_pairs * 4,
                      ob = 3 * max_batch_pairs * 4;
        E->device_bytes += 2 * wb + ab + ob;
        if (E->device_bytes > hbm_budget_bytes)
            return fail(set_error(CGM_ERR_NOMEM, "engine: slots exceed the HBM budget"));
        if (hipMalloc((void **)&s.d_qp, wb) != hipSuccess ||
            hipMalloc((void **)&s.d_rp, wb) != hipSuccess ||
            hipMalloc((void **)&s.d_arr, ab) != hipSuccess ||
            hipMalloc((void **)&s.d_out, ob) != hipSuccess ||
            hipHostMalloc((void **)&s.h_qp, wb, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s.h_rp, wb, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s.h_arr, ab, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&s.h_out, ob, hipHostMallocDefault) != hipSuccess)
            return fail(set_error(CGM_ERR_NOMEM, "engine: slot buffers"));
    }
    E->worker = std::thread([E] { E->run(); });
    *engine = E;
    return CGM_OK;
}

int cgm_engine_destroy(void *engine) {
    delete (Engine *)engine;
    return CGM_OK;
}

int cgm_engine_submit(void *engine, const uint64_t *q_packed, const uint64_t *r_packed,
                      const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                      const int32_t *r_off, int32_t n_pairs, const int32_t *kind,
                      const int32_t *band, const int32_t *max_error, double arrival_s,
                      int64_t *request_id) {
    Engine *E = (Engine *)engine;
    if (!E || !request_id || n_pairs < 1 || !q_len || !r_len || !q_off || !r_off || !kind ||
        !band || !max_error)
        return set_error(CGM_ERR_INVALID, "cgm_engine_submit: bad arguments");
    if (n_pairs > E->cap_pairs) return set_error(CGM_ERR_INVALID, "request larger than a batch");
    Request *r = new Request();
    r->n = n_pairs;
    r->ql.assign(q_len, q_len + n_pairs);
    r->rl.assign(r_len, r_len + n_pairs);
    r->kind.assign(kind, kind + n_pairs);
    r->band.assign(band, band + n_pairs);
    r->me.assign(max_error, max_error + n_pairs);
    r->qo.resize((size_t)n_pairs);
    r->ro.resize((size_t)n_pairs);
    // copy each sequence's words so the request owns compact buffers
    for (int32_t i = 0; i < n_pairs; ++i) {
        if (q_len[i] < 0 || r_len[i] < 0 || q_off[i] < 0 || r_off[i] < 0) {
            delete r;
            return set_error(CGM_ERR_INVALID, "negative length or offset");
        }
        const int32_t qw = (q_len[i] + 31) / 32, rw = (r_len[i] + 31) / 32;
        r->qo[(size_t)i] = (int32_t)r->qp.size();
        r->ro[(size_t)i] = (int32_t)r->rp.size();
        r->qp.insert(r->qp.end(), q_packed + q_off[i], q_packed + q_off[i] + qw);
        r->rp.insert(r->rp.end(), r_packed + r_off[i], r_packed + r_off[i] + rw);
    }
    if ((int64_t)r->qp.size() > E->cap_words || (int64_t)r->rp.size() > E->cap_words) {
        delete r;
        return set_error(CGM_ERR_INVALID, "request sequences exceed a batch");
    }
    r->t_arrival = arrival_s > 0 ? arrival_s : now_s();
    {
        std::lock_guard<std::mutex> lk(E->mu);
        if (E->queued_pairs + n_pairs > E->max_pending_pairs) {
            E->refused += 1;
            delete r;
            return set_error(CGM_ERR_BUSY, "engine queue full");
        }
        r->id = E->next_id++;
        E->queue.push_back(r);
        E->queued_pairs += n_pairs;
        E->live[r->id] = r;
    }
    E->cv_work.notify_one();
    *request_id = r->id;
    return CGM_OK;
}

int cgm_engine_poll(void *engine, int64_t request_id, int32_t *done) {
    Engine *E = (Engine *)engine;
    if (!E || !done) return set_error(CGM_ERR_INVALID, "cgm_engine_poll: bad arguments");
    std::lock_guard<std::mutex> lk(E->mu);
    auto it = E->live.find(request_id);
    if (it == E->live.end()) return set_error(CGM_ERR_NOTFOUND, "unknown request id");
    *done = it->second->state == 2;
    return CGM_OK;
}

int cgm_engine_wait(void *engine, int64_t request_id, int32_t timeout_ms, int32_t *score,
                    int32_t *q_end, int32_t *r_end, double *latency_ms) {
    Engine *E = (Engine *)engine;
    if (!E) return set_error(CGM_ERR_INVALID, "null engine");
    std::unique_lock<std::mutex> lk(E->mu);
    auto it = E->live.find(request_id);
    if (it == E->live.end()) return set_error(CGM_ERR_NOTFOUND, "unknown request id");
    Request *r = it->second;
    auto ready = [&] { return r->state == 2; };
    if (timeout_ms < 0) {
        E->cv_done.wait(lk, ready);
    } else if (!E->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return set_error(CGM_ERR_BUSY, "request not finished");
    }
    if (score) memcpy(score, r->score.data(), (size_t)r->n * 4);
    if (q_end) memcpy(q_end, r->qend.data(), (size_t)r->n * 4);
    if (r_end) memcpy(r_end, r->rend.data(), (size_t)r->n * 4);
    if (latency_ms) *latency_ms = (r->t_done - r->t_arrival) * 1e3;
    E->live.erase(it);
    delete r;
    return CGM_OK;
}

int cgm_engine_stats(void *engine, int64_t *out /* 8 */) {
    Engine *E = (Engine *)engine;
    if (!E || !out) return set_error(CGM_ERR_INVALID, "cgm_engine_stats: bad arguments");
    std::lock_guard<std::mutex> lk(E->mu);
    out[0] = E->batches;
    out[1] = E->batch_pairs;
    out[2] = E->done_requests;
    out[3] = E->done_pairs;
    out[4] = E->refused;
    out[5] = E->queued_pairs;
    out[6] = E->device_bytes;
    out[7] = E->worker_error;
    return CGM_OK;
}

double cgm_engine_now(void) { return now_s(); }

int cgm_serve_run(void *engine, void *gen, double rate_rps, double duration_s,
                  int32_t max_pairs_per_request, int32_t max_outstanding, uint64_t seed,
                  double *stats /* 16 */) {
    Engine *E = (Engine *)engine;
    if (!E || !gen || !stats || duration_s <= 0 || max_pairs_per_request < 1)
        return set_error(CGM_ERR_INVALID, "cgm_serve_run: bad arguments");
    int64_t n_gen = 0, qwn = 0, rwn = 0;
    if (int rc = cgm_gen_sizes(gen, &n_gen, &qwn, &rwn)) return rc;
    if (n_gen < max_pairs_per_request) return set_error(CGM_ERR_INVALID, "generator too small");
    std::vector<uint64_t> qp((size_t)qwn), rp((size_t)rwn);
    std::vector<int32_t> qo((size_t)n_gen), ql((size_t)n_gen), ro((size_t)n_gen),
        rl((size_t)n_gen), kind((size_t)n_gen), band((size_t)n_gen), me((size_t)n_gen);
    if (int rc = cgm_gen_copy(gen, qp.data(), nullptr, qo.data(), ql.data(), rp.data(), nullptr,
                              ro.data(), rl.data(), kind.data(), band.data(), me.data()))
        return rc;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_rps > 0 ? rate_rps : 1.0);
    std::uniform_int_distribution<int> sz(1, max_pairs_per_request);
    struct Sub {
        int64_t id;
        int n;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Sub> inflight;
    bool submit_done = false;
    std::vector<double> lat;
    int64_t completed = 0, pairs_done = 0, bad = 0;
    std::atomic<int64_t> outstanding{0};
    std::thread collector([&] {
        for (;;) {
            Sub s;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !inflight.empty() || submit_done; });
                if (inflight.empty()) return;
                s = inflight.front();
                inflight.pop_front();
            }
            std::vector<int32_t> sc((size_t)s.n);
            double l = 0;
            if (cgm_engine_wait(E, s.id, -1, sc.data(), nullptr, nullptr, &l) == CGM_OK) {
                lat.push_back(l);
                completed += 1;
                pairs_done += s.n;
                for (int32_t v : sc) bad += v == CGM_BAD_PAIR;
            }
            outstanding -= 1;
        }
    });
    const double t0 = now_s();
    double t_next = t0;
    int64_t cursor = 0, submitted = 0, refused = 0;
    while (true) {
        const double t = now_s();
        if (t - t0 >= duration_s) break;
        if (rate_rps > 0) {
            if (t < t_next) {
                if (t_next - t > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
                continue;
            }
        } else if (outstanding.load() >= max_outstanding) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        const int n = sz(rng);
        if (cursor + n > n_gen) cursor = 0;
        const double arrival = rate_rps > 0 ? t_next : t;
        int64_t id = 0;
        const size_t c = (size_t)cursor;
        const int rc = cgm_engine_submit(E, qp.data(), rp.data(), ql.data() + c, rl.data() + c,
                                         qo.data() + c, ro.data() + c, n, kind.data() + c,
                                         band.data() + c, me.data() + c, arrival, &id);
        cursor += n;
        if (rate_rps > 0) t_next += gap(rng);
        if (rc == CGM_OK) {
            submitted += 1;
            outstanding += 1;
            std::lock_guard<std::mutex> lk(mu);
            inflight.push_back({id, n});
            cv.notify_one();
        } else if (rc == CGM_ERR_BUSY) {
            refused += 1;
            if (rate_rps <= 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
        } else {
            break;
        }
    }
    const double t_sub_end = now_s();
    {
        std::lock_guard<std::mutex> lk(mu);
        submit_done = true;
    }
    cv.notify_all();
    collector.join();
    const double t_end = now_s();
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
        if (lat.empty()) return 0.0;
        size_t i = (size_t)std::min<double>((double)lat.size() - 1, floor(p * (lat.size() - 1) + 0.5));
        return lat[i];
    };
    double mean = 0;
    for (double v : lat) mean += v;
    if (!lat.empty()) mean /= (double)lat.size();
    int64_t st[8];
    cgm_engine_stats(E, st);
    const double elapsed = t_end - t0;
    for (int i = 0; i < 16; ++i) stats[i] = 0;
    stats[0] = (double)submitted;
    stats[1] = (double)completed;
    stats[2] = (double)refused;
    stats[3] = (double)pairs_done;
    stats[4] = elapsed;
    stats[5] = pairs_done / elapsed;
    stats[6] = completed / elapsed;
    stats[7] = pct(0.50);
    stats[8] = pct(0.90);
    stats[9] = pct(0.99);
    stats[10] = lat.empty() ? 0 : lat.back();
    stats[11] = mean;
    stats[12] = (double)st[0];
    stats[13] = st[0] ? (double)st[1] / (double)st[0] : 0;
    stats[14] = (double)bad;
    stats[15] = t_sub_end - t0;
    return CGM_OK;
}

}  // extern "C"
