// cgm_engine.hip -- continuous-batching serving engine for SW + WFA pairs.
//
// Pairs are batched by bin: bin 0 holds the SW pairs (kind 0), bin 1 the
// WFA pairs (kind 1), and each bin's batches run the mixed-kernel plan with
// that kind for every pair.  A request's pairs may go to both bins; it is
// done when all of them are.
// Each bin has its own dispatcher thread, `n_slots` slots (a stream,
// device buffers and a mixed-kernel plan each, allocated at creation and
// charged against a fixed HBM budget) and 2 n_slots + 1 pinned staging
// buffers.  cgm_engine_submit writes a request's pairs straight into the
// bin's open staging buffer (metadata and packed words, the only host
// copy); whenever a slot of the bin is free the dispatcher takes the oldest
// closed staging buffer, or the open one (in the SW bin only a closed one
// once two of its batches are in flight), and launches it: two host-to-
// device copies, a kernel that spreads the metadata into the batch arrays,
// the mixed-kernel plan, one device-to-host copy of the results and a host
// function on the stream that hands the results to their requests and
// frees the slot.  So SW requests never wait behind long WFA pairs, batches
// of both bins overlap on the GPU, requests that arrive while batches run
// are admitted as staging space and slots free up, and a request is refused
// with CGM_ERR_BUSY when its bin has no staging space left (back-pressure
// instead of unbounded queueing).  Every pair's result depends only on the
// pair, never on the batch it ran in.

#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <random>
#include <thread>
#include <vector>

#include "cgm_internal.h"

namespace cgm {

struct Engine;

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

constexpr int kBins = 2;   // 0: SW, 1: WFA (the pair kind)
constexpr int kMeta = 7;   // per pair: q_len, r_len, q_off, r_off, kind, band, max_error
constexpr int kEagerSwSlots = 2;  // SW batches in flight before only full buffers launch
constexpr int kEngineLaneMaxLen = 200;  // SW reads longer than this: one wavefront per pair

struct Request {
    std::atomic<int64_t> id{0};  // 0 while the slot is free
    int32_t slot = 0;            // index in the engine's request table
    uint32_t gen = 0;            // uses of the slot so far (part of the id)
    int32_t n = 0;
    std::vector<int32_t> score, qend, rend;
    std::atomic<int32_t> remaining{0};
    double t_arrival = 0, t_done = 0;
    int state = 0;  // 0 queued or running, 2 done
    bool has_wfa = false;
};

struct Route {
    Request *req;
    int32_t idx;
};

struct Bin;

// One batch being formed or run: pinned metadata (kMeta ints per pair),
// packed words of both sides, results, and where each pair's result goes.
struct Staging {
    int32_t *h_meta = nullptr;
    uint64_t *h_words = nullptr;
    int32_t *h_out = nullptr;
    // words of device-resident requests: (source address, destination word,
    // word count) per sequence, gathered on the device at launch
    int64_t *h_gather = nullptr;
    int64_t n = 0, words = 0, ngather = 0;
    int32_t max_band = 0, max_len = 0;  // over the batch's pairs
    std::vector<Route> routes;
    int writers = 0;  // submits copying into their reserved part (under the bin's lock)
    Bin *bin = nullptr;
    int slot = -1;
    int rc = CGM_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    MixPlan *plan = nullptr;
    int32_t *d_meta = nullptr, *d_arr = nullptr, *d_out = nullptr;
    uint64_t *d_words = nullptr;
    int64_t *d_gather = nullptr;
};

struct Bin {
    int id = 0;
    Engine *E = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Slot> slots;
    std::vector<int> free_slots;
    std::vector<Staging> stages;
    std::vector<Staging *> free_stages;
    std::deque<Staging *> ready;
    Staging *open = nullptr;
    int64_t pending_pairs = 0;
    MixTuning tune{};  // applied to the slot's plan at every launch
    int eager = 0;     // batches in flight beyond which only a closed buffer launches
    std::atomic<int64_t> batches{0}, batch_pairs{0};
    std::thread dispatcher;

    // Whether the dispatcher can launch now (caller holds mu): a slot is
    // free, and the oldest closed buffer, else the open one, holds pairs and
    // no submit is still writing into it.  Beyond `eager` batches in flight
    // only a closed (full) buffer is launched: under light load the SW bin
    // keeps few batches running, each finishing sooner, and its extra slots
    // only add concurrency once requests back up.  The WFA bin launches on
    // every slot: its batches last as long as their longest pair, and the
    // GPU fills only with several of them in flight.
    bool launchable() const {
        if (free_slots.empty()) return false;
        const int busy = (int)(slots.size() - free_slots.size());
        if (ready.empty() && busy >= eager) return false;
        const Staging *c = !ready.empty() ? ready.front() : open;
        return c->n > 0 && c->writers == 0;
    }
};

// Words of device-resident requests: one wavefront per descriptor
// (source address, destination word, word count), lanes copying words.
__global__ void engine_gather_kernel(const int64_t *desc, int64_t n, uint64_t *dst) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t d = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; d < n; d += nwaves) {
        const uint64_t *src = (const uint64_t *)desc[3 * d];
        const int64_t o = desc[3 * d + 1], cnt = desc[3 * d + 2];
        for (int64_t j = lane; j < cnt; j += 64) dst[o + j] = src[j];
    }
}

// metadata of n pairs (kMeta ints each) -> kMeta arrays of stride cap
__global__ void engine_unpack_kernel(const int32_t *meta, int64_t n, int32_t *arr, int64_t cap) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < kMeta; ++k) arr[k * cap + i] = meta[i * kMeta + k];
}

void engine_host_done(void *arg);

}  // namespace

struct Engine {
    int device = 0;
    int n_slots = 0;
    int64_t cap_pairs = 0, cap_words = 0, device_bytes = 0;
    SwScoring sc{};
    WfaPen pen{};
    Bin bins[kBins];
    std::mutex mu;  // requests' done state, cv_done
    std::condition_variable cv_done;
    // Request table: a request's id is (use count of its slot << 32) | slot;
    // slots live in chunks that are never moved or freed before the engine,
    // so a lookup by id takes no lock, and submits register a request
    // without touching `mu`.
    static constexpr int kReqChunkBits = 12, kReqChunks = 4096;
    std::atomic<Request *> req_chunks[kReqChunks];
    std::mutex req_mu;  // free_req, n_req, chunk allocation
    std::vector<int32_t> free_req;
    int32_t n_req = 0;
    std::atomic<bool> stop{false};
    std::atomic<int64_t> done_requests{0}, done_pairs{0}, refused{0};
    std::atomic<int> worker_error{0};

    Engine() {
        for (auto &c : req_chunks) c.store(nullptr, std::memory_order_relaxed);
    }

    ~Engine() {
        stop = true;
        for (Bin &B : bins) {
            {
                std::lock_guard<std::mutex> lk(B.mu);
            }
            B.cv.notify_all();
            if (B.dispatcher.joinable()) B.dispatcher.join();
        }
        (void)hipSetDevice(device);
        for (Bin &B : bins) {
            for (Slot &s : B.slots) {
                if (s.stream) (void)hipStreamSynchronize(s.stream);
                if (s.plan) mix_plan_destroy(s.plan);
                void *dev_ptrs[] = {s.d_meta, s.d_arr, s.d_out, s.d_words, s.d_gather};
                for (void *p : dev_ptrs)
                    if (p) (void)hipFree(p);
                if (s.stream) (void)hipStreamDestroy(s.stream);
            }
            for (Staging &st : B.stages) {
                void *host_ptrs[] = {st.h_meta, st.h_words, st.h_out, st.h_gather};
                for (void *p : host_ptrs)
                    if (p) (void)hipHostFree(p);
            }
        }
        for (auto &c : req_chunks) delete[] c.load();
    }

    // A free request slot (its id still 0), or null when the table is full
    // or a chunk cannot be allocated.
    Request *take_request() {
        constexpr int kChunk = 1 << kReqChunkBits;
        std::lock_guard<std::mutex> lk(req_mu);
        int32_t slot;
        if (!free_req.empty()) {
            slot = free_req.back();
            free_req.pop_back();
        } else {
            if (n_req == kReqChunks * kChunk) return nullptr;
            slot = n_req;
            const int c = slot >> kReqChunkBits;
            if (!req_chunks[c].load(std::memory_order_relaxed)) {
                Request *chunk = new (std::nothrow) Request[kChunk];
                if (!chunk) return nullptr;
                for (int i = 0; i < kChunk; ++i) chunk[i].slot = c * kChunk + i;
                req_chunks[c].store(chunk, std::memory_order_release);
            }
            n_req += 1;
        }
        Request *r = &req_chunks[slot >> kReqChunkBits].load(std::memory_order_relaxed)
                          [slot & (kChunk - 1)];
        r->gen = (r->gen + 1) & 0x7fffffffu;
        if (r->gen == 0) r->gen = 1;
        return r;
    }

    // The live request with this id, or null.
    Request *find_request(int64_t id) {
        if (id <= 0) return nullptr;
        const int64_t slot = id & 0xffffffff;
        if (slot >= (int64_t)kReqChunks << kReqChunkBits) return nullptr;
        Request *chunk = req_chunks[slot >> kReqChunkBits].load(std::memory_order_acquire);
        if (!chunk) return nullptr;
        Request *r = &chunk[slot & ((1 << kReqChunkBits) - 1)];
        return r->id.load(std::memory_order_acquire) == id ? r : nullptr;
    }

    // Returns a slot whose id the caller has set to 0.
    void release_request(Request *r) {
        std::lock_guard<std::mutex> lk(req_mu);
        free_req.push_back(r->slot);
    }

    // Launches staging `st` on slot `si` of bin B; on any failure the batch's
    // pairs are reported as refused (score CGM_BAD_PAIR, ends -1).
    void launch(Bin &B, int si, Staging *st, const MixTuning &tune) {
        Slot &S = B.slots[(size_t)si];
        mix_plan_set_tuning(S.plan, tune);
        mix_plan_set_hint(S.plan, st->max_band, st->max_len);
        hipStream_t s = S.stream;
        const int64_t p = st->n, w = st->words, cap = cap_pairs;
        st->slot = si;
        int rc = CGM_OK;
        auto ok = [&](hipError_t e) {
            if (rc == CGM_OK && e != hipSuccess) rc = CGM_ERR_HIP;
        };
        ok(hipMemcpyAsync(S.d_meta, st->h_meta, (size_t)(p * kMeta * 4), hipMemcpyHostToDevice, s));
        if (w > 0)
            ok(hipMemcpyAsync(S.d_words, st->h_words, (size_t)(w * 8), hipMemcpyHostToDevice, s));
        if (st->ngather > 0 && rc == CGM_OK) {
            // device-resident requests: their words over the (unused) parts
            // of the host copy reserved for them
            ok(hipMemcpyAsync(S.d_gather, st->h_gather, (size_t)(st->ngather * 3 * 8),
                              hipMemcpyHostToDevice, s));
            if (rc == CGM_OK) {
                const int64_t blocks = std::min<int64_t>((st->ngather + 3) / 4, 2048);
                hipLaunchKernelGGL(engine_gather_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                                   S.d_gather, st->ngather, S.d_words);
                ok(hipGetLastError());
            }
        }
        if (rc == CGM_OK) {
            hipLaunchKernelGGL(engine_unpack_kernel, dim3((unsigned)((p + 255) / 256)), dim3(256), 0,
                               s, S.d_meta, p, S.d_arr, cap);
            ok(hipGetLastError());
        }
        const int32_t *A = S.d_arr;
        Batch b = {S.d_words, S.d_words, A, A + cap, A + 2 * cap, A + 3 * cap, p, w, w};
        SwOut out = {S.d_out, S.d_out + p, S.d_out + 2 * p};
        if (rc == CGM_OK)
            rc = mix_enqueue(S.plan, b, nullptr, B.id, A + 5 * cap, 0, A + 6 * cap, 0, sc, pen, out, s);
        if (rc == CGM_OK)
            ok(hipMemcpyAsync(st->h_out, S.d_out, (size_t)(3 * p * 4), hipMemcpyDeviceToHost, s));
        if (rc == CGM_OK) ok(hipLaunchHostFunc(s, engine_host_done, st));
        B.batches += 1;
        B.batch_pairs += p;
        if (rc != CGM_OK) {
            (void)hipStreamSynchronize(s);
            worker_error = rc;
            st->rc = rc;
            engine_host_done(st);
        }
    }

    void dispatch(Bin &B) {
        (void)hipSetDevice(device);
        for (;;) {
            Staging *st = nullptr;
            int si = -1;
            MixTuning tune;
            {
                std::unique_lock<std::mutex> lk(B.mu);
                // the oldest closed staging buffer, else the open one; either
                // only once every submit writing into it has finished
                B.cv.wait(lk, [&] { return stop.load() || B.launchable(); });
                if (stop.load()) return;
                if (!B.ready.empty()) {
                    st = B.ready.front();
                    B.ready.pop_front();
                } else {
                    // a slot is free and nothing else waits, so a free staging
                    // buffer exists (2 n_slots + 1 of them)
                    st = B.open;
                    B.open = B.free_stages.back();
                    B.free_stages.pop_back();
                }
                B.pending_pairs -= st->n;
                si = B.free_slots.back();
                B.free_slots.pop_back();
                tune = B.tune;
            }
            launch(B, si, st, tune);
        }
    }
};

namespace {

// Runs once a batch's results are in pinned memory (a host function on the
// slot's stream, or directly after a failed launch): routes every pair's
// result to its request, marks the finished requests, frees the slot and
// the staging buffer.  Makes no HIP call.
void engine_host_done(void *arg) {
    Staging *st = (Staging *)arg;
    Bin &B = *st->bin;
    Engine &E = *B.E;
    const int64_t p = st->n;
    const int32_t *o = st->h_out;
    std::vector<Request *> finished;
    for (int64_t i = 0; i < p; ++i) {
        const Route &rt = st->routes[(size_t)i];
        Request *r = rt.req;
        if (st->rc == CGM_OK) {
            r->score[(size_t)rt.idx] = o[i];
            r->qend[(size_t)rt.idx] = o[p + i];
            r->rend[(size_t)rt.idx] = o[2 * p + i];
        } else {
            r->score[(size_t)rt.idx] = CGM_BAD_PAIR;
            r->qend[(size_t)rt.idx] = -1;
            r->rend[(size_t)rt.idx] = -1;
        }
        if (r->remaining.fetch_sub(1) == 1) finished.push_back(r);
    }
    if (!finished.empty()) {
        const double t = now_s();
        {
            std::lock_guard<std::mutex> lk(E.mu);
            for (Request *r : finished) {
                r->t_done = t;
                r->state = 2;
            }
        }
        E.cv_done.notify_all();
        E.done_requests += (int64_t)finished.size();
    }
    E.done_pairs += p;
    {
        std::lock_guard<std::mutex> lk(B.mu);
        B.free_slots.push_back(st->slot);
        st->n = 0;
        st->words = 0;
        st->ngather = 0;
        st->max_band = 0;
        st->max_len = 0;
        st->rc = CGM_OK;
        st->slot = -1;
        st->routes.clear();
        B.free_stages.push_back(st);
    }
    B.cv.notify_one();
}

}  // namespace

}  // namespace cgm

using namespace cgm;

extern "C" {

int cgm_engine_create(int64_t hbm_budget_bytes, int32_t n_slots, int64_t max_batch_pairs,
                      int64_t max_batch_words, int32_t max_error_cap, int32_t match,
                      int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t x, int32_t o,
                      int32_t e, void **engine) {
    if (!engine || n_slots < 1 || n_slots > 16 || max_batch_pairs < 1 || max_batch_words < 1 ||
        max_batch_pairs > (int64_t(1) << 28) || max_batch_words > (int64_t(1) << 29))
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
    E->sc = sc;
    E->pen = pen;
    const int64_t cap = max_batch_pairs, wcap = 2 * max_batch_words;
    for (int bi = 0; bi < kBins; ++bi) {
        Bin &B = E->bins[bi];
        B.id = bi;
        B.E = E;
        B.eager = bi == 0 ? kEagerSwSlots : n_slots;
        B.slots.resize((size_t)n_slots);
        for (int k = 0; k < n_slots; ++k) {
            Slot &s = B.slots[(size_t)k];
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess)
                return fail(set_error(CGM_ERR_HIP, "engine: stream"));
            // SW slots carry no WFA workspace
            int rc = mix_plan_create(max_batch_pairs, bi == 1 ? max_error_cap : 0, pen, 0, 0, &s.plan);
            if (rc) return fail(rc);
            // every kernel on the slot's own stream: the process's streams
            // share GPU_MAX_HW_QUEUES hardware queues (4 by default), and a
            // short SW batch must not queue behind a long WFA one
            mix_plan_set_fork(s.plan, false);
            if (k == 0) {
                B.tune = mix_plan_tuning(s.plan);
                // a batch lasts as long as its longest pair: SW reads beyond
                // kEngineLaneMaxLen take one wavefront each instead of a lane
                if (bi == 0) B.tune.lane_max_len = kEngineLaneMaxLen;
            }
            E->device_bytes += mix_plan_bytes(s.plan) + wcap * 8 + (2 * kMeta + 3) * cap * 4 +
                               2 * cap * 3 * 8;
            if (E->device_bytes > hbm_budget_bytes)
                return fail(set_error(CGM_ERR_NOMEM, "engine: slots exceed the HBM budget"));
            if (hipMalloc((void **)&s.d_meta, (size_t)(kMeta * cap * 4)) != hipSuccess ||
                hipMalloc((void **)&s.d_arr, (size_t)(kMeta * cap * 4)) != hipSuccess ||
                hipMalloc((void **)&s.d_out, (size_t)(3 * cap * 4)) != hipSuccess ||
                hipMalloc((void **)&s.d_words, (size_t)(wcap * 8)) != hipSuccess ||
                hipMalloc((void **)&s.d_gather, (size_t)(2 * cap * 3 * 8)) != hipSuccess)
                return fail(set_error(CGM_ERR_NOMEM, "engine: slot buffers"));
            B.free_slots.push_back(n_slots - 1 - k);
        }
        B.stages.resize((size_t)(2 * n_slots + 1));
        for (Staging &st : B.stages) {
            st.bin = &B;
            st.routes.reserve((size_t)cap);
            if (hipHostMalloc((void **)&st.h_meta, (size_t)(kMeta * cap * 4), hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc((void **)&st.h_words, (size_t)(wcap * 8), hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc((void **)&st.h_out, (size_t)(3 * cap * 4), hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc((void **)&st.h_gather, (size_t)(2 * cap * 3 * 8),
                              hipHostMallocDefault) != hipSuccess)
                return fail(set_error(CGM_ERR_NOMEM, "engine: pinned staging"));
        }
        B.open = &B.stages[0];
        for (size_t k = B.stages.size(); k-- > 1;) B.free_stages.push_back(&B.stages[k]);
    }
    for (Bin &B : E->bins) {
        Bin *pb = &B;
        B.dispatcher = std::thread([E, pb] { E->dispatch(*pb); });
    }
    *engine = E;
    return CGM_OK;
}

int cgm_engine_destroy(void *engine) {
    delete (Engine *)engine;
    return CGM_OK;
}

}  // extern "C"

namespace {
// cgm_engine_submit (device_words false: q_packed / r_packed are host
// memory, copied into the pinned staging buffer) and
// cgm_engine_submit_device (true: they are device memory, gathered on the
// device when the batch launches; the staging buffer gets descriptors).
int engine_submit(void *engine, bool device_words, const uint64_t *q_packed, int64_t q_words,
                  const uint64_t *r_packed, int64_t r_words, const int32_t *q_len,
                  const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                  int32_t n_pairs, const int32_t *kind, const int32_t *band,
                  const int32_t *max_error, double arrival_s, int64_t *request_id) {
    Engine *E = (Engine *)engine;
    if (!E || !request_id || n_pairs < 1 || !q_len || !r_len || !q_off || !r_off || !kind ||
        !band || !max_error || q_words < 0 || r_words < 0)
        return set_error(CGM_ERR_INVALID, "cgm_engine_submit: bad arguments");
    // every pair's words must lie inside the caller's buffers
    int64_t nb[kBins] = {0, 0}, wb[kBins] = {0, 0};
    int32_t mband[kBins] = {0, 0}, mlen[kBins] = {0, 0};
    for (int32_t i = 0; i < n_pairs; ++i) {
        if (q_len[i] < 0 || r_len[i] < 0 || q_off[i] < 0 || r_off[i] < 0)
            return set_error(CGM_ERR_INVALID, "cgm_engine_submit: negative length or offset");
        if (kind[i] != 0 && kind[i] != 1)
            return set_error(CGM_ERR_INVALID, "cgm_engine_submit: kind must be 0 (SW) or 1 (WFA)");
        const int64_t qw = ((int64_t)q_len[i] + 31) / 32, rw = ((int64_t)r_len[i] + 31) / 32;
        if ((qw > 0 && !q_packed) || (rw > 0 && !r_packed) || q_off[i] + qw > q_words ||
            r_off[i] + rw > r_words)
            return set_error(CGM_ERR_INVALID, "cgm_engine_submit: sequence outside its buffer");
        const int bi = kind[i];
        nb[bi] += 1;
        wb[bi] += qw + rw;
        mband[bi] = std::max(mband[bi], band[i]);
        mlen[bi] = std::max(mlen[bi], std::max(q_len[i], r_len[i]));
    }
    for (int bi = 0; bi < kBins; ++bi)
        if (nb[bi] > E->cap_pairs || wb[bi] > 2 * E->cap_words)
            return set_error(CGM_ERR_INVALID, "cgm_engine_submit: request larger than a batch");
    Request *r = E->take_request();
    if (!r) return set_error(CGM_ERR_NOMEM, "cgm_engine_submit: request table full");
    r->n = n_pairs;
    // every pair's result is written before the request is marked done
    r->score.resize((size_t)n_pairs);
    r->qend.resize((size_t)n_pairs);
    r->rend.resize((size_t)n_pairs);
    r->remaining = n_pairs;
    r->state = 0;
    r->has_wfa = nb[1] > 0;
    r->t_arrival = arrival_s > 0 ? arrival_s : now_s();
    r->t_done = 0;
    const int64_t id = ((int64_t)r->gen << 32) | r->slot;
    r->id.store(id, std::memory_order_release);
    // the locks of the bins the request uses, in bin order, so the request
    // is admitted whole or not at all
    std::unique_lock<std::mutex> lks[kBins];
    for (int bi = 0; bi < kBins; ++bi)
        if (nb[bi] > 0) lks[bi] = std::unique_lock<std::mutex>(E->bins[bi].mu);
    bool fits_open[kBins] = {true, true};
    for (int bi = 0; bi < kBins; ++bi) {
        if (nb[bi] == 0) continue;
        Bin &B = E->bins[bi];
        fits_open[bi] = B.open->n + nb[bi] <= E->cap_pairs && B.open->words + wb[bi] <= 2 * E->cap_words;
        if (!fits_open[bi] && B.free_stages.empty()) {
            for (int bj = kBins - 1; bj >= 0; --bj)
                if (lks[bj].owns_lock()) lks[bj].unlock();
            E->refused += 1;
            r->id.store(0, std::memory_order_release);
            E->release_request(r);
            return set_error(CGM_ERR_BUSY, "engine queue full");
        }
    }
    // reserve the request's part of each bin's open staging buffer under the
    // locks, copy into it after releasing them
    Staging *stg[kBins] = {nullptr, nullptr};
    int64_t n0[kBins] = {0, 0}, w0[kBins] = {0, 0}, g0[kBins] = {0, 0};
    for (int bi = 0; bi < kBins; ++bi) {
        if (nb[bi] == 0) continue;
        Bin &B = E->bins[bi];
        if (!fits_open[bi]) {
            B.ready.push_back(B.open);
            B.open = B.free_stages.back();
            B.free_stages.pop_back();
        }
        Staging *st = B.open;
        stg[bi] = st;
        n0[bi] = st->n;
        w0[bi] = st->words;
        g0[bi] = st->ngather;
        st->n += nb[bi];
        st->words += wb[bi];
        if (device_words) st->ngather += 2 * nb[bi];
        st->routes.resize((size_t)st->n);  // within the capacity reserved at creation
        st->writers += 1;
        B.pending_pairs += nb[bi];
    }
    for (int bj = kBins - 1; bj >= 0; --bj)
        if (lks[bj].owns_lock()) lks[bj].unlock();
    for (int bi = 0; bi < kBins; ++bi) {
        Staging *st = stg[bi];
        if (!st) continue;
        int64_t n = n0[bi], words = w0[bi], gd = g0[bi];
        for (int32_t i = 0; i < n_pairs; ++i) {
            if (kind[i] != bi) continue;
            const int64_t qw = ((int64_t)q_len[i] + 31) / 32, rw = ((int64_t)r_len[i] + 31) / 32;
            int32_t *m = st->h_meta + n * kMeta;
            m[0] = q_len[i];
            m[1] = r_len[i];
            m[2] = (int32_t)words;
            m[3] = (int32_t)(words + qw);
            m[4] = kind[i];
            m[5] = band[i];
            m[6] = max_error[i];
            if (device_words) {
                int64_t *g = st->h_gather + 3 * gd;
                g[0] = (int64_t)(uintptr_t)(q_packed + q_off[i]);
                g[1] = words;
                g[2] = qw;
                g[3] = (int64_t)(uintptr_t)(r_packed + r_off[i]);
                g[4] = words + qw;
                g[5] = rw;
                gd += 2;
            } else {
                if (qw) memcpy(st->h_words + words, q_packed + q_off[i], (size_t)qw * 8);
                if (rw) memcpy(st->h_words + words + qw, r_packed + r_off[i], (size_t)rw * 8);
            }
            words += qw + rw;
            st->routes[(size_t)n] = {r, i};
            n += 1;
        }
    }
    // the dispatcher only acts when it can launch; otherwise a slot's
    // completion or the last writer of the buffer it waits for wakes it, and
    // a wake-up here would cost a futex round trip per request for nothing
    for (int bi = 0; bi < kBins; ++bi) {
        Staging *st = stg[bi];
        if (!st) continue;
        Bin &B = E->bins[bi];
        bool wake;
        {
            std::lock_guard<std::mutex> lk(B.mu);
            st->max_band = std::max(st->max_band, mband[bi]);
            st->max_len = std::max(st->max_len, mlen[bi]);
            st->writers -= 1;
            wake = st->writers == 0 && B.launchable();
        }
        if (wake) B.cv.notify_one();
    }
    *request_id = id;
    return CGM_OK;
}
}  // namespace

extern "C" {

int cgm_engine_submit(void *engine, const uint64_t *q_packed, int64_t q_words,
                      const uint64_t *r_packed, int64_t r_words, const int32_t *q_len,
                      const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                      int32_t n_pairs, const int32_t *kind, const int32_t *band,
                      const int32_t *max_error, double arrival_s, int64_t *request_id) {
    return engine_submit(engine, false, q_packed, q_words, r_packed, r_words, q_len, r_len, q_off,
                         r_off, n_pairs, kind, band, max_error, arrival_s, request_id);
}

int cgm_engine_submit_device(void *engine, const uint64_t *q_packed, int64_t q_words,
                             const uint64_t *r_packed, int64_t r_words, const int32_t *q_len,
                             const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                             int32_t n_pairs, const int32_t *kind, const int32_t *band,
                             const int32_t *max_error, double arrival_s, int64_t *request_id) {
    return engine_submit(engine, true, q_packed, q_words, r_packed, r_words, q_len, r_len, q_off,
                         r_off, n_pairs, kind, band, max_error, arrival_s, request_id);
}

int cgm_engine_poll(void *engine, int64_t request_id, int32_t *done) {
    Engine *E = (Engine *)engine;
    if (!E || !done) return set_error(CGM_ERR_INVALID, "cgm_engine_poll: bad arguments");
    std::lock_guard<std::mutex> lk(E->mu);
    const Request *r = E->find_request(request_id);
    if (!r) return set_error(CGM_ERR_NOTFOUND, "unknown request id");
    *done = r->state == 2;
    return CGM_OK;
}

int cgm_engine_wait(void *engine, int64_t request_id, int32_t timeout_ms, int32_t *score,
                    int32_t *q_end, int32_t *r_end, double *latency_ms) {
    Engine *E = (Engine *)engine;
    if (!E) return set_error(CGM_ERR_INVALID, "null engine");
    std::unique_lock<std::mutex> lk(E->mu);
    Request *r = E->find_request(request_id);
    if (!r) return set_error(CGM_ERR_NOTFOUND, "unknown request id");
    // a concurrent wait on the same id may collect it first (the id then
    // no longer matches the slot)
    auto gone = [&] { return r->id.load(std::memory_order_acquire) != request_id; };
    auto ready = [&] { return r->state == 2 || gone(); };
    if (timeout_ms < 0) {
        E->cv_done.wait(lk, ready);
    } else if (!E->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return set_error(CGM_ERR_BUSY, "request not finished");
    }
    if (gone()) return set_error(CGM_ERR_NOTFOUND, "unknown request id");
    if (score) memcpy(score, r->score.data(), (size_t)r->n * 4);
    if (q_end) memcpy(q_end, r->qend.data(), (size_t)r->n * 4);
    if (r_end) memcpy(r_end, r->rend.data(), (size_t)r->n * 4);
    if (latency_ms) *latency_ms = (r->t_done - r->t_arrival) * 1e3;
    r->id.store(0, std::memory_order_release);
    lk.unlock();
    E->release_request(r);
    return CGM_OK;
}

int cgm_engine_get_tuning(void *engine, int32_t bin, int32_t *params) {
    Engine *E = (Engine *)engine;
    if (!E || !params || bin < 0 || bin >= kBins)
        return set_error(CGM_ERR_INVALID, "cgm_engine_get_tuning: bad arguments");
    Bin &B = E->bins[bin];
    std::lock_guard<std::mutex> lk(B.mu);
    tuning_to_array(B.tune, params);
    return CGM_OK;
}

int cgm_engine_set_tuning(void *engine, int32_t bin, const int32_t *params) {
    Engine *E = (Engine *)engine;
    if (!E || !params || bin < 0 || bin >= kBins)
        return set_error(CGM_ERR_INVALID, "cgm_engine_set_tuning: bad arguments");
    MixTuning t;
    if (int rc = tuning_from_array(params, &t)) return rc;
    Bin &B = E->bins[bin];
    std::lock_guard<std::mutex> lk(B.mu);
    B.tune = t;
    return CGM_OK;
}

int cgm_engine_stats(void *engine, int64_t *out /* 8 */) {
    Engine *E = (Engine *)engine;
    if (!E || !out) return set_error(CGM_ERR_INVALID, "cgm_engine_stats: bad arguments");
    int64_t batches = 0, batch_pairs = 0, queued = 0;
    for (Bin &B : E->bins) {
        std::lock_guard<std::mutex> lk(B.mu);
        batches += B.batches;
        batch_pairs += B.batch_pairs;
        queued += B.pending_pairs;
    }
    out[0] = batches;
    out[1] = batch_pairs;
    out[2] = E->done_requests;
    out[3] = E->done_pairs;
    out[4] = E->refused;
    out[5] = queued;
    out[6] = E->device_bytes;
    out[7] = E->worker_error;
    return CGM_OK;
}

double cgm_engine_now(void) { return now_s(); }

int cgm_serve_run(void *engine, void *gen, double rate_rps, double duration_s,
                  int32_t max_pairs_per_request, int32_t max_outstanding, int32_t clients,
                  int32_t device_inputs, uint64_t seed, double *stats /* 27 */) {
    Engine *E = (Engine *)engine;
    if (!E || !gen || !stats || duration_s <= 0 || max_pairs_per_request < 1 || clients < 1 ||
        clients > 64)
        return set_error(CGM_ERR_INVALID, "cgm_serve_run: bad arguments");
    int64_t n_gen = 0, qwn = 0, rwn = 0;
    if (int rc = cgm_gen_sizes(gen, &n_gen, &qwn, &rwn)) return rc;
    if (n_gen < (int64_t)max_pairs_per_request * clients)
        return set_error(CGM_ERR_INVALID, "generator too small");
    std::vector<uint64_t> qp((size_t)qwn), rp((size_t)rwn);
    std::vector<int32_t> qo((size_t)n_gen), ql((size_t)n_gen), ro((size_t)n_gen),
        rl((size_t)n_gen), kind((size_t)n_gen), band((size_t)n_gen), me((size_t)n_gen);
    if (int rc = cgm_gen_copy(gen, qp.data(), nullptr, qo.data(), ql.data(), rp.data(), nullptr,
                              ro.data(), rl.data(), kind.data(), band.data(), me.data()))
        return rc;
    // device-resident inputs: the packed words are uploaded once and every
    // request is submitted with device pointers
    uint64_t *qd = nullptr, *rd = nullptr;
    struct DevFree {
        uint64_t *&a, *&b;
        ~DevFree() {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
        }
    } dev_free{qd, rd};
    if (device_inputs) {
        CGM_HIP(hipSetDevice(E->device));
        CGM_HIP(hipMalloc((void **)&qd, (size_t)std::max<int64_t>(qwn, 1) * 8));
        CGM_HIP(hipMalloc((void **)&rd, (size_t)std::max<int64_t>(rwn, 1) * 8));
        CGM_HIP(hipMemcpy(qd, qp.data(), (size_t)qwn * 8, hipMemcpyHostToDevice));
        CGM_HIP(hipMemcpy(rd, rp.data(), (size_t)rwn * 8, hipMemcpyHostToDevice));
    }
    const uint64_t *q_src = device_inputs ? qd : qp.data();
    const uint64_t *r_src = device_inputs ? rd : rp.data();
    int64_t b0[kBins], bp0[kBins];
    for (int bi = 0; bi < kBins; ++bi) {
        b0[bi] = E->bins[bi].batches;
        bp0[bi] = E->bins[bi].batch_pairs;
    }
    struct Sub {
        int64_t id;
        int n;
        bool wfa;
    };
    // One client: a submitting thread (Poisson arrivals at rate_rps / clients,
    // or a closed loop) over its own slice of the generated pairs, and a
    // collecting thread that waits for its requests in submission order.
    struct Client {
        std::mutex mu;
        std::condition_variable cv;
        std::deque<Sub> inflight;
        bool submit_done = false;
        std::vector<double> lat, lat_sw, lat_wfa;
        int64_t completed = 0, pairs_done = 0, bad = 0, submitted = 0, refused = 0;
        int error = CGM_OK;
        double t_sub_end = 0;
        // submitter: time inside the submit call, time held back by
        // max_outstanding, sum of the in-flight count seen at each submit
        double t_submit = 0, t_throttled = 0, inflight_sum = 0;
    };
    std::vector<Client> cl((size_t)clients);
    std::atomic<int64_t> outstanding{0};
    const int64_t slice = n_gen / clients;
    const double t0 = now_s();
    auto collect = [&](Client &c) {
        std::vector<int32_t> sc((size_t)max_pairs_per_request);
        for (;;) {
            Sub sub;
            {
                std::unique_lock<std::mutex> lk(c.mu);
                c.cv.wait(lk, [&] { return !c.inflight.empty() || c.submit_done; });
                if (c.inflight.empty()) return;
                sub = c.inflight.front();
                c.inflight.pop_front();
            }
            double l = 0;
            if (cgm_engine_wait(E, sub.id, -1, sc.data(), nullptr, nullptr, &l) == CGM_OK) {
                c.lat.push_back(l);
                (sub.wfa ? c.lat_wfa : c.lat_sw).push_back(l);
                c.completed += 1;
                c.pairs_done += sub.n;
                for (int i = 0; i < sub.n; ++i) c.bad += sc[(size_t)i] == CGM_BAD_PAIR;
            }
            outstanding -= 1;
        }
    };
    auto submit = [&](int ci) {
        Client &c = cl[(size_t)ci];
        std::mt19937_64 rng(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(ci + 1));
        std::exponential_distribution<double> gap(rate_rps > 0 ? rate_rps / clients : 1.0);
        std::uniform_int_distribution<int> sz(1, max_pairs_per_request);
        const int64_t lo = slice * ci, hi = lo + slice;
        int64_t cursor = lo;
        double t_next = t0;
        for (;;) {
            const double t = now_s();
            if (t - t0 >= duration_s) break;
            if (rate_rps > 0) {
                if (t < t_next) {
                    if (t_next - t > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
                    continue;
                }
            } else if (outstanding.load() >= max_outstanding) {
                std::this_thread::sleep_for(std::chrono::microseconds(5));
                c.t_throttled += now_s() - t;
                continue;
            }
            const int n = sz(rng);
            if (cursor + n > hi) cursor = lo;
            const double arrival = rate_rps > 0 ? t_next : t;
            int64_t id = 0;
            const size_t k = (size_t)cursor;
            bool wfa = false;
            for (int i = 0; i < n; ++i) wfa |= kind[k + (size_t)i] == 1;
            const double ts = now_s();
            const int rc = engine_submit(E, device_inputs != 0, q_src, qwn, r_src, rwn,
                                         ql.data() + k, rl.data() + k, qo.data() + k,
                                         ro.data() + k, n, kind.data() + k, band.data() + k,
                                         me.data() + k, arrival, &id);
            c.t_submit += now_s() - ts;
            c.inflight_sum += (double)outstanding.load();
            cursor += n;
            if (rate_rps > 0) t_next += gap(rng);
            if (rc == CGM_OK) {
                c.submitted += 1;
                outstanding += 1;
                std::lock_guard<std::mutex> lk(c.mu);
                c.inflight.push_back({id, n, wfa});
                c.cv.notify_one();
            } else if (rc == CGM_ERR_BUSY) {
                c.refused += 1;
                if (rate_rps <= 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
            } else {
                c.error = rc;
                break;
            }
        }
        c.t_sub_end = now_s();
        {
            std::lock_guard<std::mutex> lk(c.mu);
            c.submit_done = true;
        }
        c.cv.notify_all();
    };
    std::vector<std::thread> th;
    for (int ci = 0; ci < clients; ++ci) {
        th.emplace_back(collect, std::ref(cl[(size_t)ci]));
        th.emplace_back(submit, ci);
    }
    for (std::thread &t : th) t.join();
    const double t_end = now_s();
    std::vector<double> lat, lat_sw, lat_wfa;
    int64_t submitted = 0, completed = 0, refused = 0, pairs_done = 0, bad = 0;
    double t_sub_end = t0, t_submit = 0, t_throttled = 0, inflight_sum = 0;
    int error = CGM_OK;
    for (Client &c : cl) {
        t_submit += c.t_submit;
        t_throttled += c.t_throttled;
        inflight_sum += c.inflight_sum;
        lat.insert(lat.end(), c.lat.begin(), c.lat.end());
        lat_sw.insert(lat_sw.end(), c.lat_sw.begin(), c.lat_sw.end());
        lat_wfa.insert(lat_wfa.end(), c.lat_wfa.begin(), c.lat_wfa.end());
        submitted += c.submitted;
        completed += c.completed;
        refused += c.refused;
        pairs_done += c.pairs_done;
        bad += c.bad;
        t_sub_end = std::max(t_sub_end, c.t_sub_end);
        if (c.error != CGM_OK) error = c.error;
    }
    if (error != CGM_OK) return error;
    auto pct = [](std::vector<double> &v, double p) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        const size_t i = (size_t)std::min<double>((double)v.size() - 1, floor(p * (double)(v.size() - 1) + 0.5));
        return v[i];
    };
    double mean = 0;
    for (double v : lat) mean += v;
    if (!lat.empty()) mean /= (double)lat.size();
    int64_t nb[kBins], np[kBins];
    for (int bi = 0; bi < kBins; ++bi) {
        nb[bi] = E->bins[bi].batches - b0[bi];
        np[bi] = E->bins[bi].batch_pairs - bp0[bi];
    }
    const double elapsed = t_end - t0;
    for (int i = 0; i < 27; ++i) stats[i] = 0;
    stats[0] = (double)submitted;
    stats[1] = (double)completed;
    stats[2] = (double)refused;
    stats[3] = (double)pairs_done;
    stats[4] = elapsed;
    stats[5] = pairs_done / elapsed;
    stats[6] = completed / elapsed;
    stats[7] = pct(lat, 0.50);
    stats[8] = pct(lat, 0.90);
    stats[9] = pct(lat, 0.99);
    stats[10] = lat.empty() ? 0 : lat.back();
    stats[11] = mean;
    stats[12] = (double)(nb[0] + nb[1]);
    stats[13] = nb[0] + nb[1] ? (double)(np[0] + np[1]) / (double)(nb[0] + nb[1]) : 0;
    stats[14] = (double)bad;
    stats[15] = t_sub_end - t0;
    stats[16] = (double)lat_sw.size();
    stats[17] = pct(lat_sw, 0.50);
    stats[18] = pct(lat_sw, 0.99);
    stats[19] = (double)lat_wfa.size();
    stats[20] = pct(lat_wfa, 0.50);
    stats[21] = pct(lat_wfa, 0.99);
    stats[22] = nb[0] ? (double)np[0] / (double)nb[0] : 0;
    stats[23] = nb[1] ? (double)np[1] / (double)nb[1] : 0;
    const double phase = std::max(t_sub_end - t0, 1e-9) * clients;
    stats[24] = t_submit / phase;
    stats[25] = t_throttled / phase;
    stats[26] = submitted + refused ? inflight_sum / (double)(submitted + refused) : 0;
    return CGM_OK;
}

}  // extern "C"
