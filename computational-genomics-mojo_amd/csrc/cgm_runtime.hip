// cgm_runtime.hip -- C ABI of libcgm_align.so: errors, device management,
// argument checks and dispatch of the alignment launches.

#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <set>
#include <utility>

#include "cgm_internal.h"

namespace cgm {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(e == hipErrorOutOfMemory ? CGM_ERR_NOMEM
                     : e == hipErrorNoDevice  ? CGM_ERR_NODEV
                                              : CGM_ERR_HIP,
                     m);
}

int check_score_magnitudes(const SwScoring &sc) {
    const int32_t v[4] = {sc.match, sc.mismatch, sc.gap_open, sc.gap_extend};
    for (int32_t x : v)
        if (x > kMaxScoreMagnitude || x < -kMaxScoreMagnitude)
            return set_error(CGM_ERR_INVALID, "SW scoring value out of range (|v| <= 2^20)");
    return CGM_OK;
}

int check_sw_scoring(const SwScoring &sc) {
    if (int rc = check_score_magnitudes(sc)) return rc;
    // Non-positive penalties bound every score by match * min(qlen, rlen),
    // which the kernels check per pair against int32 range.
    if (sc.mismatch > 0 || sc.gap_extend > 0 || sc.gap_open + sc.gap_extend > 0)
        return set_error(CGM_ERR_INVALID,
                         "SW penalties must be <= 0: mismatch, gap_extend and gap_open + gap_extend");
    return CGM_OK;
}

int check_wfa_pen(const WfaPen &p) {
    if (p.x < 1 || p.e < 1 || p.o < 0)
        return set_error(CGM_ERR_INVALID, "WFA penalties need x >= 1, e >= 1, o >= 0");
    if (p.x > 4096 || p.o > 4096 || p.e > 4096)
        return set_error(CGM_ERR_INVALID, "WFA penalties above 4096 are not supported");
    return CGM_OK;
}

int check_batch(const Batch &b) {
    if (b.n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (b.n_pairs > (int64_t(1) << 31) - 1) return set_error(CGM_ERR_INVALID, "n_pairs >= 2^31");
    if (b.n_pairs == 0) return CGM_OK;
    if (!b.q_len || !b.r_len || !b.q_off || !b.r_off)
        return set_error(CGM_ERR_INVALID, "null length/offset array");
    if ((b.q_words > 0 && !b.q_packed) || (b.r_words > 0 && !b.r_packed))
        return set_error(CGM_ERR_INVALID, "null packed buffer");
    if (b.q_words < 0 || b.r_words < 0) return set_error(CGM_ERR_INVALID, "negative word count");
    return CGM_OK;
}

namespace {

__global__ void fill_i32_kernel(int32_t *p, int32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void max_i32_kernel(const int32_t *a, int64_t n, int32_t *out) {
    int32_t m = INT32_MIN;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = a[i] > m ? a[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        int32_t o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void max_band_eff_kernel(Batch b, const int32_t *band, int32_t band_scalar,
                                    int32_t *out) {
    int32_t m = INT32_MIN;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < b.n_pairs;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t *q, *r;
        int ql, rl;
        if (!load_pair(b, i, q, r, ql, rl)) continue;
        const int bd = band ? band[i] : band_scalar;
        if (bd < 0) continue;
        const int e = effective_band(bd, ql, rl);
        m = e > m ? e : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        int32_t o = __shfl_down(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace

int device_max_band_eff(const Batch &b, const int32_t *band, int32_t band_scalar,
                        hipStream_t stream, int32_t *result) {
    int32_t *d = nullptr;
    CGM_HIP(hipMallocAsync((void **)&d, sizeof(int32_t), stream));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, d, INT32_MIN);
    int64_t blocks = (b.n_pairs + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(max_band_eff_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, b, band,
                       band_scalar, d);
    int32_t h = INT32_MIN;
    hipError_t e1 = hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, stream);
    hipError_t e2 = hipFreeAsync(d, stream);
    hipError_t e3 = hipStreamSynchronize(stream);
    CGM_HIP(e1);
    CGM_HIP(e2);
    CGM_HIP(e3);
    *result = h;
    return CGM_OK;
}

namespace {
std::mutex g_lds_attr_mu;
std::set<std::pair<int, const void *>> g_lds_attr_done;
}  // namespace

int set_kernel_lds(const void *kernel, int bytes) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_lds_attr_mu);
    if (g_lds_attr_done.count({dev, kernel})) return CGM_OK;
    CGM_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    g_lds_attr_done.insert({dev, kernel});
    return CGM_OK;
}

int device_max_i32(const int32_t *a, int64_t n, hipStream_t stream, int32_t *result) {
    int32_t *d = nullptr;
    CGM_HIP(hipMallocAsync((void **)&d, sizeof(int32_t), stream));
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, stream, d, INT32_MIN);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(max_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a, n, d);
    int32_t h = INT32_MIN;
    hipError_t e1 = hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, stream);
    hipError_t e2 = hipFreeAsync(d, stream);
    hipError_t e3 = hipStreamSynchronize(stream);
    CGM_HIP(e1);
    CGM_HIP(e2);
    CGM_HIP(e3);
    *result = h;
    return CGM_OK;
}

// Forked streams per (device, caller stream): each caller stream gets its
// own auxiliary streams and events, created on its first fork and kept
// until the process ends, so work forked from different streams (several
// plans, several serving slots) never queues behind each other.
namespace {
struct AuxStreams {
    hipStream_t s[kMaxAuxStreams] = {};
    hipEvent_t fork = nullptr, join[kMaxAuxStreams] = {};
};
constexpr int kMaxDevices = 64;
std::mutex g_aux_mu;
std::map<std::pair<int, hipStream_t>, AuxStreams *> g_aux;
}  // namespace

static int aux_for_stream(hipStream_t main, AuxStreams **out) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(CGM_ERR_INVALID, "device index too large");
    std::lock_guard<std::mutex> lk(g_aux_mu);
    AuxStreams *&a = g_aux[{dev, main}];
    if (!a) {
        AuxStreams *n = new AuxStreams();
        CGM_HIP(hipEventCreateWithFlags(&n->fork, hipEventDisableTiming));
        for (int i = 0; i < kMaxAuxStreams; ++i) {
            CGM_HIP(hipStreamCreateWithFlags(&n->s[i], hipStreamNonBlocking));
            CGM_HIP(hipEventCreateWithFlags(&n->join[i], hipEventDisableTiming));
        }
        a = n;
    }
    *out = a;
    return CGM_OK;
}

int fork_streams(hipStream_t main, int n, hipStream_t *aux) {
    AuxStreams *a = nullptr;
    if (int rc = aux_for_stream(main, &a)) return rc;
    CGM_HIP(hipEventRecord(a->fork, main));
    for (int i = 0; i < n && i < kMaxAuxStreams; ++i) {
        CGM_HIP(hipStreamWaitEvent(a->s[i], a->fork, 0));
        aux[i] = a->s[i];
    }
    return CGM_OK;
}

int join_streams(hipStream_t main, int n, hipStream_t *aux) {
    AuxStreams *a = nullptr;
    if (int rc = aux_for_stream(main, &a)) return rc;
    for (int i = 0; i < n && i < kMaxAuxStreams; ++i) {
        CGM_HIP(hipEventRecord(a->join[i], aux[i]));
        CGM_HIP(hipStreamWaitEvent(main, a->join[i], 0));
    }
    return CGM_OK;
}

// The host-pointer entry points run on one library stream per device and
// hold that device's lock for the whole (synchronous) call, so concurrent
// host calls cannot interleave on the stream's cached plan.
namespace {
std::mutex g_host_mu[kMaxDevices];
hipStream_t g_host_stream[kMaxDevices] = {};
}  // namespace

static int host_stream(hipStream_t *s, std::unique_lock<std::mutex> *lock) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(CGM_ERR_INVALID, "device index too large");
    *lock = std::unique_lock<std::mutex>(g_host_mu[dev]);
    if (!g_host_stream[dev])
        CGM_HIP(hipStreamCreateWithFlags(&g_host_stream[dev], hipStreamNonBlocking));
    *s = g_host_stream[dev];
    return CGM_OK;
}

// A device copy of a host batch (for the host-pointer entry points).
struct DeviceBatch {
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    ~DeviceBatch() {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <typename T>
    int upload(const T *src, int64_t count, T **dst) {
        *dst = nullptr;
        if (!src || count <= 0) return CGM_OK;
        void *p = nullptr;
        CGM_HIP(hipMalloc(&p, (size_t)count * sizeof(T)));
        allocs.push_back(p);
        CGM_HIP(hipMemcpyAsync(p, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice, stream));
        *dst = (T *)p;
        return CGM_OK;
    }
    template <typename T>
    int alloc(int64_t count, T **dst) {
        void *p = nullptr;
        CGM_HIP(hipMalloc(&p, (size_t)(count > 0 ? count : 1) * sizeof(T)));
        allocs.push_back(p);
        *dst = (T *)p;
        return CGM_OK;
    }
};

}  // namespace cgm

using namespace cgm;

extern "C" {

const char *cgm_last_error(void) { return g_last_error.c_str(); }

int cgm_version(void) { return 100; }

int cgm_device_count(int *count) {
    if (!count) return set_error(CGM_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) n = 0, e = hipSuccess;
    CGM_HIP(e);
    *count = n;
    return CGM_OK;
}

int cgm_set_device(int device) {
    CGM_HIP(hipSetDevice(device));
    return CGM_OK;
}

int cgm_device_info(int device, char *name, int name_len, int *compute_units, int *max_clock_khz,
                    int64_t *hbm_bytes, int *lds_bytes_per_cu) {
    hipDeviceProp_t prop;
    CGM_HIP(hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (max_clock_khz) *max_clock_khz = prop.clockRate;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    if (lds_bytes_per_cu) *lds_bytes_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    return CGM_OK;
}

int cgm_malloc(void **dptr, int64_t bytes) {
    if (!dptr || bytes < 0) return set_error(CGM_ERR_INVALID, "bad malloc arguments");
    CGM_HIP(hipMalloc(dptr, (size_t)(bytes > 0 ? bytes : 1)));
    return CGM_OK;
}

int cgm_free(void *dptr) {
    CGM_HIP(hipFree(dptr));
    return CGM_OK;
}

int cgm_host_malloc(void **hptr, int64_t bytes) {
    if (!hptr || bytes < 0) return set_error(CGM_ERR_INVALID, "bad host malloc arguments");
    CGM_HIP(hipHostMalloc(hptr, (size_t)(bytes > 0 ? bytes : 1), hipHostMallocDefault));
    return CGM_OK;
}

int cgm_host_free(void *hptr) {
    CGM_HIP(hipHostFree(hptr));
    return CGM_OK;
}

int cgm_memcpy_htod(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return CGM_OK;
    CGM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return CGM_OK;
}

int cgm_memcpy_dtoh(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return CGM_OK;
    CGM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return CGM_OK;
}

int cgm_memcpy_dtod(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes <= 0) return CGM_OK;
    CGM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return CGM_OK;
}

int cgm_stream_create(void **stream) {
    if (!stream) return set_error(CGM_ERR_INVALID, "null stream pointer");
    hipStream_t s;
    CGM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return CGM_OK;
}

int cgm_stream_destroy(void *stream) {
    CGM_HIP(hipStreamDestroy((hipStream_t)stream));
    return CGM_OK;
}

int cgm_stream_synchronize(void *stream) {
    CGM_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CGM_OK;
}

int cgm_device_synchronize(void) {
    CGM_HIP(hipDeviceSynchronize());
    return CGM_OK;
}

int cgm_event_create(void **event) {
    if (!event) return set_error(CGM_ERR_INVALID, "null event pointer");
    hipEvent_t e;
    CGM_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return CGM_OK;
}

int cgm_event_destroy(void *event) {
    CGM_HIP(hipEventDestroy((hipEvent_t)event));
    return CGM_OK;
}

int cgm_event_record(void *event, void *stream) {
    CGM_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return CGM_OK;
}

int cgm_event_synchronize(void *event) {
    CGM_HIP(hipEventSynchronize((hipEvent_t)event));
    return CGM_OK;
}

int cgm_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms) return set_error(CGM_ERR_INVALID, "null ms pointer");
    CGM_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return CGM_OK;
}

int cgm_sw_banded_launch(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                         const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                         int64_t n_pairs, int64_t q_words, int64_t r_words, const int32_t *band,
                         int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                         int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                         int32_t *out_r_end, int32_t layout, void *stream) {
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    if (!band && band_scalar < 0) return set_error(CGM_ERR_INVALID, "band_scalar < 0");
    switch (layout) {
        case CGM_LAYOUT_PLAIN:
            return launch_sw_plain(b, band, band_scalar, sc, out, (hipStream_t)stream);
        case CGM_LAYOUT_AUTO:
            return launch_sw_auto(b, band, band_scalar, sc, out, (hipStream_t)stream);
        case CGM_LAYOUT_LANE:
            return launch_sw_buckets(b, band, band_scalar, sc, out, (hipStream_t)stream);
        case CGM_LAYOUT_WAVE:
            return launch_sw_wave(b, band, band_scalar, sc, out, (hipStream_t)stream);
        default:
            return set_error(CGM_ERR_INVALID, "unknown SW layout");
    }
}

int cgm_wfa_affine_launch(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                          const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                          int64_t n_pairs, int64_t q_words, int64_t r_words,
                          const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                          int32_t e, int32_t *out_cost, int32_t layout, void *stream) {
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    WfaPen p = {x, o, e};
    if (n_pairs > 0 && !out_cost) return set_error(CGM_ERR_INVALID, "null output array");
    if (!max_error && max_error_scalar < 0)
        return set_error(CGM_ERR_INVALID, "max_error_scalar < 0");
    switch (layout) {
        case CGM_LAYOUT_PLAIN:
            return launch_wfa_plain(b, max_error, max_error_scalar, p, out_cost,
                                    (hipStream_t)stream);
        case CGM_LAYOUT_AUTO:
            return launch_wfa_wave(b, max_error, max_error_scalar, p, out_cost, true,
                                   (hipStream_t)stream);
        case CGM_LAYOUT_WAVE:
            return launch_wfa_wave(b, max_error, max_error_scalar, p, out_cost, false,
                                   (hipStream_t)stream);
        default:
            return set_error(CGM_ERR_INVALID, "unknown WFA layout");
    }
}

// SequenceBatch-shaped forms: each side's SoA arrays together, NMask words
// accepted and never read (the reference's decode ignores NMask,
// sequence.mojo:96-101, and align_cpu.mojo never consults it).
int cgm_sw_banded_batch(const uint64_t *q_packed, const uint64_t *q_nmask, const int32_t *q_off,
                        const int32_t *q_len, int64_t q_words, const uint64_t *r_packed,
                        const uint64_t *r_nmask, const int32_t *r_off, const int32_t *r_len,
                        int64_t r_words, int64_t n_pairs, const int32_t *band,
                        int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                        int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                        int32_t *out_r_end, void *stream) {
    (void)q_nmask;
    (void)r_nmask;
    return cgm_sw_banded_launch(q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words,
                                r_words, band, band_scalar, match, mismatch, gap_open, gap_extend,
                                out_score, out_q_end, out_r_end, CGM_LAYOUT_AUTO, stream);
}

int cgm_wfa_affine_batch(const uint64_t *q_packed, const uint64_t *q_nmask, const int32_t *q_off,
                         const int32_t *q_len, int64_t q_words, const uint64_t *r_packed,
                         const uint64_t *r_nmask, const int32_t *r_off, const int32_t *r_len,
                         int64_t r_words, int64_t n_pairs, const int32_t *max_error,
                         int32_t max_error_scalar, int32_t x, int32_t o, int32_t e,
                         int32_t *out_cost, void *stream) {
    (void)q_nmask;
    (void)r_nmask;
    return cgm_wfa_affine_launch(q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words,
                                 r_words, max_error, max_error_scalar, x, o, e, out_cost,
                                 CGM_LAYOUT_AUTO, stream);
}

int cgm_plan_create(int64_t max_pairs, int32_t max_error_cap, int32_t x, int32_t o, int32_t e,
                    int32_t waves_per_cu, int32_t lds_bytes, void **plan) {
    if (!plan) return set_error(CGM_ERR_INVALID, "null plan pointer");
    MixPlan *P = nullptr;
    int rc = mix_plan_create(max_pairs, max_error_cap, WfaPen{x, o, e}, waves_per_cu, lds_bytes,
                             &P);
    if (rc) return rc;
    *plan = P;
    return CGM_OK;
}

int cgm_plan_destroy(void *plan) {
    mix_plan_destroy((MixPlan *)plan);
    return CGM_OK;
}

int cgm_align_mixed(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                    const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                    const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                    const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                    int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                    int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                    int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                    int32_t *out_r_end, int32_t layout, void *stream) {
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    WfaPen p = {x, o, e};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    if (layout == CGM_LAYOUT_PLAIN) {
        if (int rc = check_batch(b)) return rc;
        if (int rc = check_sw_scoring(sc)) return rc;
        if (int rc = check_wfa_pen(p)) return rc;
        if (!kind) {
            if (kind_scalar == 0)
                return launch_sw_plain(b, band, band_scalar, sc, out, (hipStream_t)stream);
            return launch_wfa_plain_kind(b, max_error, max_error_scalar, p, out_score, nullptr,
                                         out_q_end, out_r_end, (hipStream_t)stream);
        }
        if (int rc = launch_sw_plain_kind(b, band, band_scalar, sc, out, kind, (hipStream_t)stream))
            return rc;
        return launch_wfa_plain_kind(b, max_error, max_error_scalar, p, out_score, kind, out_q_end,
                                     out_r_end, (hipStream_t)stream);
    }
    if (!plan) return set_error(CGM_ERR_INVALID, "the automatic mixed layout needs a plan");
    return mix_enqueue((MixPlan *)plan, b, kind, kind_scalar, band, band_scalar, max_error,
                       max_error_scalar, sc, p, out, (hipStream_t)stream);
}

int cgm_plan_capture(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                     const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                     const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                     const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                     int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                     int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                     int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                     int32_t *out_r_end, void *stream) {
    if (!plan) return set_error(CGM_ERR_INVALID, "null plan");
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    return mix_capture((MixPlan *)plan, b, kind, kind_scalar, band, band_scalar, max_error,
                       max_error_scalar, sc, WfaPen{x, o, e}, out, (hipStream_t)stream);
}

int cgm_plan_kernel_ms(void *plan, float *ms) {
    if (!plan || !ms) return set_error(CGM_ERR_INVALID, "null argument");
    return mix_plan_kernel_ms((MixPlan *)plan, ms);
}

int cgm_plan_wfa_kernel_ms(void *plan, float *ms) {
    if (!plan || !ms) return set_error(CGM_ERR_INVALID, "null argument");
    return mix_plan_wfa_ms((MixPlan *)plan, ms);
}

int cgm_plan_replay(void *plan, void *stream) {
    if (!plan) return set_error(CGM_ERR_INVALID, "null plan");
    return mix_replay((MixPlan *)plan, (hipStream_t)stream);
}

int cgm_plan_get_tuning(void *plan, int32_t *params) {
    if (!plan || !params) return set_error(CGM_ERR_INVALID, "null argument");
    tuning_to_array(mix_plan_tuning((MixPlan *)plan), params);
    return CGM_OK;
}

int cgm_plan_set_tuning(void *plan, const int32_t *params) {
    if (!plan || !params) return set_error(CGM_ERR_INVALID, "null argument");
    MixTuning t;
    if (int rc = tuning_from_array(params, &t)) return rc;
    mix_plan_set_tuning((MixPlan *)plan, t);
    return CGM_OK;
}

int cgm_tuning_get_default(int32_t *params) {
    if (!params) return set_error(CGM_ERR_INVALID, "null argument");
    tuning_to_array(mix_default_tuning(), params);
    return CGM_OK;
}

int cgm_tuning_set_default(const int32_t *params) {
    if (!params) return set_error(CGM_ERR_INVALID, "null argument");
    MixTuning t;
    if (int rc = tuning_from_array(params, &t)) return rc;
    mix_set_default_tuning(t);
    return CGM_OK;
}

int cgm_plan_autotune(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                      const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                      const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                      const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                      int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                      int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                      int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                      int32_t *out_r_end, int32_t reps, int32_t *best_params, float *best_ms,
                      float *start_ms, void *stream) {
    if (!plan || !best_params) return set_error(CGM_ERR_INVALID, "null argument");
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    MixTuning best;
    if (int rc = mix_autotune((MixPlan *)plan, b, kind, kind_scalar, band, band_scalar, max_error,
                              max_error_scalar, sc, WfaPen{x, o, e}, out, reps,
                              (hipStream_t)stream, &best, best_ms, start_ms))
        return rc;
    tuning_to_array(best, best_params);
    return CGM_OK;
}

int cgm_plan_autotune_cached(void *plan, const char *cache_path, const uint64_t *q_packed,
                             const uint64_t *r_packed, const int32_t *q_len, const int32_t *r_len,
                             const int32_t *q_off, const int32_t *r_off, int64_t n_pairs,
                             int64_t q_words, int64_t r_words, const int32_t *kind,
                             int32_t kind_scalar, const int32_t *band, int32_t band_scalar,
                             const int32_t *max_error, int32_t max_error_scalar, int32_t match,
                             int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t x,
                             int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                             int32_t *out_r_end, int32_t reps, int32_t *best_params,
                             float *best_ms, float *start_ms, int32_t *from_cache, void *stream) {
    if (!plan || !best_params || !cache_path) return set_error(CGM_ERR_INVALID, "null argument");
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    SwScoring sc = {match, mismatch, gap_open, gap_extend};
    SwOut out = {out_score, out_q_end, out_r_end};
    if (n_pairs > 0 && (!out_score || !out_q_end || !out_r_end))
        return set_error(CGM_ERR_INVALID, "null output array");
    if (int rc = check_sw_scoring(sc)) return rc;
    MixTuning best;
    bool hit = false;
    if (int rc = mix_autotune_cached((MixPlan *)plan, cache_path, b, kind, kind_scalar, band,
                                     band_scalar, max_error, max_error_scalar, sc, WfaPen{x, o, e},
                                     out, reps, (hipStream_t)stream, &best, best_ms, start_ms, &hit))
        return rc;
    tuning_to_array(best, best_params);
    if (from_cache) *from_cache = hit ? 1 : 0;
    return CGM_OK;
}

int cgm_plan_profile_key(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                         const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                         int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                         int32_t gap_extend, int32_t x, int32_t o, int32_t e, char *key,
                         int32_t key_len, void *stream) {
    if (!plan || !key || key_len < 1) return set_error(CGM_ERR_INVALID, "null argument");
    Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    std::string k;
    if (int rc = mix_profile_key((MixPlan *)plan, b, kind, kind_scalar, band, band_scalar,
                                 SwScoring{match, mismatch, gap_open, gap_extend}, WfaPen{x, o, e},
                                 (hipStream_t)stream, &k))
        return rc;
    if ((int64_t)k.size() + 1 > key_len) return set_error(CGM_ERR_INVALID, "key buffer too small");
    memcpy(key, k.c_str(), k.size() + 1);
    return CGM_OK;
}

int cgm_sw_banded_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                       const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words, const int32_t *band,
                       int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                       int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                       int32_t *out_r_end, int32_t layout) {
    if (n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (n_pairs == 0) return CGM_OK;
    DeviceBatch db;
    std::unique_lock<std::mutex> host_lock;
    if (int rc = host_stream(&db.stream, &host_lock)) return rc;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamSynchronize(s); }
    } guard{db.stream};
    uint64_t *qp, *rp;
    int32_t *ql, *rl, *qo, *ro, *bd = nullptr, *os, *oq, *orr;
    int rc = CGM_OK;
    if ((rc = db.upload(q_packed, q_words, &qp)) || (rc = db.upload(r_packed, r_words, &rp)) ||
        (rc = db.upload(q_len, n_pairs, &ql)) || (rc = db.upload(r_len, n_pairs, &rl)) ||
        (rc = db.upload(q_off, n_pairs, &qo)) || (rc = db.upload(r_off, n_pairs, &ro)) ||
        (rc = db.upload(band, band ? n_pairs : 0, &bd)) || (rc = db.alloc(n_pairs, &os)) ||
        (rc = db.alloc(n_pairs, &oq)) || (rc = db.alloc(n_pairs, &orr)))
        return rc;
    rc = cgm_sw_banded_launch(qp, rp, ql, rl, qo, ro, n_pairs, q_words, r_words, bd, band_scalar,
                              match, mismatch, gap_open, gap_extend, os, oq, orr, layout, db.stream);
    if (rc) {
        (void)hipStreamSynchronize(db.stream);
        return rc;
    }
    CGM_HIP(hipMemcpyAsync(out_score, os, n_pairs * 4, hipMemcpyDeviceToHost, db.stream));
    CGM_HIP(hipMemcpyAsync(out_q_end, oq, n_pairs * 4, hipMemcpyDeviceToHost, db.stream));
    CGM_HIP(hipMemcpyAsync(out_r_end, orr, n_pairs * 4, hipMemcpyDeviceToHost, db.stream));
    CGM_HIP(hipStreamSynchronize(db.stream));
    return CGM_OK;
}

int cgm_sw_banded_start_host(const uint64_t *q_packed, const uint64_t *r_packed,
                             const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                             const int32_t *r_off, int64_t n_pairs, int64_t q_words,
                             int64_t r_words, const int32_t *band, int32_t band_scalar,
                             int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                             int32_t *out_score, int32_t *out_q_start, int32_t *out_r_start,
                             int32_t *out_q_end, int32_t *out_r_end, int32_t layout) {
    if (n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (n_pairs == 0) return CGM_OK;
    if (!out_score || !out_q_start || !out_r_start || !out_q_end || !out_r_end)
        return set_error(CGM_ERR_INVALID, "null output array");
    DeviceBatch db;
    std::unique_lock<std::mutex> host_lock;
    if (int rc = host_stream(&db.stream, &host_lock)) return rc;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamSynchronize(s); }
    } guard{db.stream};
    uint64_t *qp, *rp;
    int32_t *ql, *rl, *qo, *ro, *bd = nullptr, *o[5];
    int rc = CGM_OK;
    if ((rc = db.upload(q_packed, q_words, &qp)) || (rc = db.upload(r_packed, r_words, &rp)) ||
        (rc = db.upload(q_len, n_pairs, &ql)) || (rc = db.upload(r_len, n_pairs, &rl)) ||
        (rc = db.upload(q_off, n_pairs, &qo)) || (rc = db.upload(r_off, n_pairs, &ro)) ||
        (rc = db.upload(band, band ? n_pairs : 0, &bd)))
        return rc;
    for (int k = 0; k < 5; ++k)
        if ((rc = db.alloc(n_pairs, &o[k]))) return rc;
    rc = cgm_sw_banded_start_launch(qp, rp, ql, rl, qo, ro, n_pairs, q_words, r_words, bd,
                                    band_scalar, match, mismatch, gap_open, gap_extend, o[0], o[1],
                                    o[2], o[3], o[4], layout, db.stream);
    if (rc) {
        (void)hipStreamSynchronize(db.stream);
        return rc;
    }
    int32_t *host[5] = {out_score, out_q_start, out_r_start, out_q_end, out_r_end};
    for (int k = 0; k < 5; ++k)
        CGM_HIP(hipMemcpyAsync(host[k], o[k], n_pairs * 4, hipMemcpyDeviceToHost, db.stream));
    CGM_HIP(hipStreamSynchronize(db.stream));
    return CGM_OK;
}

int cgm_wfa_affine_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                        const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                        int64_t n_pairs, int64_t q_words, int64_t r_words,
                        const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                        int32_t e, int32_t *out_cost, int32_t layout) {
    if (n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (n_pairs == 0) return CGM_OK;
    DeviceBatch db;
    std::unique_lock<std::mutex> host_lock;
    if (int rc = host_stream(&db.stream, &host_lock)) return rc;
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamSynchronize(s); }
    } guard{db.stream};
    uint64_t *qp, *rp;
    int32_t *ql, *rl, *qo, *ro, *me = nullptr, *oc;
    int rc = CGM_OK;
    if ((rc = db.upload(q_packed, q_words, &qp)) || (rc = db.upload(r_packed, r_words, &rp)) ||
        (rc = db.upload(q_len, n_pairs, &ql)) || (rc = db.upload(r_len, n_pairs, &rl)) ||
        (rc = db.upload(q_off, n_pairs, &qo)) || (rc = db.upload(r_off, n_pairs, &ro)) ||
        (rc = db.upload(max_error, max_error ? n_pairs : 0, &me)) || (rc = db.alloc(n_pairs, &oc)))
        return rc;
    rc = cgm_wfa_affine_launch(qp, rp, ql, rl, qo, ro, n_pairs, q_words, r_words, me,
                               max_error_scalar, x, o, e, oc, layout, db.stream);
    if (rc) {
        (void)hipStreamSynchronize(db.stream);
        return rc;
    }
    CGM_HIP(hipMemcpyAsync(out_cost, oc, n_pairs * 4, hipMemcpyDeviceToHost, db.stream));
    CGM_HIP(hipStreamSynchronize(db.stream));
    return CGM_OK;
}

}  // extern "C"
