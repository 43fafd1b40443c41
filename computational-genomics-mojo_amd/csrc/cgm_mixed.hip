// cgm_mixed.hip -- mixed SW + WFA batches: one persistent kernel pulling a
// work-sorted task queue, with every buffer preallocated in a plan so the
// whole enqueue (classify, sort, task build, persistent launch, fallbacks)
// performs no allocation and no host synchronisation and can be captured
// into a hipGraph and replayed.
//
// Pipeline (all device work; WFA pairs on a forked stream):
//  1. classify each pair (kind, effective band, lengths, max_error) into
//     0 SW register lanes (band <= reg_max_band <= 64), 1 SW packed register
//     lanes (band 128), 2 WFA, 3 SW LDS lanes, 4 SW one workgroup per pair,
//     5 bad, 6 SW one wavefront per pair (pairs longer than lane_max_len
//     with bands <= 64; longer ones with wider bands go to class 4); WFA
//     pairs also get the WFA launcher's layout class (register layouts by
//     width, team layouts beyond); bad pairs get their outputs written here,
//     WFA pairs their ends (-1);
//  2. radix-sort pair indices by (class, band, length) or (class, WFA
//     layout class, WFA work);
//  3. the WFA pairs run the WFA launcher's class kernels over their ranges
//     of the order, longest first, on two streams forked from the caller's
//     (its kernels keep their own register and LDS budgets);
//  4. build SW tasks: 256 consecutive SW pairs of one register class per
//     task, each with an estimated cost, and radix-sort them largest first;
//  5. the SW persistent kernel: CUs x blocks_per_cu four-wavefront
//     workgroups whose wavefronts first take the class-6 pairs one
//     wavefront each, longest first, then 64-pair units from a long and a
//     short queue; it is launched after the WFA kernels so that its
//     workgroups fill the CUs the WFA tail leaves idle;
//  6. classes 3 and 4 (unusual scorings, bands or lengths) run the LDS-lane
//     and one-workgroup-per-pair kernels over their ranges;
//  7. the caller's stream joins the forked one.
// The class limits and the kernel's shape are the plan's MixTuning.

#include <stdio.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <string>

#include <hipcub/hipcub.hpp>

#include "cgm_mixed_dev.h"
#include "cgm_wfa_reg.h"

namespace cgm {

namespace {

// Classification, one pair per thread in a grid-stride loop over a grid of
// at most one workgroup per CU: class counts and register-band counts are
// gathered in LDS and added to the global counters once per workgroup (a
// global atomic per wavefront and class serialised on a handful of
// addresses and dominated this kernel).
__global__ void mix_class_kernel(Batch b, const int32_t *kind, int32_t kind_scalar,
                                 const int32_t *band, int32_t band_scalar, const int32_t *me_arr,
                                 int32_t me_scalar, int32_t me_cap, SwScoring sc, WfaPen p,
                                 wfadev::WfaGeom gm, int32_t reg_pen, MixTuning tune, SwOut out,
                                 uint32_t *keys, int32_t *idx, int32_t *counts,
                                 int32_t *band_counts, int32_t *wfa_counts) {
    __shared__ int32_t hist[kRegBands];
    __shared__ int32_t cls_count[kMixClasses];
    for (int j = threadIdx.x; j < kRegBands; j += blockDim.x) hist[j] = 0;
    if (threadIdx.x < kMixClasses) cls_count[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n_round = (b.n_pairs + stride - 1) / stride * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_round; i += stride) {
        const bool in = i < b.n_pairs;
        const uint64_t *q, *r;
        int ql = 0, rl = 0;
        int cls = in ? kMixBad : -1;
        int wcls = -1;
        uint32_t payload = 0;
        if (in && load_pair(b, i, q, r, ql, rl)) {
            const int k = kind ? kind[i] : kind_scalar;
            if (k == 0) {
                const int bd = band ? band[i] : band_scalar;
                if (bd >= 0) {
                    const int be = effective_band(bd, ql, rl);
                    const int64_t hmax = (int64_t)max(sc.match, 0) * min(ql, rl);
                    const int oe = sc.gap_open + sc.gap_extend;
                    const bool reg_scoring = sc.mismatch <= 0 && sc.gap_extend <= 0 && oe <= 0;
                    const bool long_pair = max(ql, rl) > tune.lane_max_len;
                    if (hmax > kMaxScoreProduct) {
                        cls = kMixBad;
                    } else if (long_pair && be <= kWaveMaxBand && hmax < (1 << 23) &&
                               max(ql, rl) < kWaveMaxLen) {
                        cls = kMixSwWave;
                    } else if (long_pair && be <= kPlainMaxBand) {
                        cls = kMixSwPlain;
                    } else if (reg_scoring && be <= tune.reg_max_band && hmax < (1 << 23)) {
                        cls = kMixSwReg;
                    } else if (reg_scoring && be == kPackedBand && hmax + oe <= 32767 &&
                               oe >= -32768) {
                        cls = kMixSwPacked;
                    } else if (be <= kLdsLaneMaxBand) {
                        cls = kMixSwLds;
                    } else if (be <= kPlainMaxBand) {
                        cls = kMixSwPlain;
                    }
                    payload = cls == kMixSwWave
                                  ? (uint32_t)min(ql + rl, (1 << 28) - 1)  // longest first
                                  : ((uint32_t)min(be, 2047) << 17) |
                                        (uint32_t)min(max(ql, rl), (1 << 17) - 1);
                    if (cls == kMixSwReg) atomicAdd(&hist[be], 1);
                }
            } else if (k == 1) {
                const int me = me_arr ? me_arr[i] : me_scalar;
                if (me >= 0 && me <= me_cap) {
                    cls = kMixWfa;
                    wcls = wfareg::wfa_pair_class(ql, rl, me, p, gm, reg_pen != 0);
                    payload = wfareg::wfa_class_key(wcls, wfadev::wfa_work_key(me, p, gm));
                }
            }
        }
        for (int c = 0; c < kMixClasses; ++c) {
            const uint64_t m = __ballot(cls == c);
            if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&cls_count[c], __popcll(m));
        }
        if (__ballot(wcls >= 0)) {
            for (int c = 0; c < wfareg::kWfaClasses; ++c) {
                const uint64_t m = __ballot(wcls == c);
                if (m && lane == __ffsll((unsigned long long)m) - 1)
                    atomicAdd(&wfa_counts[c], __popcll(m));
            }
        }
        if (in) {
            if (cls == kMixBad || cls == kMixWfa) {
                if (cls == kMixBad) out.score[i] = CGM_BAD_PAIR;
                if (out.q_end) out.q_end[i] = -1;
                if (out.r_end) out.r_end[i] = -1;
            }
            keys[i] = ((uint32_t)(7 - cls) << 28) | (payload & 0x0fffffffu);
            idx[i] = (int32_t)i;
        }
    }
    __syncthreads();
    if (threadIdx.x < kMixClasses && cls_count[threadIdx.x])
        atomicAdd(&counts[threadIdx.x], cls_count[threadIdx.x]);
    for (int j = threadIdx.x; j < kRegBands; j += blockDim.x)
        if (hist[j]) atomicAdd(&band_counts[j], hist[j]);
}

// Class ranges in the sorted order and the task counts.  Register-lane pairs
// are sorted widest band first, so band b's pairs follow those of every
// wider band; each band gets ceil(count / kMixGroup) tasks of its own.
__global__ void mix_ranges_kernel(int32_t *small) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t *counts = small + kSmallCounts;
    int32_t *ranges = small + kSmallRanges;
    int32_t s = 0;
    for (int c = 0; c < kMixClasses; ++c) {
        ranges[2 * c] = s;
        s += counts[c];
        ranges[2 * c + 1] = s;
    }
    int32_t tasks = 0, pairs = 0;
    for (int bd = kRegBands - 1; bd >= 0; --bd) {
        const int32_t c = small[kSmallBandCount + bd];
        small[kSmallBandTask + bd] = tasks;
        small[kSmallBandPair + bd] = pairs;
        tasks += (c + kMixGroup - 1) / kMixGroup;
        pairs += c;
    }
    small[kSmallBandTask + kRegBands] = tasks;
    small[kSmallRegTasks] = tasks;
    small[kSmallNTasks] = tasks + (counts[kMixSwPacked] + kMixGroup - 1) / kMixGroup;
    // WFA layout classes inside the WFA class's range (sorted class first)
    int32_t w = ranges[2 * kMixWfa];
    for (int c = 0; c < wfareg::kWfaClasses; ++c) {
        small[kSmallWfaRanges + 2 * c] = w;
        w += small[kSmallWfaCounts + c];
        small[kSmallWfaRanges + 2 * c + 1] = w;
    }
}

__global__ void mix_tasks_kernel(const int32_t *small, const uint32_t *sorted_keys, int64_t cap,
                                 Task *tasks, uint32_t *tkeys, int32_t *tidx) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cap) return;
    const int32_t *ranges = small + kSmallRanges;
    const int32_t c1 = ranges[2 * kMixSwPacked + 1] - ranges[2 * kMixSwPacked];
    const int64_t n0 = small[kSmallRegTasks];
    Task task = {0, 0, 0, 0};
    uint32_t key = 0;
    if (t < small[kSmallNTasks]) {
        if (t < n0) {
            // band of task t: the last band (widest first) whose first task is <= t
            int lo = 0, hi = kRegBands - 1;  // band task starts decrease with the band
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (small[kSmallBandTask + mid] <= t) hi = mid;
                else lo = mid + 1;
            }
            const int bd = lo;
            const int64_t g = t - small[kSmallBandTask + bd];
            task.type = kTaskSwReg;
            task.start = ranges[2 * kMixSwReg] + small[kSmallBandPair + bd] +
                         (int32_t)(kMixGroup * g);
            task.count = (int32_t)min((int64_t)kMixGroup,
                                      small[kSmallBandCount + bd] - kMixGroup * g);
        } else {
            const int64_t g = t - n0;
            task.type = kTaskSwPacked;
            task.start = ranges[2 * kMixSwPacked] + (int32_t)(kMixGroup * g);
            task.count = (int32_t)min((int64_t)kMixGroup, c1 - kMixGroup * g);
        }
        const uint32_t payload = sorted_keys[task.start] & 0x0fffffffu;
        const uint32_t band = payload >> 17, len = payload & 0x1ffffu;
        // ~10 VALU ops x 4 cycles per wave-step of 64 cells, units of 16 cycles
        key = (uint32_t)min((uint64_t)len * (2 * band + 1) * 3 + 1, (uint64_t)0xffffffffu);
    }
    tasks[t] = task;
    tkeys[t] = key;
    tidx[t] = (int32_t)t;
}

__global__ void mix_zero_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Length-bin profile of a batch: pairs per (kind, floor(log2(max(qlen,
// rlen))), SW band bin), counted in LDS per workgroup.  Band bins of the
// effective band: 0-16, 17-64, 65-128, above (WFA pairs: bin 0); bad pairs
// are not counted.
constexpr int kProfileLenBins = 24, kProfileBandBins = 4;
constexpr int kProfileCells = 2 * kProfileLenBins * kProfileBandBins;
__global__ void mix_profile_kernel(Batch b, const int32_t *kind, int32_t kind_scalar,
                                   const int32_t *band, int32_t band_scalar, int32_t *hist) {
    __shared__ int32_t h[kProfileCells];
    for (int j = threadIdx.x; j < kProfileCells; j += blockDim.x) h[j] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < b.n_pairs; i += stride) {
        const uint64_t *q, *r;
        int ql = 0, rl = 0;
        if (!load_pair(b, i, q, r, ql, rl)) continue;
        const int k = kind ? kind[i] : kind_scalar;
        if (k != 0 && k != 1) continue;
        const int len = max(max(ql, rl), 1);
        const int lb = min(31 - __clz(len), kProfileLenBins - 1);
        int bb = 0;
        if (k == 0) {
            const int bd = band ? band[i] : band_scalar;
            if (bd < 0) continue;
            const int be = effective_band(bd, ql, rl);
            bb = be <= 16 ? 0 : be <= 64 ? 1 : be <= 128 ? 2 : 3;
        }
        atomicAdd(&h[(k * kProfileLenBins + lb) * kProfileBandBins + bb], 1);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kProfileCells; j += blockDim.x)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

}  // namespace

struct MixPlan {
    int device = 0;
    int64_t cap = 0;        // max pairs per call
    int64_t task_cap = 0;   // max tasks
    int32_t me_cap = 0;
    WfaPen pen{};
    wfadev::WfaGeom gm{};
    int cus = 0;
    int wfa_grid = 0;  // 512-thread WFA workgroups, each with a slab
    MixTuning tune{};
    // workspace
    uint32_t *keys = nullptr, *keys_out = nullptr, *tkeys = nullptr, *tkeys_out = nullptr;
    int32_t *idx = nullptr, *order = nullptr, *tidx = nullptr, *task_order = nullptr;
    int32_t *small = nullptr;  // kSmallWords words, layout above
    Task *tasks = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int32_t *slab = nullptr;
    int64_t slab_words = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t bytes = 0;  // device memory held by the plan
    bool had_wfa = false;  // the last enqueue launched the WFA kernels
    bool fork = true;      // WFA and long-SW kernels on streams forked from the caller's
    // the caller's bounds on the next batch's SW bands and read lengths
    // (-1: unknown); they let the enqueue leave out launches that cannot
    // have work
    int32_t hint_max_band = -1, hint_max_len = -1;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;  // around the SW persistent kernel
    hipEvent_t ev_wfa_start = nullptr, ev_wfa_stop = nullptr;  // around the WFA kernels
    hipEvent_t ev_wfa_mid = nullptr;  // end of the second WFA stream (no timing)
    ~MixPlan() {
        for (hipEvent_t e : {ev_start, ev_stop, ev_wfa_start, ev_wfa_stop, ev_wfa_mid})
            if (e) (void)hipEventDestroy(e);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        void *ptrs[] = {keys, keys_out, tkeys, tkeys_out, idx, order, tidx, task_order, small,
                        tasks, sort_tmp, slab};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

namespace {
std::mutex g_tune_mu;
MixTuning g_default_tuning{};
}  // namespace

void tuning_to_array(const MixTuning &t, int32_t *a) {
    a[0] = t.blocks_per_cu;
    a[1] = t.lds_bytes;
    a[2] = t.tile_steps;
    a[3] = t.wave_max_width;
    a[4] = t.reg_max_band;
    a[5] = t.lane_max_len;
    a[6] = t.sw_dual;
    a[7] = t.sw_split;
}

int tuning_from_array(const int32_t *a, MixTuning *t) {
    MixTuning v;
    v.blocks_per_cu = a[0];
    v.lds_bytes = a[1] & ~255;
    v.tile_steps = a[2];
    v.wave_max_width = a[3];
    v.reg_max_band = a[4];
    v.lane_max_len = a[5];
    v.sw_dual = a[6];
    v.sw_split = a[7];
    if (v.blocks_per_cu < 1 || v.blocks_per_cu > kMixMaxBlocksPerCu)
        return set_error(CGM_ERR_INVALID, "tuning: blocks_per_cu must be 1 or 2");
    // static LDS of the kernel (a few ints) needs room beside the dynamic part
    if (v.lds_bytes < 4096 || (int64_t)v.blocks_per_cu * (v.lds_bytes + 256) > kMixMaxLds)
        return set_error(CGM_ERR_INVALID, "tuning: lds_bytes out of range for blocks_per_cu");
    if (v.tile_steps < 1 || v.tile_steps > 1024)
        return set_error(CGM_ERR_INVALID, "tuning: tile_steps out of range");
    if (v.wave_max_width < 0) return set_error(CGM_ERR_INVALID, "tuning: wave_max_width < 0");
    if (v.reg_max_band < 0 || v.reg_max_band > 64)
        return set_error(CGM_ERR_INVALID, "tuning: reg_max_band out of range");
    if (v.lane_max_len < 1) return set_error(CGM_ERR_INVALID, "tuning: lane_max_len < 1");
    if (v.sw_dual != 0 && v.sw_dual != 1)
        return set_error(CGM_ERR_INVALID, "tuning: sw_dual must be 0 or 1");
    if (v.sw_split < 0 || v.sw_split > 100)
        return set_error(CGM_ERR_INVALID, "tuning: sw_split must be 0..100");
    *t = v;
    return CGM_OK;
}

MixTuning mix_default_tuning() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_default_tuning;
}

void mix_set_default_tuning(const MixTuning &t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_default_tuning = t;
}

MixTuning mix_plan_tuning(const MixPlan *P) { return P->tune; }

void mix_plan_set_tuning(MixPlan *P, const MixTuning &t) { P->tune = t; }
void mix_plan_set_fork(MixPlan *P, bool fork) { P->fork = fork; }
void mix_plan_set_hint(MixPlan *P, int32_t max_band, int32_t max_len) {
    P->hint_max_band = max_band;
    P->hint_max_len = max_len;
}

// waves_per_cu / lds_bytes > 0 override the default tuning's shape.
int mix_plan_create(int64_t max_pairs, int32_t me_cap, WfaPen pen, int waves_per_cu,
                    int lds_bytes, MixPlan **out) {
    if (max_pairs <= 0 || max_pairs > (int64_t(1) << 31) - 1)
        return set_error(CGM_ERR_INVALID, "plan: max_pairs out of range");
    if (me_cap < 0 || me_cap > kMaxWfaError) return set_error(CGM_ERR_INVALID, "plan: bad max_error cap");
    if (int rc = check_wfa_pen(pen)) return rc;
    if (waves_per_cu < 0 || waves_per_cu > 32) return set_error(CGM_ERR_INVALID, "plan: waves_per_cu");
    if (lds_bytes < 0 || lds_bytes > kMixMaxLds) return set_error(CGM_ERR_INVALID, "plan: lds bytes");
    MixTuning tune = mix_default_tuning();
    if (waves_per_cu > 0)
        tune.blocks_per_cu = std::min(kMixMaxBlocksPerCu, (waves_per_cu + kMixWaves - 1) / kMixWaves);
    if (lds_bytes > 0) tune.lds_bytes = std::max(4096, lds_bytes & ~255);
    if ((int64_t)tune.blocks_per_cu * (tune.lds_bytes + 256) > kMixMaxLds) tune.blocks_per_cu = 1;
    if ((int64_t)tune.lds_bytes + 256 > kMixMaxLds) tune.lds_bytes = kMixMaxLds - 256;
    MixPlan *P = new MixPlan();
    P->tune = tune;
    auto fail = [&](int rc) {
        delete P;
        return rc;
    };
    P->cap = max_pairs;
    P->task_cap = max_pairs + (max_pairs + kMixGroup - 1) / kMixGroup + 1;
    P->me_cap = me_cap;
    P->pen = pen;
    wfa_geometry(pen, &P->gm.g, &P->gm.nM, &P->gm.nI);
    if (hipGetDevice(&P->device) != hipSuccess) return fail(set_error(CGM_ERR_HIP, "hipGetDevice"));
    if (int rc = device_cu_count(&P->cus)) return fail(rc);
    const int64_t n = max_pairs, nt = P->task_cap;
#define CGM_ALLOC(ptr, nbytes)                                                     \
    do {                                                                           \
        hipError_t e_ = hipMalloc((void **)&(ptr), (size_t)(nbytes));              \
        if (e_ != hipSuccess) return fail(hip_error(e_, "plan workspace"));        \
        P->bytes += (int64_t)(nbytes);                                             \
    } while (0)
    CGM_ALLOC(P->keys, n * 4);
    CGM_ALLOC(P->keys_out, n * 4);
    CGM_ALLOC(P->idx, n * 4);
    CGM_ALLOC(P->order, n * 4);
    CGM_ALLOC(P->tkeys, nt * 4);
    CGM_ALLOC(P->tkeys_out, nt * 4);
    CGM_ALLOC(P->tidx, nt * 4);
    CGM_ALLOC(P->task_order, nt * 4);
    CGM_ALLOC(P->tasks, nt * sizeof(Task));
    CGM_ALLOC(P->small, kSmallWords * 4);
    size_t b1 = 0, b2 = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b1, P->keys, P->keys_out, P->idx,
                                                     P->order, (int)n, 0, 32) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b2, P->tkeys, P->tkeys_out, P->tidx,
                                                     P->task_order, (int)nt, 0, 32) != hipSuccess)
        return fail(set_error(CGM_ERR_HIP, "plan: sort size query"));
    P->sort_tmp_bytes = b1 > b2 ? b1 : b2;
    CGM_ALLOC(P->sort_tmp, P->sort_tmp_bytes + 256);
    // HBM wavefront slab per 512-thread WFA workgroup (within 4 GiB of
    // slabs) for WFA pairs too large for LDS.
    const int K = me_cap < reach_host(me_cap, pen) ? me_cap : reach_host(me_cap, pen);
    P->slab_words = (int64_t)(P->gm.nM + 2 * P->gm.nI) * (2 * (int64_t)K + 1);
    P->wfa_grid = P->cus * kWfaBlocksPerCu;
    while (P->wfa_grid > 1 && (int64_t)P->wfa_grid * P->slab_words * 4 > (int64_t(4) << 30))
        P->wfa_grid /= 2;
    CGM_ALLOC(P->slab, (int64_t)P->wfa_grid * P->slab_words * 4);
#undef CGM_ALLOC
    if (hipEventCreate(&P->ev_start) != hipSuccess || hipEventCreate(&P->ev_stop) != hipSuccess ||
        hipEventCreate(&P->ev_wfa_start) != hipSuccess ||
        hipEventCreate(&P->ev_wfa_stop) != hipSuccess ||
        hipEventCreateWithFlags(&P->ev_wfa_mid, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(CGM_ERR_HIP, "plan: timing events"));
    *out = P;
    return CGM_OK;
}

void mix_plan_destroy(MixPlan *P) { delete P; }

int64_t mix_plan_bytes(const MixPlan *P) { return P->bytes; }

int mix_plan_kernel_ms(MixPlan *P, float *ms) {
    CGM_HIP(hipEventSynchronize(P->ev_stop));
    CGM_HIP(hipEventElapsedTime(ms, P->ev_start, P->ev_stop));
    return CGM_OK;
}

int mix_plan_wfa_ms(MixPlan *P, float *ms) {
    if (!P->had_wfa) {
        *ms = 0.0f;
        return CGM_OK;
    }
    CGM_HIP(hipEventSynchronize(P->ev_wfa_stop));
    CGM_HIP(hipEventElapsedTime(ms, P->ev_wfa_start, P->ev_wfa_stop));
    return CGM_OK;
}

int64_t mix_plan_capacity(const MixPlan *P, int32_t *me_cap, WfaPen *pen) {
    if (me_cap) *me_cap = P->me_cap;
    if (pen) *pen = P->pen;
    return P->cap;
}

namespace {
struct PlanKey {
    int device;
    hipStream_t stream;
    bool operator<(const PlanKey &o) const {
        return device != o.device ? device < o.device : stream < o.stream;
    }
};
// One library-owned plan per (device, stream); `mu` is held from the lookup
// through the enqueue, so a concurrent call on the same stream that needs a
// larger plan cannot destroy this one while it is in use.
struct PlanSlot {
    std::mutex mu;
    MixPlan *plan = nullptr;
};
std::mutex g_plan_mu;
std::map<PlanKey, PlanSlot> g_plans;
}  // namespace

int with_cached_plan(hipStream_t stream, int64_t n_pairs, int32_t me_cap, WfaPen pen,
                     const std::function<int(MixPlan *)> &use) {
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    PlanSlot *ps = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        ps = &g_plans[PlanKey{dev, stream}];  // map nodes never move
    }
    std::lock_guard<std::mutex> lk(ps->mu);
    MixPlan *&slot = ps->plan;
    if (slot) {
        int32_t cap_me = 0;
        WfaPen cap_pen{};
        const int64_t cap = mix_plan_capacity(slot, &cap_me, &cap_pen);
        if (!(cap >= n_pairs && cap_me >= me_cap && cap_pen.x == pen.x && cap_pen.o == pen.o &&
              cap_pen.e == pen.e)) {
            CGM_HIP(hipStreamSynchronize(stream));
            mix_plan_destroy(slot);
            slot = nullptr;
        }
    }
    if (!slot) {
        // grow geometrically so a slowly growing batch size does not rebuild often
        int64_t cap = 1024;
        while (cap < n_pairs) cap *= 2;
        int rc = mix_plan_create(cap, me_cap, pen, 0, 0, &slot);
        if (rc) {
            slot = nullptr;
            return rc;
        }
    }
    return use(slot);
}

// Enqueues the whole mixed pipeline on `stream` (no allocation, no sync).
int mix_enqueue(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                SwScoring sc, WfaPen pen, SwOut out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_sw_scoring(sc)) return rc;
    if (b.n_pairs > P->cap) return set_error(CGM_ERR_INVALID, "batch larger than the plan");
    if (pen.x != P->pen.x || pen.o != P->pen.o || pen.e != P->pen.e)
        return set_error(CGM_ERR_INVALID, "WFA penalties differ from the plan's");
    if (b.n_pairs == 0) return CGM_OK;
    int dev = 0;
    CGM_HIP(hipGetDevice(&dev));
    if (dev != P->device) return set_error(CGM_ERR_INVALID, "plan belongs to another device");
    int32_t *counts = P->small + kSmallCounts, *ranges = P->small + kSmallRanges,
            *n_tasks = P->small + kSmallNTasks;
    const int64_t n = b.n_pairs, nt = P->task_cap;
    hipLaunchKernelGGL(mix_zero_kernel, dim3(1), dim3(256), 0, stream, P->small, kSmallWords);
    const MixTuning tune = P->tune;
    const unsigned class_grid = (unsigned)std::min<int64_t>((n + 255) / 256, P->cus);
    hipLaunchKernelGGL(mix_class_kernel, dim3(class_grid), dim3(256), 0, stream, b,
                       kind, kind_scalar, band, band_scalar, me, me_scalar, P->me_cap, sc, pen,
                       P->gm, (int32_t)wfa_reg_penalties(pen), tune, out, P->keys, P->idx, counts,
                       P->small + kSmallBandCount, P->small + kSmallWfaCounts);
    hipLaunchKernelGGL(mix_ranges_kernel, dim3(1), dim3(64), 0, stream, P->small);
    size_t tb = P->sort_tmp_bytes;
    CGM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(P->sort_tmp, tb, P->keys, P->keys_out,
                                                         P->idx, P->order, (int)n, 0, 32, stream));
    // WFA pairs: the WFA launcher's class kernels on two forked streams,
    // launched first so that the SW persistent kernel fills the CUs their
    // tails free
    const bool has_wfa = kind != nullptr || kind_scalar == 1;
    hipStream_t aux[2] = {stream, stream};
    const bool forked = has_wfa && P->fork;
    if (forked)
        if (int rc = fork_streams(stream, 2, aux)) return rc;
    if (has_wfa) {
        CGM_HIP(hipEventRecord(P->ev_wfa_start, aux[0]));
        if (int rc = launch_wfa_classes(b, me, me_scalar, P->me_cap, pen, true, out.score, P->order,
                                        P->small + kSmallWfaRanges, P->small + kSmallWfaQueues,
                                        P->slab, P->slab_words, P->wfa_grid, aux[0], aux[1]))
            return rc;
        // the WFA end: both streams done
        CGM_HIP(hipEventRecord(P->ev_wfa_mid, aux[1]));
        CGM_HIP(hipStreamWaitEvent(aux[0], P->ev_wfa_mid, 0));
        CGM_HIP(hipEventRecord(P->ev_wfa_stop, aux[0]));
    }
    P->had_wfa = has_wfa;
    // a batch of WFA pairs only has no SW work: none of the SW kernels is
    // launched (each would still have to be placed on the CUs, behind the
    // WFA kernels of other streams, before finding nothing to do)
    if (kind == nullptr && kind_scalar == 1) {
        CGM_HIP(hipEventRecord(P->ev_start, stream));
        CGM_HIP(hipEventRecord(P->ev_stop, stream));
        return forked ? join_streams(stream, 2, aux) : CGM_OK;
    }
    // SW tasks: at most ceil(n / kMixGroup) + kRegBands + 1 (band-aligned
    // register tasks, packed tasks)
    const int64_t tasks_needed = (n + kMixGroup - 1) / kMixGroup + kRegBands + 1;
    const int64_t tn = tasks_needed < nt ? tasks_needed : nt;
    hipLaunchKernelGGL(mix_tasks_kernel, dim3((unsigned)((tn + 255) / 256)), dim3(256), 0, stream,
                       P->small, P->keys_out, tn, P->tasks, P->tkeys, P->tidx);
    size_t tb2 = P->sort_tmp_bytes;
    CGM_HIP(hipcub::DeviceRadixSort::SortPairsDescending(P->sort_tmp, tb2, P->tkeys, P->tkeys_out,
                                                         P->tidx, P->task_order, (int)tn, 0, 32,
                                                         stream));
    CGM_HIP(hipEventRecord(P->ev_start, stream));
    // SW pairs longer than lane_max_len (bands <= 64) take the
    // wavefront-per-pair phase of the kernel
    const bool has_wave = tune.lane_max_len < kWaveMaxLen && (kind != nullptr || kind_scalar == 0);
    const dim3 grid((unsigned)(P->cus * tune.blocks_per_cu)), block(kMixThreads);
    if (has_wave) {
        if (int rc = launch_mix_persistent_wave(grid.x, stream, b, band, band_scalar, sc, out,
                                                P->order, P->tasks, P->task_order, n_tasks,
                                                P->small, tune.sw_dual, tune.sw_split))
            return rc;
    } else {
        hipLaunchKernelGGL(mix_persistent_kernel<false>, grid, block, 0, stream, b, band,
                           band_scalar, sc, out, P->order, P->tasks, P->task_order, n_tasks,
                           P->small, tune.sw_dual, tune.sw_split);
    }
    CGM_HIP(hipGetLastError());
    CGM_HIP(hipEventRecord(P->ev_stop, stream));
    // uncommon SW classes: LDS lanes, then one workgroup per pair; left out
    // when the caller's bounds send every SW pair of the batch to the
    // register lanes or the wavefront-per-pair phase (the class kernel's
    // rules with the bounds in place of each pair's band and lengths)
    const int oe = sc.gap_open + sc.gap_extend;
    const int32_t hb = P->hint_max_band, hl = P->hint_max_len;
    const bool only_reg_or_wave =
        hb >= 0 && hl >= 0 && sc.mismatch <= 0 && sc.gap_extend <= 0 && oe <= 0 &&
        hb <= tune.reg_max_band && hb <= kWaveMaxBand && hl < kWaveMaxLen &&
        (int64_t)std::max(sc.match, 0) * hl < (1 << 23);
    if (!only_reg_or_wave) {
        if (int rc = launch_sw_lds_lane(b, band, band_scalar, sc, out, P->order,
                                        ranges + 2 * kMixSwLds, P->cus * 2, stream))
            return rc;
        if (int rc = launch_sw_plain_range(b, band, band_scalar, sc, out, P->order,
                                           ranges + 2 * kMixSwPlain, P->cus * 2, stream))
            return rc;
    }
    return forked ? join_streams(stream, 2, aux) : CGM_OK;
}

int mix_enqueue_wave_direct(MixPlan *P, const Batch &b, const int32_t *band, int32_t band_scalar,
                            SwScoring sc, SwOut out, hipStream_t stream) {
    if (b.n_pairs > P->cap) return set_error(CGM_ERR_INVALID, "batch larger than the plan");
    if (b.n_pairs == 0) return CGM_OK;
    int32_t *range = P->small + kSmallRanges + 2 * kMixSwPlain;
    hipLaunchKernelGGL(mix_zero_kernel, dim3(1), dim3(256), 0, stream, P->small, kSmallWords);
    if (int rc = launch_sw_wave_direct(b, band, band_scalar, sc, out, P->order, range, stream))
        return rc;
    return launch_sw_plain_range(b, band, band_scalar, sc, out, P->order, range, P->cus * 2,
                                 stream);
}

int mix_capture(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                SwScoring sc, WfaPen pen, SwOut out, hipStream_t stream) {
    if (!stream) return set_error(CGM_ERR_INVALID, "graph capture needs a non-default stream");
    if (P->exec) {
        (void)hipGraphExecDestroy(P->exec);
        P->exec = nullptr;
    }
    if (P->graph) {
        (void)hipGraphDestroy(P->graph);
        P->graph = nullptr;
    }
    CGM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = mix_enqueue(P, b, kind, kind_scalar, band, band_scalar, me, me_scalar, sc, pen, out,
                         stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    CGM_HIP(e);
    P->graph = g;
    CGM_HIP(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0));
    return CGM_OK;
}

int mix_replay(MixPlan *P, hipStream_t stream) {
    if (!P->exec) return set_error(CGM_ERR_INVALID, "no captured graph in this plan");
    CGM_HIP(hipGraphLaunch(P->exec, stream));
    return CGM_OK;
}

int mix_autotune(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                 const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                 SwScoring sc, WfaPen pen, SwOut out, int reps, hipStream_t stream,
                 MixTuning *best, float *best_ms, float *start_ms) {
    if (reps < 1 || reps > 100) return set_error(CGM_ERR_INVALID, "autotune: reps out of range");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CGM_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(CGM_ERR_HIP, "autotune: events");
    }
    const MixTuning start = P->tune;
    // best of `reps` timed enqueues of the whole pipeline under tuning t
    auto measure = [&](const MixTuning &t, float *ms) -> int {
        P->tune = t;
        // median of the runs: two workgroups per CU vary from run to run
        // (which of two wavefronts sharing a SIMD issues first), and the best
        // run alone would favour the shape with the widest spread
        float runs[100];
        for (int i = 0; i < reps; ++i) {
            CGM_HIP(hipEventRecord(e0, stream));
            if (int rc = mix_enqueue(P, b, kind, kind_scalar, band, band_scalar, me, me_scalar, sc,
                                     pen, out, stream))
                return rc;
            CGM_HIP(hipEventRecord(e1, stream));
            CGM_HIP(hipEventSynchronize(e1));
            float x = 0;
            CGM_HIP(hipEventElapsedTime(&x, e0, e1));
            runs[i] = x;
        }
        std::sort(runs, runs + reps);
        *ms = runs[reps / 2];
        return CGM_OK;
    };
    int rc = CGM_OK;
    MixTuning cur = start;
    float cur_ms = 0;
    // warm-up enqueue (first-touch of the plan's buffers), then the start
    if (!rc) rc = measure(cur, &cur_ms);
    if (!rc) rc = measure(cur, &cur_ms);
    if (!rc && start_ms) *start_ms = cur_ms;
    // Coordinate descent: each field in turn over its candidates, keeping a
    // change only when it is at least 2 % faster.
    auto try_all = [&](const MixTuning *cands, int nc) {
        for (int i = 0; i < nc && !rc; ++i) {
            float ms = 0;
            rc = measure(cands[i], &ms);
            if (!rc && ms < 0.98f * cur_ms) {
                cur = cands[i];
                cur_ms = ms;
            }
        }
    };
    {   // four-wavefront SW workgroups per CU (two waves per SIMD at most)
        MixTuning c = cur;
        c.blocks_per_cu = cur.blocks_per_cu == 1 ? 2 : 1;
        try_all(&c, 1);
    }
    {   // band split between register lanes and LDS lanes
        const int v[] = {32, 48, 64};
        MixTuning c[3];
        int nc = 0;
        for (int x : v)
            if (x != cur.reg_max_band) (c[nc] = cur, c[nc++].reg_max_band = x);
        try_all(c, nc);
    }
    {   // switch-over between one lane and one workgroup per SW pair
        const int v[] = {1024, 4096, 1 << 30};
        MixTuning c[3];
        int nc = 0;
        for (int x : v)
            if (x != cur.lane_max_len) (c[nc] = cur, c[nc++].lane_max_len = x);
        try_all(c, nc);
    }
    {   // register SW lanes: dual-row 16-bit (half precision or int16) or int32
        MixTuning c = cur;
        c.sw_dual = 1 - cur.sw_dual;
        try_all(&c, 1);
    }
    {   // SW-only batches: share of the units in the long queue
        const int v[] = {40, 50, 55, 65};
        MixTuning c[4];
        int nc = 0;
        for (int x : v)
            if (x != cur.sw_split) (c[nc] = cur, c[nc++].sw_split = x);
        try_all(c, nc);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->tune = rc ? start : cur;
    if (rc) return rc;
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return CGM_OK;
}

int mix_profile_key(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                    const int32_t *band, int32_t band_scalar, SwScoring sc, WfaPen pen,
                    hipStream_t stream, std::string *key) {
    if (int rc = check_batch(b)) return rc;
    int32_t hist[kProfileCells];
    hipLaunchKernelGGL(mix_zero_kernel, dim3(1), dim3(256), 0, stream, P->small, kProfileCells);
    if (b.n_pairs > 0) {
        const unsigned grid = (unsigned)std::min<int64_t>((b.n_pairs + 255) / 256, P->cus);
        hipLaunchKernelGGL(mix_profile_kernel, dim3(grid), dim3(256), 0, stream, b, kind,
                           kind_scalar, band, band_scalar, P->small);
    }
    CGM_HIP(hipGetLastError());
    CGM_HIP(hipMemcpyAsync(hist, P->small, sizeof(hist), hipMemcpyDeviceToHost, stream));
    CGM_HIP(hipStreamSynchronize(stream));
    char dev[160] = {0};
    if (int rc = cgm_device_info(P->device, dev, (int)sizeof(dev), nullptr, nullptr, nullptr,
                                 nullptr))
        return rc;
    int64_t total = 0;
    for (int32_t v : hist) total += v;
    int nbin = 0;
    while (nbin < 62 && (int64_t(1) << (nbin + 1)) <= b.n_pairs) ++nbin;
    char buf[256];
    snprintf(buf, sizeof(buf), "dev=%s;sw=%d,%d,%d,%d;wfa=%d,%d,%d;n=2^%d", dev, sc.match,
             sc.mismatch, sc.gap_open, sc.gap_extend, pen.x, pen.o, pen.e, nbin);
    std::string k = buf;
    for (int c = 0; c < kProfileCells; ++c) {
        if (!hist[c]) continue;
        const int share = (int)std::max<int64_t>(1, (hist[c] * 32 + total / 2) / std::max<int64_t>(total, 1));
        const int kd = c / (kProfileLenBins * kProfileBandBins);
        const int lb = (c / kProfileBandBins) % kProfileLenBins, bb = c % kProfileBandBins;
        snprintf(buf, sizeof(buf), ";%s%d.%d=%d", kd ? "wfa" : "sw", lb, bb, share);
        k += buf;
    }
    for (char &ch : k)
        if (ch == '\t' || ch == '\n') ch = ' ';
    *key = k;
    return CGM_OK;
}

namespace {
std::mutex g_tune_cache_mu;
}  // namespace

int mix_autotune_cached(MixPlan *P, const char *cache_path, const Batch &b, const int32_t *kind,
                        int32_t kind_scalar, const int32_t *band, int32_t band_scalar,
                        const int32_t *me, int32_t me_scalar, SwScoring sc, WfaPen pen, SwOut out,
                        int reps, hipStream_t stream, MixTuning *best, float *best_ms,
                        float *start_ms, bool *from_cache) {
    if (!cache_path || !*cache_path) return set_error(CGM_ERR_INVALID, "autotune: no cache path");
    std::string key;
    if (int rc = mix_profile_key(P, b, kind, kind_scalar, band, band_scalar, sc, pen, stream, &key))
        return rc;
    std::lock_guard<std::mutex> lk(g_tune_cache_mu);
    bool found = false;
    MixTuning t{};
    float t_ms = 0;
    if (FILE *f = fopen(cache_path, "r")) {
        std::string line;
        char chunk[1024];
        while (fgets(chunk, sizeof(chunk), f)) {
            line += chunk;
            if (line.empty() || line.back() != '\n') continue;  // long line: keep reading
            line.pop_back();
            const size_t tab = line.find('\t');
            if (tab != std::string::npos && line.compare(0, tab, key) == 0 && tab == key.size()) {
                int32_t a[CGM_TUNE_N];
                float ms = 0;
                if (sscanf(line.c_str() + tab + 1, "%d,%d,%d,%d,%d,%d,%d,%d\t%f", &a[0], &a[1],
                           &a[2], &a[3], &a[4], &a[5], &a[6], &a[7], &ms) == 9) {
                    MixTuning v;
                    if (tuning_from_array(a, &v) == CGM_OK) {
                        t = v;
                        t_ms = ms;
                        found = true;
                    }
                }
            }
            line.clear();
        }
        fclose(f);
    }
    *from_cache = found;
    if (found) {
        P->tune = t;
        // the outputs hold the batch's results, as after a measured autotune
        if (int rc = mix_enqueue(P, b, kind, kind_scalar, band, band_scalar, me, me_scalar, sc, pen,
                                 out, stream))
            return rc;
        CGM_HIP(hipStreamSynchronize(stream));
        *best = t;
        if (best_ms) *best_ms = t_ms;
        if (start_ms) *start_ms = -1.0f;
        return CGM_OK;
    }
    float bm = 0;
    if (int rc = mix_autotune(P, b, kind, kind_scalar, band, band_scalar, me, me_scalar, sc, pen,
                              out, reps, stream, best, &bm, start_ms))
        return rc;
    if (best_ms) *best_ms = bm;
    FILE *f = fopen(cache_path, "a");
    if (!f) return set_error(CGM_ERR_INVALID, "autotune: cannot write the cache file");
    int32_t a[CGM_TUNE_N];
    tuning_to_array(*best, a);
    fprintf(f, "%s\t%d,%d,%d,%d,%d,%d,%d,%d\t%.4f\n", key.c_str(), a[0], a[1], a[2], a[3], a[4],
            a[5], a[6], a[7], bm);
    const bool ok = fclose(f) == 0;
    if (!ok) return set_error(CGM_ERR_INVALID, "autotune: cannot write the cache file");
    return CGM_OK;
}

}  // namespace cgm
