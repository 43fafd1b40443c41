// cgm_internal.h -- shared declarations of libcgm_align.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>

#include "../../include/cgm_align.h"

namespace cgm {

// ---------------------------------------------------------------- errors
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);
#define CGM_HIP(call)                                        \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return ::cgm::hip_error(_e, #call); \
    } while (0)

// Limits enforced by every entry point (documented in cgm_align.h).
constexpr int32_t kMaxScoreMagnitude = 1 << 20;  // |match|,|mismatch|,|open|,|extend|
constexpr int32_t kMaxBand = 1 << 15;            // SW band upper limit
constexpr int32_t kMaxWfaError = 1 << 20;        // WFA max_error upper limit
constexpr int64_t kMaxScoreProduct = int64_t(1) << 30;  // match * min(qlen, rlen)

// Pair batch in device memory.  pair i aligns query i with reference i.
struct Batch {
    const uint64_t *q_packed;
    const uint64_t *r_packed;
    const int32_t *q_len;
    const int32_t *r_len;
    const int32_t *q_off;
    const int32_t *r_off;
    int64_t n_pairs;
    int64_t q_words;
    int64_t r_words;
};

struct SwScoring {
    int32_t match, mismatch, gap_open, gap_extend;
};

struct SwOut {
    int32_t *score;
    int32_t *q_end;
    int32_t *r_end;
};

struct WfaPen {
    int32_t x, o, e;
};

// ------------------------------------------------------------ device code
#ifdef __HIPCC__
// Base `pos` of a packed sequence (sequence.mojo:96-101, dna.mojo:97-100).
__device__ __forceinline__ int base_at(const uint64_t *w, int pos) {
    return (int)((w[pos >> 5] >> (62 - 2 * (pos & 31))) & 3ull);
}

// Loads and bounds-checks pair i.  Returns false when the pair's lengths or
// offsets fall outside the packed buffers; nothing may then be read.
__device__ __forceinline__ bool load_pair(const Batch &b, int64_t i, const uint64_t *&q,
                                          const uint64_t *&r, int &qlen, int &rlen) {
    int ql = b.q_len[i], rl = b.r_len[i], qo = b.q_off[i], ro = b.r_off[i];
    if (ql < 0 || rl < 0 || qo < 0 || ro < 0) return false;
    int64_t qw = ((int64_t)ql + 31) >> 5, rw = ((int64_t)rl + 31) >> 5;
    if ((int64_t)qo + qw > b.q_words || (int64_t)ro + rw > b.r_words) return false;
    q = b.q_packed + qo;
    r = b.r_packed + ro;
    qlen = ql;
    rlen = rl;
    return true;
}
#endif

// ------------------------------------------------------------- launchers
// All launchers validate scalar arguments, return a CGM_* status and run
// asynchronously on `stream`.
int launch_sw_plain(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                    SwOut out, hipStream_t stream);
int launch_wfa_plain(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                     WfaPen p, int32_t *out, hipStream_t stream);
// order: optional device permutation of pair indices (work-sorted); NULL = 0..n-1.
// reg: register layouts where they apply (else the team layouts only).
int launch_wfa_wave(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                    WfaPen p, int32_t *out, bool reg, hipStream_t stream);
// The pieces of the automatic WFA layout, on caller-provided buffers (no
// allocation): classify every pair into the 8 layout classes (counts: 8
// zeroed ints; bad pairs get CGM_BAD_PAIR), class ranges of the sorted
// order from the counts (start: device offset of the first class, or
// null for 0; ranges: 16 ints), and the class kernels over those ranges
// (queues: 8 zeroed counters; slab: grid_big x slab_words int32 for the
// largest pairs), split over two streams.
bool wfa_reg_penalties(const WfaPen &p);
int wfa_slab_words(int32_t me_cap, const WfaPen &p);
int launch_wfa_classify(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                        int32_t me_cap, WfaPen p, bool reg, int32_t *out, uint32_t *keys,
                        int32_t *idx, int32_t *counts, hipStream_t stream);
int launch_wfa_class_ranges(const int32_t *counts, const int32_t *start, int32_t *ranges,
                            hipStream_t stream);
int launch_wfa_classes(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                       int32_t me_cap, WfaPen p, bool reg, int32_t *out, const int32_t *order,
                       const int32_t *ranges, int32_t *queues, int32_t *slab, int64_t slab_words,
                       int grid_big, hipStream_t s_long, hipStream_t s_short);

int wfa_geometry(const WfaPen &p, int *g, int *nM, int *nI);
// One-workgroup-per-pair baselines restricted to one kind of a mixed batch
// (kind[i] == 0: SW, 1: WFA); kind == NULL means every pair.
int launch_sw_plain_kind(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                         SwOut out, const int32_t *kind, hipStream_t stream);
int launch_wfa_plain_kind(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                          WfaPen p, int32_t *out, const int32_t *kind, int32_t *q_end,
                          int32_t *r_end, hipStream_t stream);

// Mixed SW + WFA batches (cgm_mixed.hip).
struct MixPlan;
int mix_plan_create(int64_t max_pairs, int32_t me_cap, WfaPen pen, int waves_per_cu,
                    int lds_bytes, MixPlan **out);
void mix_plan_destroy(MixPlan *P);
int64_t mix_plan_bytes(const MixPlan *P);
int mix_plan_kernel_ms(MixPlan *P, float *ms);
// Duration of the last call's WFA kernels (0 when it had no WFA pair).
int mix_plan_wfa_ms(MixPlan *P, float *ms);
int mix_enqueue(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                SwScoring sc, WfaPen pen, SwOut out, hipStream_t stream);
int mix_capture(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                SwScoring sc, WfaPen pen, SwOut out, hipStream_t stream);
int mix_replay(MixPlan *P, hipStream_t stream);

// Layout choices of the persistent mixed kernel, per plan (values and
// indices as CGM_TUNE_* in cgm_align.h).  None changes a result: every
// layout computes the same recurrences exactly.
struct MixTuning {
    int32_t blocks_per_cu = 2;         // four-wavefront workgroups per CU (1 or 2)
    int32_t lds_bytes = 64 * 1024;     // LDS per workgroup
    int32_t tile_steps = 32;           // kept for the ABI; WFA pairs run the WFA launcher's layout
    int32_t wave_max_width = 1 << 30;  // kept for the ABI; as tile_steps
    int32_t reg_max_band = 64;         // SW bands up to this: register lanes, wider: LDS lanes
    int32_t lane_max_len = 1 << 30;    // SW pairs longer than this: one workgroup per pair
    int32_t sw_dual = 1;               // register SW lanes: dual-row f16 / int16 where exact (1) or int32 (0)
    int32_t sw_split = 55;             // SW-only: percent of the 64-pair units in the long queue
};
constexpr int kMixTuneN = 8;
void tuning_to_array(const MixTuning &t, int32_t *a);
int tuning_from_array(const int32_t *a, MixTuning *t);  // validates
MixTuning mix_default_tuning();
void mix_set_default_tuning(const MixTuning &t);
MixTuning mix_plan_tuning(const MixPlan *P);
void mix_plan_set_tuning(MixPlan *P, const MixTuning &t);
// false: every kernel of an enqueue on the caller's stream (no forked streams)
void mix_plan_set_fork(MixPlan *P, bool fork);
// Bounds the next enqueues' SW pairs keep to: every band <= max_band and
// every read length <= max_len (-1: no bound known).
void mix_plan_set_hint(MixPlan *P, int32_t max_band, int32_t max_len);
// The SW persistent kernel's instance with the wavefront-per-pair phase
// (cgm_mixed_wave.hip), `grid` workgroups on `stream`; tasks: the plan's
// task records.
// launch_sw_wave on a plan's workspace: zeroes its counters, runs the
// wavefront-per-pair kernel and the one-workgroup-per-pair kernel over the
// pairs that kernel left.
int mix_enqueue_wave_direct(MixPlan *P, const Batch &b, const int32_t *band, int32_t band_scalar,
                            SwScoring sc, SwOut out, hipStream_t stream);
int launch_mix_persistent_wave(unsigned grid, hipStream_t stream, const Batch &b,
                               const int32_t *band, int32_t band_scalar, SwScoring sc, SwOut out,
                               const int32_t *order, const void *tasks, const int32_t *task_order,
                               const int32_t *n_tasks, int32_t *small, int32_t sw_dual,
                               int32_t sw_split);
// Measures candidate tunings on one batch (coordinate descent over the
// fields, `reps` timed enqueues each, best kept), leaves the best in the
// plan and returns it with its time and the starting tuning's time.
int mix_autotune(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                 const int32_t *band, int32_t band_scalar, const int32_t *me, int32_t me_scalar,
                 SwScoring sc, WfaPen pen, SwOut out, int reps, hipStream_t stream,
                 MixTuning *best, float *best_ms, float *start_ms);
// The batch's length-bin profile as a text key: device, scoring, the log2
// bucket of the pair count and, per (kind, log2 length bin, SW band bin),
// the share of the pairs in 32nds (nonzero cells only).  Reads counts back
// (synchronises `stream`).
int mix_profile_key(MixPlan *P, const Batch &b, const int32_t *kind, int32_t kind_scalar,
                    const int32_t *band, int32_t band_scalar, SwScoring sc, WfaPen pen,
                    hipStream_t stream, std::string *key);
// mix_autotune with the choice persisted per profile key in a text file
// (one line per key: key, tuning, best ms; the last line of a key wins): a
// key found there sets the plan's tuning without measuring; otherwise the
// autotuner runs and its choice is appended.  Either way the outputs hold the
// batch's results.
int mix_autotune_cached(MixPlan *P, const char *cache_path, const Batch &b, const int32_t *kind,
                        int32_t kind_scalar, const int32_t *band, int32_t band_scalar,
                        const int32_t *me, int32_t me_scalar, SwScoring sc, WfaPen pen, SwOut out,
                        int reps, hipStream_t stream, MixTuning *best, float *best_ms,
                        float *start_ms, bool *from_cache);
// A plan owned by the library for the drop-in launchers, one per (device,
// stream), recreated larger when a batch exceeds it (the old plan is freed
// after the stream drains).  `use` runs with the plan locked, so calls on
// one stream from several threads take turns.
int with_cached_plan(hipStream_t stream, int64_t n_pairs, int32_t me_cap, WfaPen pen,
                     const std::function<int(MixPlan *)> &use);
int64_t mix_plan_capacity(const MixPlan *P, int32_t *me_cap, WfaPen *pen);


// WFA workgroup-per-pair layout: pairs whose branch-free LDS layout fits
// kWfaSmallLdsBytes run on 256-thread workgroups with that budget,
// kWfaSmallBlocksPerCu per CU; the rest on 512-thread workgroups with
// kWfaLdsBytes, kWfaBlocksPerCu per CU (twice the wavefronts per pair keeps
// the CU as full as the small layout does).
constexpr int kWfaLdsBytes = 52 * 1024;
constexpr int kWfaBlocksPerCu = 3;
constexpr int kWfaSmallLdsBytes = 20 * 1024;
constexpr int kWfaSmallBlocksPerCu = 8;
// Largest diagonal |k| reachable at score s: (s - o) / e from o + e on.
inline int reach_host(int s, const WfaPen &p) { return s >= p.o + p.e ? (s - p.o) / p.e : 0; }

// Device-side max of an int32 array (blocking; used by launchers that size
// LDS or scratch from a per-pair device array).
int device_max_i32(const int32_t *a, int64_t n, hipStream_t stream, int32_t *result);
// Max over valid pairs of the effective band min(band, max(qlen, rlen))
// (blocking).  A band wider than both sequences adds no cell and changes no
// boundary value, so every SW layout aligns with the effective band.
int device_max_band_eff(const Batch &b, const int32_t *band, int32_t band_scalar,
                        hipStream_t stream, int32_t *result);

#ifdef __HIPCC__
__device__ __forceinline__ int effective_band(int band, int qlen, int rlen) {
    const int m = qlen > rlen ? qlen : rlen;
    return band < m ? band : m;
}
#endif

// Work ordering (cgm_sort.hip): order[] = pair indices sorted by a
// descending key.
int sort_desc_by_key(uint32_t *keys, int32_t *idx, int32_t *order_out, int64_t n,
                     hipStream_t stream);
// Bytes of temporary storage the radix sort of n keys needs (no allocation).
int sort_tmp_bytes(int64_t n, size_t *bytes);
// The same sort on caller-provided storage (keys_out: n entries; tmp:
// tmp_bytes >= sort_tmp_bytes(n)); allocates nothing.
int sort_desc_by_key_ws(const uint32_t *keys, uint32_t *keys_out, const int32_t *idx,
                        int32_t *order_out, int64_t n, void *tmp, size_t tmp_bytes,
                        hipStream_t stream);

// SW pair classes of the automatic layout.  0..3: register lane kernels for
// effective bands [0,16], [17,32], [33,48], [49,64] (non-positive penalties,
// h < 2^23); 4: LDS lane kernel (band <= kLdsLaneMaxBand); 5: one workgroup
// per pair (band <= kPlainMaxBand); 6: not alignable (reported bad).
constexpr int kSwClasses = 7;
constexpr int kSwClassLds = 4, kSwClassPlain = 5, kSwClassBad = 6;
constexpr int kLdsLaneMaxBand = 150;
constexpr int kPlainMaxBand = 1600;
// Classifies, writes bad-pair outputs, sorts pair indices by (class, band,
// length) and fills ranges[2*c], ranges[2*c+1] = [start, end) of class c
// in order[] -- all on the device.
int sw_classify(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                SwOut out, int32_t *order, int32_t *ranges, hipStream_t stream);
// Device workspace of the classification (n = n_pairs entries each).
struct SwClassWs {
    int32_t *order = nullptr;      // n: pair indices, sorted
    int32_t *ranges = nullptr;     // 2 * kSwClasses
    uint32_t *keys = nullptr;      // n
    uint32_t *keys_out = nullptr;  // n
    int32_t *idx = nullptr;        // n
    int32_t *counts = nullptr;     // 16 (zeroed by the classification itself)
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
};
// sw_classify on caller-provided workspace; allocates nothing.
int sw_classify_ws(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, const SwClassWs &ws, hipStream_t stream);
// fork: run the four band buckets side by side on forked streams (else one
// after another on `stream`, creating nothing).
int launch_sw_reg_buckets(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                          SwOut out, const int32_t *order, const int32_t *ranges, int grid,
                          hipStream_t stream, bool fork = true);
int launch_sw_lds_lane(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                       SwOut out, const int32_t *order, const int32_t *range, int grid,
                       hipStream_t stream);
int launch_sw_plain_range(const Batch &b, const int32_t *band, int32_t band_scalar,
                          SwScoring sc, SwOut out, const int32_t *order, const int32_t *range,
                          int grid, hipStream_t stream);
// One wavefront per pair (cgm_sw_wave_dev.h, run inside the SW persistent
// kernel): pairs with effective band <= kWaveMaxBand, both lengths below
// kWaveMaxLen and match * min(qlen, rlen) < 2^23.
constexpr int kWaveMaxBand = 64;
constexpr int kWaveMaxLen = 1 << 20;
int launch_sw_auto(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, hipStream_t stream);
// Every SW pair of effective band <= 64 on one wavefront of its own (pair i
// on wavefront i, no sort), the rest on the one-workgroup-per-pair kernel.
int launch_sw_wave(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, hipStream_t stream);
// The wavefront-per-pair kernel of launch_sw_wave (cgm_mixed_wave.hip): pairs
// it does not take are appended to rest[], their count added to
// rest_range[1] (both zeroed by the caller).
int launch_sw_wave_direct(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                          SwOut out, int32_t *rest, int32_t *rest_range, hipStream_t stream);
int launch_sw_buckets(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                      SwOut out, hipStream_t stream);
// The launches of the band-bucket layout after classification (order and
// ranges from sw_classify); fork: the register buckets side by side on
// forked streams.
int sw_bucket_kernels(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                      SwOut out, const int32_t *order, const int32_t *ranges, bool fork,
                      hipStream_t stream);
// Start coordinates of SW alignments whose score and ends are already in
// score / q_end / r_end (cgm_sw_start.hip).
int launch_sw_start(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                    const int32_t *score, const int32_t *q_end, const int32_t *r_end,
                    int32_t *q_start, int32_t *r_start, hipStream_t stream);
int device_cu_count(int *cus);
// Raises a kernel's dynamic LDS limit to `bytes` once per device
// (thread-safe; every launch of that kernel uses the same limit).
int set_kernel_lds(const void *kernel, int bytes);

// Fork/join of auxiliary streams (default priority, non-blocking) so that
// independent launches of one call overlap on the GPU; both sides are
// event-ordered, so the pattern also records correctly under graph capture.
constexpr int kMaxAuxStreams = 4;
int fork_streams(hipStream_t main, int n, hipStream_t *aux);
int join_streams(hipStream_t main, int n, hipStream_t *aux);

// SW scoring: |value| <= 2^20 and non-positive penalties (mismatch,
// gap_extend, gap_open + gap_extend <= 0), as AlignConfig documents
// (align_cpu.mojo:18-25).
int check_sw_scoring(const SwScoring &sc);
// Magnitudes only (banded NW bounds its values per pair instead).
int check_score_magnitudes(const SwScoring &sc);
int check_wfa_pen(const WfaPen &p);
int check_batch(const Batch &b);

}  // namespace cgm
