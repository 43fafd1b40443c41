/*
 * cgm_align.h -- C ABI of the MI355X (gfx950) pairwise-alignment library
 * (libcgm_align.so): banded affine Smith-Waterman and gap-affine WFA over
 * 2-bit packed SequenceBatch structure-of-arrays buffers.
 *
 * Every entry point is plain C: pointers, sizes and int status codes, no
 * C++ or torch types, so a Mojo `external_call`, Python ctypes or a C caller
 * can bind it directly.  Citations name the reference (Computational-
 * Genomics-Mojo) interface each entry point replaces.
 *
 * Batch layout (genomics/core/sequence.mojo:23-75 and the launcher arguments
 * of genomics/gpu/align_gpu.mojo:460-470):
 *   q_packed / r_packed  uint64 words, 32 bases per word, base i of a word at
 *                        bits (63-2i)..(62-2i) (MSB first), A=0 C=1 G=2 T=3.
 *                        N and other non-ACGT bytes are stored as A; their
 *                        NMask is not consulted when scoring, exactly as the
 *                        reference does (sequence.mojo:96-101).
 *   q_off / r_off        int32 word offset of each sequence in its buffer.
 *   q_len / r_len        int32 length in bases.
 *   pair i aligns query i against reference i; q_words / r_words are the
 *   lengths of the two packed buffers (the reference's `total_words`), used
 *   for bounds checks.
 *
 * Outputs are caller-allocated device arrays of n_pairs int32.  A pair whose
 * lengths/offsets fall outside its buffer (or whose band / max_error is
 * negative or above the supported maximum) is not aligned: it gets
 * score = CGM_BAD_PAIR and ends = -1, and nothing is read for it.
 *
 * Status: 0 = CGM_OK; otherwise an error code, with a message available from
 * cgm_last_error() (thread-local).  Launch entry points are asynchronous on
 * the given stream (NULL = the default stream) and thread-safe per stream.
 */
#ifndef CGM_ALIGN_H
#define CGM_ALIGN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGM_OK 0
#define CGM_ERR_INVALID 1   /* bad argument */
#define CGM_ERR_HIP 2       /* HIP runtime error */
#define CGM_ERR_NOMEM 3     /* allocation or budget refused */
#define CGM_ERR_BUSY 4      /* queue full (serving engine back-pressure) */
#define CGM_ERR_NODEV 5     /* no GPU visible */
#define CGM_ERR_NOTFOUND 6  /* unknown request id */

#define CGM_BAD_PAIR ((int32_t)0x80000000)

/* Kernel layouts for the launch entry points. */
#define CGM_LAYOUT_AUTO 0   /* SW: persistent task-queue kernel, work-sorted, one pair per
                               lane, on a plan the library keeps per (device, stream);
                               batches of at most 16 pairs per CU as CGM_LAYOUT_WAVE;
                               WFA: register layout (wavefronts in VGPRs, one diagonal per
                               lane and block) where the penalties are x : o+e : e =
                               2 : 4 : 1 and the pair's diagonals fit, else as
                               CGM_LAYOUT_WAVE; longest pairs first */
#define CGM_LAYOUT_PLAIN 1  /* one workgroup per pair (baseline) */
#define CGM_LAYOUT_LANE 2   /* SW: one pair per lane, one kernel per band bucket */
#define CGM_LAYOUT_WAVE 3   /* SW: one wavefront per pair sweeping the band's anti-diagonals
                               (effective bands <= 64; wider bands as CGM_LAYOUT_AUTO);
                               WFA: team layouts only: one workgroup per pair, wavefronts in
                               LDS (or tiled over an HBM slab), longest pairs first */

/* ------------------------------------------------------------ errors, info */

/* Message of the last failed call on this thread ("" if none). */
const char *cgm_last_error(void);
/* Library version as 10000*major + 100*minor + patch. */
int cgm_version(void);

/* ---------------------------------------------------------------- device
 * Replace GenomicsDevice (genomics/gpu/device.mojo:21-138): device context,
 * typed allocation, host<->device copies, synchronisation. */
int cgm_device_count(int *count);
int cgm_set_device(int device);
/* name: at least name_len bytes; any output pointer may be NULL. */
int cgm_device_info(int device, char *name, int name_len, int *compute_units,
                    int *max_clock_khz, int64_t *hbm_bytes, int *lds_bytes_per_cu);
int cgm_malloc(void **dptr, int64_t bytes);
int cgm_free(void *dptr);
/* Pinned host memory (for asynchronous copies). */
int cgm_host_malloc(void **hptr, int64_t bytes);
int cgm_host_free(void *hptr);
int cgm_memcpy_htod(void *dst, const void *src, int64_t bytes, void *stream);
int cgm_memcpy_dtoh(void *dst, const void *src, int64_t bytes, void *stream);
int cgm_memcpy_dtod(void *dst, const void *src, int64_t bytes, void *stream);
int cgm_stream_create(void **stream);
int cgm_stream_destroy(void *stream);
int cgm_stream_synchronize(void *stream);
int cgm_device_synchronize(void);
int cgm_event_create(void **event);
int cgm_event_destroy(void *event);
int cgm_event_record(void *event, void *stream);
int cgm_event_synchronize(void *event);
int cgm_event_elapsed_ms(void *start, void *stop, float *ms);

/* -------------------------------------------------------- banded SW batch
 * Replaces launch_banded_sw[band] (genomics/gpu/align_gpu.mojo:460-493),
 * with the semantics of smith_waterman_banded (genomics/cpu/align_cpu.mojo:
 * 229-309) per pair: local alignment restricted to |ri - qi| <= band,
 * out_score = best H, out_q_end/out_r_end = first cell reaching it in
 * row-major order (query_start = ref_start = 0 and the CIGAR is empty, as in
 * the reference).  band: per-pair device array, or NULL to use band_scalar
 * for every pair.  Scoring: |value| <= 2^20 and penalties <= 0 as AlignConfig
 * documents (align_cpu.mojo:18-25): mismatch > 0, gap_extend > 0 or
 * gap_open + gap_extend > 0 is refused with CGM_ERR_INVALID.  A pair with
 * match * min(qlen, rlen) > 2^30 or an effective band min(band, max(qlen,
 * rlen)) above 1600 is reported with CGM_BAD_PAIR.  The reference's band is
 * a compile-time parameter; here it is a runtime value per pair. */
int cgm_sw_banded_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len,
                         const int32_t *q_off, const int32_t *r_off,
                         int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *band, int32_t band_scalar,
                         int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                         int32_t *out_score, int32_t *out_q_end, int32_t *out_r_end,
                         int32_t layout, void *stream);

/* -------------------------------------------- banded SW, allocation-free
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory during a call (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): cgm_sw_banded_launch with
 * layout CGM_LAYOUT_LANE, same arguments and results, but every temporary
 * (sort keys, pair order, class counts, radix-sort storage) lives in one
 * static 12 MiB device array per GPU, held until the process ends, and no
 * stream is created.  Limit: n_pairs <= 262144 per call (CGM_ERR_INVALID
 * above).  Calls on one GPU share that array: calls that overlap in time
 * on one GPU (two streams, or two host threads) race and must not be
 * issued; consecutive calls on one stream are safe. */
int cgm_sw_banded_launch_noalloc(const uint64_t *q_packed, const uint64_t *r_packed,
                                 const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                                 const int32_t *r_off, int64_t n_pairs, int64_t q_words,
                                 int64_t r_words, const int32_t *band, int32_t band_scalar,
                                 int32_t match, int32_t mismatch, int32_t gap_open,
                                 int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                                 int32_t *out_r_end, void *stream);

/* ------------------------------------------- banded SW with start coordinates
 * cgm_sw_banded_launch plus the start of each alignment, which the
 * reference reports as query_start = ref_start = 0 (smith_waterman_banded,
 * genomics/cpu/align_cpu.mojo:302-305) and does not define.  Here: among
 * the alignments of maximal score that end in H at (out_q_end, out_r_end),
 * the start of the one whose first diagonal step (the first query and
 * reference base it pairs) is latest -- largest query start, then largest
 * reference start.  Score and ends are exactly those of
 * cgm_sw_banded_launch (same layouts, same arguments).  Starts are (0, 0)
 * when the score is 0, -1 for a bad pair, and -1 when
 * (q_end + 2) * (2 * band_eff + 1) exceeds 2^32.  Runs a second pass over
 * rows 0 .. q_end of each pair; asynchronous on `stream`. */
int cgm_sw_banded_start_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                               const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                               const int32_t *r_off, int64_t n_pairs, int64_t q_words,
                               int64_t r_words, const int32_t *band, int32_t band_scalar,
                               int32_t match, int32_t mismatch, int32_t gap_open,
                               int32_t gap_extend, int32_t *out_score, int32_t *out_q_start,
                               int32_t *out_r_start, int32_t *out_q_end, int32_t *out_r_end,
                               int32_t layout, void *stream);
/* Host-pointer form (copies in, runs on an internal stream, copies out). */
int cgm_sw_banded_start_host(const uint64_t *q_packed, const uint64_t *r_packed,
                             const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                             const int32_t *r_off, int64_t n_pairs, int64_t q_words,
                             int64_t r_words, const int32_t *band, int32_t band_scalar,
                             int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                             int32_t *out_score, int32_t *out_q_start, int32_t *out_r_start,
                             int32_t *out_q_end, int32_t *out_r_end, int32_t layout);

/* ------------------------------------------------------- banded NW batch
 * The batch form of needleman_wunsch_banded (genomics/cpu/align_cpu.mojo:
 * 312-377), which the reference has on the CPU only: global score within
 * |ri - qi| <= band (every gap base costs gap_open + gap_extend inside the
 * matrix; the row before the first charges gap_open + c * gap_extend for a
 * leading reference gap of c bases; no leading query gap), out_score =
 * that score or -1000000 when |rlen - qlen| > band, out_q_end = qlen - 1,
 * out_r_end = rlen - 1.  The reference allocates its two rows with
 * List(capacity=width) and indexes them without appending (align_cpu.mojo:
 * 324-337); the rows are taken to hold 2*band+1 entries.  Arguments as
 * cgm_sw_banded_launch; layout CGM_LAYOUT_AUTO (one pair per lane for
 * effective bands <= 64, one workgroup per pair above) or CGM_LAYOUT_PLAIN
 * (one workgroup per pair).  Pairs whose values could leave int32 range,
 * (qlen + rlen + 2 band + 4) * (|match| + |mismatch| + |open| + |extend|)
 * + 10^6 > 2^30, or with an effective band above 1600, are reported with
 * CGM_BAD_PAIR. */
int cgm_nw_banded_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                         const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *band, int32_t band_scalar, int32_t match,
                         int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                         int32_t *out_score, int32_t *out_q_end, int32_t *out_r_end,
                         int32_t layout, void *stream);
/* Host-pointer form (copies in, runs on an internal stream, copies out). */
int cgm_nw_banded_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                       const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words, const int32_t *band,
                       int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                       int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                       int32_t *out_r_end, int32_t layout);

/* ------------------------------------------------------ gap-affine WFA batch
 * Replaces launch_wfa_affine[max_error, x, o, e] (genomics/gpu/align_gpu.mojo:
 * 496-535) with the semantics of wfa_affine_cpu (align_cpu.mojo:75-226):
 * out_cost = minimum global cost (mismatch x, gap of length l = o + l*e,
 * match 0) or -1 when it exceeds max_error.  max_error: per-pair device
 * array or NULL for max_error_scalar.  Requires x >= 1, e >= 1, o >= 0 (the
 * reference GPU kernel asserts e >= 1, align_gpu.mojo:288).  With a per-pair
 * max_error array, CGM_LAYOUT_AUTO and CGM_LAYOUT_WAVE read its maximum back
 * to size their wavefront storage (one host synchronisation). */
int cgm_wfa_affine_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                          const int32_t *q_len, const int32_t *r_len,
                          const int32_t *q_off, const int32_t *r_off,
                          int64_t n_pairs, int64_t q_words, int64_t r_words,
                          const int32_t *max_error, int32_t max_error_scalar,
                          int32_t x, int32_t o, int32_t e,
                          int32_t *out_cost, int32_t layout, void *stream);

/* ----------------------------------------- SequenceBatch-shaped batch forms
 * cgm_sw_banded_launch / cgm_wfa_affine_launch with layout CGM_LAYOUT_AUTO,
 * taking each side as the SequenceBatch structure of arrays it is held in
 * (genomics/core/sequence.mojo:23-36: packed words, NMask words, word
 * offsets, lengths, and the buffer's word count), as the launchers of
 * genomics/gpu/align_gpu.mojo:460-535 receive them after
 * GenomicsDevice.upload_* (genomics/gpu/device.mojo:60-108).  q_nmask /
 * r_nmask may be NULL; they are accepted and never read: an N (or any
 * non-ACGT base) is scored as the A it is packed as, exactly as the
 * reference scores it (its decode ignores NMask, sequence.mojo:96-101).
 * Results, limits and errors as for the launch entry points. */
int cgm_sw_banded_batch(const uint64_t *q_packed, const uint64_t *q_nmask, const int32_t *q_off,
                        const int32_t *q_len, int64_t q_words, const uint64_t *r_packed,
                        const uint64_t *r_nmask, const int32_t *r_off, const int32_t *r_len,
                        int64_t r_words, int64_t n_pairs, const int32_t *band,
                        int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                        int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                        int32_t *out_r_end, void *stream);
int cgm_wfa_affine_batch(const uint64_t *q_packed, const uint64_t *q_nmask, const int32_t *q_off,
                         const int32_t *q_len, int64_t q_words, const uint64_t *r_packed,
                         const uint64_t *r_nmask, const int32_t *r_off, const int32_t *r_len,
                         int64_t r_words, int64_t n_pairs, const int32_t *max_error,
                         int32_t max_error_scalar, int32_t x, int32_t o, int32_t e,
                         int32_t *out_cost, void *stream);

/* ------------------------------------------- gap-affine WFA with traceback
 * Replaces wfa_affine_cigar_cpu (genomics/cpu/align_cpu.mojo:380-623) for a
 * batch: out_cost as cgm_wfa_affine_launch, plus the CIGAR that function's
 * backtrace builds (in M: the match run, then mismatch before insertion
 * before deletion; in I and D: gap opening before extension), as
 * run-length operations run = (count << 2) | op, op 0 '=', 1 'X', 2 'I'
 * (query base), 3 'D' (reference base), first operation first, written to
 * cigar[cigar_off[i] .. cigar_off[i + 1]) (cigar_off: n_pairs + 1 int64 on
 * the device).  cigar_len[i] = number of runs; 0 for cost -1, an invalid
 * pair or two empty sequences; -2 when the pair's capacity was too small.
 * A CIGAR of cost c has at most 2 c / min(x, o + e) + 1 runs.  Computes
 * the costs first, then re-runs each pair with its whole wavefront history
 * in HBM (longest pairs first, in chunks of at most half the free device
 * memory, 8-96 GiB) and walks it back; synchronises with the host before
 * returning.  The history workspace (the largest one a call has needed so
 * far) stays allocated per device between calls, so repeated calls do not
 * pay for a tens-of-GB allocation each; calls on one device take turns on
 * it. */
int cgm_wfa_cigar_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                         const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                         int32_t e, int32_t *out_cost, uint32_t *cigar, const int64_t *cigar_off,
                         int32_t *cigar_len, void *stream);
/* Host-pointer form (cigar_off is a host array here). */
int cgm_wfa_cigar_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                       const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words,
                       const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                       int32_t e, int32_t *out_cost, uint32_t *cigar, const int64_t *cigar_off,
                       int32_t *cigar_len);

/* ------------------------------------------------ mixed SW + WFA batches
 * The batched persistent form of the two launchers above for a batch that
 * mixes SW pairs (kind[i] == 0) and WFA pairs (kind[i] == 1): one call
 * aligns every pair, writing the SW score/ends or the WFA cost (ends -1).
 * SW pairs run one persistent kernel over a work-sorted task queue; WFA
 * pairs run the two longest-first kernels of cgm_wfa_affine_launch's
 * automatic layout on a stream forked from the call's (and joined back),
 * over the plan's buffers.
 * No reference counterpart takes a mixed batch; per pair the results are
 * those of launch_banded_sw / launch_wfa_affine (align_gpu.mojo:460-535),
 * i.e. of smith_waterman_banded / wfa_affine_cpu.
 *
 * A plan owns every buffer the automatic layout needs (sort keys, task
 * queue, HBM wavefront slabs for WFA pairs up to max_error_cap), so
 * cgm_align_mixed performs no allocation and no host synchronisation and
 * can be captured: cgm_plan_capture records one call into a hipGraph kept
 * in the plan, cgm_plan_replay launches it again (same pointers; refill
 * their contents between replays).  waves_per_cu / lds_bytes: 0 = the
 * default tuning's (below), else they set its workgroup shape.  A plan
 * serves one device and one call at a time. */
int cgm_plan_create(int64_t max_pairs, int32_t max_error_cap, int32_t x, int32_t o, int32_t e,
                    int32_t waves_per_cu, int32_t lds_bytes, void **plan);
int cgm_plan_destroy(void *plan);
/* layout CGM_LAYOUT_AUTO: the persistent task-queue kernel (plan required);
 * CGM_LAYOUT_PLAIN: one workgroup per pair for both kinds (plan ignored). */
int cgm_align_mixed(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                    const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                    const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                    const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                    int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                    int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                    int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                    int32_t *out_r_end, int32_t layout, void *stream);
int cgm_plan_capture(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                     const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                     const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                     const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                     int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                     int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                     int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                     int32_t *out_r_end, void *stream);
int cgm_plan_replay(void *plan, void *stream);
/* Duration of the SW persistent kernel of the plan's last call (or
 * replay), from events recorded around it on the call's stream; waits for
 * it. */
int cgm_plan_kernel_ms(void *plan, float *ms);
/* Duration of the last call's WFA kernels (the WFA pairs run the kernels
 * of cgm_wfa_affine_launch's automatic layout on a stream forked from the
 * call's and joined back), from events around them; 0 when the call had no
 * WFA pair; waits for them. */
int cgm_plan_wfa_kernel_ms(void *plan, float *ms);

/* ------------------------------------------------------------ autotuning
 * Layout choices of the persistent kernel, as an int32 array of
 * CGM_TUNE_N entries (no reference counterpart: the reference has one
 * launch shape per kernel, align_gpu.mojo:24-38).  None changes a result.
 *   BLOCKS_PER_CU   four-wavefront SW workgroups per CU (1 or 2)
 *   LDS_BYTES       validated (blocks_per_cu x (lds + 256) <= 160 KiB) and
 *                   kept for the ABI; the SW persistent kernel uses no LDS
 *   TILE_STEPS      kept for the ABI (1..1024); WFA pairs run the WFA
 *                   launcher's own layout
 *   WAVE_MAX_WIDTH  kept for the ABI (>= 0); as TILE_STEPS
 *   REG_MAX_BAND    band split: SW bands up to this (<= 64) run in register
 *                   lanes, wider ones in LDS lanes (band 128: packed lanes)
 *   LANE_MAX_LEN    switch-over: SW pairs longer than this run one workgroup
 *                   per pair instead of one lane per pair
 *   SW_DUAL         1: register SW lanes hold the band frame's two halves,
 *                   one row apart, in the two 16-bit halves of each register:
 *                   IEEE half precision wherever every value is an exact
 *                   multiple of t (match * min(qlen, rlen) + match <=
 *                   2048 t, t the largest power of two <= 8 dividing every
 *                   score), else packed int16 wherever every value fits
 *                   (match * min(qlen, rlen) + match <= 32767, penalties
 *                   >= -8191); 0: int32 lanes only
 *   SW_SPLIT        SW-only batches: the longest-first list of 64-pair units
 *                   is cut after this percent (0..100) of the units; the
 *                   first wavefront on each SIMD takes the long part, then
 *                   the short part, the second wavefront only the short part,
 *                   so the last units of the launch are the shortest
 * cgm_plan_autotune times the whole pipeline on the given batch for
 * candidate values of each field in turn (coordinate descent, median of
 * `reps` runs per candidate), keeps a change only if it is >= 2 % faster,
 * leaves the best tuning in the plan and returns it with its time and the
 * starting tuning's time; the outputs hold the batch's results.  New plans
 * (and the library's own plans behind CGM_LAYOUT_AUTO and the serving
 * engine) start from the process default tuning. */
#define CGM_TUNE_BLOCKS_PER_CU 0
#define CGM_TUNE_LDS_BYTES 1
#define CGM_TUNE_TILE_STEPS 2
#define CGM_TUNE_WAVE_MAX_WIDTH 3
#define CGM_TUNE_REG_MAX_BAND 4
#define CGM_TUNE_LANE_MAX_LEN 5
#define CGM_TUNE_SW_DUAL 6
#define CGM_TUNE_SW_SPLIT 7
#define CGM_TUNE_N 8
int cgm_plan_get_tuning(void *plan, int32_t *params);
int cgm_plan_set_tuning(void *plan, const int32_t *params);
int cgm_tuning_get_default(int32_t *params);
int cgm_tuning_set_default(const int32_t *params);
int cgm_plan_autotune(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                      const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                      const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                      const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                      int32_t band_scalar, const int32_t *max_error, int32_t max_error_scalar,
                      int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                      int32_t x, int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                      int32_t *out_r_end, int32_t reps, int32_t *best_params, float *best_ms,
                      float *start_ms, void *stream);

/* The tuning chosen per batch profile and kept across processes.  A batch's
 * profile key (cgm_plan_profile_key, NUL-terminated, at most key_len bytes)
 * names the device, the scoring, the log2 bucket of the pair count and,
 * per (kind, log2 bin of the longer length, SW effective-band bin 0-16 /
 * 17-64 / 65-128 / wider), the share of the pairs in 32nds.
 * cgm_plan_autotune_cached looks the key up in the text file cache_path
 * (one line per choice: key, tab, the CGM_TUNE_N fields, tab, best ms; the
 * file's last line for a key wins): when found it sets the plan's tuning
 * without measuring (*from_cache = 1, *start_ms = -1); otherwise it runs
 * cgm_plan_autotune and appends the choice (*from_cache = 0).  Either way
 * the outputs hold the batch's results.  Both read the profile back from
 * the device (they synchronise `stream`); calls in one process take turns
 * on the file. */
int cgm_plan_profile_key(void *plan, const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                         const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *kind, int32_t kind_scalar, const int32_t *band,
                         int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                         int32_t gap_extend, int32_t x, int32_t o, int32_t e, char *key,
                         int32_t key_len, void *stream);
int cgm_plan_autotune_cached(void *plan, const char *cache_path, const uint64_t *q_packed,
                             const uint64_t *r_packed, const int32_t *q_len, const int32_t *r_len,
                             const int32_t *q_off, const int32_t *r_off, int64_t n_pairs,
                             int64_t q_words, int64_t r_words, const int32_t *kind,
                             int32_t kind_scalar, const int32_t *band, int32_t band_scalar,
                             const int32_t *max_error, int32_t max_error_scalar, int32_t match,
                             int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t x,
                             int32_t o, int32_t e, int32_t *out_score, int32_t *out_q_end,
                             int32_t *out_r_end, int32_t reps, int32_t *best_params,
                             float *best_ms, float *start_ms, int32_t *from_cache, void *stream);

/* ------------------------------------------------------ serving engine
 * Continuous batching for a stream of alignment requests (no reference
 * counterpart; the reference aligns one pair per call).  Pairs are batched
 * by bin: SW pairs (kind 0) and WFA pairs (kind 1), so SW requests never wait behind
 * long WFA pairs.  Each bin preallocates n_slots slots (a stream, device
 * buffers and one mixed-kernel plan each) and 2 n_slots + 1 pinned staging
 * buffers of max_batch_pairs pairs and 2 max_batch_words packed words;
 * creation fails with CGM_ERR_NOMEM if the slots exceed hbm_budget_bytes.
 * Each bin has its own dispatcher thread, which launches the next staging
 * buffer whenever one of the bin's slots is free; a host function on the
 * slot's stream returns the results, so requests that arrive while others
 * run are admitted as slots free up and batches of both bins overlap on
 * the GPU.  max_batch_words <= 2^29, max_batch_pairs <= 2^28.
 *
 * cgm_engine_submit copies the request's pairs (kind/band/max_error per
 * pair, sequences at word offsets inside q_packed[0, q_words) and
 * r_packed[0, r_words)) into the open staging buffers and returns at once;
 * it refuses with CGM_ERR_INVALID a pair whose words lie outside those
 * buffers or whose kind is not 0 or 1, and with CGM_ERR_BUSY a request for which a bin has no staging
 * space left (back-pressure); a request is admitted whole or not at all.
 * arrival_s (cgm_engine_now() clock, <= 0 = now) is the time latency is
 * measured from.  Results do not depend on which requests share a batch;
 * a batch whose launch fails reports CGM_BAD_PAIR scores and -1 ends. */
int cgm_engine_create(int64_t hbm_budget_bytes, int32_t n_slots, int64_t max_batch_pairs,
                      int64_t max_batch_words, int32_t max_error_cap, int32_t match,
                      int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t x, int32_t o,
                      int32_t e, void **engine);
int cgm_engine_destroy(void *engine);
int cgm_engine_submit(void *engine, const uint64_t *q_packed, int64_t q_words,
                      const uint64_t *r_packed, int64_t r_words, const int32_t *q_len,
                      const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                      int32_t n_pairs, const int32_t *kind, const int32_t *band,
                      const int32_t *max_error, double arrival_s, int64_t *request_id);
/* cgm_engine_submit for a request whose packed words are already in device
 * memory (q_packed / r_packed device pointers, q_words / r_words their
 * lengths, e.g. a batch encoded on the device by cgm_encode_ascii_launch);
 * lengths, offsets, kind, band and max_error stay host arrays.  The words
 * are gathered on the device into the batch when it launches, so only the
 * metadata crosses PCIe; the caller keeps the device buffers unchanged until
 * the request is done. */
int cgm_engine_submit_device(void *engine, const uint64_t *q_packed, int64_t q_words,
                             const uint64_t *r_packed, int64_t r_words, const int32_t *q_len,
                             const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                             int32_t n_pairs, const int32_t *kind, const int32_t *band,
                             const int32_t *max_error, double arrival_s, int64_t *request_id);
/* done = 1 once the request's results are ready; CGM_ERR_NOTFOUND for an
 * id never issued or already collected. */
int cgm_engine_poll(void *engine, int64_t request_id, int32_t *done);
/* Blocks (timeout_ms < 0: forever) until the request is done, copies its
 * n_pairs results (any pointer may be NULL) and releases the request: its
 * id is unknown from then on (of two waits on one id, one gets the results
 * and the other CGM_ERR_NOTFOUND). */
int cgm_engine_wait(void *engine, int64_t request_id, int32_t timeout_ms, int32_t *score,
                    int32_t *q_end, int32_t *r_end, double *latency_ms);
/* Tuning of one bin's batches (bin 0: SW pairs, 1: WFA pairs) as a
 * CGM_TUNE_N array (fields as for cgm_plan_get_tuning); a new value applies
 * from the bin's next launch.  Bin 0 starts from the process default with
 * LANE_MAX_LEN = 200: SW reads longer than that run one wavefront per pair
 * inside the persistent kernel instead of one lane, which shortens a small
 * batch's longest serial chain. */
int cgm_engine_get_tuning(void *engine, int32_t bin, int32_t *params);
int cgm_engine_set_tuning(void *engine, int32_t bin, const int32_t *params);
/* out[8]: batches, batched pairs, finished requests, finished pairs,
 * refused requests, queued pairs, device bytes held, last worker error. */
int cgm_engine_stats(void *engine, int64_t *out);
double cgm_engine_now(void);
/* Load generator: `clients` client threads (1..64), each submitting
 * requests of U{1..max_pairs_per_request} consecutive pairs from its own
 * slice of a generated batch (cgm_gen_create), with Poisson arrivals at
 * rate_rps / clients each for duration_s (rate_rps <= 0: a closed loop with
 * at most max_outstanding requests in flight over all clients), and each
 * collecting its results in submission order.  device_inputs = 1: the
 * batch's packed words are uploaded once and requests go through
 * cgm_engine_submit_device (sequences resident in HBM); 0: through
 * cgm_engine_submit (host words copied per request).  stats[27]: submitted,
 * completed, refused, pairs completed, elapsed s, pairs/s, requests/s, p50,
 * p90, p99, max, mean latency (ms), batches, mean pairs per batch, bad
 * results, submit phase s, then SW-only requests: count, p50, p99 (ms);
 * requests with a WFA pair: count, p50, p99 (ms); mean pairs per SW batch,
 * per WFA batch; the submitting threads' fraction of the submit phase spent
 * inside the submit call and held back by max_outstanding; mean requests in
 * flight at a submit. */
int cgm_serve_run(void *engine, void *gen, double rate_rps, double duration_s,
                  int32_t max_pairs_per_request, int32_t max_outstanding, int32_t clients,
                  int32_t device_inputs, uint64_t seed, double *stats);

/* -------------------------------------------- host-pointer conveniences
 * Same as the launch entry points but every pointer is a host pointer; the
 * call copies in, runs on an internal stream and copies out before it
 * returns.  These back the per-pair CPU signatures smith_waterman_banded /
 * wfa_affine_cpu (align_cpu.mojo:229-233, :75-79) in the Mojo and Python
 * mirrors. */
int cgm_sw_banded_host(const uint64_t *q_packed, const uint64_t *r_packed,
                       const int32_t *q_len, const int32_t *r_len,
                       const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words,
                       const int32_t *band, int32_t band_scalar,
                       int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                       int32_t *out_score, int32_t *out_q_end, int32_t *out_r_end,
                       int32_t layout);
int cgm_wfa_affine_host(const uint64_t *q_packed, const uint64_t *r_packed,
                        const int32_t *q_len, const int32_t *r_len,
                        const int32_t *q_off, const int32_t *r_off,
                        int64_t n_pairs, int64_t q_words, int64_t r_words,
                        const int32_t *max_error, int32_t max_error_scalar,
                        int32_t x, int32_t o, int32_t e,
                        int32_t *out_cost, int32_t layout);

/* ------------------------------------------------------- ingest on the device
 * Replaces SequenceBatch.add_sequence (genomics/core/sequence.mojo:38-66)
 * with ascii_to_2bit (genomics/core/dna.mojo:26-40) for a whole batch at
 * once: sequence i is the seq_len[i] ASCII bytes at byte offset
 * seq_start[i] of `ascii` (ascii_bytes long).  Writes, in the order the
 * reference appends them: word_off[i] = running sum of ceil(len / 32)
 * (int32, as GenomicsDevice narrows offsets, genomics/gpu/device.mojo:
 * 82-108), lengths[i], and the packed / NMask words (nmask may be NULL).
 * A/a=0 C/c=1 G/g=2 T/t=3; every other byte is stored as 0 and flagged in
 * NMask (LSB = first base of the word); bits past a sequence's end are 0.
 * A sequence whose bytes fall outside the buffer gets lengths[i] = -1 and no
 * words.  *total_words (device int64, may be NULL) receives the number of
 * words the batch needs; only words below words_cap are written, so a call
 * with *total_words > words_cap did not fit (words_cap = n_seqs +
 * ascii_bytes / 32 always fits sequences that do not overlap).  Uses a
 * stream-ordered workspace (hipMallocAsync / hipFreeAsync on `stream`). */
int cgm_encode_ascii_launch(const uint8_t *ascii, int64_t ascii_bytes, const int64_t *seq_start,
                            const int32_t *seq_len, int64_t n_seqs, uint64_t *packed,
                            uint64_t *nmask, int64_t words_cap, int32_t *word_off,
                            int32_t *lengths, int64_t *total_words, void *stream);
/* Host-pointer form of the above (copies in and out; total_words is a
 * host pointer here). */
int cgm_encode_ascii_host(const uint8_t *ascii, int64_t ascii_bytes, const int64_t *seq_start,
                          const int32_t *seq_len, int64_t n_seqs, uint64_t *packed,
                          uint64_t *nmask, int64_t words_cap, int32_t *word_off, int32_t *lengths,
                          int64_t *total_words);

/* ------------------------------------------------------- synthetic batches
 * Seeded workload generator (replaces make_random_batch,
 * benchmarks/harness.mojo:55-67, whose queries equal their references).
 * Every pair is generated from splitmix64 keyed by (seed, workload, pair
 * index), so the batch does not depend on thread count or call order.
 * workload: one of the CGM_WL_* ids below.
 * kind[i]: 0 = SW pair, 1 = WFA pair.  For WFA pairs max_error[i] = planted
 * edit cost + x * (#N in the query) + 8, an upper bound of the optimum. */
enum {
    CGM_WL_SW_150BP = 1,     /* SW, 150 bp, band 32 */
    CGM_WL_SW_100_300BP = 2, /* batched SW, targets 100-300 bp, band 16-64 */
    CGM_WL_WFA_1_10KB = 3,   /* WFA, 1-10 kb */
    CGM_WL_MIXED = 4,        /* SW and WFA pairs, WFA up to 20 kb */
    CGM_WL_SERVING = 5       /* serving pair mix: SW band 64 below 2 kb, WFA above */
};
int cgm_gen_create(int32_t workload, int64_t n_pairs, uint64_t seed, void **handle);
int cgm_gen_sizes(void *handle, int64_t *n_pairs, int64_t *q_words, int64_t *r_words);
/* Copies the batch into host arrays (any pointer may be NULL to skip it):
 * packed words/NMask words (q_words / r_words entries), offsets and lengths
 * (n_pairs), kind, band and max_error (n_pairs). */
int cgm_gen_copy(void *handle, uint64_t *q_packed, uint64_t *q_nmask, int32_t *q_off,
                 int32_t *q_len, uint64_t *r_packed, uint64_t *r_nmask, int32_t *r_off,
                 int32_t *r_len, int32_t *kind, int32_t *band, int32_t *max_error);
int cgm_gen_destroy(void *handle);

#ifdef __cplusplus
}
#endif
#endif /* CGM_ALIGN_H */
