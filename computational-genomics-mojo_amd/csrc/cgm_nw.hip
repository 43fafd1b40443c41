// cgm_nw.hip -- banded Needleman-Wunsch (global) kernels for gfx950.
//
// Semantics of needleman_wunsch_banded (genomics/cpu/align_cpu.mojo:
// 312-377), with its two rows taken to hold 2*band+1 entries (the function
// allocates them with List(capacity=width) and indexes them directly).
// Row qi covers ri = qi - band + j, j in [0, 2 band]; the row before the
// first is H(-1, c) = 0 for c = -1, gap_open + (c + 1) gap_extend for
// 0 <= c < band and -1000000 otherwise.  An in-matrix cell computes
//     h = max(H(qi-1, ri-1) + sub, H(qi-1, ri) + go + ge, H(qi, ri-1) + go + ge)
// (sub = match or mismatch; every gap base costs go + ge), cells with ri < 0
// or ri >= rlen hold exactly -1000000 and neighbours outside the frame read
// -1000000.  The result is H(qlen-1, rlen-1) when |rlen - qlen| <= band,
// else -1000000, with query_end = qlen - 1 and ref_end = rlen - 1.  A band
// wider than both sequences adds no cell and no first-row entry any cell
// reads, so the kernels use min(band, max(qlen, rlen)).
//
// Layouts: one pair per lane with the row in VGPRs (band <= 64, four band
// buckets), and one 128-thread workgroup per pair sweeping anti-diagonals
// (any band up to kPlainMaxBand; also the baseline layout).

#include "cgm_sw_dev.h"

namespace cgm {

namespace {

constexpr int kNwNeg = -1000000;  // align_cpu.mojo:322
constexpr int kNwClasses = 6;     // 0..3 register buckets, 4 plain, 5 bad
constexpr int kNwPlain = 4, kNwBad = 5;
constexpr int kNwPlainThreads = 128;

// Every value a cell can take stays within (qlen + rlen + 2 band + 4) *
// (|match| + |mismatch| + |go| + |ge|) of 0 or of -1000000; the kernels
// compute in int32, so pairs beyond that bound are reported bad.
__device__ __forceinline__ bool nw_pair_ok(int qlen, int rlen, int band, const SwScoring &sc) {
    if (band < 0 || band > kPlainMaxBand) return false;
    const int64_t mag = (int64_t)abs(sc.match) + abs(sc.mismatch) + abs(sc.gap_open) +
                        abs(sc.gap_extend);
    return ((int64_t)qlen + rlen + 2 * (int64_t)band + 4) * mag + 1000000 <= (int64_t(1) << 30);
}

// H(-1, c) of the row before the first, c = -1 .. band - 1.
__device__ __forceinline__ int nw_first_row(int c, int band, const SwScoring &sc) {
    if (c < -1 || c >= band) return kNwNeg;
    return c == -1 ? 0 : sc.gap_open + (c + 1) * sc.gap_extend;
}

__device__ __forceinline__ void nw_write_bad(const SwOut &o, int64_t i) {
    o.score[i] = CGM_BAD_PAIR;
    o.q_end[i] = -1;
    o.r_end[i] = -1;
}

// Sign-extended bit u of a 32-bit mask (0 / -1) masked to `val`, one
// v_bfe_i32 with the offset in an SGPR (a constant offset would let the
// compiler turn the extraction into a compare and select).
__device__ __forceinline__ int mask_bit_sel(uint32_t m, int u, int val) {
    int v;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(v) : "v"(m), "s"(u));
    return v & val;
}

// ---------------------------------------------------------------------------
// Register lanes: one pair per lane, row-major like the reference, the row
// kept as G = H + (go + ge) in VGPRs.  Per cell: the substitution from the
// row's 32-base match mask (the reference held as 1-bit planes, one bit per
// base), max3 over the three predecessors, and the in-matrix mask that
// forces -1000000 on cells the reference skips.  A lane whose
// query is exhausted skips the remaining rows of its wavefront (divergent
// branch), so its row still holds row qlen - 1 at the end.
// ---------------------------------------------------------------------------
template <int B>
__device__ __forceinline__ int nw_reg_lane(const uint64_t *q, int qlen, const uint64_t *r,
                                           int rlen, int rows, const SwScoring &sc) {
    constexpr int W = 2 * B + 1;
    constexpr int NCH = (W + 31) / 32;
    const int gap = sc.gap_open + sc.gap_extend;
    const int cmis = sc.mismatch - gap;
    const int diff = sc.match - sc.mismatch;
    const int negG = kNwNeg + gap;
    const int qwords = (qlen + 31) >> 5, rwords = (rlen + 31) >> 5;
    int G[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const int ri = j - B;  // row -1 entry read as diagonal of (0, ri)
        G[j] = (ri < 0 ? kNwNeg : ri == 0 ? 0 : sc.gap_open + ri * sc.gap_extend) + gap;
    }
    uint64_t qwin = 0;
    swdev::AlignedPlanes<NCH + 1> rw;
    rw.load(r, (-B) >> 5, rwords);
    for (int qi = 0; qi < rows; ++qi) {
        const int rs0 = qi - B;
        if (qi > 0 && (rs0 & 31) == 0) rw.advance(r, rs0 >> 5, rwords);
        if ((qi & 31) == 0) qwin = ((qi >> 5) < qwords) ? q[qi >> 5] : 0ull;
        const int qb = (int)(qwin >> 62);
        qwin <<= 2;
        const uint32_t qlo = (qb & 1) ? ~0u : 0u, qhi = (qb & 2) ? ~0u : 0u;
        const int o = rs0 & 31;
        if (qi < qlen) {
            int Gl = negG;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const int rs = rs0 + 32 * c;
                const uint32_t range = swdev::range_mask32(-rs, rlen - rs);
                uint32_t wl, wh;
                rw.window(c, o, wl, wh);
                const uint32_t m = ~((wl ^ qlo) | (wh ^ qhi)) & range;
#pragma unroll
                for (int u = 0; u < 32; ++u) {
                    const int j = 32 * c + u;
                    if (j < W) {
                        const int sel = mask_bit_sel(m, u, diff);
                        const int inside = mask_bit_sel(range, u, -1);
                        const int Gu = (j + 1 < W) ? G[j + 1] : negG;
                        const int dterm = G[j] + cmis + sel;
                        // ordered (volatile) max3: otherwise the substitution and
                        // mask extractions of the whole row are hoisted ahead of
                        // the dependency chain and spill
                        int h = swdev::max3_chain(dterm, Gu, Gl);
                        h = (h & inside) | (kNwNeg & ~inside);
                        G[j] = h + gap;
                        Gl = G[j];
                    }
                }
            }
        }
    }
    // H(qlen-1, rlen-1) sits at j = rlen - qlen + B of row qlen - 1
    const int jf = rlen - qlen + B;
    int score = kNwNeg;
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (j == jf) score = G[j] - gap;
    return score;
}

template <int LO, int HI>
__device__ __forceinline__ int nw_reg_dispatch(int band, const uint64_t *q, int qlen,
                                               const uint64_t *r, int rlen, int rows,
                                               const SwScoring &sc) {
    if constexpr (LO == HI) {
        return nw_reg_lane<LO>(q, qlen, r, rlen, rows, sc);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (band <= MID) return nw_reg_dispatch<LO, MID>(band, q, qlen, r, rlen, rows, sc);
        return nw_reg_dispatch<MID + 1, HI>(band, q, qlen, r, rlen, rows, sc);
    }
}

constexpr uint32_t kNwSplitPercent = 55;  // groups in the long queue
constexpr int kNwCtlSlots = 22;           // ctl offset of the per-SIMD counters
constexpr int kNwSimdKeys = 8192;         // XCC (3 bits) x CU (8) x SIMD (2)

// Register lanes over order[range[0] .. range[1]) (every pair there has an
// effective band <= 64, sorted by band then length, largest first): each
// wavefront takes groups of 64 entries by itself from a long queue (the first
// kNwSplitPercent % of the groups) and a short queue (the rest), both
// largest first.  Of two wavefronts on one SIMD the older one gets most
// issue slots, so the first one started on a SIMD takes the long queue and
// then the short one, the second only the short one: the widest groups start
// first, never on a starved wavefront, and the last groups of the launch are
// the smallest.  `ctl` holds the 64-bit counter (long-queue claims in the low
// half, short-queue claims in the high half) and, from ctl + 22, one
// wavefront counter per SIMD; a group holding two bands runs once per band.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) void nw_queue_kernel(
    Batch b, const int32_t *band_arr, int32_t band_scalar, SwScoring sc, SwOut out,
    const int32_t *order, const int32_t *range, int32_t *ctl) {
    const int64_t start = range[0], end = range[1];
    const int lane = threadIdx.x & 63;
    const uint32_t units = (uint32_t)((end - start + 63) / 64);
    const uint32_t split = (uint32_t)(((uint64_t)units * kNwSplitPercent) / 100u);
    unsigned long long *ends = (unsigned long long *)ctl;
    int slot = 0;
    if (lane == 0) {
        const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);    // HW_ID
        const uint32_t xcc = __builtin_amdgcn_s_getreg((15 << 11) | 20);  // XCC_ID
        const uint32_t key = ((xcc & 7u) << 10) | (((hw >> 8) & 0xffu) << 2) | ((hw >> 4) & 3u);
        slot = atomicAdd(ctl + kNwCtlSlots + key, 1);
    }
    const bool second = (__shfl(slot, 0, 64) & 1) != 0;
    for (;;) {
        int u = -1;
        if (lane == 0) {
            if (!second) {
                const uint32_t a = (uint32_t)atomicAdd(ends, 1ull);
                if (a < split) u = (int)a;
            }
            if (u < 0) {
                const uint32_t c = (uint32_t)(atomicAdd(ends, 1ull << 32) >> 32);
                if ((uint64_t)split + c < units) u = (int)(split + c);
            }
        }
        u = __shfl(u, 0, 64);
        if (u < 0) break;
        const int64_t base = start + 64 * (int64_t)u;
        const int count = (int)min((int64_t)64, end - base);
        const bool active = lane < count;
        const int64_t pair = active ? order[base + lane] : 0;
        const uint64_t *q = b.q_packed, *r = b.r_packed;
        int qlen = 0, rlen = 0, band = -1;
        if (active && load_pair(b, pair, q, r, qlen, rlen))
            band = effective_band(band_arr ? band_arr[pair] : band_scalar, qlen, rlen);
        uint64_t pending = __ballot(band >= 0);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int bsel = __shfl(band, lead, 64);
            const bool mine = band == bsel;
            pending &= ~__ballot(mine);
            int rows = mine ? qlen : 0;
            for (int o = 32; o > 0; o >>= 1) rows = max(rows, __shfl_xor(rows, o, 64));
            const int s = nw_reg_dispatch<0, 64>(bsel, q, mine ? qlen : 0, r, mine ? rlen : 0,
                                                 rows, sc);
            if (mine) {
                out.score[pair] = s;
                out.q_end[pair] = qlen - 1;
                out.r_end[pair] = rlen - 1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// One workgroup per pair: anti-diagonal sweep of the band (threads own band
// diagonals j = ri - qi + band), three anti-diagonals of H in LDS, cells
// outside the matrix written as -1000000, row 0 reading the row before the
// first from its closed form.
// ---------------------------------------------------------------------------
__device__ void nw_plain_pair(const Batch &b, int64_t pair, const int32_t *band_arr,
                              int32_t band_scalar, const SwScoring &sc, const SwOut &out,
                              int32_t *lds, int *result) {
    const int tid = threadIdx.x;
    const uint64_t *q, *r;
    int qlen, rlen;
    const bool ok = load_pair(b, pair, q, r, qlen, rlen);
    int band = band_arr ? band_arr[pair] : band_scalar;
    if (ok && band >= 0) band = effective_band(band, qlen, rlen);
    if (!ok || !nw_pair_ok(qlen, rlen, band, sc)) {
        if (tid == 0) nw_write_bad(out, pair);
        return;
    }
    const int W = 2 * band + 1;
    const int gap = sc.gap_open + sc.gap_extend;
    int32_t *H[3] = {lds, lds + W, lds + 2 * W};
    for (int j = tid; j < 3 * W; j += blockDim.x) lds[j] = kNwNeg;
    if (tid == 0) *result = qlen == 0 ? nw_first_row(rlen - 1, band, sc) : kNwNeg;
    __syncthreads();
    const int steps = (qlen > 0 && rlen > 0) ? qlen + rlen - 1 : 0;
    for (int t = 0; t < steps; ++t) {
        int32_t *Hc = H[t % 3], *Hp = H[(t + 2) % 3], *Hpp = H[(t + 1) % 3];
        for (int j = tid; j < W; j += blockDim.x) {
            const int k = j - band;
            if ((t + k) & 1) continue;  // no cell of this diagonal on anti-diagonal t
            const int qi = (t - k) >> 1, ri = (t + k) >> 1;
            if (qi < 0 || qi >= qlen || ri < 0 || ri >= rlen) {
                Hc[j] = kNwNeg;
                continue;
            }
            const int sub = base_at(q, qi) == base_at(r, ri) ? sc.match : sc.mismatch;
            const int hd = qi == 0 ? nw_first_row(ri - 1, band, sc) : Hpp[j];
            const int hu = qi == 0 ? nw_first_row(ri, band, sc) : (j + 1 < W ? Hp[j + 1] : kNwNeg);
            const int hl = j > 0 ? Hp[j - 1] : kNwNeg;
            const int h = max(hd + sub, max(hu + gap, hl + gap));
            Hc[j] = h;
            if (qi == qlen - 1 && ri == rlen - 1) *result = h;
        }
        __syncthreads();
    }
    if (tid == 0) {
        out.score[pair] = *result;
        out.q_end[pair] = qlen - 1;
        out.r_end[pair] = rlen - 1;
    }
    __syncthreads();
}

// Pairs order[range[0] + i] (or pair i < n_pairs when order is NULL), one
// workgroup per pair, grid-stride.
__global__ __launch_bounds__(kNwPlainThreads) void nw_plain_kernel(Batch b, const int32_t *band_arr,
                                                                   int32_t band_scalar, SwScoring sc,
                                                                   SwOut out, const int32_t *order,
                                                                   const int32_t *range) {
    extern __shared__ int32_t lds[];
    __shared__ int result;
    const int64_t start = order ? range[0] : 0, end = order ? range[1] : b.n_pairs;
    for (int64_t i = start + blockIdx.x; i < end; i += gridDim.x)
        nw_plain_pair(b, order ? order[i] : i, band_arr, band_scalar, sc, out, lds, &result);
}

__global__ void nw_class_kernel(Batch b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                                SwOut out, uint32_t *keys, int32_t *idx, int32_t *counts) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in = i < b.n_pairs;
    const uint64_t *q, *r;
    int ql = 0, rl = 0, be = 0;
    int cls = in ? kNwBad : -1;
    if (in && load_pair(b, i, q, r, ql, rl)) {
        const int bd = band ? band[i] : band_scalar;
        if (bd >= 0) {
            be = effective_band(bd, ql, rl);
            if (nw_pair_ok(ql, rl, be, sc))
                cls = be <= 16 ? 0 : be <= 32 ? 1 : be <= 48 ? 2 : be <= 64 ? 3 : kNwPlain;
        }
    }
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < kNwClasses; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&counts[c], __popcll(m));
    }
    if (!in) return;
    if (cls == kNwBad) nw_write_bad(out, i);
    const uint32_t len = (uint32_t)min(max(ql, rl), (1 << 17) - 1);
    // the register classes share the top key, so they sort by band then
    // length, widest first (classes 3, 2, 1, 0), then plain, then bad
    const uint32_t top = cls <= 3 ? 15u : (uint32_t)(18 - cls);
    keys[i] = (top << 28) | ((uint32_t)min(be, 2047) << 17) | len;
    idx[i] = (int32_t)i;
}

// Class ranges in the sorted order: register classes 3, 2, 1, 0, then
// plain, then bad.
__global__ void nw_ranges_kernel(int32_t *counts, int32_t *ranges) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int seq[kNwClasses] = {3, 2, 1, 0, kNwPlain, kNwBad};
    int32_t s = 0;
    for (int i = 0; i < kNwClasses; ++i) {
        const int c = seq[i];
        ranges[2 * c] = s;
        s += counts[c];
        ranges[2 * c + 1] = s;
    }
}

// [start, end) of the register classes 3..0, which follow each other in the
// sorted order.
__global__ void nw_range_join_kernel(const int32_t *ranges, int32_t *reg_range) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    reg_range[0] = ranges[2 * 3];
    reg_range[1] = ranges[2 * 0 + 1];
}

__global__ void nw_zero_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

int launch_nw_plain(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                    SwOut out, const int32_t *order, const int32_t *range, int grid,
                    int32_t band_cap, hipStream_t stream) {
    const size_t lds_bytes = (size_t)3 * (2 * band_cap + 1) * sizeof(int32_t);
    hipLaunchKernelGGL(nw_plain_kernel, dim3(grid), dim3(kNwPlainThreads), lds_bytes, stream, b,
                       band, band_scalar, sc, out, order, range);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

// Classify, sort by (class, band, length), then the register buckets side
// by side on forked streams and the plain layout for wider bands.
int launch_nw_auto(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                   SwOut out, hipStream_t stream) {
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    const int64_t n = b.n_pairs;
    uint32_t *keys = nullptr;
    int32_t *idx = nullptr, *order = nullptr, *small = nullptr;
    CGM_HIP(hipMallocAsync((void **)&keys, (size_t)n * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&idx, (size_t)n * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&order, (size_t)n * 4, stream));
    // counts, ranges, the joined register range (40, 41), the queue control
    // block (from 42: 64-bit counter, then one wavefront counter per SIMD)
    const int small_words = 42 + kNwCtlSlots + kNwSimdKeys;
    CGM_HIP(hipMallocAsync((void **)&small, (size_t)small_words * 4, stream));
    int32_t *counts = small, *ranges = small + 16;
    hipLaunchKernelGGL(nw_zero_kernel, dim3(1), dim3(256), 0, stream, small, small_words);
    hipLaunchKernelGGL(nw_class_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, b,
                       band, band_scalar, sc, out, keys, idx, counts);
    hipLaunchKernelGGL(nw_ranges_kernel, dim3(1), dim3(64), 0, stream, counts, ranges);
    int rc = sort_desc_by_key(keys, idx, order, n, stream);
    // register classes 3, 2, 1, 0 (widest band first) are one contiguous
    // range of the sorted order: one queue over all of them
    const int64_t waves = (n + 63) / 64;
    const int grid = (int)(waves < (int64_t)cus * 8 ? waves : (int64_t)cus * 8);
    if (!rc) {
        hipLaunchKernelGGL(nw_range_join_kernel, dim3(1), dim3(64), 0, stream, ranges, small + 40);
        hipLaunchKernelGGL(nw_queue_kernel, dim3(grid), dim3(64), 0, stream, b, band, band_scalar,
                           sc, out, order, small + 40, small + 42);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_error(e, "nw register kernel");
    }
    if (!rc)
        rc = launch_nw_plain(b, band, band_scalar, sc, out, order, ranges + 2 * kNwPlain, cus * 2,
                             kPlainMaxBand, stream);
    (void)hipFreeAsync(keys, stream);
    (void)hipFreeAsync(idx, stream);
    (void)hipFreeAsync(order, stream);
    (void)hipFreeAsync(small, stream);
    return rc;
}

}  // namespace

}  // namespace cgm

using namespace cgm;

extern "C" {

int cgm_nw_banded_launch(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                         const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                         int64_t n_pairs, int64_t q_words, int64_t r_words, const int32_t *band,
                         int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                         int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                         int32_t *out_r_end, int32_t layout, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    const Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    const SwScoring sc = {match, mismatch, gap_open, gap_extend};
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_score_magnitudes(sc)) return rc;
    if (n_pairs == 0) return CGM_OK;
    if (!out_score || !out_q_end || !out_r_end)
        return set_error(CGM_ERR_INVALID, "nw: null output buffer");
    const SwOut out = {out_score, out_q_end, out_r_end};
    if (layout == CGM_LAYOUT_PLAIN) {
        int32_t max_band = 0;
        if (int rc = device_max_band_eff(b, band, band_scalar, stream, &max_band)) return rc;
        max_band = max_band < 0 ? 0 : max_band > kPlainMaxBand ? kPlainMaxBand : max_band;
        return launch_nw_plain(b, band, band_scalar, sc, out, nullptr, nullptr,
                               (int)(n_pairs < (1 << 30) ? n_pairs : (1 << 30)), max_band, stream);
    }
    if (layout != CGM_LAYOUT_AUTO && layout != CGM_LAYOUT_LANE)
        return set_error(CGM_ERR_INVALID, "nw: unknown layout");
    return launch_nw_auto(b, band, band_scalar, sc, out, stream);
}

int cgm_nw_banded_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                       const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words, const int32_t *band,
                       int32_t band_scalar, int32_t match, int32_t mismatch, int32_t gap_open,
                       int32_t gap_extend, int32_t *out_score, int32_t *out_q_end,
                       int32_t *out_r_end, int32_t layout) {
    if (n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (n_pairs == 0) return CGM_OK;
    hipStream_t s = nullptr;
    CGM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        void *bufs[10];
        int n;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            for (int i = 0; i < n; ++i) (void)hipFree(bufs[i]);
            (void)hipStreamDestroy(s);
        }
    } g{s, {}, 0};
    auto up = [&](const void *src, size_t bytes, void **d) -> int {
        CGM_HIP(hipMalloc(d, bytes > 0 ? bytes : 8));
        g.bufs[g.n++] = *d;
        if (src && bytes) CGM_HIP(hipMemcpyAsync(*d, src, bytes, hipMemcpyHostToDevice, s));
        return CGM_OK;
    };
    void *qp, *rp, *ql, *rl, *qo, *ro, *bd = nullptr, *os, *oq, *orr;
    const size_t n4 = (size_t)n_pairs * 4;
    int rc = CGM_OK;
    if ((rc = up(q_packed, (size_t)q_words * 8, &qp)) || (rc = up(r_packed, (size_t)r_words * 8, &rp)) ||
        (rc = up(q_len, n4, &ql)) || (rc = up(r_len, n4, &rl)) || (rc = up(q_off, n4, &qo)) ||
        (rc = up(r_off, n4, &ro)) || (band && (rc = up(band, n4, &bd))) ||
        (rc = up(nullptr, n4, &os)) || (rc = up(nullptr, n4, &oq)) || (rc = up(nullptr, n4, &orr)))
        return rc;
    if ((rc = cgm_nw_banded_launch((const uint64_t *)qp, (const uint64_t *)rp, (const int32_t *)ql,
                                   (const int32_t *)rl, (const int32_t *)qo, (const int32_t *)ro,
                                   n_pairs, q_words, r_words, (const int32_t *)bd, band_scalar,
                                   match, mismatch, gap_open, gap_extend, (int32_t *)os,
                                   (int32_t *)oq, (int32_t *)orr, layout, s)))
        return rc;
    CGM_HIP(hipMemcpyAsync(out_score, os, n4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipMemcpyAsync(out_q_end, oq, n4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipMemcpyAsync(out_r_end, orr, n4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipStreamSynchronize(s));
    return CGM_OK;
}

}  // extern "C"
