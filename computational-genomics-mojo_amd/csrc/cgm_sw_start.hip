// cgm_sw_start.hip -- start coordinates of banded SW alignments (gfx950).
//
// smith_waterman_banded reports query_start = ref_start = 0
// (genomics/cpu/align_cpu.mojo:302-305).  This file gives the start of the
// alignment whose end the reference reports: among the alignments of
// maximal score that end in H at (q_end, r_end), the one whose first
// diagonal step (the first query / reference base it pairs) is latest --
// largest query start, then largest reference start.
//
// Second pass after any SW layout has produced score and ends.  Every cell
// carries one packed value v = score * 2^K + key, key = qs * W + js, the
// start (qs, js) of the path the score belongs to in band-frame coordinates
// (W = 2 band + 1, reference start rs = qs - band + js).  A maximum over
// packed values is then the maximum over (score, qs, js): the score
// recurrence of align_cpu.mojo:271-289 with ties resolved towards the later
// start, in the same instructions as the score alone.  A cell's restart
// value "score 0, the path starts at the next diagonal step" is
// R(qi, j) = (qi + 1) * W + j, which beats every real path of score 0 there
// (their starts are at most qi) and turns into the start (qi + 1, j) when a
// diagonal step leaves it.  Leading gaps never matter: with
// mismatch <= 0, gap_extend <= 0 and gap_open + gap_extend <= 0 (required
// by every SW entry point) a path that begins with a gap scores at most 0
// until its first diagonal step and loses to the restart value.  Only cells
// up to (q_end, r_end) feed the end cell, so a pair stops at row q_end.
//
// Layouts: one pair per lane with the band-frame row in VGPRs (one exact
// frame per band 0..64, int32 packing when the key and score fit: (q_end + 2) W <= 2^K,
// score < 2^(30-K) - 1, |mismatch|, |open + extend|, |extend| <= 2^(29-K));
// one 128-thread workgroup per pair with int64 packing (K = 32) otherwise.

#include "cgm_internal.h"
#include "cgm_sw_dev.h"

namespace cgm {

namespace {

struct StartIO {
    const int32_t *score, *q_end, *r_end;
    int32_t *q_start, *r_start;
};

// 0-3: lanes for effective bands [0,16], [17,32], [33,48], [49,64];
// 4: one workgroup per pair
constexpr int kStartClasses = 5;
constexpr int kStartPlain = 4;
constexpr int kStartNone = 5;  // nothing to compute (written by the classifier)

// Key bits of the int32 packing for a pair, or -1 when it does not fit.
__device__ __forceinline__ int start_key_bits32(int score, int qe, int band, const SwScoring &sc) {
    const int64_t W = 2 * (int64_t)band + 1;
    const int64_t keymax = ((int64_t)qe + 2) * W;
    int K = 0;
    while (K < 31 && (int64_t(1) << K) < keymax) ++K;
    if (K > 28) return -1;
    // packed values stay below 2^30 + 2^29: score < 2^(30-K) - 1, so that
    // (match - mismatch) * 2^K fits too (match <= score for any score > 0)
    const int64_t smax = (int64_t(1) << (30 - K)) - 1;
    if ((int64_t)score >= smax) return -1;
    const int64_t pmax = int64_t(1) << (29 - K);
    const int64_t oe = (int64_t)sc.gap_open + sc.gap_extend;
    if (-(int64_t)sc.mismatch > pmax || -oe > pmax || -(int64_t)sc.gap_extend > pmax) return -1;
    return K;
}

__global__ void sw_start_class_kernel(Batch b, const int32_t *band_arr, int32_t band_scalar,
                                      SwScoring sc, StartIO io, uint32_t *keys, int32_t *idx,
                                      int32_t *counts) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in = i < b.n_pairs;
    int cls = kStartNone, be = 0, qe = 0;
    if (in) {
        const int s = io.score[i];
        const uint64_t *q, *r;
        int ql = 0, rl = 0;
        if (s == CGM_BAD_PAIR || !load_pair(b, i, q, r, ql, rl)) {
            io.q_start[i] = -1;
            io.r_start[i] = -1;
        } else if (s <= 0) {
            io.q_start[i] = 0;
            io.r_start[i] = 0;
        } else {
            be = effective_band(band_arr ? band_arr[i] : band_scalar, ql, rl);
            qe = io.q_end[i];
            if (be <= 64 && start_key_bits32(s, qe, be, sc) >= 0) {
                cls = be <= 16 ? 0 : be <= 32 ? 1 : be <= 48 ? 2 : 3;
            } else if (((int64_t)qe + 2) * (2 * (int64_t)be + 1) <= (int64_t(1) << 32)) {
                cls = kStartPlain;
            } else {
                io.q_start[i] = -1;  // start key beyond 32 bits (documented limit)
                io.r_start[i] = -1;
            }
        }
    }
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < kStartClasses; ++c) {
        const uint64_t m = __ballot(in && cls == c);
        if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&counts[c], __popcll(m));
    }
    if (!in) return;
    // classes in order, each by band then end row, largest first
    keys[i] = cls == kStartNone ? 0u
                                : ((uint32_t)(kStartClasses - cls) << 28) |
                                      ((uint32_t)min(be, 2047) << 17) |
                                      (uint32_t)min(qe, (1 << 17) - 1);
    idx[i] = (int32_t)i;
}

__global__ void start_ranges_kernel(const int32_t *counts, int32_t *ranges) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t s = 0;
    for (int c = 0; c < kStartClasses; ++c) {
        ranges[2 * c] = s;
        s += counts[c];
        ranges[2 * c + 1] = s;
    }
}

__global__ void zero_counts_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__device__ __forceinline__ int wave_max_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_min_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// v_max3_i32 with the restart value from an SGPR, as an opaque step: it
// keeps the cells of a row in order, so the scheduler cannot hoist the
// match bits of a whole chunk and run out of registers.
__device__ __forceinline__ int max3_restart(int a, int b, int r_uniform) {
    int d;
    asm volatile("v_max3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(r_uniform));
    return d;
}

// Packed scoring constants of one lane (all zero for a lane that only
// computes along).
struct StartConsts {
    int oe_sh, ext_sh, cmis_sh, diff_sh;
};

// One pass of the lane layout for band L (compile time): frame cells
// j = 0 .. 2L of row qi are (qi, qi - L + j), the whole frame in VGPRs
// (G = H + (open + extend), packed).  Cells outside the matrix get a forced
// mismatch; that is exact as in the register score layout (left of the
// matrix every value keeps score 0; cells right of it or below q_end are
// never read by the cells that feed the end cell), and G[2L + 1], the
// out-of-band upper neighbour of cell 2L, keeps score 0.  Returns the packed
// value of the lane's end cell (row qe, frame cell jend).
template <int L>
__device__ __forceinline__ int start_lane_pass(const uint64_t *q, int qlen, const uint64_t *r,
                                               int rlen, int rows, int rmin, int qe, int jend,
                                               bool act, const StartConsts &k) {
    using namespace swdev;
    constexpr int W = 2 * L + 1;
    constexpr int NCH = (W + 31) / 32;
    const int qwords = act ? (qlen + 31) >> 5 : 0;
    const int rwords = act ? (rlen + 31) >> 5 : 0;
    int G[W + 1];
#pragma unroll
    for (int j = 0; j <= W; ++j) G[j] = j + k.oe_sh;  // R(-1, j): next step starts (0, j)
    uint64_t qwin = 0;
    AlignedWords<NCH + 1> rw;
    rw.load(r, (-L) >> 5, rwords);
    int cap = 0;
    for (int qi = 0; qi < rows; ++qi) {
        const int rs0 = qi - L;
        if (qi > 0 && (rs0 & 31) == 0) rw.advance(r, rs0 >> 5, rwords);
        if ((qi & 31) == 0) qwin = ((qi >> 5) < qwords) ? q[qi >> 5] : 0ull;
        const uint64_t qrep = (qwin >> 62) * 0x5555555555555555ull;
        qwin <<= 2;
        const bool row_ok = act && qi < qlen;
        const int sh = 2 * (rs0 & 31);
        const int rbase = (qi + 1) * W;
        int Gl = k.oe_sh, El = -(1 << 30);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int rs = rs0 + 32 * c;
            const uint64_t m = chunk_mask(rw.window(c, sh), qrep, rs, rlen, row_ok, rmin);
            const uint32_t mh = (uint32_t)(m >> 32), ml = (uint32_t)m;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int j = 32 * c + u;
                if (j < W) {
                    const int sel = match_sel(u < 16 ? mh : ml, u, k.diff_sh);
                    const int dterm = G[j] + k.cmis_sh + sel;
                    const int e = max(El + k.ext_sh, G[j + 1]);
                    const int h = max3_restart(dterm, max(e, Gl), rbase + j);
                    G[j] = h + k.oe_sh;
                    Gl = G[j];
                    El = e;
                }
            }
        }
        const bool ends_here = act && qi == qe;
        if (__ballot(ends_here)) {  // rare: at most one row per lane
            const int jc = ends_here ? jend : -1;
#pragma unroll
            for (int j = 0; j < W; ++j) cap = (j == jc) ? G[j] : cap;
        }
    }
    return cap - k.oe_sh;
}

template <int LO, int HI>
__device__ __forceinline__ int start_lane_dispatch(int L, const uint64_t *q, int qlen,
                                                   const uint64_t *r, int rlen, int rows, int rmin,
                                                   int qe, int jend, bool act,
                                                   const StartConsts &k) {
    if constexpr (LO == HI) {
        return start_lane_pass<LO>(q, qlen, r, rlen, rows, rmin, qe, jend, act, k);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (L <= MID)
            return start_lane_dispatch<LO, MID>(L, q, qlen, r, rlen, rows, rmin, qe, jend, act, k);
        return start_lane_dispatch<MID + 1, HI>(L, q, qlen, r, rlen, rows, rmin, qe, jend, act, k);
    }
}

// One pair per lane, bands LO..HI.  The lanes of a wavefront are swept
// together once per distinct effective band among them (the sorted order
// makes that one pass almost always); lanes of another band compute along
// with zero scores and write nothing.
template <int LO, int HI, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void sw_start_lane_kernel(Batch b, const int32_t *band_arr,
                                                           int32_t band_scalar, SwScoring sc,
                                                           StartIO io, const int32_t *order,
                                                           const int32_t *range) {
    const int64_t start = range[0], end = range[1];
    for (int64_t base = start + (int64_t)blockIdx.x * 64; base < end;
         base += (int64_t)gridDim.x * 64) {
        const int64_t item = base + threadIdx.x;
        const bool valid = item < end;
        int64_t pair = 0;
        const uint64_t *q = b.q_packed, *r = b.r_packed;
        int qlen = 0, rlen = 0, L = -1, qe = 0, re = 0, K = 0, s = 0;
        if (valid) {
            pair = order[item];
            load_pair(b, pair, q, r, qlen, rlen);  // checked by the classifier
            L = effective_band(band_arr ? band_arr[pair] : band_scalar, qlen, rlen);
            qe = io.q_end[pair];
            re = io.r_end[pair];
            s = io.score[pair];
            K = start_key_bits32(s, qe, L, sc);
        }
        uint64_t pending = __ballot(valid);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int Lw = __builtin_amdgcn_readfirstlane(__shfl(L, lead, 64));
            const bool act = valid && L == Lw;
            pending &= ~__ballot(act);
            const int rows = wave_max_i32(act ? qe + 1 : 0);
            const int rmin = wave_min_i32(act && qlen > 0 ? rlen : INT32_MAX);
            const int Ks = act ? K : 0;
            StartConsts k;
            k.oe_sh = act ? (sc.gap_open + sc.gap_extend) * (1 << Ks) : 0;
            k.ext_sh = act ? sc.gap_extend * (1 << Ks) : 0;
            k.cmis_sh = act ? sc.mismatch * (1 << Ks) - k.oe_sh : 0;
            k.diff_sh = act ? (sc.match - sc.mismatch) * (1 << Ks) : 0;
            const int v = start_lane_dispatch<LO, HI>(Lw, q, qlen, r, rlen, rows, rmin, qe,
                                                      re - qe + Lw, act, k);
            if (act) {
                const int W = 2 * Lw + 1;
                const int key = v & ((1 << K) - 1);
                const int qs = key / W;
                const int js = key - qs * W;
                const bool ok = (v >> K) == s;
                io.q_start[pair] = ok ? qs : -1;
                io.r_start[pair] = ok ? qs - Lw + js : -1;
            }
        }
    }
}

// One 128-thread workgroup per pair, int64 packing (K = 32), anti-diagonal
// sweep over the rectangle [0, q_end] x [0, r_end] of the band: three
// anti-diagonals of H and two of E in LDS, thread j owning band diagonal
// j - band.  Cells outside the matrix hold the restart value and E = 0, as
// the reference's skipped cells hold H = 0, E = 0.
constexpr int kStartPlainThreads = 128;

__global__ __launch_bounds__(kStartPlainThreads) void sw_start_plain_kernel(
    Batch b, const int32_t *band_arr, int32_t band_scalar, SwScoring sc, StartIO io,
    const int32_t *order, const int32_t *range) {
    extern __shared__ int64_t slds[];
    const int tid = threadIdx.x;
    const int64_t start = range[0], end = range[1];
    const int64_t sh32 = int64_t(1) << 32;
    const int64_t oe_sh = (int64_t)(sc.gap_open + sc.gap_extend) * sh32;
    const int64_t ext_sh = (int64_t)sc.gap_extend * sh32;
    const int64_t neg = -(int64_t(1) << 62);
    for (int64_t item = start + blockIdx.x; item < end; item += gridDim.x) {
        const int64_t pair = order[item];
        const uint64_t *q, *r;
        int qlen, rlen;
        load_pair(b, pair, q, r, qlen, rlen);  // checked by the classifier
        const int band = effective_band(band_arr ? band_arr[pair] : band_scalar, qlen, rlen);
        const int qe = io.q_end[pair], re = io.r_end[pair];
        const int W = 2 * band + 1;
        int64_t *H[3] = {slds, slds + W, slds + 2 * W};
        int64_t *E[2] = {slds + 3 * W, slds + 4 * W};
        __syncthreads();  // the previous pair's last reads are done
        for (int j = tid; j < W; j += kStartPlainThreads) {
            // anti-diagonals -1 and -2: cells (qi, ri) with qi = -1 or ri = -1
            for (int t = -2; t <= -1; ++t) {
                const int k = j - band;
                const int qi = (t - k) >> 1;
                H[(t + 3) % 3][j] = (int64_t)max(qi + 1, 0) * W + j;
            }
            E[1][j] = 0;
        }
        __syncthreads();
        const int steps = qe + re + 1;
        int64_t endv = 0;
        for (int t = 0; t < steps; ++t) {
            int64_t *Hc = H[t % 3], *Hp = H[(t + 2) % 3], *Hpp = H[(t + 1) % 3];
            int64_t *Ec = E[t & 1], *Ep = E[(t + 1) & 1];
            for (int j = tid; j < W; j += kStartPlainThreads) {
                const int k = j - band;
                if ((t + k) & 1) continue;
                const int qi = (t - k) >> 1, ri = (t + k) >> 1;
                if (qi < 0 || qi >= qlen || ri < 0 || ri >= rlen || qi > qe || ri > re) {
                    Hc[j] = (int64_t)max(qi + 1, 0) * W + j;
                    Ec[j] = 0;
                    continue;
                }
                const int sub = base_at(q, qi) == base_at(r, ri) ? sc.match : sc.mismatch;
                const int64_t hd = Hpp[j] + (int64_t)sub * sh32;
                const int64_t hu = (j + 1 < W) ? Hp[j + 1] : 0;
                const int64_t hl = (j > 0) ? Hp[j - 1] : 0;
                const int64_t el = (j > 0) ? Ep[j - 1] + ext_sh : neg;
                const int64_t e = max(el, hu + oe_sh);
                const int64_t rv = (int64_t)(qi + 1) * W + j;
                const int64_t h = max(max(rv, hd), max(e, hl + oe_sh));
                Hc[j] = h;
                Ec[j] = e;
                if (qi == qe && ri == re) endv = h;
            }
            __syncthreads();
        }
        const int jend = re - qe + band;
        if (tid == (jend & (kStartPlainThreads - 1)) && jend >= 0) {
            const int64_t key = endv & 0xffffffffll;
            const int64_t qs = key / W, js = key - qs * W;
            const bool ok = (endv >> 32) == io.score[pair];
            io.q_start[pair] = ok ? (int32_t)qs : -1;
            io.r_start[pair] = ok ? (int32_t)(qs - band + js) : -1;
        }
    }
}

}  // namespace

int launch_sw_start(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                    const int32_t *score, const int32_t *q_end, const int32_t *r_end,
                    int32_t *q_start, int32_t *r_start, hipStream_t stream) {
    if (b.n_pairs == 0) return CGM_OK;
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    const StartIO io = {score, q_end, r_end, q_start, r_start};
    uint32_t *keys = nullptr;
    int32_t *idx = nullptr, *counts = nullptr, *order = nullptr;
    CGM_HIP(hipMallocAsync((void **)&keys, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&idx, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&order, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&counts, 64 + 2 * kStartClasses * 4, stream));
    int32_t *ranges = counts + 16;
    hipLaunchKernelGGL(zero_counts_kernel, dim3(1), dim3(64), 0, stream, counts, 16);
    const unsigned blocks = (unsigned)((b.n_pairs + 255) / 256);
    hipLaunchKernelGGL(sw_start_class_kernel, dim3(blocks), dim3(256), 0, stream, b, band,
                       band_scalar, sc, io, keys, idx, counts);
    hipLaunchKernelGGL(start_ranges_kernel, dim3(1), dim3(64), 0, stream, counts, ranges);
    int rc = sort_desc_by_key(keys, idx, order, b.n_pairs, stream);
    if (!rc) {
        const int64_t waves = (b.n_pairs + 63) / 64;
        const int grid = (int)(waves < (int64_t)cus * 8 ? waves : (int64_t)cus * 8);
        // the band buckets are independent: side by side on forked streams,
        // each held to the waves per SIMD its registers allow
        hipStream_t aux[4];
        rc = fork_streams(stream, 4, aux);
        if (!rc) {
#define CGM_START(K, LO, HI, WPE)                                                              \
    hipLaunchKernelGGL((sw_start_lane_kernel<LO, HI, WPE>), dim3(grid), dim3(64), 0, aux[3 - (K)], \
                       b, band, band_scalar, sc, io, order, ranges + 2 * (K));
            CGM_START(3, 49, 64, 2)
            CGM_START(2, 33, 48, 3)
            CGM_START(1, 17, 32, 4)
            CGM_START(0, 0, 16, 5)
#undef CGM_START
            rc = join_streams(stream, 4, aux);
        }
        const size_t lds = (size_t)5 * (2 * kPlainMaxBand + 1) * sizeof(int64_t);
        if (!rc) rc = set_kernel_lds((const void *)sw_start_plain_kernel, (int)lds);
        if (!rc) {
            hipLaunchKernelGGL(sw_start_plain_kernel, dim3(cus * 2), dim3(kStartPlainThreads), lds,
                               stream, b, band, band_scalar, sc, io, order,
                               ranges + 2 * kStartPlain);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = hip_error(e, "sw_start kernels");
        }
    }
    hipError_t f1 = hipFreeAsync(keys, stream), f2 = hipFreeAsync(idx, stream);
    hipError_t f3 = hipFreeAsync(order, stream), f4 = hipFreeAsync(counts, stream);
    if (rc) return rc;
    CGM_HIP(f1);
    CGM_HIP(f2);
    CGM_HIP(f3);
    CGM_HIP(f4);
    return CGM_OK;
}

}  // namespace cgm

using namespace cgm;

extern "C" int cgm_sw_banded_start_launch(
    const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
    const int32_t *r_len, const int32_t *q_off, const int32_t *r_off, int64_t n_pairs,
    int64_t q_words, int64_t r_words, const int32_t *band, int32_t band_scalar, int32_t match,
    int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t *out_score,
    int32_t *out_q_start, int32_t *out_r_start, int32_t *out_q_end, int32_t *out_r_end,
    int32_t layout, void *stream) {
    if (n_pairs > 0 && (!out_q_start || !out_r_start))
        return set_error(CGM_ERR_INVALID, "null start output array");
    if (int rc = cgm_sw_banded_launch(q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs,
                                      q_words, r_words, band, band_scalar, match, mismatch,
                                      gap_open, gap_extend, out_score, out_q_end, out_r_end,
                                      layout, stream))
        return rc;
    const Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    const SwScoring sc = {match, mismatch, gap_open, gap_extend};
    return launch_sw_start(b, band, band_scalar, sc, out_score, out_q_end, out_r_end, out_q_start,
                           out_r_start, (hipStream_t)stream);
}
