// cgm_sw_lane.hip -- register-resident one-pair-per-lane banded SW (gfx950).
//
// Each lane aligns one pair in the reference's own row-major order
// (genomics/cpu/align_cpu.mojo:259-300), with the band as a compile-time
// parameter B so the whole previous row lives in VGPRs (W = 2B+1 values, no
// LDS).  Pairs of one wavefront are sorted to share B and have similar
// lengths; a wavefront whose lanes carry several bands runs once per band.
//
// Per row, the lane sweeps the full band frame j = 0..W-1 (ri = qi - B + j)
// without per-cell bounds tests.  Cells outside the matrix get a forced
// mismatch through the 32-base match mask; with mismatch <= 0,
// gap_extend <= 0 and gap_open + gap_extend <= 0 this is exact:
//  * left of ri = 0 every value stays H = 0 and E <= 0, and an E boundary
//    value v with v + gap_extend <= 0 leaves every H unchanged;
//  * right of ri = rlen - 1 and below qi = qlen - 1 every H is reached from
//    an earlier cell through non-positive steps, so it never exceeds the
//    running maximum (ties lose to earlier cells: strict `>`), and no
//    in-matrix cell reads those cells.
// The row keeps G = H + (gap_open + gap_extend) so the up and left terms
// need no add; the row maximum is found with one max per cell on the key
// (h << 8) | (64 - j mod 64) per group of 64 cells (largest h, then
// smallest j; the low byte stays an inline constant), then compared with
// strict `>` across rows.  Requires h < 2^23.

#include "cgm_internal.h"

namespace cgm {

namespace {

constexpr int kNegE = -1000000;  // align_cpu.mojo:282

// Bases pos .. pos+31 MSB-first; words outside [0, nwords) read as 0.
__device__ __forceinline__ uint64_t window_signed(const uint64_t *s, int pos, int nwords) {
    const int w = pos >> 5;  // floor division, also for pos < 0
    const int sh = 2 * (pos & 31);
    const uint64_t lo = (w >= 0 && w < nwords) ? s[w] : 0ull;
    if (!sh) return lo;
    const uint64_t hi = (w + 1 >= 0 && w + 1 < nwords) ? s[w + 1] : 0ull;
    return (lo << sh) | (hi >> (64 - sh));
}

// Even-bit mask (base u at bit 62 - 2u) of window bases u in [lo, hi).
__device__ __forceinline__ uint64_t base_range_mask(int lo, int hi) {
    lo = min(max(lo, 0), 32);
    hi = min(max(hi, 0), 32);
    if (lo >= hi) return 0ull;
    const uint64_t m1 = ~0ull >> (2 * lo);
    const uint64_t m2 = hi == 32 ? ~0ull : (~0ull << (64 - 2 * hi));
    return m1 & m2 & 0x5555555555555555ull;
}

struct LaneResult {
    int score, q_end, r_end;
};

template <int B>
__device__ __forceinline__ LaneResult sw_reg_lane(const uint64_t *q, int qlen, const uint64_t *r,
                                                  int rlen, int rows, const SwScoring &sc) {
    constexpr int W = 2 * B + 1;
    constexpr int NCH = (W + 31) / 32;
    const int oe = sc.gap_open + sc.gap_extend;
    const int ext = sc.gap_extend;
    const int cmis = sc.mismatch - oe;
    const int diff = sc.match - sc.mismatch;
    const int qwords = (qlen + 31) >> 5, rwords = (rlen + 31) >> 5;
    int G[W];
#pragma unroll
    for (int j = 0; j < W; ++j) G[j] = oe;
    int best = 0, bq = 0, br = 0;
    uint64_t qwin = 0;
    constexpr int NG = (W + 63) / 64;  // 64-cell key groups
    for (int qi = 0; qi < rows; ++qi) {
        if ((qi & 31) == 0) qwin = ((qi >> 5) < qwords) ? q[qi >> 5] : 0ull;
        const uint64_t qrep = (qwin >> 62) * 0x5555555555555555ull;
        qwin <<= 2;
        const bool row_ok = qi < qlen;
        int Gl = oe, El = kNegE;
        // Row maximum per 64-cell group as (h << 8) | (64 - (j & 63)): the
        // low byte is an inline constant, larger means smaller j.
        int rowkey[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) rowkey[g] = INT32_MIN;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int rs = qi - B + 32 * c;
            const uint64_t x = window_signed(r, rs, rwords) ^ qrep;
            const uint64_t m = ~(x | (x >> 1)) & (row_ok ? base_range_mask(-rs, rlen - rs) : 0ull);
            const uint32_t mh = (uint32_t)(m >> 32), ml = (uint32_t)m;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int j = 32 * c + u;
                if (j < W) {
                    const uint32_t half = u < 16 ? mh : ml;
                    // 0 or -1 from the match bit, masked to the match bonus
                    const int sel = __builtin_amdgcn_sbfe((int)half, 30 - 2 * (u & 15), 1) & diff;
                    const int Gu = (j + 1 < W) ? G[j + 1] : oe;
                    const int dterm = G[j] + cmis + sel;
                    const int e = max(El + ext, Gu);
                    const int h = max(max(dterm, 0), max(e, Gl));
                    rowkey[j >> 6] = max(rowkey[j >> 6], (h << 8) | (64 - (j & 63)));
                    G[j] = h + oe;
                    Gl = G[j];
                    El = e;
                }
            }
        }
        int hrow = INT32_MIN, jrow = 0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int hg = rowkey[g] >> 8;
            if (hg > hrow) {
                hrow = hg;
                jrow = 64 * g + 64 - (rowkey[g] & 255);
            }
        }
        if (hrow > best) {
            best = hrow;
            bq = qi;
            br = qi - B + jrow;
        }
    }
    return {best, bq, br};
}

template <int LO, int HI>
__device__ __forceinline__ LaneResult sw_reg_dispatch(int band, const uint64_t *q, int qlen,
                                                      const uint64_t *r, int rlen, int rows,
                                                      const SwScoring &sc) {
    if constexpr (LO == HI) {
        return sw_reg_lane<LO>(q, qlen, r, rlen, rows, sc);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (band <= MID) return sw_reg_dispatch<LO, MID>(band, q, qlen, r, rlen, rows, sc);
        return sw_reg_dispatch<MID + 1, HI>(band, q, qlen, r, rlen, rows, sc);
    }
}

}  // namespace

// Register lane kernel for bands LO..HI over order[range[0] .. range[1]),
// one wavefront per 64 consecutive entries, persistent over the range.
// Every pair it receives was classified register-eligible (valid, effective
// band in LO..HI, h < 2^23, non-positive penalties).
template <int LO, int HI>
__global__ __launch_bounds__(64) void sw_reg_kernel(Batch b, const int32_t *band_arr,
                                                    int32_t band_scalar, SwScoring sc, SwOut out,
                                                    const int32_t *order, const int32_t *range) {
    const int64_t start = range[0], end = range[1];
    for (int64_t base = start + (int64_t)blockIdx.x * 64; base < end;
         base += (int64_t)gridDim.x * 64) {
        const int64_t item = base + threadIdx.x;
        const bool active = item < end;
        const int64_t pair = active ? order[item] : 0;
        const uint64_t *q = b.q_packed, *r = b.r_packed;
        int qlen = 0, rlen = 0, band = -1;
        if (active && load_pair(b, pair, q, r, qlen, rlen)) {
            band = effective_band(band_arr ? band_arr[pair] : band_scalar, qlen, rlen);
            if (band < LO || band > HI) band = -1;
        }
        uint64_t pending = __ballot(band >= 0);
        while (pending) {
            const int lead = __ffsll((unsigned long long)pending) - 1;
            const int bsel = __shfl(band, lead, 64);
            const bool mine = band == bsel;
            pending &= ~__ballot(mine);
            int rows = mine ? qlen : 0;
            for (int off = 32; off > 0; off >>= 1) rows = max(rows, __shfl_xor(rows, off, 64));
            const LaneResult res = sw_reg_dispatch<LO, HI>(bsel, q, mine ? qlen : 0, r,
                                                           mine ? rlen : 0, rows, sc);
            if (mine) {
                out.score[pair] = res.score;
                out.q_end[pair] = res.q_end;
                out.r_end[pair] = res.r_end;
            }
        }
    }
}

// Band buckets of the register layout, one kernel each so small bands are
// not held to the register budget of the widest: [0,16], [17,32], [33,48],
// [49,64] (the classifier in cgm_sort.hip uses the same bounds).
int launch_sw_reg_buckets(const Batch &b, const int32_t *band, int32_t band_scalar, SwScoring sc,
                          SwOut out, const int32_t *order, const int32_t *ranges, int grid,
                          hipStream_t stream) {
    // The buckets are independent: run them side by side on forked streams
    // so that together they fill the GPU.
    hipStream_t aux[4];
    if (int rc = fork_streams(stream, 4, aux)) return rc;
#define CGM_REG(K, LO, HI)                                                                   \
    hipLaunchKernelGGL((sw_reg_kernel<LO, HI>), dim3(grid), dim3(64), 0, aux[3 - (K)], b,     \
                       band, band_scalar, sc, out, order, ranges + 2 * (K));
    CGM_REG(3, 49, 64)
    CGM_REG(2, 33, 48)
    CGM_REG(1, 17, 32)
    CGM_REG(0, 0, 16)
#undef CGM_REG
    CGM_HIP(hipGetLastError());
    return join_streams(stream, 4, aux);
}


}  // namespace cgm
