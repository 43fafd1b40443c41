// cgm_sw_dev.h -- device code of the register-resident one-pair-per-lane
// banded SW layout (gfx950), shared by the per-class and persistent kernels.
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

#pragma once

#include "cgm_internal.h"

namespace cgm {
namespace swdev {

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

// v_max3_i32 as an opaque step, so the running row maximum stays a short
// sequential chain instead of being re-associated into a tree that keeps
// every cell's key live until the end of the row.
__device__ __forceinline__ int max3_chain(int a, int b, int c) {
    int d;
    asm volatile("v_max3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// Match bonus of window base u: bit (30 - 2*(u % 16)) of the 32-bit half of
// the match mask sign-extended (0 / -1) and masked to (match - mismatch).
__device__ __forceinline__ int match_sel(uint32_t half, int u, int diff) {
    int v;
    asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(v) : "v"(half), "s"(30 - 2 * (u & 15)));
    return v & diff;
}

// Reference words r[w0 .. w0 + N - 1] of the band window, 32-base aligned,
// kept in registers; the row's first window position rs0 = qi - B moves one
// base per row, so a new word is loaded once per 32 rows (wave-uniform).
template <int N>
struct AlignedWords {
    uint64_t a[N];
    __device__ __forceinline__ void load(const uint64_t *r, int w0, int rwords) {
#pragma unroll
        for (int c = 0; c < N; ++c) a[c] = (w0 + c >= 0 && w0 + c < rwords) ? r[w0 + c] : 0ull;
    }
    __device__ __forceinline__ void advance(const uint64_t *r, int w0_new, int rwords) {
#pragma unroll
        for (int c = 0; c + 1 < N; ++c) a[c] = a[c + 1];
        const int w = w0_new + N - 1;
        a[N - 1] = (w >= 0 && w < rwords) ? r[w] : 0ull;
    }
    // 32 bases starting sh/2 bases into word c
    __device__ __forceinline__ uint64_t window(int c, int sh) const {
        return sh ? (a[c] << sh) | (a[c + 1] >> (64 - sh)) : a[c];
    }
};

// Match mask of one 32-base chunk: bit 62-2u set when base u of the chunk
// equals the row's query base and lies inside [0, rlen) on a valid row.
__device__ __forceinline__ uint64_t chunk_mask(uint64_t win, uint64_t qrep, int rs, int rlen,
                                               bool row_ok) {
    const uint64_t x = win ^ qrep;
    return ~(x | (x >> 1)) & (row_ok ? base_range_mask(-rs, rlen - rs) : 0ull);
}

template <int B>
__device__ __forceinline__ LaneResult sw_reg_lane(const uint64_t *q, int qlen, const uint64_t *r,
                                                  int rlen, int rows, const SwScoring &sc) {
    constexpr int W = 2 * B + 1;
    constexpr int NCH = (W + 31) / 32;
    constexpr int NG = (W + 63) / 64;
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
    AlignedWords<NCH + 1> rw;
    rw.load(r, (-B) >> 5, rwords);
    for (int qi = 0; qi < rows; ++qi) {
        const int rs0 = qi - B;
        if (qi > 0 && (rs0 & 31) == 0) rw.advance(r, rs0 >> 5, rwords);
        if ((qi & 31) == 0) qwin = ((qi >> 5) < qwords) ? q[qi >> 5] : 0ull;
        const uint64_t qrep = (qwin >> 62) * 0x5555555555555555ull;
        qwin <<= 2;
        const bool row_ok = qi < qlen;
        const int sh = 2 * (rs0 & 31);
        int Gl = oe, El = kNegE;
        // Row maximum per 64-cell group as (h << 8) | (64 - (j & 63)): the
        // low byte is an inline constant, larger means smaller j.
        int rowkey[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) rowkey[g] = INT32_MIN;
        int kpend = INT32_MIN;  // key of an even cell waiting for its pair
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int rs = rs0 + 32 * c;
            const uint64_t m = chunk_mask(rw.window(c, sh), qrep, rs, rlen, row_ok);
            const uint32_t mh = (uint32_t)(m >> 32), ml = (uint32_t)m;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int j = 32 * c + u;
                if (j < W) {
                    const int sel = match_sel(u < 16 ? mh : ml, u, diff);
                    const int Gu = (j + 1 < W) ? G[j + 1] : oe;
                    const int dterm = G[j] + cmis + sel;
                    const int e = max(El + ext, Gu);
                    const int h = max(max(dterm, 0), max(e, Gl));
                    const int key = (h << 8) | (64 - (j & 63));
                    if ((j & 1) == 0 && j + 1 < W && ((j + 1) >> 6) == (j >> 6)) {
                        kpend = key;
                    } else {
                        rowkey[j >> 6] = max3_chain(rowkey[j >> 6], kpend, key);
                        kpend = INT32_MIN;
                    }
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

// The same row sweep for wide bands with the row stored as packed int16
// pairs P[t] = G[2t] | G[2t+1] << 16 (halves the VGPRs: band 128 needs 129
// instead of 257).  One sign-extension per cell unpacks the up value (the
// diagonal value is the previous cell's up value) and one v_perm per two
// cells packs the new row; the left value of the next cell is read back
// from the packed word, which keeps the packing on the critical path.
// Requires oe <= G <= 32767, i.e. match * min(qlen, rlen) +
// (gap_open + gap_extend) <= 32767 and gap_open + gap_extend >= -32768.
template <int B>
__device__ __forceinline__ LaneResult sw_reg_lane_packed(const uint64_t *q, int qlen,
                                                         const uint64_t *r, int rlen, int rows,
                                                         const SwScoring &sc) {
    constexpr int W = 2 * B + 1;
    constexpr int NP = (W + 1) / 2;
    constexpr int NCH = (W + 31) / 32;
    constexpr int NG = (W + 63) / 64;
    const int oe = sc.gap_open + sc.gap_extend;
    const int ext = sc.gap_extend;
    const int cmis = sc.mismatch - oe;
    const int diff = sc.match - sc.mismatch;
    const int qwords = (qlen + 31) >> 5, rwords = (rlen + 31) >> 5;
    const uint32_t oe_pair = ((uint32_t)oe & 0xffffu) | ((uint32_t)oe << 16);
    uint32_t P[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) P[t] = oe_pair;
    int best = 0, bq = 0, br = 0;
    uint64_t qwin = 0;
    AlignedWords<NCH + 1> rw;
    rw.load(r, (-B) >> 5, rwords);
    for (int qi = 0; qi < rows; ++qi) {
        const int rs0 = qi - B;
        if (qi > 0 && (rs0 & 31) == 0) rw.advance(r, rs0 >> 5, rwords);
        if ((qi & 31) == 0) qwin = ((qi >> 5) < qwords) ? q[qi >> 5] : 0ull;
        const uint64_t qrep = (qwin >> 62) * 0x5555555555555555ull;
        qwin <<= 2;
        const bool row_ok = qi < qlen;
        const int sh = 2 * (rs0 & 31);
        int Gl = oe, El = kNegE;
        int rowkey[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) rowkey[g] = INT32_MIN;
        int kpend = INT32_MIN;
        int Gd = __builtin_amdgcn_sbfe((int)P[0], 0, 16);
        int lo_new = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int rs = rs0 + 32 * c;
            const uint64_t m = chunk_mask(rw.window(c, sh), qrep, rs, rlen, row_ok);
            const uint32_t mh = (uint32_t)(m >> 32), ml = (uint32_t)m;
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int j = 32 * c + u;
                if (j < W) {
                    const int sel = match_sel(u < 16 ? mh : ml, u, diff);
                    int Gu = oe;
                    if (j + 1 < W) {
                        Gu = ((j + 1) & 1) ? ((int)P[(j + 1) >> 1] >> 16)
                                           : __builtin_amdgcn_sbfe((int)P[(j + 1) >> 1], 0, 16);
                    }
                    const int dterm = Gd + cmis + sel;
                    const int e = max(El + ext, Gu);
                    const int h = max(max(dterm, 0), max(e, Gl));
                    const int key = (h << 8) | (64 - (j & 63));
                    if ((j & 1) == 0 && j + 1 < W && ((j + 1) >> 6) == (j >> 6)) {
                        kpend = key;
                    } else {
                        rowkey[j >> 6] = max3_chain(rowkey[j >> 6], kpend, key);
                        kpend = INT32_MIN;
                    }
                    const int Gn = h + oe;
                    if (j & 1) {
                        P[j >> 1] = __builtin_amdgcn_perm((uint32_t)Gn, (uint32_t)lo_new, 0x05040100u);
                        Gl = (int)P[j >> 1] >> 16;
                    } else {
                        if (j == W - 1) P[j >> 1] = ((uint32_t)Gn & 0xffffu) | (P[j >> 1] & 0xffff0000u);
                        lo_new = Gn;
                        Gl = Gn;
                    }
                    Gd = Gu;
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

// One wavefront aligns order[start + lane] for lane < count with the
// register layout: lanes are served once per distinct band among them
// (pairs arrive sorted by band, so normally once).  PACKED selects the
// packed-int16 instantiation for band 128, otherwise bands LO..HI.
template <int LO, int HI, bool PACKED>
__device__ __forceinline__ void sw_reg_group(const Batch &b, const int32_t *band_arr,
                                             int32_t band_scalar, const SwScoring &sc,
                                             const SwOut &out, const int32_t *order, int64_t start,
                                             int count) {
    const int lane = threadIdx.x & 63;
    const bool active = lane < count;
    const int64_t pair = active ? order[start + lane] : 0;
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
        LaneResult res;
        if constexpr (PACKED) {
            res = sw_reg_lane_packed<HI>(q, mine ? qlen : 0, r, mine ? rlen : 0, rows, sc);
        } else {
            res = sw_reg_dispatch<LO, HI>(bsel, q, mine ? qlen : 0, r, mine ? rlen : 0, rows, sc);
        }
        if (mine) {
            out.score[pair] = res.score;
            out.q_end[pair] = res.q_end;
            out.r_end[pair] = res.r_end;
        }
    }
}

}  // namespace swdev
}  // namespace cgm
