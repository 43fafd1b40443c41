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

// Even-bit mask (base u at bit 62 - 2u) of window bases u in [lo, hi),
// without branches: bases >= lo AND bases < hi (empty when lo >= hi).
__device__ __forceinline__ uint64_t base_range_mask(int lo, int hi) {
    lo = min(max(lo, 0), 32);
    hi = min(max(hi, 0), 32);
    const uint64_t m1 = lo == 32 ? 0ull : (~0ull >> (2 * lo));
    const uint64_t m2 = hi == 0 ? 0ull : (~0ull << (64 - 2 * hi));
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

// Reference words as two 1-bit planes per 32-base word, LSB-first: bit u
// of lo / hi is the low / high bit of base u of the word.  A 64-bit word
// holds base u at bits 63-2u (high) and 62-2u (low); reversed, at bits 2u
// (high) and 2u + 1 (low), so the planes are its even and odd bits.
__device__ __forceinline__ uint32_t compress_even64(uint64_t x) {
    x &= 0x5555555555555555ull;
    x = (x | (x >> 1)) & 0x3333333333333333ull;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
    return (uint32_t)x;
}
template <int N>
struct AlignedPlanes {
    uint32_t lo[N], hi[N];
    __device__ __forceinline__ void set(int c, uint64_t w) {
        const uint64_t v = __builtin_bitreverse64(w);
        hi[c] = compress_even64(v);
        lo[c] = compress_even64(v >> 1);
    }
    __device__ __forceinline__ void load(const uint64_t *r, int w0, int rwords) {
#pragma unroll
        for (int c = 0; c < N; ++c) set(c, (w0 + c >= 0 && w0 + c < rwords) ? r[w0 + c] : 0ull);
    }
    __device__ __forceinline__ void advance(const uint64_t *r, int w0_new, int rwords) {
#pragma unroll
        for (int c = 0; c + 1 < N; ++c) {
            lo[c] = lo[c + 1];
            hi[c] = hi[c + 1];
        }
        const int w = w0_new + N - 1;
        set(N - 1, (w >= 0 && w < rwords) ? r[w] : 0ull);
    }
    // bases o .. o + 31 from word c on (bit i: base o + i), o in [0, 32)
    __device__ __forceinline__ void window(int c, int o, uint32_t &wl, uint32_t &wh) const {
        wl = __builtin_amdgcn_alignbit(lo[c + 1], lo[c], (uint32_t)o);
        wh = __builtin_amdgcn_alignbit(hi[c + 1], hi[c], (uint32_t)o);
    }
};

// Bits [lo, hi) of a 32-bit mask (empty when lo >= hi).
__device__ __forceinline__ uint32_t range_mask32(int lo, int hi) {
    lo = min(max(lo, 0), 32);
    hi = min(max(hi, 0), 32);
    const uint32_t m1 = lo == 32 ? 0u : (~0u << lo);
    const uint32_t m2 = hi == 32 ? ~0u : ((1u << hi) - 1u);
    return m1 & m2;
}

// Match mask of a 32-base plane window: bit i set when base i equals the
// row's query base (qlo, qhi: its bits replicated) and lies inside
// [0, rlen) on a valid row (rmin as in chunk_mask).
__device__ __forceinline__ uint32_t plane_mask(uint32_t wl, uint32_t wh, uint32_t qlo,
                                               uint32_t qhi, int rs, int rlen, bool row_ok,
                                               int rmin) {
    uint32_t eq = ~((wl ^ qlo) | (wh ^ qhi));
    if (rs < 0 || rs + 32 > rmin) eq &= range_mask32(-rs, rlen - rs);
    return row_ok ? eq : 0u;
}

// Match mask of one 32-base chunk: bit 62-2u set when base u of the chunk
// equals the row's query base and lies inside [0, rlen) on a valid row.
// rmin (wave-uniform) is the shortest reference among the wavefront's lanes
// that have query bases: a chunk inside [0, rmin) needs no range mask, and
// that test is one scalar branch (the other lanes never have a valid row).
__device__ __forceinline__ uint64_t chunk_mask(uint64_t win, uint64_t qrep, int rs, int rlen,
                                               bool row_ok, int rmin) {
    const uint64_t x = win ^ qrep;
    const uint64_t eq = ~(x | (x >> 1));
    uint64_t range = 0x5555555555555555ull;
    if (rs < 0 || rs + 32 > rmin) range = base_range_mask(-rs, rlen - rs);
    return eq & (row_ok ? range : 0ull);
}

// Shortest reference among the lanes that have query bases (wave-uniform).
__device__ __forceinline__ int wave_min_ref(int qlen, int rlen) {
    int v = qlen > 0 ? rlen : INT32_MAX;
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
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
    const int rmin = wave_min_ref(qlen, rlen);
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
            const uint64_t m = chunk_mask(rw.window(c, sh), qrep, rs, rlen, row_ok, rmin);
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
    const int rmin = wave_min_ref(qlen, rlen);
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
            const uint64_t m = chunk_mask(rw.window(c, sh), qrep, rs, rlen, row_ok, rmin);
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

// ------------------------------------------------ dual-row 16-bit lanes
// The band frame of W = 2B + 1 cells is cut into a left half (j in [0, S))
// and a right half (j in [S, 2S)), S = B + 1, and both halves are swept in
// one pass over the row registers: register u holds cell u of the left half
// in its low 16 bits and cell S + u of the right half in its high 16 bits,
// the right half one row behind the left one.  Step t computes row t of the
// left half and row t - 1 of the right half with packed 16-bit instructions
// (two cells per instruction; the encodings are DualArith below).  The right
// half's first cell reads its left neighbour (row t - 1, cell B) from the
// left half's previous step; the left half's last cell reads its upper
// neighbour (row t - 1, cell B + 1) from the right half's first cell of this
// step.  Cell 2S - 1 = W lies past the frame: it is forced to H = 0 so that
// cell W - 1 reads the zero upper neighbour the frame edge gives.  Each
// row's match masks come from the reference held as 1-bit planes
// (AlignedPlanes): per half and 32-cell chunk two funnel shifts per plane
// and one compare against the query base's bits, one mask bit per cell.
//
// Exactness: the per-cell recurrence, the forced mismatches outside the
// matrix and the first-cell boundary values are those of sw_reg_lane, in
// 16-bit lanes.  Every value stays in [-16384 + gap_extend, H_max + match]
// with H_max = match * min(qlen, rlen), so dual_ok() admits a pair only when
// that range fits int16; the E boundary value -16384 then lies below every
// value it is compared with (gap_open >= -16384), as -1000000 does in the
// reference.  Tie-break: per half, the first row whose maximum beats the
// half's best (strict) is kept whole; its first cell holding the best gives
// the column.  The two halves' candidates are then ordered by score, row,
// and left half first, which is the reference's first-in-row-major-order.
__device__ __forceinline__ bool dual_scoring_ok(const SwScoring &sc) {
    return sc.match <= 16383 && sc.mismatch >= -8191 && sc.gap_open >= -8191 &&
           sc.gap_extend >= -8191;
}
__device__ __forceinline__ bool dual_ok(const SwScoring &sc, int qlen, int rlen) {
    return (int64_t)max(sc.match, 0) * min(qlen, rlen) + sc.match <= 32767;
}

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bit_select(uint32_t m, uint32_t a, uint32_t b) {
    return (a & m) | (b & ~m);
}

// Dual-row cell arithmetic.  Two 16-bit encodings of the same integer
// recurrence: int16 (v_pk_add_u16 / v_pk_max_i16), and IEEE half precision
// (v_pk_add_f16 / v_pk_maximum3_f16, three-input maximum), which is exact
// while every value is a multiple of t in [-2048 t, 2048 t], t the largest
// power of two (at most 8) dividing all four scores: such values have at
// most 11 significant bits, so every sum and maximum is the integer one.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
template <bool F16>
struct DualArith;
template <>
struct DualArith<false> {
    using V = s16x2;
    static constexpr uint32_t kNegE = 0xC000u;  // -16384
    __device__ static V add(V a, V b) { return a + b; }
    __device__ static V mx(V a, V b) { return __builtin_elementwise_max(a, b); }
    __device__ static uint32_t enc(int x) { return (uint32_t)x & 0xffffu; }
    __device__ static int dec(uint32_t bits) { return (int)(short)(bits & 0xffffu); }
};
template <>
struct DualArith<true> {
    using V = f16x2;
    static constexpr uint32_t kNegE = 0xF400u;  // -16384
    __device__ static V add(V a, V b) { return a + b; }
    __device__ static V mx(V a, V b) { return __builtin_elementwise_maximum(a, b); }
    __device__ static uint32_t enc(int x) {
        return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
    }
    __device__ static int dec(uint32_t bits) {
        return (int)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xffffu));
    }
};
template <class V>
__device__ __forceinline__ V dual_bits(uint32_t v) { return __builtin_bit_cast(V, v); }
template <class V>
__device__ __forceinline__ uint32_t dual_u32(V v) { return __builtin_bit_cast(uint32_t, v); }
// Running row maximum kept in sequence (for the same reason as max3_chain)
// by an empty asm on its result.  Unlike an asm instruction this leaves the
// hazard recognizer free to schedule around it: on gfx950 a packed
// instruction that reads the result of the packed instruction just before it
// needs a wait state, and an asm instruction is always assumed to.
template <class V>
__device__ __forceinline__ V dual_seq(V v) {
    uint32_t d = dual_u32(v);
    asm("" : "+v"(d));
    return dual_bits<V>(d);
}
// t of the half-precision encoding: largest power of two <= 8 dividing all scores
__device__ __forceinline__ int64_t f16_unit(const SwScoring &sc) {
    const uint64_t m = (uint64_t)llabs(sc.match) | (uint64_t)llabs(sc.mismatch) |
                       (uint64_t)llabs(sc.gap_open) | (uint64_t)llabs(sc.gap_extend);
    const uint64_t low = m ? (m & (0ull - m)) : 8ull;
    return (int64_t)(low < 8 ? low : 8);
}
// every constant and every value bound of the recurrence within 2048 t:
// SUBM = match - oe, SUBX = mismatch - oe, oe, oe + gap_extend, mismatch
__device__ __forceinline__ bool f16_scoring_ok(const SwScoring &sc) {
    const int64_t lim = 2048 * f16_unit(sc);
    const int64_t oe = (int64_t)sc.gap_open + sc.gap_extend;
    return sc.match <= lim && sc.match - oe <= lim && sc.mismatch - oe <= lim &&
           sc.mismatch - oe >= -lim && sc.mismatch >= -lim && oe >= -lim &&
           oe + sc.gap_extend >= -lim;
}
// H_max + match within 2048 t (H_max = match * min(qlen, rlen))
__device__ __forceinline__ bool f16_ok(const SwScoring &sc, int qlen, int rlen) {
    return (int64_t)max(sc.match, 0) * min(qlen, rlen) + sc.match <= 2048 * f16_unit(sc);
}

template <int B, bool F16>
__device__ __forceinline__ LaneResult sw_dual_lane(const uint64_t *q, int qlen, const uint64_t *r,
                                                   int rlen, int rows, const SwScoring &sc) {
    using T = DualArith<F16>;
    using V = typename T::V;
    constexpr int S = B + 1;
    constexpr int NCH = (S + 31) / 32;
    const int oe = sc.gap_open + sc.gap_extend;
    const uint32_t OE = T::enc(oe) * 0x10001u;
    const uint32_t EXT = T::enc(sc.gap_extend) * 0x10001u;
    // the diagonal term's addend: SUBX (mismatch - oe) or SUBM (match - oe),
    // selected per half as SUBX + bit * DSUB on the 16-bit patterns
    const uint32_t SUBX = T::enc(sc.mismatch - oe) * 0x10001u;
    const uint32_t DSUB = ((T::enc(sc.match - oe) - T::enc(sc.mismatch - oe)) & 0xffffu) * 0x10001u;
    constexpr uint32_t kNegE16 = T::kNegE;  // -16384
    const int qwords = (qlen + 31) >> 5, rwords = (rlen + 31) >> 5;
    const int rmin = wave_min_ref(qlen, rlen);
    V P[S], SV[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        P[u] = dual_bits<V>(OE);
        SV[u] = dual_bits<V>(OE);
    }
    int best_lo = 0, row_lo = 0, best_hi = 0, row_hi = 0;
    uint64_t qwin = 0;
    int qb_prev = 0;
    uint32_t el_carry = kNegE16;  // E of the left half's last cell, previous step
    AlignedPlanes<NCH + 1> rwl, rwh;
    rwl.load(r, (-B) >> 5, rwords);
    rwh.load(r, 0, rwords);
    uint32_t p0_hi = OE >> 16;  // right half's first cell, G of its latest row
    for (int t = 0; t <= rows; ++t) {
        const int rs_lo = t - B, rs_hi = t;  // reference base of each half's first cell
        if (t > 0 && (rs_lo & 31) == 0) rwl.advance(r, rs_lo >> 5, rwords);
        if (t > 0 && (rs_hi & 31) == 0) rwh.advance(r, rs_hi >> 5, rwords);
        if ((t & 31) == 0) qwin = ((t >> 5) < qwords) ? q[t >> 5] : 0ull;
        const int qb = (int)(qwin >> 62);
        qwin <<= 2;
        const uint32_t ql_lo = (qb & 1) ? ~0u : 0u, qh_lo = (qb & 2) ? ~0u : 0u;
        const uint32_t ql_hi = (qb_prev & 1) ? ~0u : 0u, qh_hi = (qb_prev & 2) ? ~0u : 0u;
        qb_prev = qb;
        const bool ok_lo = t < qlen, ok_hi = t >= 1 && t - 1 < qlen;
        const int o_lo = rs_lo & 31, o_hi = rs_hi & 31;
        // left neighbours of the two first cells: H = 0 and the E boundary
        // value for the left half; cell B of row t - 1 for the right half
        V Gl = dual_bits<V>((dual_u32(P[S - 1]) << 16) | (OE & 0xffffu));
        V El = dual_bits<V>((el_carry << 16) | kNegE16);
        V rowmax = dual_bits<V>(0u), hprev = dual_bits<V>(0u);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            uint32_t al, ah, bl, bh;
            rwl.window(c, o_lo, al, ah);
            rwh.window(c, o_hi, bl, bh);
            const uint32_t ml = plane_mask(al, ah, ql_lo, qh_lo, rs_lo + 32 * c, rlen, ok_lo, rmin);
            const uint32_t mh = plane_mask(bl, bh, ql_hi, qh_hi, rs_hi + 32 * c, rlen, ok_hi, rmin);
            // cells 0-15 / 16-31 of the chunk: the left half's bits in the low
            // 16 bits, the right half's in the high 16 bits
            const uint32_t x0 = __builtin_amdgcn_perm(mh, ml, 0x05040100u);
            const uint32_t x1 = __builtin_amdgcn_perm(mh, ml, 0x07060302u);
#pragma unroll
            for (int v = 0; v < 32; ++v) {
                const int u = 32 * c + v;
                if (u < S) {
                    // each half's bit v % 16 as 0 / 1, then SUBX + bit (SUBM - SUBX)
                    // on the halves' 16-bit patterns (exact modulo 2^16, for both
                    // encodings): one shift, one and, one v_pk_mad_u16
                    const uint32_t b01 = ((v < 16 ? x0 : x1) >> (v & 15)) & 0x00010001u;
                    const u16x2 bm = __builtin_bit_cast(u16x2, b01);
                    const u16x2 sv = bm * __builtin_bit_cast(u16x2, DSUB) +
                                     __builtin_bit_cast(u16x2, SUBX);
                    const uint32_t sel = __builtin_bit_cast(uint32_t, sv);
                    // upper neighbour: register u + 1 of the previous step;
                    // for the last cell, the right half's first cell of this
                    // step (low) and H = 0 past the frame (high)
                    const V Gu = (u + 1 < S) ? P[u + 1]
                                             : dual_bits<V>((OE & 0xffff0000u) | p0_hi);
                    const V zero = dual_bits<V>(0u);
                    const V dterm = T::add(P[u], dual_bits<V>(sel));
                    const V ea = T::add(El, dual_bits<V>(EXT));
                    const V e = T::mx(ea, Gu);
                    V h;
                    if constexpr (F16) {
                        // max(dterm, 0, e, Gl) as two three-input maxima,
                        // the first off both dependency chains
                        const V x = T::mx(T::mx(dterm, zero), Gu);
                        h = T::mx(T::mx(x, ea), Gl);
                    } else {
                        h = T::mx(T::mx(T::mx(dterm, zero), e), Gl);
                    }
                    if (u == S - 1) h = dual_bits<V>(dual_u32(h) & 0xffffu);
                    if constexpr (F16) {
                        // two cells per three-input maximum
                        if (u & 1)
                            rowmax = dual_seq(T::mx(T::mx(rowmax, hprev), h));
                        else if (u == S - 1)
                            rowmax = dual_seq(T::mx(rowmax, h));
                        else
                            hprev = h;
                    } else {
                        rowmax = dual_seq(T::mx(rowmax, h));
                    }
                    const V pn = T::add(h, dual_bits<V>(OE));
                    P[u] = pn;
                    Gl = pn;
                    El = e;
                    if (u == 0) p0_hi = dual_u32(pn) >> 16;
                }
            }
        }
        el_carry = dual_u32(El) & 0xffffu;
        const int rm_lo = T::dec(dual_u32(rowmax));
        const int rm_hi = T::dec(dual_u32(rowmax) >> 16);
        const bool up_lo = rm_lo > best_lo, up_hi = rm_hi > best_hi;
        if (__ballot(up_lo || up_hi)) {
            const uint32_t m = (up_lo ? 0xffffu : 0u) | (up_hi ? 0xffff0000u : 0u);
#pragma unroll
            for (int u = 0; u < S; ++u)
                SV[u] = dual_bits<V>(bit_select(m, dual_u32(P[u]), dual_u32(SV[u])));
            if (up_lo) {
                best_lo = rm_lo;
                row_lo = t;
            }
            if (up_hi) {
                best_hi = rm_hi;
                row_hi = t - 1;
            }
        }
    }
    // first cell of each half's kept row that holds the half's best
    int ulo = 0, uhi = 0;
#pragma unroll
    for (int u = S - 1; u >= 0; --u) {
        const uint32_t v = dual_u32(SV[u]);
        if (T::dec(v) - oe == best_lo) ulo = u;
        if (T::dec(v >> 16) - oe == best_hi) uhi = u;
    }
    LaneResult lo = {best_lo, row_lo, row_lo - B + ulo};
    LaneResult hi = {best_hi, row_hi, row_hi + 1 + uhi};  // ri = row - B + S + u
    if (best_lo == 0) lo = {0, 0, 0};
    if (best_hi == 0) hi = {0, 0, 0};
    if (hi.score > lo.score || (hi.score == lo.score && hi.score > 0 && hi.q_end < lo.q_end))
        return hi;
    return lo;
}

template <int LO, int HI, bool F16>
__device__ __forceinline__ LaneResult sw_dual_dispatch(int band, const uint64_t *q, int qlen,
                                                       const uint64_t *r, int rlen, int rows,
                                                       const SwScoring &sc) {
    if constexpr (LO == HI) {
        return sw_dual_lane<LO, F16>(q, qlen, r, rlen, rows, sc);
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (band <= MID) return sw_dual_dispatch<LO, MID, F16>(band, q, qlen, r, rlen, rows, sc);
        return sw_dual_dispatch<MID + 1, HI, F16>(band, q, qlen, r, rlen, rows, sc);
    }
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
// packed-int16 instantiation for band 128, otherwise bands LO..HI.  DUAL
// compiles the dual-row lanes in as well (half-precision and int16
// encodings); they serve a pass when `dual` is set and every lane of the
// pass fits one of them (f16_ok, else dual_ok).
template <int LO, int HI, bool PACKED, bool DUAL = false>
__device__ __forceinline__ void sw_reg_group(const Batch &b, const int32_t *band_arr,
                                             int32_t band_scalar, const SwScoring &sc,
                                             const SwOut &out, const int32_t *order, int64_t start,
                                             int count, bool dual = false) {
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
        } else if constexpr (DUAL) {
            // half precision where every lane's values stay exact in it,
            // else int16 where they fit, else int32
            const bool use_f16 =
                dual && f16_scoring_ok(sc) && __all(!mine || f16_ok(sc, qlen, rlen));
            const bool use_i16 = !use_f16 && dual && dual_scoring_ok(sc) &&
                                 __all(!mine || dual_ok(sc, qlen, rlen));
            if (use_f16)
                res = sw_dual_dispatch<LO, HI, true>(bsel, q, mine ? qlen : 0, r,
                                                     mine ? rlen : 0, rows, sc);
            else if (use_i16)
                res = sw_dual_dispatch<LO, HI, false>(bsel, q, mine ? qlen : 0, r,
                                                      mine ? rlen : 0, rows, sc);
            else
                res = sw_reg_dispatch<LO, HI>(bsel, q, mine ? qlen : 0, r, mine ? rlen : 0, rows,
                                              sc);
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
