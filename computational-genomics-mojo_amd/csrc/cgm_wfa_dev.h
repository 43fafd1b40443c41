// cgm_wfa_dev.h -- device code of the one-wavefront-per-pair gap-affine WFA
// layout (gfx950), shared by the WFA and persistent mixed kernels.
#pragma once

#include "cgm_internal.h"

namespace cgm {
namespace wfadev {

// ---------------------------------------------------------------------------
// Wave layout: one 64-lane wavefront per pair, persistent workgroups pulling
// pair indices from an atomic queue (optionally through a work-sorted
// order[]).  Exact reductions of the reference loop:
//  * only scores that are multiples of g = gcd(x, o, e) can hold a wavefront
//    (every recurrence step costs x, o+e or e), so the others are skipped;
//  * diagonal k needs a gap opening and |k| extensions, so at score s only
//    |k| <= R(s) = (s - o) / e (0 below o + e) can be reached; the range is
//    [max(-rlen, -me, -R(s)), min(qlen, me, R(s))] and widening it to the
//    target diagonal, as the reference does, only adds unreachable cells;
//  * ranges only grow with s, so score slots are initialised once and never
//    cleared: a slot's entries outside its range were never written;
//  * NEXT and EXTEND fuse per diagonal: EXTEND of diagonal k touches only
//    M[s,k] and NEXT never reads the current score.
// Extension compares 32 bases per step: two funnel-shifted 64-bit windows
// XORed, leading zeros / 2 = matching bases, capped at the sequence ends.
// A whole workgroup works on one pair; sequences and int16 wavefronts sit
// in LDS when they fit the workgroup's budget, otherwise the wavefronts live
// in an HBM slab (int16 or int32 rows).
// ---------------------------------------------------------------------------
constexpr int kWaveThreads = 64;
constexpr int kNegRow = -(1 << 30);
constexpr int kTileSteps = 32;      // default score steps per block of the tiled layout
constexpr int kTileMinWidth = 256;  // narrowest tile worth the halo

__device__ __forceinline__ uint64_t window64(const uint64_t *s, int pos, int nwords) {
    const int w = pos >> 5, sh = 2 * (pos & 31);
    const uint64_t lo = s[w];
    if (!sh) return lo;
    const uint64_t hi = (w + 1 < nwords) ? s[w + 1] : 0ull;
    return (lo << sh) | (hi >> (64 - sh));
}

// Furthest row reached from (row, row - k) following matching bases.
__device__ __forceinline__ int extend_diag(const uint64_t *q, int qlen, int qw, const uint64_t *r,
                                           int rlen, int rw, int row, int k) {
    int col = row - k;
    if (col < 0) return row;
    while (row < qlen && col < rlen) {
        const uint64_t x = window64(q, row, qw) ^ window64(r, col, rw);
        const int n = x ? (__clzll(x) >> 1) : 32;
        const int lim = min(qlen - row, rlen - col);
        if (n >= lim) return row + lim;
        row += n;
        col += n;
        if (n < 32) break;
    }
    return row;
}

struct WfaGeom {
    int g, nM, nI;  // score step, M slots, I/D slots
};

__device__ __forceinline__ int reach(int s, const WfaPen &p) {
    return s >= p.o + p.e ? (s - p.o) / p.e : 0;
}

// Per-score bookkeeping of one pair, advanced one score step (g) at a time
// with no division: the step index si = s / g, the M slot si % nM and I/D
// slot si % nI of score s, the reach R(s), and the slots of s - x, s - o - e
// (M) and s - e (I/D), or -1 before those scores exist.
struct WfaSteps {
    int dx, dO, de;  // x / g, (o + e) / g, e / g
    int si, ms, is, R;

    __device__ __forceinline__ void init(const WfaPen &p, const WfaGeom &gm) {
        dx = p.x / gm.g;
        dO = (p.o + p.e) / gm.g;
        de = p.e / gm.g;
        si = ms = is = R = 0;
    }
    __device__ __forceinline__ void next(const WfaPen &p, const WfaGeom &gm) {
        ++si;
        if (++ms == gm.nM) ms = 0;
        if (++is == gm.nI) is = 0;
        const int s = si * gm.g;
        while ((R + 1) * p.e + p.o <= s) ++R;  // at most once: g <= e
    }
    __device__ __forceinline__ static int back(int slot, int d, int n) {
        return slot >= d ? slot - d : slot - d + n;
    }
    __device__ __forceinline__ int mx(const WfaGeom &gm) const {
        return si >= dx ? back(ms, dx, gm.nM) : -1;
    }
    __device__ __forceinline__ int mo(const WfaGeom &gm) const {
        return si >= dO ? back(ms, dO, gm.nM) : -1;
    }
    __device__ __forceinline__ int ie(const WfaGeom &gm) const {
        return si >= de ? back(is, de, gm.nI) : -1;
    }
};

// Storage of wavefront rows: int16 when every row fits (see rows_fit_int16),
// int32 otherwise.  Invalid entries hold the type's most negative value; any
// stored value < 0 is invalid.
template <typename T>
struct RowNeg;
template <>
struct RowNeg<int32_t> {
    static constexpr int32_t v = kNegRow;
};
template <>
struct RowNeg<int16_t> {
    static constexpr int16_t v = -32768;
};

// Rows grow past qlen only by +1 steps (a mismatch costs x, an insertion e),
// so every row stays below qlen + me / min(x, e) + 2.
__device__ __forceinline__ bool rows_fit_int16(int qlen, int me, const WfaPen &p) {
    return (int64_t)qlen + me / min(p.x, p.e) + 2 <= 32767;
}

// One pair aligned by the whole workgroup (blockDim.x threads, any number
// of wavefronts): threads stride over the diagonals of each score, and one
// barrier per score also carries the termination test.
// q / r may point to LDS copies of the sequences or to global memory.
template <int U, typename T>
__device__ int wfa_pair_team(const uint64_t *q, int qlen, const uint64_t *r, int rlen, int me,
                             const WfaPen &p, const WfaGeom &gm, T *wf, int W, int K) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int slots = gm.nM + 2 * gm.nI;
    const T neg = RowNeg<T>::v;
    for (int i = tid; i < slots * W; i += nt) wf[i] = neg;
    __syncthreads();
    T *M = wf;
    T *I = wf + gm.nM * W;
    T *D = I + gm.nI * W;
    const int target = qlen - rlen;
    __shared__ int done_at;  // score at which the target was reached
    if (tid == 0) {
        M[K] = (T)extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
        done_at = -1;
    }
    __syncthreads();
    if (target == 0 && (int)M[K] >= qlen) return 0;
    WfaSteps st;
    st.init(p, gm);
    for (int s = gm.g; s <= me; s += gm.g) {
        st.next(p, gm);
        const int ms = st.ms * W, is = st.is * W;
        const int Rs = st.R;
        const int lo = max(max(-rlen, -me), -Rs), hi = min(min(qlen, me), Rs);
        const int mx = st.mx(gm) >= 0 ? st.mx(gm) * W : -1;
        const int mo = st.mo(gm) >= 0 ? st.mo(gm) * W : -1;
        const int ie = st.ie(gm) >= 0 ? st.ie(gm) * W : -1;
        int found = 0;
        // U diagonals per thread and pass: every read of older score slots
        // is issued before the first write of the current ones (different
        // slots, so no aliasing), which overlaps the memory latencies of the
        // diagonals instead of chaining them.
        for (int k0 = lo + tid; k0 <= hi; k0 += U * nt) {
            int iv[U], dv[U], mv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * nt, di = k + K;
                int a = kNegRow, c = kNegRow, bb = kNegRow, d = kNegRow, m = kNegRow;
                if (k <= hi) {
                    if (mo >= 0) {
                        if (di > 0) a = (int)M[mo + di - 1];
                        if (di < W - 1) bb = (int)M[mo + di + 1];
                    }
                    if (ie >= 0) {
                        if (di > 0) c = (int)I[ie + di - 1];
                        if (di < W - 1) d = (int)D[ie + di + 1];
                    }
                    if (mx >= 0) m = (int)M[mx + di];
                }
                iv[u] = max(a, c) + 1;
                dv[u] = max(bb, d);
                mv[u] = max(max(m + 1, iv[u]), dv[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * nt, di = k + K;
                if (k <= hi) {
                    I[is + di] = iv[u] < 0 ? neg : (T)iv[u];
                    D[is + di] = dv[u] < 0 ? neg : (T)dv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * nt, di = k + K;
                if (k <= hi) {
                    int m = mv[u];
                    if (m >= 0) {
                        m = extend_diag(q, qlen, qw, r, rlen, rw, m, k);
                        found |= (k == target && m >= qlen);
                    }
                    M[ms + di] = m < 0 ? neg : (T)m;
                }
            }
        }
        // One barrier ends the score: a thread that reached the target
        // records s before it; scores differ per step, so the flag needs
        // no reset, and the pair ends at the first s recorded.
        if (found) done_at = s;
        __syncthreads();
        if (done_at == s) return s;
    }
    return -1;
}

// ---------------------------------------------------------------------------
// Branch-free team layout for pairs whose sequences and int16 wavefronts fit
// LDS (the common case).  Same recurrence and schedule as wfa_pair_team; what
// changes is the data layout, so that a diagonal costs no bounds or slot
// tests and its first extension step no branch:
//  * sequences are staged as 32-bit words with base i at bits 2i..2i+1 of
//    word i / 16 (each half of a packed word bit-reversed, which also swaps
//    the two bits of a base: harmless, bases are only compared), followed by
//    zero words.  The 32 bases from any position are then two alignbit
//    funnel shifts of three consecutive words, and the matching prefix of
//    two windows is the trailing-zero count of their XOR over two;
//  * every wavefront row has a guard entry on each side that stays invalid,
//    so diagonals k - 1 and k + 1 need no range test, and I and D of one
//    diagonal sit side by side (one address for I[k - 1] and D[k + 1]);
//  * the slot of s - x, s - o - e or s - e before that score exists is a
//    ring slot never written so far (slots 0..s/g only), so it reads as
//    invalid without the existence test;
//  * diagonals are taken in wavefront-wide groups: a group past the score's
//    range is skipped as a whole (a scalar test), and lanes of a partial
//    group past `hi` compute on whatever their rows hold but write nothing.
//    `kFastSlack` bytes after the rows keep those reads inside the layout.
// A valid M/I/D row on diagonal k is always >= max(k, 0) (each step adds to
// the row and column or keeps both), so the column row - k of a valid cell
// is never negative; clamping positions only matters for invalid lanes,
// whose results are discarded.
// ---------------------------------------------------------------------------
constexpr int kFastSlack = 512;

__device__ __forceinline__ int64_t team_fast_bytes(int qlen, int rlen, int W, const WfaGeom &gm) {
    const int64_t nq = 2 * (int64_t)((qlen + 31) >> 5) + 4, nr = 2 * (int64_t)((rlen + 31) >> 5) + 4;
    const int64_t WP = W + 2;
    const int64_t m_ents = (gm.nM * WP + 1) & ~(int64_t)1;
    return 4 * (nq + nr) + 2 * m_ents + 4 * gm.nI * WP + kFastSlack;
}

// Whether a pair runs on the small workgroups of the WFA launchers: int16
// rows and the branch-free LDS layout within kWfaSmallLdsBytes.  The other
// pairs take the 512-thread workgroups (52 KB of LDS, tiled over an HBM
// slab beyond that).
__device__ __forceinline__ bool wfa_fits_small(int qlen, int rlen, int me, const WfaPen &p,
                                               const WfaGeom &gm) {
    const int K = min(min(me, max(qlen, rlen)), reach(me, p));
    return rows_fit_int16(qlen, me, p) &&
           team_fast_bytes(qlen, rlen, 2 * K + 1, gm) <= kWfaSmallLdsBytes;
}

// Work key of a WFA pair for the longest-first order: diagonal count summed
// over the scores up to max_error (an upper bound of the cost), < 2^27.
__device__ __forceinline__ uint32_t wfa_work_key(int me, const WfaPen &p, const WfaGeom &gm) {
    const int64_t R = me >= p.o + p.e ? (me - p.o) / p.e : 0;
    const int64_t steps = (me > 0 ? me / gm.g : 0) * (2 * R + 1);
    return (uint32_t)min(steps, (int64_t)((1 << 27) - 1));
}

// Bases pos .. pos + 31 of an LSB-first sequence as two words (base pos + i
// at bits 2i of w0 for i < 16, of w1 for i >= 16).
__device__ __forceinline__ void window_lsb(const uint32_t *s, int pos, uint32_t &w0, uint32_t &w1) {
    const uint32_t *p = s + (pos >> 4);
    const uint32_t sh = 2u * (uint32_t)(pos & 15);
    const uint32_t a = p[0], b = p[1], c = p[2];
    w0 = __builtin_amdgcn_alignbit(b, a, sh);
    w1 = __builtin_amdgcn_alignbit(c, b, sh);
}

// Index of the lowest set bit, 0xFFFFFFFF for 0 (the instruction's own
// result for 0, which the compiler's count-trailing-zeros does not promise).
__device__ __forceinline__ uint32_t lowest_bit(uint32_t x) {
    uint32_t r;
    asm("v_ffbl_b32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Middle value of three (one instruction; the compiler builds it only for
// constant bounds): clamps x to [lo, hi] when lo <= hi.
__device__ __forceinline__ int med3(int x, int a, int b) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(a), "v"(b));
    return r;
}

// Leading bases (0..32) on which two 32-base windows agree: the lowest
// differing bit over 2, 64 (-> 32) when the windows are equal.
__device__ __forceinline__ int match32(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    const uint32_t t1 = min(lowest_bit(a1 ^ b1), 32u);
    return (int)(min(lowest_bit(a0 ^ b0), 32u + t1) >> 1);
}

// Furthest row reached from (row, col), col >= 0, following matching bases.
__device__ __forceinline__ int extend_lsb(const uint32_t *q, int qlen, const uint32_t *r, int rlen,
                                          int row, int col) {
    while (row < qlen && col < rlen) {
        uint32_t a0, a1, b0, b1;
        window_lsb(q, row, a0, a1);
        window_lsb(r, col, b0, b1);
        const int n = match32(a0, a1, b0, b1);
        const int lim = min(qlen - row, rlen - col);
        if (n >= lim) return row + lim;
        row += n;
        col += n;
        if (n < 32) break;
    }
    return row;
}

// One pair aligned by the whole workgroup of NT threads in the layout above
// (`lds` holds at least team_fast_bytes; the caller checks rows_fit_int16).
template <int U, int NT>
__device__ int wfa_pair_team_fast(const uint64_t *qg, int qlen, const uint64_t *rg, int rlen,
                                  int me, const WfaPen &p, const WfaGeom &gm, uint8_t *lds, int W,
                                  int K) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int nq = 2 * qw + 4, nr = 2 * rw + 4;
    uint32_t *qs = (uint32_t *)lds;
    uint32_t *rs = qs + nq;
    for (int i = tid; i < nq + nr; i += NT) {
        const bool isq = i < nq;
        const int j = isq ? i : i - nq;
        uint32_t v = 0;
        if ((j >> 1) < (isq ? qw : rw)) {
            const uint64_t w = (isq ? qg : rg)[j >> 1];
            v = __builtin_bitreverse32((j & 1) ? (uint32_t)w : (uint32_t)(w >> 32));
        }
        qs[i] = v;
    }
    const int WP = W + 2;
    const int m_ents = (gm.nM * WP + 1) & ~1;
    int16_t *Mb = (int16_t *)(lds + 4 * (nq + nr));
    int16_t *IDb = Mb + m_ents;
    const int16_t neg = RowNeg<int16_t>::v;
    {
        // rows and the slack after them, two entries per store
        const int words = (m_ents + 2 * gm.nI * WP + kFastSlack / 2) >> 1;
        uint32_t *w32 = (uint32_t *)Mb;
        for (int i = tid; i < words; i += NT) w32[i] = 0x80008000u;
    }
    // done_at[si & 1]: score step si that reached the target, or -1.  The
    // flag of step si is read after that step's barrier and acted on after
    // the next one, so the read's latency hides under the next step; one
    // extra step is computed, harmlessly.  Two slots keep the read of step
    // si apart from writes of step si + 1 (a wavefront may run ahead into
    // that step before another has read the flag).
    __shared__ int done_at[2];
    if (tid == 0) done_at[0] = done_at[1] = -1;
    __syncthreads();
    if (tid == 0) Mb[K + 1] = (int16_t)extend_lsb(qs, qlen, rs, rlen, 0, 0);
    __syncthreads();
    const int target = qlen - rlen;
    if (target == 0 && (int)Mb[K + 1] >= qlen) return 0;
    // first thread of this wavefront, as a scalar so that group tests branch
    const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);
    WfaSteps st;
    st.init(p, gm);
    int seen = -1;  // done_at of the previous step
    for (int s = gm.g; s <= me; s += gm.g) {
        st.next(p, gm);
        const int lo = max(max(-rlen, -me), -st.R), hi = min(min(qlen, me), st.R);
        const int16_t *Mo = Mb + WfaSteps::back(st.ms, st.dO, gm.nM) * WP;
        const int16_t *Mx = Mb + WfaSteps::back(st.ms, st.dx, gm.nM) * WP;
        const int16_t *IDe = IDb + WfaSteps::back(st.is, st.de, gm.nI) * 2 * WP;
        int16_t *Ms = Mb + st.ms * WP;
        int16_t *IDs = IDb + st.is * 2 * WP;
        int found = 0;
        for (int kb = lo + wave0; kb <= hi; kb += U * NT) {
            // di = k + K: row entry di + 1 holds diagonal k
            const int di0 = kb + lane + K;
            const int ng = min(U, (hi - kb) / NT + 1);  // groups of this pass in range
            int mv[U], mr[U];
            // NEXT: M, I, D of every group from the older scores
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (u >= ng) break;
                const int di = di0 + u * NT;
                // M first: every read has returned before the store below,
                // so nothing after it waits on the store
                const int m = Mx[di + 1];
                __builtin_amdgcn_sched_barrier(0);  // keep the M read first
                const int a = Mo[di], b = Mo[di + 2];
                const int c = IDe[2 * di], d = IDe[2 * di + 5];
                const int iv = max(a, c) + 1;
                const int dv = max(b, d);  // negative only when both are `neg`
                mv[u] = max(max(m + 1, iv), dv);
                if (di - K <= hi)  // I and D of diagonal k in one 32-bit store
                    *(uint32_t *)(IDs + 2 * di + 2) =
                        (uint32_t)(uint16_t)(iv < 0 ? neg : (int16_t)iv) | ((uint32_t)dv << 16);
            }
            // EXTEND, first 32 bases of every group: no branch
            uint64_t more = 0;  // bit u: some lane of group u matched all 32
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (u >= ng) break;
                const int k = di0 + u * NT - K;
                const int row = mv[u], col = row - k;
                uint32_t a0, a1, b0, b1;
                window_lsb(qs, med3(row, 0, qlen), a0, a1);
                window_lsb(rs, med3(col, 0, rlen), b0, b1);
                const int lim = min(qlen - row, rlen - col);
                const int n = med3(match32(a0, a1, b0, b1), lim, 0);  // match >= 0
                mr[u] = row + n;
                if (__ballot(row >= 0 && n == 32 && lim > 32)) more |= 1ull << u;
            }
            // the rest of long matches (few lanes)
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!(more >> u & 1)) continue;
                    const int k = di0 + u * NT - K;
                    const int row = mv[u], col = row - k;
                    if (row >= 0 && mr[u] == row + 32 && min(qlen - row, rlen - col) > 32 && k <= hi)
                        mr[u] = extend_lsb(qs, qlen, rs, rlen, row + 32, col + 32);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (u >= ng) break;
                const int di = di0 + u * NT, k = di - K;
                if (k <= hi) {
                    const int row = mv[u];
                    found |= (k == target && row >= 0 && mr[u] >= qlen);
                    Ms[di + 1] = row < 0 ? neg : (int16_t)mr[u];
                }
            }
        }
        if (found) done_at[st.si & 1] = s;
        __syncthreads();
        if (seen >= 0) return seen;
        seen = done_at[st.si & 1];
    }
    return seen;
}

// Trapezoid tiling of the team layout for pairs whose wavefronts exceed
// LDS.  Scores advance in blocks of up to TB steps.  Each block's widest
// range is cut into tiles of TW = LW - 2 TB diagonals; a tile copies every
// slot of its diagonals plus TB on each side from slab set `cur` into LDS
// (entries outside [-K, K] read as invalid), runs the block's steps there on
// a range that narrows by one diagonal per step on each side (a cell reads
// diagonals k-1..k+1 of earlier scores only, and e >= g), and writes its own
// diagonals of every slot to slab set `cur ^ 1`.  The slab is thus read and
// written once per TB scores instead of every score.  The first step whose
// M row reaches the target ends the pair, in whichever tile computes it.
// Entries a step does not compute keep what the slot held, which is invalid
// outside the score's range because ranges only grow.  `slab` holds two sets
// of slots * W int16 entries; `loc` holds slots * LW int16 entries.
template <int U>
__device__ int wfa_pair_tiled(const uint64_t *q, int qlen, const uint64_t *r, int rlen, int me,
                              const WfaPen &p, const WfaGeom &gm, int16_t *slab, int W, int K,
                              int16_t *loc, int LW, int TB) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int slots = gm.nM + 2 * gm.nI;
    const int16_t neg = RowNeg<int16_t>::v;
    const int64_t set_len = (int64_t)slots * W;
    for (int64_t i = tid; i < 2 * set_len; i += nt) slab[i] = neg;
    __syncthreads();
    const int target = qlen - rlen;
    __shared__ int done_at;  // score at which the target was reached
    if (tid == 0) {
        slab[K] = (int16_t)extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
        done_at = -1;
    }
    __syncthreads();
    if (target == 0 && (int)slab[K] >= qlen) return 0;
    const int TW = LW - 2 * TB;
    WfaSteps blk;  // state at the start of the current block of scores
    blk.init(p, gm);
    int16_t *M = loc;
    int16_t *I = loc + gm.nM * LW;
    int16_t *D = I + gm.nI * LW;
    int cur = 0;
    for (int s0 = 0; s0 + gm.g <= me;) {
        const int nsteps = min(TB, (me - s0) / gm.g);
        const int s_end = s0 + nsteps * gm.g;
        const int Re = reach(s_end, p);
        const int lo_e = max(max(-rlen, -me), -Re), hi_e = min(min(qlen, me), Re);
        const int16_t *src = slab + cur * set_len;
        int16_t *dst = slab + (cur ^ 1) * set_len;
        for (int a = lo_e; a <= hi_e; a += TW) {
            const int b = min(a + TW - 1, hi_e);
            const int base = a - TB;  // diagonal of local index 0
            for (int sl = 0; sl < slots; ++sl)
                for (int j = tid; j < LW; j += nt) {
                    const int k = base + j;
                    loc[sl * LW + j] = (k >= -K && k <= K) ? src[(int64_t)sl * W + k + K] : neg;
                }
            __syncthreads();
            WfaSteps st = blk;
            for (int t = 1; t <= nsteps; ++t) {
                const int s = s0 + t * gm.g;
                st.next(p, gm);
                const int ms = st.ms * LW, is = st.is * LW;
                const int Rs = st.R;
                const int shrink = nsteps - t;
                const int klo = max(max(max(-rlen, -me), -Rs), a - shrink);
                const int khi = min(min(min(qlen, me), Rs), b + shrink);
                const int mx = st.mx(gm) >= 0 ? st.mx(gm) * LW : -1;
                const int mo = st.mo(gm) >= 0 ? st.mo(gm) * LW : -1;
                const int ie = st.ie(gm) >= 0 ? st.ie(gm) * LW : -1;
                int found = 0;
                // klo - 1 >= base and khi + 1 < base + LW: no bounds tests
                for (int k0 = klo + tid; k0 <= khi; k0 += U * nt) {
                    int iv[U], dv[U], mv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * nt, di = k - base;
                        int a_ = kNegRow, c = kNegRow, bb = kNegRow, d = kNegRow, m = kNegRow;
                        if (k <= khi) {
                            if (mo >= 0) {
                                a_ = M[mo + di - 1];
                                bb = M[mo + di + 1];
                            }
                            if (ie >= 0) {
                                c = I[ie + di - 1];
                                d = D[ie + di + 1];
                            }
                            if (mx >= 0) m = M[mx + di];
                        }
                        iv[u] = max(a_, c) + 1;
                        dv[u] = max(bb, d);
                        mv[u] = max(max(m + 1, iv[u]), dv[u]);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * nt, di = k - base;
                        if (k <= khi) {
                            I[is + di] = iv[u] < 0 ? neg : (int16_t)iv[u];
                            D[is + di] = dv[u] < 0 ? neg : (int16_t)dv[u];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int k = k0 + u * nt, di = k - base;
                        if (k <= khi) {
                            int m = mv[u];
                            if (m >= 0) {
                                m = extend_diag(q, qlen, qw, r, rlen, rw, m, k);
                                found |= (k == target && m >= qlen);
                            }
                            M[ms + di] = m < 0 ? neg : (int16_t)m;
                        }
                    }
                }
                // the first score that reaches the target ends the pair
                // (the flag is set once, then every thread returns)
                if (found) done_at = s;
                __syncthreads();
                if (done_at == s) return s;
            }
            const int wd = b - a + 1;
            for (int sl = 0; sl < slots; ++sl)
                for (int j = tid; j < wd; j += nt)
                    dst[(int64_t)sl * W + a + j + K] = loc[sl * LW + (a + j - base)];
            __syncthreads();
        }
        for (int t = 0; t < nsteps; ++t) blk.next(p, gm);
        cur ^= 1;
        s0 = s_end;
    }
    return -1;
}

// Tile width (diagonals, even) of the branch-free tiled layout in
// `lds_bytes`: LSB-first sequence words, then nM M rows and nI I/D rows
// (I and D of a diagonal side by side) of LW entries, then kFastSlack bytes.
__device__ __forceinline__ int tiled_fast_width(int qlen, int rlen, const WfaGeom &gm,
                                                int lds_bytes) {
    const int64_t seq = 4 * (2 * (int64_t)((qlen + 31) >> 5) + 4 + 2 * (int64_t)((rlen + 31) >> 5) + 4);
    const int64_t left = lds_bytes - seq - kFastSlack;
    return left > 0 ? (int)((left / (2 * (gm.nM + 2 * gm.nI))) & ~(int64_t)1) : 0;
}

// The trapezoid tiling of wfa_pair_tiled (same blocks of TB score steps,
// same tiles, same slab sets and copies) with the tile in the branch-free
// layout of wfa_pair_team_fast.  A step reads diagonals k - 1 .. k + 1 of its
// range only, which stays inside [base, base + LW), so tile rows need no
// guards; lanes of a group past the step's range read into the next row or
// the slack and write nothing.
template <int U, int NT>
__device__ int wfa_pair_tiled_fast(const uint64_t *qg, int qlen, const uint64_t *rg, int rlen,
                                   int me, const WfaPen &p, const WfaGeom &gm, int16_t *slab,
                                   int W, int K, uint8_t *lds, int LW, int TB) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int nq = 2 * qw + 4, nr = 2 * rw + 4;
    uint32_t *qs = (uint32_t *)lds;
    uint32_t *rs = qs + nq;
    for (int i = tid; i < nq + nr; i += NT) {
        const bool isq = i < nq;
        const int j = isq ? i : i - nq;
        uint32_t v = 0;
        if ((j >> 1) < (isq ? qw : rw)) {
            const uint64_t w = (isq ? qg : rg)[j >> 1];
            v = __builtin_bitreverse32((j & 1) ? (uint32_t)w : (uint32_t)(w >> 32));
        }
        qs[i] = v;
    }
    int16_t *Mt = (int16_t *)(lds + 4 * (nq + nr));  // nM rows of LW
    int16_t *IDt = Mt + gm.nM * LW;                     // nI rows of 2 LW
    const int16_t neg = RowNeg<int16_t>::v;
    const int slots = gm.nM + 2 * gm.nI;
    const int64_t set_len = (int64_t)slots * W;
    for (int64_t i = tid; i < 2 * set_len; i += NT) slab[i] = neg;
    for (int i = tid; i < kFastSlack / 2; i += NT) IDt[2 * gm.nI * LW + i] = neg;
    __syncthreads();
    const int target = qlen - rlen;
    __shared__ int done_at;  // score at which the target was reached
    if (tid == 0) {
        slab[K] = (int16_t)extend_lsb(qs, qlen, rs, rlen, 0, 0);
        done_at = -1;
    }
    __syncthreads();
    if (target == 0 && (int)slab[K] >= qlen) return 0;
    const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);
    const int TW = LW - 2 * TB;
    WfaSteps blk;  // state at the start of the current block of scores
    blk.init(p, gm);
    int cur = 0;
    for (int s0 = 0; s0 + gm.g <= me;) {
        const int nsteps = min(TB, (me - s0) / gm.g);
        const int s_end = s0 + nsteps * gm.g;
        const int Re = reach(s_end, p);
        const int lo_e = max(max(-rlen, -me), -Re), hi_e = min(min(qlen, me), Re);
        const int16_t *src = slab + cur * set_len;
        int16_t *dst = slab + (cur ^ 1) * set_len;
        for (int a = lo_e; a <= hi_e; a += TW) {
            const int b = min(a + TW - 1, hi_e);
            const int base = a - TB;  // diagonal of local index 0
            for (int sl = 0; sl < gm.nM; ++sl)
                for (int j = tid; j < LW; j += NT) {
                    const int k = base + j;
                    Mt[sl * LW + j] = (k >= -K && k <= K) ? src[(int64_t)sl * W + k + K] : neg;
                }
            for (int sl = 0; sl < gm.nI; ++sl)
                for (int j = tid; j < LW; j += NT) {
                    const int k = base + j;
                    const bool in = k >= -K && k <= K;
                    const int16_t iv = in ? src[(int64_t)(gm.nM + sl) * W + k + K] : neg;
                    const int16_t dv = in ? src[(int64_t)(gm.nM + gm.nI + sl) * W + k + K] : neg;
                    *(uint32_t *)(IDt + sl * 2 * LW + 2 * j) =
                        (uint32_t)(uint16_t)iv | ((uint32_t)(uint16_t)dv << 16);
                }
            __syncthreads();
            WfaSteps st = blk;
            for (int t = 1; t <= nsteps; ++t) {
                const int s = s0 + t * gm.g;
                st.next(p, gm);
                const int shrink = nsteps - t;
                const int klo = max(max(max(-rlen, -me), -st.R), a - shrink);
                const int khi = min(min(min(qlen, me), st.R), b + shrink);
                const int16_t *Mo = Mt + WfaSteps::back(st.ms, st.dO, gm.nM) * LW;
                const int16_t *Mx = Mt + WfaSteps::back(st.ms, st.dx, gm.nM) * LW;
                const int16_t *IDe = IDt + WfaSteps::back(st.is, st.de, gm.nI) * 2 * LW;
                int16_t *Ms = Mt + st.ms * LW;
                int16_t *IDs = IDt + st.is * 2 * LW;
                int found = 0;
                for (int kb = klo + wave0; kb <= khi; kb += U * NT) {
                    const int j0 = kb + lane - base;  // local index of the lane's diagonal
                    int mv[U], mr[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (kb + u * NT > khi) break;
                        const int j = j0 + u * NT;
                        const int m = Mx[j];
                        __builtin_amdgcn_sched_barrier(0);  // keep the M read first
                        const int a_ = Mo[j - 1], b_ = Mo[j + 1];
                        const int c = IDe[2 * j - 2], d = IDe[2 * j + 3];
                        const int iv = max(a_, c) + 1;
                        const int dv = max(b_, d);  // negative only when both are `neg`
                        mv[u] = max(max(m + 1, iv), dv);
                        if (j + base <= khi)
                            *(uint32_t *)(IDs + 2 * j) =
                                (uint32_t)(uint16_t)(iv < 0 ? neg : (int16_t)iv) |
                                ((uint32_t)dv << 16);
                    }
                    uint64_t more = 0;
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (kb + u * NT > khi) break;
                        const int k = j0 + u * NT + base;
                        const int row = mv[u], col = row - k;
                        uint32_t a0, a1, b0, b1;
                        window_lsb(qs, med3(row, 0, qlen), a0, a1);
                        window_lsb(rs, med3(col, 0, rlen), b0, b1);
                        const int lim = min(qlen - row, rlen - col);
                        const int n = med3(match32(a0, a1, b0, b1), lim, 0);  // match >= 0
                        mr[u] = row + n;
                        if (__ballot(row >= 0 && n == 32 && lim > 32)) more |= 1ull << u;
                    }
                    if (more) {
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            if (!(more >> u & 1)) continue;
                            const int k = j0 + u * NT + base;
                            const int row = mv[u], col = row - k;
                            if (row >= 0 && mr[u] == row + 32 && min(qlen - row, rlen - col) > 32 &&
                                k <= khi)
                                mr[u] = extend_lsb(qs, qlen, rs, rlen, row + 32, col + 32);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        if (kb + u * NT > khi) break;
                        const int j = j0 + u * NT, k = j + base;
                        if (k <= khi) {
                            const int row = mv[u];
                            found |= (k == target && row >= 0 && mr[u] >= qlen);
                            Ms[j] = row < 0 ? neg : (int16_t)mr[u];
                        }
                    }
                }
                // the first score that reaches the target ends the pair
                if (found) done_at = s;
                __syncthreads();
                if (done_at == s) return s;
            }
            const int wd = b - a + 1;
            for (int sl = 0; sl < gm.nM; ++sl)
                for (int j = tid; j < wd; j += NT)
                    dst[(int64_t)sl * W + a + j + K] = Mt[sl * LW + (a + j - base)];
            for (int sl = 0; sl < gm.nI; ++sl)
                for (int j = tid; j < wd; j += NT) {
                    const uint32_t v = *(const uint32_t *)(IDt + sl * 2 * LW + 2 * (a + j - base));
                    dst[(int64_t)(gm.nM + sl) * W + a + j + K] = (int16_t)(v & 0xffffu);
                    dst[(int64_t)(gm.nM + gm.nI + sl) * W + a + j + K] = (int16_t)(v >> 16);
                }
            __syncthreads();
        }
        for (int t = 0; t < nsteps; ++t) blk.next(p, gm);
        cur ^= 1;
        s0 = s_end;
    }
    return -1;
}

// Orders the LDS accesses of one wavefront's lanes: a wavefront's LDS
// operations complete in issue order, so only the compiler must be kept
// from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS bytes the one-wavefront layout needs for a pair: both sequences and
// the int16 wavefront slots (0 when the pair ends before any score).
__device__ __forceinline__ int64_t wave_pair_bytes(int qlen, int rlen, int me, const WfaPen &p,
                                                   const WfaGeom &gm) {
    const int K = min(min(me, max(qlen, rlen)), reach(me, p));
    const int64_t W = 2 * (int64_t)K + 1;
    return 8 * (int64_t)(((qlen + 31) >> 5) + ((rlen + 31) >> 5)) +
           2 * (int64_t)(gm.nM + 2 * gm.nI) * W;
}

// One pair per wavefront: the 64 lanes stride over the diagonals of each
// score with the team recurrence, sequences and int16 wavefronts in the
// wavefront's own LDS part (`part`, at least wave_pair_bytes; the caller
// checks rows_fit_int16).  Steps are ordered by wave_sync and the pair ends
// on the wavefront's vote, so small pairs pay no workgroup barrier.
template <int U>
__device__ int wfa_pair_wave(const uint64_t *qg, int qlen, const uint64_t *rg, int rlen, int me,
                             const WfaPen &p, const WfaGeom &gm, uint8_t *part) {
    const int lane = threadIdx.x & 63;
    if (qlen == 0 && rlen == 0) return 0;
    const int target = qlen - rlen;
    const int K = min(min(me, max(qlen, rlen)), reach(me, p));
    const int W = 2 * K + 1;
    if (target != 0 && (target > K || -target > K)) return -1;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int slots = gm.nM + 2 * gm.nI;
    uint64_t *seq = (uint64_t *)part;
    for (int i = lane; i < qw + rw; i += kWaveThreads) seq[i] = i < qw ? qg[i] : rg[i - qw];
    const uint64_t *q = seq, *r = seq + qw;
    int16_t *M = (int16_t *)(part + 8 * (qw + rw));
    int16_t *I = M + gm.nM * W;
    int16_t *D = I + gm.nI * W;
    const int16_t neg = RowNeg<int16_t>::v;
    for (int i = lane; i < slots * W; i += kWaveThreads) M[i] = neg;
    wave_sync();
    if (lane == 0) M[K] = (int16_t)extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
    wave_sync();
    if (target == 0 && (int)M[K] >= qlen) return 0;
    WfaSteps st;
    st.init(p, gm);
    for (int s = gm.g; s <= me; s += gm.g) {
        st.next(p, gm);
        const int ms = st.ms * W, is = st.is * W;
        const int lo = max(max(-rlen, -me), -st.R), hi = min(min(qlen, me), st.R);
        const int mx = st.mx(gm) >= 0 ? st.mx(gm) * W : -1;
        const int mo = st.mo(gm) >= 0 ? st.mo(gm) * W : -1;
        const int ie = st.ie(gm) >= 0 ? st.ie(gm) * W : -1;
        int found = 0;
        for (int k0 = lo + lane; k0 <= hi; k0 += U * kWaveThreads) {
            int iv[U], dv[U], mv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * kWaveThreads, di = k + K;
                int a = kNegRow, c = kNegRow, bb = kNegRow, d = kNegRow, m = kNegRow;
                if (k <= hi) {
                    if (mo >= 0) {
                        if (di > 0) a = M[mo + di - 1];
                        if (di < W - 1) bb = M[mo + di + 1];
                    }
                    if (ie >= 0) {
                        if (di > 0) c = I[ie + di - 1];
                        if (di < W - 1) d = D[ie + di + 1];
                    }
                    if (mx >= 0) m = M[mx + di];
                }
                iv[u] = max(a, c) + 1;
                dv[u] = max(bb, d);
                mv[u] = max(max(m + 1, iv[u]), dv[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * kWaveThreads, di = k + K;
                if (k <= hi) {
                    I[is + di] = iv[u] < 0 ? neg : (int16_t)iv[u];
                    D[is + di] = dv[u] < 0 ? neg : (int16_t)dv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * kWaveThreads, di = k + K;
                if (k <= hi) {
                    int m = mv[u];
                    if (m >= 0) {
                        m = extend_diag(q, qlen, qw, r, rlen, rw, m, k);
                        found |= (k == target && m >= qlen);
                    }
                    M[ms + di] = m < 0 ? neg : (int16_t)m;
                }
            }
        }
        if (__ballot(found)) return s;
        wave_sync();
    }
    return -1;
}

// Everything one workgroup of NT threads needs to align a WFA pair: the
// branch-free LDS layout when sequences and int16 wavefronts fit it;
// otherwise sequences staged to LDS when they fit `lds_bytes`, int16
// wavefronts in LDS when they fit the rest, else in the workgroup's HBM slab
// (int16 rows when they fit int16, int32 otherwise).  Returns the cost, -1,
// or CGM_BAD_PAIR.
template <int U, int NT>
__device__ __forceinline__ int wfa_pair_block(const uint64_t *qg, int qlen, const uint64_t *rg,
                                              int rlen, int me, const WfaPen &p,
                                              const WfaGeom &gm, uint8_t *lds, int lds_bytes,
                                              int32_t *slab, int64_t slab_words,
                                              int tile_steps = kTileSteps) {
    if (qlen == 0 && rlen == 0) return 0;
    const int target = qlen - rlen;
    const int K = min(min(me, max(qlen, rlen)), reach(me, p));
    const int W = 2 * K + 1;
    if (target != 0 && (target > K || -target > K)) return -1;  // target never reached
    const int slots = gm.nM + 2 * gm.nI;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int64_t seq_bytes = 8 * (int64_t)(qw + rw);
    const int64_t wf16_bytes = 2 * (int64_t)slots * W;
    const bool fit16 = rows_fit_int16(qlen, me, p);
    if (fit16 && team_fast_bytes(qlen, rlen, W, gm) <= lds_bytes)
        return wfa_pair_team_fast<U, NT>(qg, qlen, rg, rlen, me, p, gm, lds, W, K);
    const int LWf = tiled_fast_width(qlen, rlen, gm, lds_bytes);
    if (fit16 && LWf - 2 * tile_steps >= kTileMinWidth && (int64_t)slots * W <= slab_words)
        return wfa_pair_tiled_fast<U, NT>(qg, qlen, rg, rlen, me, p, gm, (int16_t *)slab, W, K,
                                          lds, LWf, tile_steps);
    const uint64_t *q = qg, *r = rg;
    if (seq_bytes <= lds_bytes) {
        // Sequences go to LDS whenever they fit, also when the wavefronts
        // do not: the extension reads them on every diagonal.
        uint64_t *seq = (uint64_t *)lds;
        for (int i = threadIdx.x; i < qw + rw; i += blockDim.x) seq[i] = i < qw ? qg[i] : rg[i - qw];
        q = seq;
        r = seq + qw;
        // the team function's first barrier orders these copies
        if (fit16 && seq_bytes + wf16_bytes <= lds_bytes)
            return wfa_pair_team<U>(q, qlen, r, rlen, me, p, gm, (int16_t *)(lds + seq_bytes), W, K);
        // Tiled: two int16 slab sets (slots * W int32 words) and LDS tiles.
        const int LW = (int)((lds_bytes - seq_bytes) / (2 * slots));
        if (fit16 && LW - 2 * tile_steps >= kTileMinWidth && (int64_t)slots * W <= slab_words)
            return wfa_pair_tiled<U>(q, qlen, r, rlen, me, p, gm, (int16_t *)slab, W, K,
                                     (int16_t *)(lds + seq_bytes), LW, tile_steps);
    }
    // Untiled HBM slab of slab_words int32 words: int16 rows when they fit
    // (half the traffic), int32 otherwise.
    if (fit16 && (int64_t)slots * W <= 2 * slab_words)
        return wfa_pair_team<U>(q, qlen, r, rlen, me, p, gm, (int16_t *)slab, W, K);
    if ((int64_t)slots * W <= slab_words)
        return wfa_pair_team<U>(q, qlen, r, rlen, me, p, gm, slab, W, K);
    return CGM_BAD_PAIR;
}

}  // namespace wfadev
}  // namespace cgm
