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
// in LDS when they fit the workgroup's budget, otherwise the sequences are
// read from global memory and int32 wavefronts live in an HBM slab.
// ---------------------------------------------------------------------------
constexpr int kWaveThreads = 64;
constexpr int kNegRow = -(1 << 30);

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

// Storage of wavefront rows: int16 in LDS when every row fits (see
// rows_fit_int16), int32 in an HBM slab otherwise.  Invalid entries hold the
// type's most negative value; any stored value < 0 is invalid.
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
// __syncthreads_or per score is both the barrier and the termination test.
// q / r may point to LDS copies of the sequences or to global memory.
template <typename T>
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
    if (tid == 0) M[K] = (T)extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
    __syncthreads();
    if (target == 0 && (int)M[K] >= qlen) return 0;
    for (int s = gm.g; s <= me; s += gm.g) {
        const int si = s / gm.g;
        const int ms = (si % gm.nM) * W, is = (si % gm.nI) * W;
        const int Rs = reach(s, p);
        const int lo = max(max(-rlen, -me), -Rs), hi = min(min(qlen, me), Rs);
        const int mx = s >= p.x ? (((s - p.x) / gm.g) % gm.nM) * W : -1;
        const int mo = s >= p.o + p.e ? (((s - p.o - p.e) / gm.g) % gm.nM) * W : -1;
        const int ie = s >= p.e ? (((s - p.e) / gm.g) % gm.nI) * W : -1;
        int found = 0;
        for (int k = lo + tid; k <= hi; k += nt) {
            const int di = k + K;
            int iv = kNegRow, dv = kNegRow, mv = kNegRow;
            if (mo >= 0) {
                if (di > 0) iv = (int)M[mo + di - 1] + 1;
                if (di < W - 1) dv = (int)M[mo + di + 1];
            }
            if (ie >= 0) {
                if (di > 0) iv = max(iv, (int)I[ie + di - 1] + 1);
                if (di < W - 1) dv = max(dv, (int)D[ie + di + 1]);
            }
            if (mx >= 0) mv = (int)M[mx + di] + 1;
            mv = max(max(mv, iv), dv);
            I[is + di] = iv < 0 ? neg : (T)iv;
            D[is + di] = dv < 0 ? neg : (T)dv;
            if (mv >= 0) {
                mv = extend_diag(q, qlen, qw, r, rlen, rw, mv, k);
                M[ms + di] = (T)mv;
                found |= (k == target && mv >= qlen);
            } else {
                M[ms + di] = neg;
            }
        }
        if (__syncthreads_or(found)) return s;
    }
    return -1;
}

// Everything one workgroup needs to align a WFA pair: sequences staged to
// LDS and int16 wavefronts in LDS when both fit `lds_bytes`, otherwise the
// sequences are read from global memory and the wavefronts live in the
// workgroup's HBM slab (int32).  Returns the cost, -1, or CGM_BAD_PAIR.
__device__ __forceinline__ int wfa_pair_block(const uint64_t *qg, int qlen, const uint64_t *rg,
                                              int rlen, int me, const WfaPen &p,
                                              const WfaGeom &gm, uint8_t *lds, int lds_bytes,
                                              int32_t *slab, int64_t slab_words) {
    if (qlen == 0 && rlen == 0) return 0;
    const int target = qlen - rlen;
    const int K = min(min(me, max(qlen, rlen)), reach(me, p));
    const int W = 2 * K + 1;
    if (target != 0 && (target > K || -target > K)) return -1;  // target never reached
    const int slots = gm.nM + 2 * gm.nI;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int64_t seq_bytes = 8 * (int64_t)(qw + rw);
    const int64_t wf16_bytes = 2 * (int64_t)slots * W;
    if (rows_fit_int16(qlen, me, p) && seq_bytes + wf16_bytes <= lds_bytes) {
        uint64_t *seq = (uint64_t *)lds;
        for (int i = threadIdx.x; i < qw + rw; i += blockDim.x) seq[i] = i < qw ? qg[i] : rg[i - qw];
        // the team function's first barrier orders these copies
        return wfa_pair_team(seq, qlen, seq + qw, rlen, me, p, gm,
                             (int16_t *)(lds + seq_bytes), W, K);
    }
    if ((int64_t)slots * W <= slab_words)
        return wfa_pair_team(qg, qlen, rg, rlen, me, p, gm, slab, W, K);
    return CGM_BAD_PAIR;
}

}  // namespace wfadev
}  // namespace cgm
