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
// Wavefront slots live in LDS when 4 * slots * W bytes fit the block's LDS
// budget, otherwise in a per-block HBM slab.
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

template <typename Ptr>
__device__ int wfa_pair_wave(const uint64_t *q, int qlen, const uint64_t *r, int rlen, int me,
                             const WfaPen &p, const WfaGeom &gm, Ptr wf, int W, int K) {
    const int lane = threadIdx.x;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int slots = gm.nM + 2 * gm.nI;
    for (int i = lane; i < slots * W; i += kWaveThreads) wf[i] = kNegRow;
    __syncthreads();
    Ptr M = wf;
    Ptr I = wf + gm.nM * W;
    Ptr D = I + gm.nI * W;
    const int target = qlen - rlen;
    if (lane == 0) M[K] = extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
    __syncthreads();
    if (target == 0 && M[K] >= qlen) return 0;
    for (int s = gm.g; s <= me; s += gm.g) {
        const int si = s / gm.g;
        const int ms = (si % gm.nM) * W, is = (si % gm.nI) * W;
        const int Rs = reach(s, p);
        const int lo = max(max(-rlen, -me), -Rs), hi = min(min(qlen, me), Rs);
        const int mx = s >= p.x ? (((s - p.x) / gm.g) % gm.nM) * W : -1;
        const int mo = s >= p.o + p.e ? (((s - p.o - p.e) / gm.g) % gm.nM) * W : -1;
        const int ie = s >= p.e ? (((s - p.e) / gm.g) % gm.nI) * W : -1;
        bool found = false;
        for (int k = lo + lane; k <= hi; k += kWaveThreads) {
            const int di = k + K;
            int iv = kNegRow, dv = kNegRow, mv = kNegRow;
            if (mo >= 0) {
                if (di > 0) iv = M[mo + di - 1] + 1;
                if (di < W - 1) dv = M[mo + di + 1];
            }
            if (ie >= 0) {
                if (di > 0) iv = max(iv, I[ie + di - 1] + 1);
                if (di < W - 1) dv = max(dv, D[ie + di + 1]);
            }
            if (mx >= 0) mv = M[mx + di] + 1;
            iv = iv < 0 ? kNegRow : iv;
            dv = dv < 0 ? kNegRow : dv;
            mv = max(max(mv, iv), dv);
            I[is + di] = iv;
            D[is + di] = dv;
            if (mv >= 0) {
                mv = extend_diag(q, qlen, qw, r, rlen, rw, mv, k);
            } else {
                mv = kNegRow;
            }
            M[ms + di] = mv;
            found |= (k == target && mv >= qlen);
        }
        if (__any(found)) return s;
        __syncthreads();
    }
    return -1;
}

}  // namespace wfadev
}  // namespace cgm
