// cgm_sw_wave_dev.h -- banded SW with one 64-lane wavefront per pair: an
// anti-diagonal sweep with every cell's inputs in registers (gfx950).
//
// Same results as smith_waterman_banded (align_cpu.mojo:229-309).  The band
// of effective width B <= 64 has diagonals d = ri - qi in [-B, B]; a cell of
// diagonal d lies on anti-diagonal a = qi + ri when d = a (mod 2).  Lane l
// owns diagonals de = -B + 2l and do = -B + 2l + 1 and computes one cell per
// anti-diagonal: on de when a = B (mod 2) ("even" steps), on do otherwise.
// A cell's inputs lie on the two previous anti-diagonals: diag H(qi-1, ri-1)
// is the lane's own last value on the same diagonal, up H(qi-1, ri) and left
// H, E(qi, ri-1) are diagonals d + 1 and d - 1 of the previous step -- the
// lane's other diagonal, or the neighbour lane's one DPP shift away (wave
// shifts: down for do's up, up for de's left), with the reference's band
// edge values entering at lanes 0 and 63 (left of d = -B: H 0 and no E;
// above d = B: H 0).  At B = 64 diagonal d = 64 has no lane of its own: its
// cell is computed in a third register set of every lane, of which lane 63's
// is the one that counts (its inputs -- diagonal 63's H and E, and the
// diagonal's own previous H -- are lane 63's own values).
//
// Bases move the same way: every step of a parity shifts the query bases one
// lane up and the reference bases one lane down, one new base of each
// entering at the edge.  The new bases come from two feed registers, one
// base per lane: the query feed holds the next 64 query bases with the next
// one in lane 0 and moves one lane down per step, the reference feed holds
// the next 64 reference bases with the next one in lane 63 and moves one
// lane up; every 64 steps both are refilled from words loaded 64 steps
// earlier.  So a step makes no scalar read of sequence data and no branch
// beyond the refill test.  Cells outside the matrix hold H = 0 and E = 0,
// like the reference's skipped cells.
// Each diagonal keeps its first strict maximum; the end cell is the
// largest of them with the smallest query row, then reference column,
// which is the reference's first row-major maximum.
//
// Pinned on the CPU by tests/test_sw_wave_schedule.py (this data flow
// restated in Python equals the oracle at 4, 5 and 64 lanes; planted slips
// in the extra diagonal, the diag source, the up side, the maximum's tie
// rule and the base shift each fail; the feeds' positions, refills and the
// extra diagonal's bases are restated there and equal the direct reads).
//
// The layout needs 2 (qlen + rlen) steps per pair but keeps every step's
// 64 cells in one instruction stream: the low-latency layout for long reads,
// where one pair per lane would hold a lane for the whole band area.
#pragma once

#include "cgm_internal.h"

namespace cgm {
namespace swwave {

constexpr int kWaveNeg = -(1 << 30);

__device__ __forceinline__ int shift_up(int edge, int v) {  // lane l <- l - 1, lane 0 <- edge
    return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int shift_down(int edge, int v) {  // lane l <- l + 1, lane 63 <- edge
    return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}

// The packed word holding base `pos` (0 outside [0, len), so its base reads 0).
__device__ __forceinline__ uint64_t feed_word(const uint64_t *w, int len, int pos) {
    return (unsigned)pos < (unsigned)len ? w[pos >> 5] : 0ull;
}
__device__ __forceinline__ int feed_base(uint64_t word, int pos) {
    return (int)((word >> (62 - 2 * (pos & 31))) & 3ull);
}
__device__ __forceinline__ int base_or0(const uint64_t *w, int len, int pos) {
    return feed_base(feed_word(w, len, pos), pos);
}

// 64-bit key of a candidate end cell: larger score first, then smaller
// query row, then smaller reference column (score < 2^23, rows < 2^20).
__device__ __forceinline__ uint64_t end_key(int v, int qi, int ri) {
    return v > 0 ? ((uint64_t)v << 40) | ((uint64_t)(0xFFFFF - qi) << 20) | (uint64_t)(0xFFFFF - ri)
                 : 0ull;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)k, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), off);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        k = o > k ? o : k;
    }
    return k;
}

// One pair by one wavefront (every lane active): B = effective band <= 64,
// match * min(qlen, rlen) < 2^23, qlen and rlen < 2^20 (checked by the
// classifier).  The results are valid in every lane.
__device__ __forceinline__ void sw_wave_pair(const uint64_t *q, int ql, const uint64_t *r, int rl,
                                             int B, const SwScoring &sc, int &score, int &qend,
                                             int &rend) {
    const int lane = threadIdx.x & 63;
    score = qend = rend = 0;
    if (ql == 0 || rl == 0) return;
    const int match = sc.match, mismatch = sc.mismatch, ge = sc.gap_extend;
    const int oe = sc.gap_open + sc.gap_extend;
    const bool extra = B == 64;
    int He = 0, Ee = 0, Ho = 0, Eo = 0;
    int bve = 0, bae = 0, bvo = 0, bao = 0;
    int Hx = 0, bvx = 0, bax = 0;  // diagonal d = 64 (lane 63's copy counts)
    const int a_first = (B & 1) - 1;
    // bases as they stand before the first step of each parity
    int qbo = base_or0(q, ql, ((a_first - 2 + B - 1) >> 1) - lane);
    int rbo = base_or0(r, rl, ((a_first - 2 - B + 1) >> 1) + lane);
    int qbe = base_or0(q, ql, ((a_first - 1 + B) >> 1) - lane);
    int rbe = base_or0(r, rl, ((a_first - 1 - B) >> 1) + lane);
    // the last anti-diagonal holding a cell of the band
    const int a_end = (ql - 1) + min(rl - 1, ql - 1 + B);
    // feeds: the step starting at anti-diagonal a (c1 = (a + B - 1) / 2,
    // c2 = (a - B + 1) / 2) takes query base c1 + 1 from lane 0 of fq and
    // reference base c2 + 63 from lane 63 of fr; the odd step's query base
    // c1 is the previous step's (qlast, lane 0)
    const int c1_0 = (a_first + B - 1) >> 1, c2_0 = (a_first - B + 1) >> 1;
    int pq = c1_0 + 1 + lane;       // this lane's query position in the feed block
    int pr = c2_0 + 63 + 63 - lane;  // this lane's reference position
    int fq = base_or0(q, ql, pq), fr = base_or0(r, rl, pr);
    uint64_t nq = feed_word(q, ql, pq + 64), nr = feed_word(r, rl, pr + 64);
    int qlast = base_or0(q, ql, c1_0);
    int fill = 0;
    // loop-invariant VGPR copies of the substitution scores
    int match_v, mismatch_v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(match_v) : "s"(match));
    asm volatile("v_mov_b32 %0, %1" : "=v"(mismatch_v) : "s"(mismatch));
    for (int a = a_first; a <= a_end; a += 2) {
        // uniform per step: the lanes inside the matrix and the band, [lo, lo + n)
        const int c1 = (a + B - 1) >> 1;  // odd step: lane l's query row is c1 - l
        const int c2 = (a - B + 1) >> 1;  // both steps: lane l's reference column is c2 + l
        const int lo_o = max(max(c1 - ql + 1, -c2), 0);
        const unsigned n_o = (unsigned)max(min(min(c1, rl - 1 - c2), B - 1) - lo_o + 1, 0);
        const int lo_e = max(max(c1 + 2 - ql, -c2), 0);
        const unsigned n_e = (unsigned)max(min(min(c1 + 1, rl - 1 - c2), min(B, 63)) - lo_e + 1, 0);
        // ---- odd step a: diagonal do(l)
        {
            qbo = shift_up(qlast, qbo);
            rbo = shift_down(fr, rbo);
            const int up = shift_down(Hx, He);  // Hx stays 0 unless B = 64
            const int sub = qbo == rbo ? match_v : mismatch_v;
            int e = max(Ee + ge, up + oe);
            int h = max(max(max(Ho + sub, e), He + oe), 0);
            const bool in = (unsigned)(lane - lo_o) < n_o;
            h = in ? h : 0;
            e = in ? e : 0;
            if (h > bvo) {
                bvo = h;
                bao = a;
            }
            Ho = h;
            Eo = e;
        }
        // ---- even step a + 1: diagonal de(l)
        const int qx = qbe;  // lane 63: query base of the d = 64 cell
        {
            qbe = shift_up(fq, qbe);
            rbe = shift_down(fr, rbe);
            const int hl = shift_up(0, Ho), el = shift_up(kWaveNeg, Eo);
            const int sub = qbe == rbe ? match_v : mismatch_v;
            int e = max(el + ge, Ho + oe);
            int h = max(max(max(He + sub, e), hl + oe), 0);
            const bool in = (unsigned)(lane - lo_e) < n_e;
            h = in ? h : 0;
            e = in ? e : 0;
            if (h > bve) {
                bve = h;
                bae = a + 1;
            }
            He = h;
            Ee = e;
        }
        // ---- the feeds move one base; every 64 steps the next block enters
        qlast = fq;
        fq = shift_down(0, fq);
        fr = shift_up(0, fr);
        if (++fill == 64) {
            fill = 0;
            pq += 64;
            pr += 64;
            fq = feed_base(nq, pq);
            fr = feed_base(nr, pr);
            nq = feed_word(q, ql, pq + 64);
            nr = feed_word(r, rl, pr + 64);
        }
        // ---- d = 64 at step a + 1 beside lane 63: left = lane 63's do cell
        // of step a, up = outside the band (H 0), reference base c2 + 64
        // (now in lane 63 of fr)
        if (extra) {
            const int qi = (a + 1 - 64) >> 1, ri = (a + 1 + 64) >> 1;
            const bool in = qi >= 0 && qi < ql && ri < rl;
            const int sub = qx == fr ? match_v : mismatch_v;
            int hx = max(max(max(Hx + sub, max(Eo + ge, oe)), Ho + oe), 0);
            hx = in ? hx : 0;
            if (hx > bvx) {
                bvx = hx;
                bax = a + 1;
            }
            Hx = hx;
        }
    }
    // the end cell: each diagonal's first maximum, merged over the wavefront
    uint64_t k = end_key(bve, ((bae + B) >> 1) - lane, ((bae - B) >> 1) + lane);
    const uint64_t ko = end_key(bvo, ((bao + B - 1) >> 1) - lane, ((bao - B + 1) >> 1) + lane);
    k = ko > k ? ko : k;
    if (extra && lane == 63) {
        const uint64_t kx = end_key(bvx, (bax - 64) >> 1, (bax + 64) >> 1);
        k = kx > k ? kx : k;
    }
    k = wave_max_u64(k);
    if (k) {
        score = (int)(k >> 40);
        qend = 0xFFFFF - (int)((k >> 20) & 0xFFFFF);
        rend = 0xFFFFF - (int)(k & 0xFFFFF);
    }
}

}  // namespace swwave
}  // namespace cgm
