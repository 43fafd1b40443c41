// cgm_wfa_reg.h -- register-resident gap-affine WFA layout (gfx950).
//
// Same recurrence and exact reductions as the wave layout (cgm_wfa_dev.h:
// scores that are multiples of g only, diagonals |k| <= R(s) within
// [-rlen, qlen], slots never cleared, NEXT and EXTEND fused), one more exact
// reduction (cells that max_error rules out, reg_span below), and a
// different home for the wavefronts.  Each lane owns one diagonal per
// block: block gb holds the 64 diagonals kbase + 64 gb .. + 63 of the
// pair's alive span [kbase, kmax], and the lane keeps that diagonal's M, I
// and D history in VGPRs as rings of NI = DE + 1 and NM >= DM + 1 slots
// (DM = max(x, o + e) / g, DE = e / g, NM a multiple of NI).  The score loop
// is unrolled by the rings' common period NM, so every slot index is a
// compile-time register; a slot is read without an existence test because
// before its score it still holds the invalid value it started with.
//   I[s,k] = max(M[s-o-e,k-1], I[s-e,k-1]) + 1   -> own max, DPP wave_shr:1
//   D[s,k] = max(M[s-o-e,k+1], D[s-e,k+1])       -> own max, DPP wave_shl:1
//   M[s,k] = extend(max(M[s-x,k] + 1, I, D))
// so NEXT costs no memory access at all.  Lane 0 and lane 63 take the
// neighbour across the block edge from the adjacent block's edge value of
// the step (max(M_o, I_e) of its lane 63, max(M_o, D_e) of its lane 0),
// published through LDS at the end of the previous step (two buffers by
// step parity, one barrier per step; the two outer edges read the invalid
// value).  Blocks are interleaved over the workgroup's NW wavefronts
// (gb = b NW + w), so every wavefront gets work from the first scores on.
// Per score the alive diagonals form one range [lo(s), hi(s)]: lanes outside
// it are held invalid, a block wholly outside it skips the step (a scalar
// test) once its slots and published edges hold the invalid value, and a
// block that leaves the range writes the invalid value for NM more scores
// before it skips.  Sequences sit in LDS as LSB-first 32-bit words, and the
// extension is the wave layout's 32-bases-per-step comparison.
//
// Pinned on the CPU by tests/test_wfa_reg_schedule.py (this data flow
// restated in Python equals the oracle, also at tight max_error; planted
// slips in the edge side, the ring offset, the outer edge value, the block
// skip, the gap budget and the leaving blocks' flush each fail; the span's
// bisection equals a scan over all scores).

#pragma once

#include "cgm_wfa_dev.h"

namespace cgm {
namespace wfareg {

using wfadev::extend_lsb;
using wfadev::match32;
using wfadev::med3;
using wfadev::window_lsb;

constexpr int kNeg = -(1 << 30);

__host__ __device__ constexpr int gcd_c(int a, int b) { return b ? gcd_c(b, a % b) : a; }
__host__ __device__ constexpr int pos_mod(int a, int n) { return ((a % n) + n) % n; }

// Penalties in units of g: x / g, (o + e) / g, e / g.  The instantiated
// set is x:o+e:e = 2:4:1 (the reference's defaults 4, 6, 2 and every
// multiple of them); other penalties take the wave layout.
constexpr int kRegDX = 2, kRegDO = 4, kRegDE = 1;

// LDS bytes a pair needs: the two sequences as LSB-first words, the edge
// buffers of nblk_cap blocks (+2 outer entries) and two done flags.
__host__ __device__ inline int64_t reg_lds_bytes(int qlen, int rlen, int nblk_cap) {
    const int64_t nq = 2 * (int64_t)((qlen + 31) >> 5) + 4, nr = 2 * (int64_t)((rlen + 31) >> 5) + 4;
    return 4 * (nq + nr) + 4 * 4 * ((int64_t)nblk_cap + 2) + 16;
}

// Exact pruning by max_error: from a cell at score s on diagonal k every
// completion still needs |k - t| gap characters (t = qlen - rlen) at e or
// more each, so a cell with e |k - t| > me - s lies on no alignment within
// max_error and is held invalid.  Diagonal k is then alive at score s for
// lo(s) <= k <= hi(s), lo(s) = max(-rlen, -R(s), t - P(s)),
// hi(s) = min(qlen, R(s), t + P(s)), P(s) = (me - s) / e; and a pair with
// e |t| > me has no alignment within max_error at all.
//
// G(u) = max over s = 0, g, .. <= me of min(R(s), u + (me - s) / e), for
// u + me / e >= 0: R rises and the second term falls with s, so the
// maximum sits where they cross, found by bisection on s / g.
__host__ __device__ inline int reg_span_side(int u, int me, int o, int e, int g) {
    auto R = [&](int s) { return s >= o + e ? (s - o) / e : 0; };
    auto H = [&](int s) { return u + (me - s) / e; };
    int a = 0, b = me / g;  // largest j with R(j g) <= H(j g); j = 0 holds
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (R(m * g) <= H(m * g)) a = m;
        else b = m - 1;
    }
    const int v = R(a * g);
    return a < me / g ? max(v, H((a + 1) * g)) : v;
}

// The alive diagonal span [lo, hi] over all scores of a pair with
// e |qlen - rlen| <= me.
__host__ __device__ inline void reg_span(int qlen, int rlen, int me, int o, int e, int g,
                                         int &lo, int &hi) {
    const int t = qlen - rlen;
    hi = min(qlen, reg_span_side(t, me, o, e, g));
    lo = max(-rlen, -reg_span_side(-t, me, o, e, g));
}

// The number of 64-diagonal blocks of a pair (1 for a pair with no
// alignment within max_error, which ends before its first step).
__host__ __device__ inline int reg_blocks(int qlen, int rlen, int me, int o, int e, int g) {
    const int64_t t = (int64_t)qlen - rlen;
    if ((t < 0 ? -t : t) * e > me) return 1;
    int lo, hi;
    reg_span(qlen, rlen, me, o, e, g, lo, hi);
    return (hi - lo + 64) / 64;
}

// Layout classes of a WFA pair, used by every WFA launcher: 0..6 the
// register layout with (NB blocks per wavefront, NW wavefronts) =
// (2, 1), (4, 1), (8, 1), (8, 2), (8, 4), (8, 8), (8, 16), each within an
// LDS budget (a pair of 9-16 blocks on two wavefronts rather than four: half
// the wavefronts meet at each score's barrier, and the workgroup's block
// slots are mostly in use); 7 the 512-thread team layout (52 KB, tiled over
// an HBM slab beyond it), 8 the 256-thread team layout (int16 rows within
// 20 KB).
constexpr int kWfaClasses = 9;
constexpr int kWfaRegClasses = 7;
constexpr int kWfaTeamBig = 7, kWfaTeamSmall = 8;
__host__ __device__ constexpr int reg_class_nb(int c) { return c == 0 ? 2 : c == 1 ? 4 : 8; }
__host__ __device__ constexpr int reg_class_nw(int c) {
    return c <= 2 ? 1 : c == 3 ? 2 : c == 4 ? 4 : c == 5 ? 8 : 16;
}
__host__ __device__ constexpr int reg_class_lds(int c) {
    return c == 0 ? 4096 : c == 1 ? 6144 : c == 2 ? 8192 : c == 3 ? 12288 : c == 4 ? 16384
         : c == 5 ? 24576 : 32768;
}
// waves per SIMD each class's registers allow (kept by amdgpu_waves_per_eu)
__host__ __device__ constexpr int reg_class_wpe(int c) { return c == 0 ? 8 : c == 1 ? 6 : 4; }

__device__ __forceinline__ int wfa_pair_class(int qlen, int rlen, int me, const WfaPen &p,
                                              const wfadev::WfaGeom &gm, bool reg_pen) {
    if (reg_pen) {
        const int nblk = reg_blocks(qlen, rlen, me, p.o, p.e, gm.g);
#pragma unroll
        for (int c = 0; c < kWfaRegClasses; ++c)
            if (nblk <= reg_class_nb(c) * reg_class_nw(c) &&
                reg_lds_bytes(qlen, rlen, reg_class_nb(c) * reg_class_nw(c)) <= reg_class_lds(c))
                return c;
    }
    return wfadev::wfa_fits_small(qlen, rlen, me, p, gm) ? kWfaTeamSmall : kWfaTeamBig;
}

// Sort key of a WFA pair: its class (first classes first in a descending
// sort; bits 24-27, below the mixed path's own class bits), then its work
// (longest first, 24 bits); never 0, the key of excluded pairs.
__device__ __forceinline__ uint32_t wfa_class_key(int cls, uint32_t work) {
    return ((uint32_t)(kWfaClasses - 1 - cls) << 24) | (min(work, (uint32_t)((1 << 24) - 2)) + 1);
}

// Bases pos .. pos + 15 of an LSB-first sequence as one word (base pos + i
// at bits 2i).
__device__ __forceinline__ uint32_t window16_lsb(const uint32_t *s, int pos) {
    const uint32_t *w = s + (pos >> 4);
    return __builtin_amdgcn_alignbit(w[1], w[0], 2u * (uint32_t)(pos & 15));
}

// Leading bases (0..16) on which two 16-base windows agree.
__device__ __forceinline__ int match16(uint32_t a, uint32_t b) {
    return (int)(min(wfadev::lowest_bit(a ^ b), 32u) >> 1);
}

// Ring sizes: NI = DE + 1 I and D slots, and M slots rounded up from
// max(DX, DO) + 1 to a multiple of NI, so that both rings repeat with the
// M ring's period P (a shorter unrolled score loop; the extra M slot is
// never read).
template <int DX, int DO, int DE>
struct RegRing {
    static constexpr int NI = DE + 1;
    static constexpr int NM = ((DX > DO ? DX : DO) + 1 + NI - 1) / NI * NI;
    static constexpr int P = NM;
};

// One pair by a workgroup of NW wavefronts with NB blocks per wavefront
// (nblk <= NB * NW).  lds: at least reg_lds_bytes(qlen, rlen, NB * NW).
template <int NB, int NW, int DX, int DO, int DE>
__device__ int wfa_reg_pair(const uint64_t *qg, int qlen, const uint64_t *rg, int rlen, int me,
                            const WfaPen &p, int g, uint8_t *lds) {
    using Ring = RegRing<DX, DO, DE>;
    constexpr int NM = Ring::NM, NI = Ring::NI, P = Ring::P;
    constexpr int NBLK = NB * NW;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int nq = 2 * qw + 4, nr = 2 * rw + 4;
    uint32_t *qs = (uint32_t *)lds;
    uint32_t *rs = qs + nq;
    // edge buffers, two by step parity: E[buf][gb + 1] = {max(M_o, I_e) of
    // block gb's lane 63, max(M_o, D_e) of its lane 0}; entries 0 and
    // nblk + 1 are the outer edges and stay invalid
    int2 *E = (int2 *)(rs + nr);
    int *done = (int *)(E + 2 * (NBLK + 2));
    for (int i = tid; i < nq + nr; i += NW * 64) {
        const bool isq = i < nq;
        const int j = isq ? i : i - nq;
        uint32_t v = 0;
        if ((j >> 1) < (isq ? qw : rw)) {
            const uint64_t wd = (isq ? qg : rg)[j >> 1];
            v = __builtin_bitreverse32((j & 1) ? (uint32_t)wd : (uint32_t)(wd >> 32));
        }
        qs[i] = v;
    }
    for (int i = tid; i < 2 * (NBLK + 2); i += NW * 64) E[i] = make_int2(kNeg, kNeg);
    if (tid == 0) done[0] = done[1] = -1;
    __syncthreads();

    const int target = qlen - rlen;
    if ((int64_t)(target < 0 ? -target : target) * p.e > me) return -1;
    int kbase, kmax;
    reg_span(qlen, rlen, me, p.o, p.e, g, kbase, kmax);
    const int nblk = (kmax - kbase + 64) / 64;
    // the block and lane of the target diagonal
    const int tb = (target - kbase) >> 6, tl = (target - kbase) & 63;

    int Mr[NB][NM], Ir[NB][NI], Dr[NB][NI];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int j = 0; j < NM; ++j) Mr[b][j] = kNeg;
#pragma unroll
        for (int j = 0; j < NI; ++j) Ir[b][j] = Dr[b][j] = kNeg;
    }
    // score 0: M[0, 0] = the matching prefix, in the lane owning diagonal 0
    const int g0 = -kbase >> 6, l0 = -kbase & 63;
#pragma unroll
    for (int b = 0; b < NB; ++b)
        if (b * NW + w == g0 && lane == l0) Mr[b][0] = extend_lsb(qs, qlen, rs, rlen, 0, 0);
    // score 0 alone reaches the target only on diagonal 0
    if (target == 0) {
        int m0 = kNeg;
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (b * NW + w == g0 && lane == l0) m0 = Mr[b][0];
        if (__ballot(m0 >= qlen)) done[0] = 0;
        __syncthreads();
        if (done[0] == 0) return 0;
    }
    // edges for step 1, from the block holding diagonal 0
    {
        constexpr int mo1 = pos_mod(1 - DO, NM), ie1 = pos_mod(1 - DE, NI);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b * NW + w == g0) {
                const int ei = __builtin_amdgcn_readlane(max(Mr[b][mo1], Ir[b][ie1]), 63);
                const int ed = max(Mr[b][mo1], Dr[b][ie1]);
                if (lane == 0) E[(NBLK + 2) + g0 + 1] = make_int2(ei, ed);
            }
        }
    }
    __syncthreads();
    // loop-invariant VGPR copies of the window clamps' bounds
    int zero_v, qlen_v, rlen_v;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero_v));
    asm volatile("v_mov_b32 %0, %1" : "=v"(qlen_v) : "s"(qlen));
    asm volatile("v_mov_b32 %0, %1" : "=v"(rlen_v) : "s"(rlen));
    // with o + e = DO g and e = DE g exactly, every per-score quantity is a
    // function of si = s / g: R(s) = (si - DO + DE) / DE from si = DO on,
    // P(s) = (mg - si) / DE with mg = me / g, and s <= me iff si <= mg
    const int mg = me / g;
    // blocks are gb = b NW + w: the first b of this wavefront at or after
    // block index x is (x - w + NW - 1) >> LNW, the last at or before it
    // (x - w) >> LNW (arithmetic shifts: floor and ceiling of x / NW)
    constexpr int LNW = NW == 1 ? 0 : NW == 2 ? 1 : NW == 4 ? 2 : NW == 8 ? 3 : 4;
    static_assert((1 << LNW) == NW, "NW: 1, 2, 4, 8 or 16");
    // this wavefront's b of the target diagonal's block, or -1
    const int tbb = ((tb - w) & (NW - 1)) == 0 && tb >= 0 ? (tb - w) >> LNW : -1;
    int seen = -1;
    for (int base = 0;; base += P) {
#pragma unroll
        for (int ph = 1; ph <= P; ++ph) {
            const int si = base + ph;
            if (si > mg) return seen;
            const int s = si * g;
            const int Rs = si >= DO ? (si - DO + DE) / DE : 0;
            const int Ps = (mg - si) / DE;
            const int Pl = (mg - max(si - NM, 0)) / DE;  // P(s - NM g), P(0) before it
            // the alive diagonals of this score
            const int lo_s = max(max(-rlen, -Rs), target - Ps);
            const int hi_s = min(min(qlen, Rs), target + Ps);
            // the blocks that run: every block that may hold a cell alive at
            // one of the scores s - NM g .. s (so a block that leaves the
            // alive range writes the invalid value for NM more scores and
            // publishes edges from its older slots meanwhile): this
            // wavefront's b in [b_lo, b_lo + nrun)
            const int run_lo = max(max(-rlen, -Rs), target - Pl);
            const int run_hi = min(min(qlen, Rs), target + Pl);
            const int b_lo = (((run_lo - kbase) >> 6) - w + NW - 1) >> LNW;
            const int b_hi = ((min((run_hi - kbase) >> 6, nblk - 1)) - w) >> LNW;
            const unsigned nrun = (unsigned)max(b_hi - b_lo + 1, 0);
            const int ms = ph % NM, is = ph % NI;
            const int mx = pos_mod(ph - DX, NM);
            const int mo = pos_mod(ph - DO, NM), ie = pos_mod(ph - DE, NI);
            const int mo2 = pos_mod(ph + 1 - DO, NM), ie2 = pos_mod(ph + 1 - DE, NI);
            const int cur = (si & 1) * (NBLK + 2), nxt = (NBLK + 2) - cur;
            int found = 0;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int gb = b * NW + w, kb0 = kbase + 64 * gb;
                if ((unsigned)(b - b_lo) < nrun) {
                    // block gb - 1's lane 63 and block gb + 1's lane 0
                    const int ti_in = E[cur + gb].x, td_in = E[cur + gb + 2].y;
                    // each lane's max(M_o, I_e) and max(M_o, D_e), shifted
                    // one lane up and one lane down
                    const int ti = max(Mr[b][mo], Ir[b][ie]), td = max(Mr[b][mo], Dr[b][ie]);
                    int iv = __builtin_amdgcn_update_dpp(ti_in, ti, 0x138, 0xf, 0xf, false) + 1;
                    int dv = __builtin_amdgcn_update_dpp(td_in, td, 0x130, 0xf, 0xf, false);
                    int mv = max(max(Mr[b][mx] + 1, iv), dv);
                    const int k = kb0 + lane;
                    if (kb0 < lo_s || kb0 + 63 > hi_s) {  // a block across the range's ends
                        if (k < lo_s || k > hi_s) iv = dv = mv = kNeg;
                    }
                    // EXTEND: the first 16 bases without a branch (one
                    // window word per side), the rest in 32-base windows for
                    // the few lanes that matched all of them
                    const int row = mv, col = row - k;
                    const uint32_t a = window16_lsb(qs, med3(row, zero_v, qlen_v));
                    const uint32_t c = window16_lsb(rs, med3(col, zero_v, rlen_v));
                    const int lim = min(qlen - row, rlen - col);
                    const int n = med3(match16(a, c), lim, zero_v);
                    int mr = row >= 0 ? row + n : kNeg;
                    if (row >= 0 && n == 16 && lim > 16)
                        mr = extend_lsb(qs, qlen, rs, rlen, row + 16, col + 16);
                    Ir[b][is] = iv;
                    Dr[b][is] = dv;
                    Mr[b][ms] = mr;
                    if (b == tbb) found = __ballot(lane == tl && mr >= qlen) != 0;
                    // this block's edge values for the next step
                    const int ei = __builtin_amdgcn_readlane(max(Mr[b][mo2], Ir[b][ie2]), 63);
                    const int ed = max(Mr[b][mo2], Dr[b][ie2]);
                    if (lane == 0) E[nxt + gb + 1] = make_int2(ei, ed);
                }
            }
            if (NW == 1) {
                // one wavefront: LDS accesses of a wavefront are ordered, no
                // barrier; the pair ends as soon as its target lane reaches
                // the query end
                if (found) return s;
            } else {
                // one barrier per step; the flag of step si is acted on after
                // the next step's barrier (one extra step, harmless)
                if (found && lane == 0) done[si & 1] = s;
                __syncthreads();
                if (seen >= 0) return seen;
                seen = done[si & 1];
            }
        }
    }
}

}  // namespace wfareg
}  // namespace cgm
