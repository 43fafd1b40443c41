/*
 * cgm_oracle.c -- CPU restatement of the reference alignment algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  Nothing in the shipped library links or calls it.
 *
 * What it restates (reference = Computational-Genomics-Mojo, Mojo sources):
 *   or_get_base          genomics/core/sequence.mojo:96-101, genomics/core/dna.mojo:97-100
 *   or_pack_ascii        genomics/core/sequence.mojo:38-66, genomics/core/dna.mojo:26-40
 *   or_sw_banded         genomics/cpu/align_cpu.mojo:229-309  (smith_waterman_banded)
 *   or_wfa_affine        genomics/cpu/align_cpu.mojo:75-226   (wfa_affine_cpu)
 *   or_wfa_affine_cigar  genomics/cpu/align_cpu.mojo:380-623  (wfa_affine_cigar_cpu)
 *   or_nw_banded         genomics/cpu/align_cpu.mojo:312-377  (needleman_wunsch_banded)
 *   or_sw_banded_start   or_sw_banded plus start coordinates, which the
 *                        reference hard-wires to 0 (align_cpu.mojo:302-305);
 *                        the definition is the library's own (see below)
 * Independent cross-checks (not restatements; used to pin the above):
 *   or_sw_antidiag       the same SW recurrence swept by anti-diagonals with
 *                        an explicit (H desc, qi asc, ri asc) reduction
 *   or_gotoh_global      textbook Gotoh global alignment (mismatch x, match 0,
 *                        gap of length l costs o + l*e), full matrix
 *
 * Pinning status (what checks each function; the tests live in tests/):
 *   or_pack_ascii / or_get_base: reference known answers of test_dna.mojo and
 *     test_sequence.mojo (MSB-first, N -> A + NMask bit, 33 bases -> 2 words).
 *   or_wfa_affine: reference known answers test_alignment.mojo:87-138, closed
 *     forms, equality with or_gotoh_global (itself pinned by exhaustive
 *     alignment enumeration on tiny inputs) on random and exhaustive inputs.
 *   or_wfa_affine_cigar: reference known answers test_alignment.mojo:201-279
 *     (exact CIGAR strings), its cost equal to or_wfa_affine, and per CIGAR:
 *     cost recomputed from the operations, query / reference bases consumed,
 *     '=' on equal and 'X' on different bases.  Which of several co-optimal
 *     CIGARs it returns (mismatch before insertion before deletion in M,
 *     gap opening before extension in I and D) follows align_cpu.mojo:
 *     520-600 line by line and is pinned by the known answers only.
 *   or_nw_banded: reference known answers test_alignment.mojo:61-80, closed
 *     forms (identical pair, one substitution, band too narrow for the
 *     length difference, empty inputs) and a brute force over every banded
 *     path on tiny inputs.  The reference allocates its two rows with
 *     List(capacity=width) and indexes them without appending
 *     (align_cpu.mojo:324-337); this restatement takes the rows to hold
 *     width entries, which is what the indexing intends.
 *   or_sw_banded: reference known answers test_alignment.mojo:24-58, closed
 *     forms (identical pair, no common base, empty input, band 0, one
 *     substitution, a deletion or insertion of l bases planted in an
 *     identical pair), invariants (band monotonicity, single-diagonal Kadane
 *     lower bound), equality with or_sw_antidiag (visiting order and
 *     tie-break), and equality with an explicit enumeration of every local
 *     alignment of tiny pairs under the path model of align_cpu.mojo:277-289
 *     (E opens with a step down and extends to the right, F is one step
 *     right that never extends).  Copies of the loop with a planted mistake
 *     (textbook Gotoh, an extending F, a dropped term, a transposed index, a
 *     tie-break by >=) fail those checks (tests/test_oracle_sw_gaps.py).
 *   or_sw_banded_start: the enumeration above, reporting for the end cell
 *     the latest first diagonal step among its best alignments; closed forms
 *     (an identical k-mer between flanks that share no base, band 0 =
 *     Kadane's latest start); a planted earliest-start tie-break fails them.
 *
 * Integer widths: the reference computes in Mojo Int (64-bit); so does this.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define OR_NEG_INF (-1000000LL) /* align_cpu.mojo:107 and :283 use -1000000 */

/* ---------------------------------------------------------------- packing */

/* dna.mojo:26-40 -- A/a=0 C/c=1 G/g=2 T/t=3, anything else is N (0xFF). */
static int or_ascii_to_2bit(unsigned char c) {
    if (c == 65 || c == 97) return 0;
    if (c == 67 || c == 99) return 1;
    if (c == 71 || c == 103) return 2;
    if (c == 84 || c == 116) return 3;
    return 0xFF;
}

/* sequence.mojo:38-66 -- base i of word w sits at bits (62 - 2i)..(63 - 2i);
 * an N is stored as code 0 and flagged at bit i of that word's NMask
 * (dna.mojo:116-122).  Returns the number of words written. */
int64_t or_pack_ascii(const char *s, int64_t len, uint64_t *words, uint64_t *nmask) {
    int64_t nw = (len + 31) / 32;
    for (int64_t w = 0; w < nw; ++w) {
        int64_t base = w * 32;
        int64_t n = len - base < 32 ? len - base : 32;
        uint64_t word = 0, mask = 0;
        for (int64_t i = 0; i < n; ++i) {
            int b = or_ascii_to_2bit((unsigned char)s[base + i]);
            if (b == 0xFF) {
                mask |= (uint64_t)1 << i;
                b = 0;
            }
            word |= (uint64_t)b << (62 - 2 * i);
        }
        words[w] = word;
        if (nmask) nmask[w] = mask;
    }
    return nw;
}

/* sequence.mojo:96-101 + dna.mojo:97-100 */
static inline int or_get_base(const uint64_t *w, int64_t pos) {
    return (int)((w[pos / 32] >> (62 - 2 * (pos % 32))) & 3);
}

int or_get_base_public(const uint64_t *w, int64_t pos) { return or_get_base(w, pos); }

/* ------------------------------------------------------- banded SW (exact) */

static inline int64_t or_max(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t or_min(int64_t a, int64_t b) { return a < b ? a : b; }

/* align_cpu.mojo:229-309.  out[0]=score out[1]=query_end out[2]=ref_end
 * (query_start and ref_start are always 0, cigar always empty). */
void or_sw_banded(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                  int64_t match, int64_t mismatch, int64_t gap_open, int64_t gap_extend,
                  int64_t band, int64_t out[3]) {
    int64_t width = 2 * band + 1; /* :239 */
    int64_t *H_prev = (int64_t *)calloc((size_t)(width > 0 ? width : 1), sizeof(int64_t));
    int64_t *H_curr = (int64_t *)calloc((size_t)(width > 0 ? width : 1), sizeof(int64_t));
    int64_t *E_curr = (int64_t *)calloc((size_t)(width > 0 ? width : 1), sizeof(int64_t));
    int64_t best = 0, bq = 0, br = 0; /* :255-257 */
    for (int64_t qi = 0; qi < qlen; ++qi) {
        for (int64_t j = 0; j < width; ++j) { H_curr[j] = 0; E_curr[j] = 0; } /* :261-263 */
        for (int64_t j = 0; j < width; ++j) { /* :265 */
            int64_t ri = qi - band + j;
            if (ri < 0 || ri >= rlen) continue; /* :267-268 */
            int same = or_get_base(q, qi) == or_get_base(r, ri); /* :271 */
            int64_t sub = same ? match : mismatch;
            int64_t h_diag = H_prev[j]; /* :277 (j always in range) */
            int64_t h_up = (j + 1 < width) ? H_prev[j + 1] : 0; /* :278 */
            int64_t h_left = (j > 0) ? H_curr[j - 1] : 0; /* :279 */
            int64_t e_val = or_max((j > 0) ? E_curr[j - 1] + gap_extend : OR_NEG_INF,
                                   h_up + gap_open + gap_extend); /* :282-283 */
            int64_t f_val = h_left + gap_open + gap_extend; /* :287 */
            int64_t cell = or_max(0, or_max(h_diag + sub, or_max(e_val, f_val))); /* :289 */
            H_curr[j] = cell;
            E_curr[j] = e_val;
            if (cell > best) { best = cell; bq = qi; br = ri; } /* :293-296 */
        }
        for (int64_t j = 0; j < width; ++j) H_prev[j] = H_curr[j]; /* :299-300 */
    }
    free(H_prev); free(H_curr); free(E_curr);
    out[0] = best; out[1] = bq; out[2] = br;
}

/* SW start coordinates.  The reference reports query_start = ref_start = 0
 * (align_cpu.mojo:302-305) and has no definition of its own, so this one is
 * the library's: among the alignments of maximal score that end in H at the
 * cell or_sw_banded reports, the start is the first diagonal step (the
 * first query and reference base the alignment pairs) of the one whose
 * start is latest -- largest query start, then largest reference start.
 * Computed by the row-major loop of or_sw_banded carrying, next to each H
 * and E value, the start of the path that value belongs to, with ties
 * between equal values resolved towards the later start; a path with no
 * diagonal step yet (a fresh start, or gaps only) carries no start and
 * ranks above every real start of the same value.  Requires the penalties
 * of every configuration (mismatch, gap_extend, gap_open + gap_extend <= 0),
 * under which leading gaps never raise a score.  out[0..4] = score,
 * query_start, ref_start, query_end, ref_end ((0,0,0,0,0) when no cell is
 * positive). */
typedef struct { int64_t v, qs, rs; } or_cand;
#define OR_NO_START INT64_MAX

static inline int or_cand_gt(or_cand a, or_cand b) {
    if (a.v != b.v) return a.v > b.v;
    if (a.qs != b.qs) return a.qs > b.qs;
    return a.rs > b.rs;
}
static inline or_cand or_cand_max(or_cand a, or_cand b) { return or_cand_gt(b, a) ? b : a; }
static inline or_cand or_cand_add(or_cand a, int64_t d) { a.v += d; return a; }

void or_sw_banded_start(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                        int64_t match, int64_t mismatch, int64_t gap_open, int64_t gap_extend,
                        int64_t band, int64_t out[5]) {
    int64_t width = 2 * band + 1;
    size_t n = (size_t)(width > 0 ? width : 1);
    or_cand *Hp = (or_cand *)malloc(n * sizeof(or_cand));
    or_cand *Hc = (or_cand *)malloc(n * sizeof(or_cand));
    or_cand *Ec = (or_cand *)malloc(n * sizeof(or_cand));
    const or_cand zero = {0, OR_NO_START, OR_NO_START};
    for (int64_t j = 0; j < width; ++j) Hp[j] = zero;
    or_cand best = zero;
    int64_t bq = 0, br = 0;
    const int64_t oe = gap_open + gap_extend;
    for (int64_t qi = 0; qi < qlen; ++qi) {
        for (int64_t j = 0; j < width; ++j) { Hc[j] = zero; Ec[j] = zero; }
        for (int64_t j = 0; j < width; ++j) {
            int64_t ri = qi - band + j;
            if (ri < 0 || ri >= rlen) continue;
            int64_t sub = or_get_base(q, qi) == or_get_base(r, ri) ? match : mismatch;
            or_cand d = or_cand_add(Hp[j], sub);
            if (Hp[j].qs == OR_NO_START) { d.qs = qi; d.rs = ri; } /* first diagonal step */
            or_cand up = (j + 1 < width) ? Hp[j + 1] : zero;
            or_cand left = (j > 0) ? Hc[j - 1] : zero;
            or_cand el = (j > 0) ? or_cand_add(Ec[j - 1], gap_extend)
                                 : (or_cand){OR_NEG_INF, OR_NO_START, OR_NO_START};
            or_cand e = or_cand_max(el, or_cand_add(up, oe));
            or_cand f = or_cand_add(left, oe);
            or_cand h = or_cand_max(or_cand_max(zero, d), or_cand_max(e, f));
            if (h.v == 0) h = zero; /* a zero-valued path restarts here */
            Hc[j] = h;
            Ec[j] = e;
            if (h.v > best.v) { best = h; bq = qi; br = ri; } /* or_sw_banded's end cell */
        }
        or_cand *t = Hp; Hp = Hc; Hc = t;
    }
    free(Hp); free(Hc); free(Ec);
    out[0] = best.v;
    out[1] = best.v > 0 ? best.qs : 0;
    out[2] = best.v > 0 ? best.rs : 0;
    out[3] = bq;
    out[4] = br;
}

/* The same recurrence evaluated anti-diagonal by anti-diagonal over a full
 * (qlen x rlen) matrix restricted to |ri - qi| <= band, with the maximum
 * chosen by key (H desc, qi asc, ri asc).  Used only to check that the
 * row-major restatement's tie-break and boundary handling do not depend on
 * the visiting order. */
void or_sw_antidiag(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                    int64_t match, int64_t mismatch, int64_t gap_open, int64_t gap_extend,
                    int64_t band, int64_t out[3]) {
    int64_t best = 0, bq = 0, br = 0;
    if (qlen <= 0 || rlen <= 0) { out[0] = 0; out[1] = 0; out[2] = 0; return; }
    int64_t *H = (int64_t *)calloc((size_t)(qlen * rlen), sizeof(int64_t));
    int64_t *E = (int64_t *)calloc((size_t)(qlen * rlen), sizeof(int64_t));
#define IN_BAND(a, b) ((b) - (a) <= band && (a) - (b) <= band)
    for (int64_t t = 0; t <= qlen + rlen - 2; ++t) {
        for (int64_t qi = or_max(0, t - (rlen - 1)); qi <= or_min(qlen - 1, t); ++qi) {
            int64_t ri = t - qi;
            if (!IN_BAND(qi, ri)) continue;
            int64_t sub = or_get_base(q, qi) == or_get_base(r, ri) ? match : mismatch;
            int64_t hd = (qi > 0 && ri > 0 && IN_BAND(qi - 1, ri - 1)) ? H[(qi - 1) * rlen + ri - 1] : 0;
            int64_t hu = (qi > 0 && IN_BAND(qi - 1, ri)) ? H[(qi - 1) * rlen + ri] : 0;
            int64_t hl = (ri > 0 && IN_BAND(qi, ri - 1)) ? H[qi * rlen + ri - 1] : 0;
            int64_t el;
            if (ri - qi == -band) el = OR_NEG_INF;          /* leftmost band slot */
            else if (ri == 0) el = 0;                        /* skipped cell left of ri=0 */
            else el = E[qi * rlen + ri - 1];
            int64_t e = or_max(el + gap_extend, hu + gap_open + gap_extend);
            int64_t f = hl + gap_open + gap_extend;
            int64_t h = or_max(0, or_max(hd + sub, or_max(e, f)));
            H[qi * rlen + ri] = h;
            E[qi * rlen + ri] = e;
            if (h > best || (h == best && h > 0 && (qi < bq || (qi == bq && ri < br)))) {
                best = h; bq = qi; br = ri;
            }
        }
    }
#undef IN_BAND
    free(H); free(E);
    out[0] = best; out[1] = bq; out[2] = br;
}

/* ------------------------------------------------------ gap-affine WFA (exact) */

/* align_cpu.mojo:75-226.  Returns the cost, or -1 when it exceeds max_err. */
int64_t or_wfa_affine(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                      int64_t x, int64_t o, int64_t e, int64_t max_err) {
    if (qlen == 0 && rlen == 0) return 0; /* :100-101 */
    int64_t n_diags = 2 * max_err + 1; /* :110 */
    int64_t diag_offset = max_err;
    int64_t m_hist = or_max(x, o + e) + 1; /* :114 */
    int64_t ie_hist = e + 1; /* :115 */
    int64_t *M = (int64_t *)malloc((size_t)(m_hist * n_diags) * sizeof(int64_t));
    int64_t *I = (int64_t *)malloc((size_t)(ie_hist * n_diags) * sizeof(int64_t));
    int64_t *D = (int64_t *)malloc((size_t)(ie_hist * n_diags) * sizeof(int64_t));
    for (int64_t i = 0; i < m_hist * n_diags; ++i) M[i] = OR_NEG_INF;
    for (int64_t i = 0; i < ie_hist * n_diags; ++i) { I[i] = OR_NEG_INF; D[i] = OR_NEG_INF; }
    int64_t result = -1;

    M[diag_offset] = 0; /* :130 */
    { /* :133-142 extend score 0 along diagonal 0 */
        int64_t row0 = M[diag_offset], col0 = row0;
        while (row0 < qlen && col0 >= 0 && col0 < rlen) {
            if (or_get_base(q, row0) != or_get_base(r, col0)) break;
            ++row0; ++col0;
        }
        M[diag_offset] = row0;
    }
    int64_t target_d = qlen - rlen; /* :145 */
    if (target_d >= -max_err && target_d <= max_err) {
        if (M[diag_offset + target_d] >= qlen) { result = 0; goto done; }
    }
    for (int64_t s = 1; s <= max_err; ++s) { /* :148 */
        int64_t m_slot = s % m_hist, ie_slot = s % ie_hist;
        for (int64_t j = 0; j < n_diags; ++j) { /* :153-156 */
            M[m_slot * n_diags + j] = OR_NEG_INF;
            I[ie_slot * n_diags + j] = OR_NEG_INF;
            D[ie_slot * n_diags + j] = OR_NEG_INF;
        }
        int64_t lo = or_max(-rlen, -s), hi = or_min(qlen, s); /* :158-164 */
        if (target_d < lo || target_d > hi) { lo = or_min(lo, target_d); hi = or_max(hi, target_d); }
        lo = or_max(lo, -max_err);
        hi = or_min(hi, max_err);
        for (int64_t d = lo; d <= hi; ++d) { /* :166-205 */
            int64_t di = d + diag_offset;
            int64_t i_val = OR_NEG_INF;
            if (di > 0) {
                if (s >= o + e) {
                    int64_t mv = M[((s - o - e) % m_hist) * n_diags + di - 1];
                    if (mv != OR_NEG_INF) i_val = or_max(i_val, mv + 1);
                }
                if (s >= e) {
                    int64_t iv = I[((s - e) % ie_hist) * n_diags + di - 1];
                    if (iv != OR_NEG_INF) i_val = or_max(i_val, iv + 1);
                }
            }
            I[ie_slot * n_diags + di] = i_val;
            int64_t d_val = OR_NEG_INF;
            if (di < n_diags - 1) {
                if (s >= o + e) {
                    int64_t mv = M[((s - o - e) % m_hist) * n_diags + di + 1];
                    if (mv != OR_NEG_INF) d_val = or_max(d_val, mv);
                }
                if (s >= e) {
                    int64_t dv = D[((s - e) % ie_hist) * n_diags + di + 1];
                    if (dv != OR_NEG_INF) d_val = or_max(d_val, dv);
                }
            }
            D[ie_slot * n_diags + di] = d_val;
            int64_t m_val = OR_NEG_INF;
            if (s >= x) {
                int64_t mv = M[((s - x) % m_hist) * n_diags + di];
                if (mv != OR_NEG_INF) m_val = or_max(m_val, mv + 1);
            }
            if (i_val != OR_NEG_INF) m_val = or_max(m_val, i_val);
            if (d_val != OR_NEG_INF) m_val = or_max(m_val, d_val);
            M[m_slot * n_diags + di] = m_val;
        }
        for (int64_t d = lo; d <= hi; ++d) { /* :208-219 */
            int64_t di = d + diag_offset;
            int64_t row = M[m_slot * n_diags + di];
            if (row == OR_NEG_INF || row < 0) continue;
            int64_t col = row - d;
            while (row < qlen && col >= 0 && col < rlen) {
                if (or_get_base(q, row) != or_get_base(r, col)) break;
                ++row; ++col;
            }
            M[m_slot * n_diags + di] = row;
        }
        if (target_d >= lo && target_d <= hi) { /* :222-224 */
            if (M[m_slot * n_diags + diag_offset + target_d] >= qlen) { result = s; goto done; }
        }
    }
done:
    free(M); free(I); free(D);
    return result; /* :226 returns -1 when max_err is exceeded */
}

/* ---------------------------------------------- gap-affine WFA with CIGAR */

/* align_cpu.mojo:380-623.  Full wavefront history (scores 0..max_err, every
 * diagonal in +-max_err), then the reference's backtrace.  The CIGAR is
 * written as run-length operations run = (count << 2) | op with op 0 '=',
 * 1 'X', 2 'I' (query base), 3 'D' (reference base), first operation first;
 * *n_runs receives the number of runs, or -2 if more than cap were needed.
 * Returns the cost, or -1 (no CIGAR, *n_runs = 0) above max_err. */
int64_t or_wfa_affine_cigar(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                            int64_t x, int64_t o, int64_t e, int64_t max_err, uint32_t *runs,
                            int64_t cap, int64_t *n_runs) {
    *n_runs = 0;
    if (qlen == 0 && rlen == 0) return 0; /* :396-397, empty CIGAR */
    const int64_t n_diags = 2 * max_err + 1, offset = max_err; /* :404-405 */
    const int64_t total = (max_err + 1) * n_diags;           /* :410 */
    int32_t *M_pre = (int32_t *)malloc((size_t)total * 4);
    int32_t *M_full = (int32_t *)malloc((size_t)total * 4);
    int32_t *I_arr = (int32_t *)malloc((size_t)total * 4);
    int32_t *D_arr = (int32_t *)malloc((size_t)total * 4);
    const int32_t NEG = (int32_t)OR_NEG_INF;
    for (int64_t i = 0; i < total; ++i) M_pre[i] = M_full[i] = I_arr[i] = D_arr[i] = NEG;
    M_pre[offset] = 0; /* :423-428 */
    int64_t r0 = 0;
    while (r0 < qlen && r0 < rlen && or_get_base(q, r0) == or_get_base(r, r0)) ++r0;
    M_full[offset] = (int32_t)r0;
    const int64_t target_d = qlen - rlen; /* :430 */
    int64_t opt_score = -1;
    if (target_d >= -max_err && target_d <= max_err)
        if (M_full[offset + target_d] >= qlen) opt_score = 0;
    if (opt_score == -1) { /* :437-505 */
        for (int64_t s = 1; s <= max_err; ++s) {
            int64_t lo = or_max(-rlen, -s), hi = or_min(qlen, s);
            if (target_d < lo) lo = or_min(lo, target_d);
            if (target_d > hi) hi = or_max(hi, target_d);
            lo = or_max(lo, -max_err);
            hi = or_min(hi, max_err);
            for (int64_t d = lo; d <= hi; ++d) {
                const int64_t di = d + offset;
                int64_t i_val = OR_NEG_INF;
                if (di > 0) {
                    if (s >= o + e) {
                        int64_t mv = M_full[(s - o - e) * n_diags + di - 1];
                        if (mv != OR_NEG_INF) i_val = or_max(i_val, mv + 1);
                    }
                    if (s >= e) {
                        int64_t iv = I_arr[(s - e) * n_diags + di - 1];
                        if (iv != OR_NEG_INF) i_val = or_max(i_val, iv + 1);
                    }
                }
                I_arr[s * n_diags + di] = (int32_t)i_val;
                int64_t d_val = OR_NEG_INF;
                if (di < n_diags - 1) {
                    if (s >= o + e) {
                        int64_t mv = M_full[(s - o - e) * n_diags + di + 1];
                        if (mv != OR_NEG_INF) d_val = or_max(d_val, mv);
                    }
                    if (s >= e) {
                        int64_t dv = D_arr[(s - e) * n_diags + di + 1];
                        if (dv != OR_NEG_INF) d_val = or_max(d_val, dv);
                    }
                }
                D_arr[s * n_diags + di] = (int32_t)d_val;
                int64_t m_val = OR_NEG_INF;
                if (s >= x) {
                    int64_t mv = M_full[(s - x) * n_diags + di];
                    if (mv != OR_NEG_INF) m_val = or_max(m_val, mv + 1);
                }
                if (i_val != OR_NEG_INF) m_val = or_max(m_val, i_val);
                if (d_val != OR_NEG_INF) m_val = or_max(m_val, d_val);
                M_pre[s * n_diags + di] = (int32_t)m_val;
                if (m_val != OR_NEG_INF) { /* :491-499 extend */
                    int64_t row = m_val, col = row - d;
                    while (row < qlen && col >= 0 && col < rlen) {
                        if (or_get_base(q, row) != or_get_base(r, col)) break;
                        ++row; ++col;
                    }
                    M_full[s * n_diags + di] = (int32_t)row;
                }
            }
            if (target_d >= lo && target_d <= hi) /* :502-505 */
                if (M_full[s * n_diags + offset + target_d] >= qlen) { opt_score = s; break; }
        }
    }
    if (opt_score == -1) goto out; /* :507-508 */
    { /* :510-600 backtrack, operations collected last first */
        int64_t ops_cap = qlen + rlen + 2 * (opt_score + 1) + 16, n_ops = 0;
        uint8_t *ops = (uint8_t *)malloc((size_t)ops_cap);
        int64_t cur_s = opt_score, cur_d = target_d, cur_row = M_full[opt_score * n_diags + target_d + offset];
        int cur_matrix = 0;
#define OR_PUSH(op) do { if (n_ops < ops_cap) ops[n_ops] = (uint8_t)(op); ++n_ops; } while (0)
        for (;;) {
            const int64_t di = cur_d + offset;
            if (cur_matrix == 0) { /* M */
                const int64_t pre_r = M_pre[cur_s * n_diags + di];
                for (int64_t k = 0; k < cur_row - pre_r; ++k) OR_PUSH(0);
                cur_row = pre_r;
                if (cur_s == 0 && cur_d == 0) break;
                if (cur_s >= x) {
                    const int64_t prev_m = M_full[(cur_s - x) * n_diags + di];
                    if (prev_m != OR_NEG_INF && prev_m + 1 == cur_row) {
                        OR_PUSH(1);
                        cur_s -= x;
                        cur_row = prev_m;
                        continue;
                    }
                }
                const int64_t i_here = I_arr[cur_s * n_diags + di];
                if (i_here != OR_NEG_INF && i_here == cur_row) { cur_matrix = 1; continue; }
                const int64_t d_here = D_arr[cur_s * n_diags + di];
                if (d_here != OR_NEG_INF && d_here == cur_row) { cur_matrix = 2; continue; }
                break;
            } else if (cur_matrix == 1) { /* I */
                OR_PUSH(2);
                const int64_t prev_d = cur_d - 1, prev_di = prev_d + offset;
                if (cur_s >= o + e && prev_di >= 0) {
                    const int64_t mv = M_full[(cur_s - o - e) * n_diags + prev_di];
                    if (mv != OR_NEG_INF && mv + 1 == cur_row) {
                        cur_s -= o + e; cur_d = prev_d; cur_row = mv; cur_matrix = 0;
                        continue;
                    }
                }
                if (cur_s >= e && prev_di >= 0) {
                    const int64_t iv = I_arr[(cur_s - e) * n_diags + prev_di];
                    if (iv != OR_NEG_INF && iv + 1 == cur_row) {
                        cur_s -= e; cur_d = prev_d; cur_row = iv; cur_matrix = 1;
                        continue;
                    }
                }
                break;
            } else { /* D */
                OR_PUSH(3);
                const int64_t prev_d = cur_d + 1, prev_di = prev_d + offset;
                if (cur_s >= o + e && prev_di < n_diags) {
                    const int64_t mv = M_full[(cur_s - o - e) * n_diags + prev_di];
                    if (mv != OR_NEG_INF && mv == cur_row) {
                        cur_s -= o + e; cur_d = prev_d; cur_row = mv; cur_matrix = 0;
                        continue;
                    }
                }
                if (cur_s >= e && prev_di < n_diags) {
                    const int64_t dv = D_arr[(cur_s - e) * n_diags + prev_di];
                    if (dv != OR_NEG_INF && dv == cur_row) {
                        cur_s -= e; cur_d = prev_d; cur_row = dv; cur_matrix = 2;
                        continue;
                    }
                }
                break;
            }
        }
#undef OR_PUSH
        /* :602-621 run-length encode, first operation first */
        int64_t nr = 0;
        for (int64_t i = n_ops - 1; i >= 0;) {
            const uint8_t op = ops[i];
            int64_t cnt = 1;
            while (i - cnt >= 0 && ops[i - cnt] == op) ++cnt;
            if (nr < cap) runs[nr] = (uint32_t)((cnt << 2) | op);
            ++nr;
            i -= cnt;
        }
        *n_runs = nr <= cap && n_ops <= ops_cap ? nr : -2;
        free(ops);
    }
out:
    free(M_pre); free(M_full); free(I_arr); free(D_arr);
    return opt_score;
}

typedef struct {
    const uint64_t *qp, *rp;
    const int32_t *qo, *ro, *ql, *rl, *me;
    int32_t me_scalar;
    int64_t x, o, e, n;
    int32_t *score;
    uint32_t *runs;
    const int64_t *run_off;
    int32_t *n_runs;
    int64_t next; pthread_mutex_t *lock;
} or_cigar_job;

static void *or_cigar_worker(void *arg) {
    or_cigar_job *j = (or_cigar_job *)arg;
    for (;;) {
        pthread_mutex_lock(j->lock);
        int64_t i = j->next++;
        pthread_mutex_unlock(j->lock);
        if (i >= j->n) break;
        int64_t nr = 0;
        int64_t me = j->me ? j->me[i] : j->me_scalar;
        j->score[i] = (int32_t)or_wfa_affine_cigar(j->qp + j->qo[i], j->ql[i], j->rp + j->ro[i],
                                                   j->rl[i], j->x, j->o, j->e, me,
                                                   j->runs + j->run_off[i],
                                                   j->run_off[i + 1] - j->run_off[i], &nr);
        j->n_runs[i] = (int32_t)nr;
    }
    return NULL;
}

/* or_wfa_affine_cigar over a batch (layout as or_batch); pair i's runs go to
 * runs[run_off[i] .. run_off[i + 1]). */
int or_wfa_cigar_batch(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_off,
                       const int32_t *r_off, const int32_t *q_len, const int32_t *r_len,
                       int64_t n_pairs, const int32_t *max_err, int32_t max_err_scalar,
                       int32_t x, int32_t o, int32_t e, int32_t *score, uint32_t *runs,
                       const int64_t *run_off, int32_t *n_runs, int n_threads) {
    pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    or_cigar_job j;
    memset(&j, 0, sizeof(j));
    j.qp = q_packed; j.rp = r_packed; j.qo = q_off; j.ro = r_off; j.ql = q_len; j.rl = r_len;
    j.me = max_err; j.me_scalar = max_err_scalar; j.x = x; j.o = o; j.e = e; j.n = n_pairs;
    j.score = score; j.runs = runs; j.run_off = run_off; j.n_runs = n_runs;
    j.next = 0; j.lock = &lock;
    if (n_threads < 1) n_threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
    for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, or_cigar_worker, &j);
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    free(th);
    return 0;
}

/* ------------------------------------------------------- banded NW (exact) */

/* align_cpu.mojo:312-377.  Global score within the band |ri - qi| <= band;
 * out[0] = score, out[1] = query_end = qlen - 1, out[2] = ref_end = rlen - 1. */
void or_nw_banded(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                  int64_t match, int64_t mismatch, int64_t gap_open, int64_t gap_extend,
                  int64_t band, int64_t *out) {
    const int64_t width = 2 * band + 1;
    int64_t *H_prev = (int64_t *)malloc((size_t)width * sizeof(int64_t));
    int64_t *H_curr = (int64_t *)malloc((size_t)width * sizeof(int64_t));
    for (int64_t j = 0; j < width; ++j) { /* :327-334 */
        const int64_t ri = -band + j;
        if (ri < 0) H_prev[j] = OR_NEG_INF;
        else if (ri == 0) H_prev[j] = 0;
        else H_prev[j] = gap_open + ri * gap_extend;
    }
    for (int64_t qi = 0; qi < qlen; ++qi) { /* :336-361 */
        for (int64_t j = 0; j < width; ++j) H_curr[j] = OR_NEG_INF;
        for (int64_t ro = 0; ro < width; ++ro) {
            const int64_t ri = qi - band + ro;
            if (ri < 0 || ri >= rlen) continue;
            const int64_t sub = or_get_base(q, qi) == or_get_base(r, ri) ? match : mismatch;
            const int64_t h_diag = H_prev[ro];
            const int64_t h_up = ro + 1 < width ? H_prev[ro + 1] : OR_NEG_INF;
            const int64_t h_left = ro > 0 ? H_curr[ro - 1] : OR_NEG_INF;
            H_curr[ro] = or_max(h_diag + sub, or_max(h_up + gap_open + gap_extend,
                                                     h_left + gap_open + gap_extend));
        }
        for (int64_t j = 0; j < width; ++j) H_prev[j] = H_curr[j];
    }
    const int64_t final_offset = (rlen - 1) - (qlen - 1) + band; /* :363-367 */
    int64_t final_score = OR_NEG_INF;
    if (final_offset >= 0 && final_offset < width) final_score = H_prev[final_offset];
    out[0] = final_score;
    out[1] = qlen - 1;
    out[2] = rlen - 1;
    free(H_prev); free(H_curr);
}

/* ---------------------------------------------------- textbook Gotoh (check) */

/* Global alignment, minimum cost: match 0, mismatch x, a gap of length l
 * costs o + l*e (in either sequence).  Full O(qlen*rlen) matrices.
 * Returns the cost if it is <= max_err, else -1 (max_err < 0: no cap). */
int64_t or_gotoh_global(const uint64_t *q, int64_t qlen, const uint64_t *r, int64_t rlen,
                        int64_t x, int64_t o, int64_t e, int64_t max_err) {
    const int64_t INF = (int64_t)1 << 50;
    int64_t W = rlen + 1;
    int64_t *Mm = (int64_t *)malloc((size_t)((qlen + 1) * W) * sizeof(int64_t));
    int64_t *Ig = (int64_t *)malloc((size_t)((qlen + 1) * W) * sizeof(int64_t)); /* ends in a query-base gap */
    int64_t *Dg = (int64_t *)malloc((size_t)((qlen + 1) * W) * sizeof(int64_t)); /* ends in a ref-base gap */
    for (int64_t i = 0; i <= qlen; ++i) {
        for (int64_t j = 0; j <= rlen; ++j) {
            int64_t m = INF, ig = INF, dg = INF;
            if (i == 0 && j == 0) m = 0;
            if (i > 0 && j > 0) {
                int64_t p = Mm[(i - 1) * W + j - 1];
                int64_t pi = Ig[(i - 1) * W + j - 1], pd = Dg[(i - 1) * W + j - 1];
                int64_t best = or_min(p, or_min(pi, pd));
                int64_t c = (or_get_base(q, i - 1) == or_get_base(r, j - 1)) ? 0 : x;
                m = best + c;
            }
            if (i > 0) {
                int64_t from_open = or_min(Mm[(i - 1) * W + j], Dg[(i - 1) * W + j]) + o + e;
                int64_t from_ext = Ig[(i - 1) * W + j] + e;
                ig = or_min(from_open, from_ext);
            }
            if (j > 0) {
                int64_t from_open = or_min(Mm[i * W + j - 1], Ig[i * W + j - 1]) + o + e;
                int64_t from_ext = Dg[i * W + j - 1] + e;
                dg = or_min(from_open, from_ext);
            }
            Mm[i * W + j] = m; Ig[i * W + j] = ig; Dg[i * W + j] = dg;
        }
    }
    int64_t c = or_min(Mm[qlen * W + rlen], or_min(Ig[qlen * W + rlen], Dg[qlen * W + rlen]));
    free(Mm); free(Ig); free(Dg);
    if (max_err >= 0 && c > max_err) return -1;
    return c;
}

/* ------------------------------------------------------------ batch drivers */

/* Batch layout (same as the GPU entry points): pair i aligns query i
 * (words q_packed[q_off[i] ...], q_len[i] bases) with reference i. */
typedef struct {
    const uint64_t *qp, *rp;
    const int32_t *qo, *ro, *ql, *rl;
    const int32_t *band; int32_t band_scalar;
    const int32_t *max_err; int32_t max_err_scalar;
    const int32_t *kind;
    int64_t p0, p1, n;
    int64_t match, mismatch, gap_open, gap_extend;
    int64_t x, o, e;
    int32_t *score, *q_end, *r_end;
    int32_t *q_start, *r_start; /* mode 3 only */
    int64_t next_pair; pthread_mutex_t *lock; int mode;
} or_job;

static void or_do_pair(const or_job *j, int64_t i) {
    const uint64_t *q = j->qp + j->qo[i];
    const uint64_t *r = j->rp + j->ro[i];
    if (j->mode == 2) {
        int64_t band = j->band ? j->band[i] : j->band_scalar;
        int64_t out[3];
        or_nw_banded(q, j->ql[i], r, j->rl[i], j->match, j->mismatch, j->gap_open, j->gap_extend,
                     band, out);
        j->score[i] = (int32_t)out[0];
        if (j->q_end) j->q_end[i] = (int32_t)out[1];
        if (j->r_end) j->r_end[i] = (int32_t)out[2];
        return;
    }
    if (j->mode == 3) {
        int64_t band = j->band ? j->band[i] : j->band_scalar;
        int64_t out[5];
        or_sw_banded_start(q, j->ql[i], r, j->rl[i], j->match, j->mismatch, j->gap_open,
                           j->gap_extend, band, out);
        j->score[i] = (int32_t)out[0];
        j->q_start[i] = (int32_t)out[1];
        j->r_start[i] = (int32_t)out[2];
        j->q_end[i] = (int32_t)out[3];
        j->r_end[i] = (int32_t)out[4];
        return;
    }
    int is_wfa = j->kind ? j->kind[i] != 0 : j->mode == 1;
    if (is_wfa) {
        int64_t me = j->max_err ? j->max_err[i] : j->max_err_scalar;
        int64_t c = or_wfa_affine(q, j->ql[i], r, j->rl[i], j->x, j->o, j->e, me);
        j->score[i] = (int32_t)c;
        if (j->q_end) j->q_end[i] = -1;
        if (j->r_end) j->r_end[i] = -1;
    } else {
        int64_t band = j->band ? j->band[i] : j->band_scalar;
        int64_t out[3];
        or_sw_banded(q, j->ql[i], r, j->rl[i], j->match, j->mismatch, j->gap_open, j->gap_extend, band, out);
        j->score[i] = (int32_t)out[0];
        if (j->q_end) j->q_end[i] = (int32_t)out[1];
        if (j->r_end) j->r_end[i] = (int32_t)out[2];
    }
}

static void *or_worker(void *arg) {
    or_job *j = (or_job *)arg;
    for (;;) {
        pthread_mutex_lock(j->lock);
        int64_t i = j->next_pair;
        j->next_pair += 1;
        pthread_mutex_unlock(j->lock);
        if (i >= j->n) break;
        or_do_pair(j, i);
    }
    return NULL;
}

static int or_run_job(or_job *j, int64_t n_pairs, int n_threads) {
    if (n_threads <= 1) {
        for (int64_t i = 0; i < n_pairs; ++i) or_do_pair(j, i);
        return 0;
    }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
    for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, or_worker, j);
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    free(th);
    return 0;
}

/* Runs every pair of a batch with n_threads POSIX threads pulling pair
 * indices from a shared counter.  kind[i] (0 = SW, 1 = WFA) may be NULL, in
 * which case mode selects: 0 = all SW, 1 = all WFA; mode 2 = all banded NW
 * (kind ignored); mode 3 = all SW with start coordinates
 * (or_sw_start_batch). */
int or_batch(int mode,
             const uint64_t *q_packed, const uint64_t *r_packed,
             const int32_t *q_off, const int32_t *r_off,
             const int32_t *q_len, const int32_t *r_len, int64_t n_pairs,
             const int32_t *kind,
             const int32_t *band, int32_t band_scalar,
             int32_t match, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
             const int32_t *max_err, int32_t max_err_scalar,
             int32_t x, int32_t o, int32_t e,
             int32_t *score, int32_t *q_end, int32_t *r_end, int n_threads) {
    pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    or_job j;
    memset(&j, 0, sizeof(j));
    j.qp = q_packed; j.rp = r_packed; j.qo = q_off; j.ro = r_off; j.ql = q_len; j.rl = r_len;
    j.band = band; j.band_scalar = band_scalar;
    j.max_err = max_err; j.max_err_scalar = max_err_scalar;
    j.kind = kind;
    j.n = n_pairs;
    j.match = match; j.mismatch = mismatch; j.gap_open = gap_open; j.gap_extend = gap_extend;
    j.x = x; j.o = o; j.e = e;
    j.score = score; j.q_end = q_end; j.r_end = r_end;
    j.next_pair = 0; j.lock = &lock; j.mode = mode;
    return or_run_job(&j, n_pairs, n_threads);
}

/* SW with start coordinates (or_sw_banded_start) over a whole batch. */
int or_sw_start_batch(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_off,
                      const int32_t *r_off, const int32_t *q_len, const int32_t *r_len,
                      int64_t n_pairs, const int32_t *band, int32_t band_scalar, int32_t match,
                      int32_t mismatch, int32_t gap_open, int32_t gap_extend, int32_t *score,
                      int32_t *q_start, int32_t *r_start, int32_t *q_end, int32_t *r_end,
                      int n_threads) {
    pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
    or_job j;
    memset(&j, 0, sizeof(j));
    j.qp = q_packed; j.rp = r_packed; j.qo = q_off; j.ro = r_off; j.ql = q_len; j.rl = r_len;
    j.band = band; j.band_scalar = band_scalar;
    j.n = n_pairs;
    j.match = match; j.mismatch = mismatch; j.gap_open = gap_open; j.gap_extend = gap_extend;
    j.score = score; j.q_end = q_end; j.r_end = r_end; j.q_start = q_start; j.r_start = r_start;
    j.next_pair = 0; j.lock = &lock; j.mode = 3;
    return or_run_job(&j, n_pairs, n_threads);
}
