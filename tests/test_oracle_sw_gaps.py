"""Pins the gap recurrence of the SW oracle (oracle/cgm_oracle.c, or_sw_banded)
-- no GPU needed.

smith_waterman_banded (genomics/cpu/align_cpu.mojo:229-309) is not textbook
Gotoh.  Its gap states, read from align_cpu.mojo:277-289, are:
  * E(qi, ri) = max(E(qi, ri-1) + extend, H(qi-1, ri) + open + extend):
    E opens with a DOWN step (a query base against no reference base) and
    then extends with RIGHT steps (reference bases against no query base);
  * F(qi, ri) = H(qi, ri-1) + open + extend: a single RIGHT step out of H
    that never extends;
  * E at the band's left edge reads -1000000 (:282), E one column left of
    the matrix reads the row's reset value 0 (:261-263), H outside the
    band or the matrix reads 0 (:277-279).

Three kinds of check follow, none of which re-evaluates that recurrence:
  1. an explicit enumeration of every local alignment of tiny pairs under
     the path model above (each alignment scored as a sum of its steps),
     whose per-cell maximum and first row-major argmax must equal the
     oracle;
  2. closed forms for a deletion or an insertion of l bases planted in an
     identical pair, derived from the path model by hand;
  3. planted mistakes: copies of the row-major loop with one plausible slip
     each (textbook Gotoh, an extending F, a dropped term, a transposed
     index, a tie-break by >=) must each fail 1, and every slip in the
     scoring also 2, while the oracle passes both.
"""
import random

import numpy as np
import pytest

from conftest import load_oracle

O = load_oracle()
CONFIG_SW = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)
DEFAULT_SW = dict(match=2, mismatch=-3, gap_open=-5, gap_extend=-2)   # align_cpu.mojo:46-53
SCORINGS = [
    CONFIG_SW,
    DEFAULT_SW,
    dict(match=3, mismatch=-1, gap_open=-1, gap_extend=-1),
    dict(match=1, mismatch=-1, gap_open=0, gap_extend=-1),
    dict(match=5, mismatch=-2, gap_open=-7, gap_extend=0),
    # gaps cheap against matches, so two-base gaps between short flanks win
    dict(match=10, mismatch=-3, gap_open=-3, gap_extend=-1),
    # positive values: the reference accepts any Int, and these make the
    # boundary values (E = 0 left of the matrix, -1000000 at the band edge)
    # decide results, so the enumeration pins them too
    dict(match=2, mismatch=-3, gap_open=-4, gap_extend=1),
    dict(match=1, mismatch=1, gap_open=-1, gap_extend=1),
]


def rseq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(alphabet) for _ in range(n))


# ------------------------------------------------- 1. alignment enumeration

def sw_by_alignment_enumeration(q, r, band, match, mismatch, gap_open, gap_extend):
    """Best local alignment score and its first row-major end cell (see
    enumerate_alignments)."""
    best = enumerate_alignments(q, r, band, match, mismatch, gap_open, gap_extend)
    out = (0, 0, 0)
    for a in range(len(q)):
        for b in range(len(r)):
            h = max(0, best.get((a, b), (0,))[0])
            if h > out[0]:
                out = (h, a, b)
    return out


def enumerate_alignments(q, r, band, match, mismatch, gap_open, gap_extend):
    """For every cell, the best (score, start) over the alignments ending
    there in H, found by listing every alignment of the path model
    explicitly (depth-first, each alignment's score the sum of its steps; no
    table of sub-results).  start = (qi, ri) of the alignment's first
    diagonal step; among alignments of equal score the later start wins,
    and an alignment with no diagonal step ranks above every start.

    Steps, with cells (qi, ri) that must lie in the matrix and in the band
    |ri - qi| <= band:
      diagonal   (a, b) -> (a+1, b+1) in H, + match or mismatch
      right      (a, b) -> (a, b+1)   H -> H, + open + extend
      down       (a, b) -> (a+1, b)   H -> E, + open + extend
      E right    (a, b) -> (a, b+1)   E -> E, + extend
      close      E at a cell -> H at the same cell, + 0
    An alignment starts in H with score 0 just before its first step, at any
    position (the matrix, its row -1 / column -1, or outside the band), or
    in E with score 0 at column -1 of a row qi <= band - 1 (the cell left of
    the matrix that the row's reset leaves at 0 inside the band frame).
    """
    qlen, rlen = len(q), len(r)
    oe = gap_open + gap_extend
    best = {}
    NONE = (10 ** 9, 10 ** 9)   # no diagonal step yet

    def ok(a, b):
        return 0 <= a < qlen and 0 <= b < rlen and abs(b - a) <= band

    def walk_h(a, b, v, st):
        if ok(a, b) and (v, st) > best.get((a, b), (-10 ** 9,)):
            best[(a, b)] = (v, st)
        if ok(a + 1, b + 1):
            walk_h(a + 1, b + 1, v + (match if q[a + 1] == r[b + 1] else mismatch),
                   (a + 1, b + 1) if st == NONE else st)
        if a >= 0 and ok(a, b + 1):
            walk_h(a, b + 1, v + oe, st)
        if ok(a + 1, b):
            walk_e(a + 1, b, v + oe, st)

    def walk_e(a, b, v, st):
        if ok(a, b):
            walk_h(a, b, v, st)
        if ok(a, b + 1):
            walk_e(a, b + 1, v + gap_extend, st)

    for a in range(-1, qlen):
        for b in range(-1, rlen):
            walk_h(a, b, 0, NONE)
    for a in range(0, min(qlen, band)):
        if ok(a, 0):
            walk_e(a, 0, gap_extend, NONE)   # E = 0 at (a, -1), one E step right
    return best


def sw_start_by_alignment_enumeration(q, r, band, **sc):
    """(score, query_start, ref_start, query_end, ref_end): the first
    row-major best cell and the latest start among its best alignments."""
    score, qe, re_ = sw_by_alignment_enumeration(q, r, band, **sc)
    if score == 0:
        return (0, 0, 0, 0, 0)
    st = enumerate_alignments(q, r, band, **sc)[(qe, re_)][1]
    return (score, st[0], st[1], qe, re_)


def _tiny_cases(seed, n):
    """Random tiny pairs, half of them one sequence with 1-3 bases removed
    or added (so that gaps between flanks are common)."""
    rng = random.Random(seed)
    cases = []
    for i in range(n):
        alpha = rng.choice(["AC", "ACG", "ACGT"])
        if i % 2 == 0:
            q = rseq(rng, rng.randint(0, 5), alpha)
            r = rseq(rng, rng.randint(0, 5), alpha)
        else:
            s = rseq(rng, rng.randint(2, 5), alpha)
            p, l = rng.randint(0, len(s)), rng.randint(1, 3)
            t = s[:p] + rseq(rng, l, alpha) + s[p:]
            q, r = (s, t) if rng.random() < 0.5 else (t, s)
        cases.append((q, r, rng.randint(0, 5), rng.choice(SCORINGS)))
    return cases


TINY = _tiny_cases(2024, 800)


def _check_enumeration(sw, cases=TINY):
    """First case where `sw` differs from the enumeration, or None."""
    for q, r, band, sc in cases:
        exp = sw_by_alignment_enumeration(q, r, band, **sc)
        got = tuple(sw(q, r, band=band, **sc))
        if got != exp:
            return (q, r, band, sc, got, exp)
    return None


def test_enumeration_closed_forms():
    # The enumeration itself, on answers that need no program: identical
    # pairs, no common base, and one planted deletion worked out by hand.
    assert sw_by_alignment_enumeration("ACGT", "ACGT", 2, **CONFIG_SW) == (8, 3, 3)
    assert sw_by_alignment_enumeration("AAA", "TTTT", 3, **CONFIG_SW) == (0, 0, 0)
    # query CC|GG, reference CC AA GG.  Bridging the AA takes two right
    # steps out of H at open + extend = -6 each (down once and three right
    # steps in E costs -6 - 6 and loses a match).  Every other diagonal run
    # is at most one flank.  Match 2: 8 - 12 < flank 4, first flank (1, 1);
    # match 5: 20 - 12 < 10; match 7: 28 - 12 = 16 > 14, ending at (3, 5)
    # (textbook Gotoh would give 28 - 4 - 2 - 2 = 20).
    assert sw_by_alignment_enumeration("CCGG", "CCAAGG", 3, **CONFIG_SW) == (4, 1, 1)
    assert sw_by_alignment_enumeration("CCGG", "CCAAGG", 3, **dict(CONFIG_SW, match=5)) == (10, 1, 1)
    assert sw_by_alignment_enumeration("CCGG", "CCAAGG", 3, **dict(CONFIG_SW, match=7)) == (16, 3, 5)


def test_oracle_equals_alignment_enumeration():
    bad = _check_enumeration(O.sw_banded)
    assert bad is None, bad


# ------------------------------------------------------- 2. closed forms

def _planted(L, p, l, kind, seed):
    """An identical pair of length L with l bases deleted from (kind 'del')
    or inserted into (kind 'ins') the query at position p."""
    rng = random.Random(seed)
    ref = rseq(rng, L)
    if kind == "del":
        return ref[:p] + ref[p + l:], ref
    return ref[:p] + rseq(rng, l) + ref[p:], ref


def closed_form_deletion(L, p, l, match, mismatch, gap_open, gap_extend):
    """Query = reference minus l bases at p (query length L - l).  Skipping
    l reference bases takes either l right steps out of H, or one down step
    (the query base after the gap left unaligned) then l + 1 right steps in
    E; or the alignment keeps one flank only."""
    oe = gap_open + gap_extend
    via_f = match * (L - l) + l * oe
    via_e = match * (L - l - 1) + oe + (l + 1) * gap_extend
    return max(via_f, via_e, match * p, match * (L - l - p))


def closed_form_insertion(L, p, l, match, mismatch, gap_open, gap_extend):
    """Query = reference plus l bases at p.  Down steps never extend (E
    extends to the right only), so each inserted base costs open + extend;
    or the alignment keeps one flank only."""
    oe = gap_open + gap_extend
    return max(match * L + l * oe, match * p, match * (L - p))


PLANTED = [(120, 60, 1), (120, 60, 2), (120, 60, 3), (120, 60, 5), (150, 40, 4),
           (100, 70, 6), (90, 45, 8), (200, 100, 2)]


def _check_closed_forms(sw):
    """First planted case where `sw` differs from the closed form, or None."""
    for i, (L, p, l) in enumerate(PLANTED):
        for sc in (CONFIG_SW, DEFAULT_SW, dict(match=3, mismatch=-6, gap_open=-2,
                                                gap_extend=-1)):
            for kind in ("del", "ins"):
                q, r = _planted(L, p, l, kind, seed=1000 * i + len(kind))
                if kind == "del":
                    exp_s = closed_form_deletion(L, p, l, **sc)
                    full = max(sc["match"] * (L - l) + l * (sc["gap_open"] + sc["gap_extend"]),
                               sc["match"] * (L - l - 1) + sc["gap_open"]
                               + (l + 2) * sc["gap_extend"])
                else:
                    exp_s = closed_form_insertion(L, p, l, **sc)
                    full = sc["match"] * L + l * (sc["gap_open"] + sc["gap_extend"])
                got = tuple(sw(q, r, band=16, **sc))
                ok = got[0] == exp_s
                if ok and full == exp_s:
                    # the whole alignment wins: it ends at the last cell
                    ok = got[1:] == (len(q) - 1, len(r) - 1)
                if not ok:
                    return (L, p, l, kind, sc, got, exp_s)
    return None


def test_closed_forms_from_the_reference_recurrence():
    # 2/-4/-4/-2, L = 120, gap at 60: a 2-base deletion scores 224 (two
    # right steps out of H), a 3-base deletion 218 (down once, then four
    # right steps in E), a 2-base insertion 228 (two down steps); textbook
    # Gotoh would give 228, 224 and 232.
    assert closed_form_deletion(120, 60, 2, **CONFIG_SW) == 224
    assert closed_form_deletion(120, 60, 3, **CONFIG_SW) == 218
    assert closed_form_insertion(120, 60, 2, **CONFIG_SW) == 228
    assert O.sw_banded(*_planted(120, 60, 2, "del", 5), band=16, **CONFIG_SW) == (224, 117, 119)
    assert O.sw_banded(*_planted(120, 60, 3, "del", 5), band=16, **CONFIG_SW) == (218, 116, 119)
    assert O.sw_banded(*_planted(120, 60, 2, "ins", 5), band=16, **CONFIG_SW)[0] == 228
    bad = _check_closed_forms(O.sw_banded)
    assert bad is None, bad


# ------------------------------------------------------ 3. planted mistakes

def sw_row_major(q, r, band, match, mismatch, gap_open, gap_extend, mutant=None):
    """The row-major loop of align_cpu.mojo:258-300, with one optional
    planted mistake.  Only the mutants are used; the unmutated form is here
    so that each mutant differs from it in one place."""
    w = 2 * band + 1
    NEG = -1000000
    hp = [0] * w
    ep = [0] * w          # previous row's E (textbook mutant only)
    best = (0, 0, 0)
    for qi in range(len(q)):
        hc, ec, fc = [0] * w, [0] * w, [0] * w
        for j in range(w):
            ri = qi - band + j
            if ri < 0 or ri >= len(r):
                continue
            sub = match if q[qi] == r[ri] else mismatch
            hd = hp[j]
            hu = hp[j + 1] if j + 1 < w else 0
            hl = hc[j - 1] if j > 0 else 0
            if mutant == "swap_up_diag":
                hd, hu = hu, hd
            oe = gap_open + gap_extend
            if mutant == "open_without_extend":
                oe = gap_open
            eleft = ec[j - 1] + gap_extend if j > 0 else NEG
            if mutant == "textbook_gotoh":
                eup = ep[j + 1] + gap_extend if j + 1 < w else NEG
                e = max(eup, hu + oe)
            elif mutant == "e_never_extends":
                e = hu + oe
            else:
                e = max(eleft, hu + oe)
            f = hl + oe
            if mutant in ("f_extends", "textbook_gotoh"):
                fleft = fc[j - 1] + gap_extend if j > 0 else NEG
                f = max(fleft, hl + oe)
            h = max(0, hd + sub, e, f)
            hc[j], ec[j], fc[j] = h, e, f
            if h > best[0] or (mutant == "tie_last" and h == best[0] and h > 0):
                best = (h, qi, ri)
        hp, ep = hc, ec
    return best


MUTANTS = ["textbook_gotoh", "f_extends", "e_never_extends", "open_without_extend",
           "swap_up_diag", "tie_last"]


def test_unmutated_loop_equals_oracle():
    # The harness for the mutants is itself faithful when nothing is planted.
    for q, r, band, sc in TINY[:200]:
        assert sw_row_major(q, r, band, **sc) == O.sw_banded(q, r, band=band, **sc)


@pytest.mark.parametrize("mutant", MUTANTS)
def test_planted_mistake_fails_the_enumeration(mutant):
    def sw(q, r, band, **sc):
        return sw_row_major(q, r, band, mutant=mutant, **sc)
    assert _check_enumeration(sw) is not None, f"mutant {mutant} equals the enumeration"


@pytest.mark.parametrize("mutant", [m for m in MUTANTS if m != "tie_last"])
def test_planted_gap_mistake_fails_the_closed_forms(mutant):
    # every mistake in the scoring recurrence also changes a planted-gap score
    def sw(q, r, band, **sc):
        return sw_row_major(q, r, band, mutant=mutant, **sc)
    assert _check_closed_forms(sw) is not None, f"mutant {mutant} meets every closed form"


# ---------------------------------------------------- start coordinates

NONPOS = [sc for sc in SCORINGS if sc["mismatch"] <= 0 and sc["gap_extend"] <= 0
          and sc["gap_open"] + sc["gap_extend"] <= 0]
TINY_NONPOS = [c for c in TINY if c[3] in NONPOS]


def sw_start_row_major(q, r, band, match, mismatch, gap_open, gap_extend, mutant=None):
    """The oracle's start-carrying loop (or_sw_banded_start) with one planted
    mistake: 'earliest' resolves equal values towards the earlier start,
    'keep_zero' lets a path whose value fell to 0 keep its start instead
    of restarting, 'start_before' records the cell before the first
    diagonal step."""
    w = 2 * band + 1
    NO = 10 ** 9
    zero = (0, NO, NO)
    sign = -1 if mutant == "earliest" else 1

    def rank(c):
        return (c[0], sign * c[1], sign * c[2])

    def cmax(*cs):
        return max(cs, key=rank)
    hp = [zero] * w
    best, bq, br = zero, 0, 0
    for qi in range(len(q)):
        hc, ec = [zero] * w, [zero] * w
        for j in range(w):
            ri = qi - band + j
            if ri < 0 or ri >= len(r):
                continue
            sub = match if q[qi] == r[ri] else mismatch
            d = (hp[j][0] + sub,) + hp[j][1:]
            if hp[j][1] == NO:
                d = (d[0], qi - 1, ri - 1) if mutant == "start_before" else (d[0], qi, ri)
            up = hp[j + 1] if j + 1 < w else zero
            left = hc[j - 1] if j > 0 else zero
            el = (ec[j - 1][0] + gap_extend,) + ec[j - 1][1:] if j > 0 else (-10 ** 6, NO, NO)
            oe = gap_open + gap_extend
            e = cmax(el, (up[0] + oe,) + up[1:])
            f = (left[0] + oe,) + left[1:]
            h = cmax(zero, d, e, f)
            if h[0] == 0:
                h = zero
            if mutant == "keep_zero":
                c = cmax(d, e, f)
                if c[0] == 0 and c[1] != NO:
                    h = c
            hc[j], ec[j] = h, e
            if h[0] > best[0]:
                best, bq, br = h, qi, ri
        hp = hc
    if best[0] <= 0:
        return (0, 0, 0, 0, 0)
    return (best[0], best[1], best[2], bq, br)


def _check_start_enumeration(sw_start, cases=TINY_NONPOS):
    for q, r, band, sc in cases:
        exp = sw_start_by_alignment_enumeration(q, r, band, **sc)
        got = tuple(sw_start(q, r, band=band, **sc))
        if got != exp:
            return (q, r, band, sc, got, exp)
    return None


def _check_start_closed_forms(sw_start):
    """An identical k-mer over {G, T} between flanks over {A} (query) and
    {C} (reference), which share no base with it or with each other: the
    alignment is exactly the k-mer.  And band 0: Kadane's maximum run on
    the main diagonal, started after the last prefix sum <= 0."""
    rng = random.Random(77)
    for _ in range(60):
        k = rng.randint(1, 12)
        x, u = rng.randint(0, 15), rng.randint(0, 15)
        kmer = rseq(rng, k, "GT")
        q = "A" * x + kmer + "A" * rng.randint(0, 10)
        r = "C" * u + kmer + "C" * rng.randint(0, 10)
        sc = rng.choice(NONPOS)
        band = abs(x - u) + rng.randint(0, 4)
        exp = (sc["match"] * k, x, u, x + k - 1, u + k - 1)
        if tuple(sw_start(q, r, band=band, **sc)) != exp:
            return ("kmer", q, r, band, sc)
    for _ in range(200):
        n = rng.randint(1, 30)
        q, r = rseq(rng, n, "AC"), rseq(rng, n, "AC")
        sc = rng.choice(NONPOS)
        cur, st, best = 0, 0, (0, 0, 0)
        for i in range(n):
            if cur <= 0:
                cur, st = 0, i
            cur += sc["match"] if q[i] == r[i] else sc["mismatch"]
            if cur > best[0]:
                best = (cur, st, i)
        exp = (best[0], best[1], best[1], best[2], best[2]) if best[0] > 0 else (0, 0, 0, 0, 0)
        if tuple(sw_start(q, r, band=0, **sc)) != exp:
            return ("kadane", q, r, sc)
    return None


def test_start_enumeration_by_hand():
    # GGTT shared by AAAGGTTAA and CCGGTTCC: starts (3, 2), ends (6, 5)
    assert sw_start_by_alignment_enumeration("AAGGTTA", "CGGTTCC", 3, **CONFIG_SW) == \
        (8, 2, 1, 5, 4)
    # A C A vs A C A with mismatch 0 would tie; with match 2 / mismatch -2
    # "ACGAC" vs "ACTAC": 2+2-2+2+2 = 6 from (0,0), or 4 from (3,3): the
    # whole diagonal wins; a zero-sum lead-in "AT" + "GG" vs "AA" + "GG" with
    # mismatch -2: 2 - 2 = 0 before the GG, so the later start (2, 2) wins
    sc = dict(match=2, mismatch=-2, gap_open=-4, gap_extend=-2)
    assert sw_start_by_alignment_enumeration("ACGAC", "ACTAC", 0, **sc) == (6, 0, 0, 4, 4)
    assert sw_start_by_alignment_enumeration("ATGG", "AAGG", 0, **sc) == (4, 2, 2, 3, 3)


def test_oracle_start_equals_alignment_enumeration():
    bad = _check_start_enumeration(O.sw_banded_start)
    assert bad is None, bad


def test_oracle_start_closed_forms_and_same_end_as_reference_loop():
    assert _check_start_closed_forms(O.sw_banded_start) is None
    for q, r, band, sc in TINY_NONPOS[:300]:
        s = O.sw_banded_start(q, r, band=band, **sc)
        assert (s[0], s[3], s[4]) == O.sw_banded(q, r, band=band, **sc)


def test_start_loop_copy_equals_oracle():
    for q, r, band, sc in TINY_NONPOS[:200]:
        assert sw_start_row_major(q, r, band, **sc) == O.sw_banded_start(q, r, band=band, **sc)


@pytest.mark.parametrize("mutant", ["earliest", "keep_zero", "start_before"])
def test_planted_start_mistake_fails_a_check(mutant):
    def sw(q, r, band, **sc):
        return sw_start_row_major(q, r, band, mutant=mutant, **sc)
    assert _check_start_enumeration(sw) is not None
    if mutant != "earliest":
        # (ties between two positive-valued paths with different starts
        # need a gap, which neither closed form has)
        assert _check_start_closed_forms(sw) is not None


def test_oracle_start_batch_matches_per_pair_calls():
    rng = random.Random(31)
    pairs = [(rseq(rng, rng.randint(0, 80)), rseq(rng, rng.randint(0, 80))) for _ in range(60)]
    bands = [rng.randint(0, 40) for _ in pairs]

    def cat(seqs):
        ws, offs, lens, o = [], [], [], 0
        for x in seqs:
            w, _ = O.pack_ascii(x)
            ws.append(w)
            offs.append(o)
            lens.append(len(x))
            o += w.size
        return np.concatenate(ws), np.array(offs, np.int32), np.array(lens, np.int32)
    qp, qo, ql = cat([p[0] for p in pairs])
    rp, ro, rl = cat([p[1] for p in pairs])
    got = O.sw_start_batch(qp, rp, qo, ro, ql, rl, band=np.array(bands, np.int32), threads=4,
                           **CONFIG_SW)
    for i, (x, y) in enumerate(pairs):
        assert tuple(int(g[i]) for g in got) == O.sw_banded_start(x, y, band=bands[i], **CONFIG_SW)
