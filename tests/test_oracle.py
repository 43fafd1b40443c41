"""Pins the CPU oracle (oracle/cgm_oracle.c) -- no GPU needed.

Each oracle function is checked against the reference's own known answers
and against what the mathematics fixes independently of the restatement:
closed forms, invariants, a textbook Gotoh DP, an exhaustive enumeration of
alignments, Kadane's maximum-subarray on band 0, and order independence.
"""
import functools
import itertools
import json
import os
import random

import numpy as np
import pytest

from conftest import load_oracle

O = load_oracle()
DEFAULT_SW = dict(match=2, mismatch=-3, gap_open=-5, gap_extend=-2)   # align_cpu.mojo:46-53
CONFIG_SW = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "alignment_vectors.json")


def rseq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(alphabet) for _ in range(n))


# ------------------------------------------------------------------ packing

def test_pack_msb_first_and_zero_fill():
    # test_dna.mojo:86-97 and test_sequence.mojo:30-38: ACGT at positions 0-3, rest A
    w, m = O.pack_ascii("ACGT")
    assert w.size == 1 and m[0] == 0
    assert int(w[0]) == (0 << 62) | (1 << 60) | (2 << 58) | (3 << 56)
    assert [O.lib().or_get_base_public(O._ptr(w), i) for i in range(5)] == [0, 1, 2, 3, 0]


def test_pack_33_bases_spans_two_words():
    # test_sequence.mojo:40-47
    w, _ = O.pack_ascii("ACGT" * 8 + "A")
    assert w.size == 2
    assert O.lib().or_get_base_public(O._ptr(w), 32) == 0


def test_pack_n_is_a_plus_mask_bit():
    # test_sequence.mojo:50-60 (ACNT: N at position 2), dna.mojo:26-40 ('-' -> N)
    w, m = O.pack_ascii("ACNT")
    assert int(m[0]) == 1 << 2
    assert O.lib().or_get_base_public(O._ptr(w), 2) == 0
    w2, m2 = O.pack_ascii("acgt-")
    assert int(m2[0]) == 1 << 4
    assert [O.lib().or_get_base_public(O._ptr(w2), i) for i in range(4)] == [0, 1, 2, 3]


# ------------------------------------------------------------ SW: known answers

def test_sw_reference_known_answers():
    # tests/test_alignment.mojo:24-58
    assert O.sw_banded("ACGTACGT", "ACGTACGT", band=8, **DEFAULT_SW)[0] == 16
    s = O.sw_banded("ACGTACGT", "ACGAACGT", band=4, **DEFAULT_SW)[0]
    assert s > 0 and s >= 8
    assert O.sw_banded("AAAA", "TTTT", band=4, **DEFAULT_SW)[0] == 0


def test_sw_self_alignment_100bp_is_200():
    # test_real_data.mojo:132-140 asks 100 * match for a 100-bp self alignment
    rng = random.Random(7)
    for _ in range(5):
        s = rseq(rng, 100)
        assert O.sw_banded(s, s, band=32, **DEFAULT_SW) == (200, 99, 99)


@pytest.mark.parametrize("L", [1, 2, 31, 32, 33, 150, 333])
@pytest.mark.parametrize("band", [0, 1, 5, 32, 64])
def test_sw_identical_closed_form(L, band):
    s = rseq(random.Random(L * 131 + band), L)
    for sc in (DEFAULT_SW, CONFIG_SW):
        assert O.sw_banded(s, s, band=band, **sc) == (sc["match"] * L, L - 1, L - 1)


def test_sw_no_common_base_and_empty():
    assert O.sw_banded("AAAAAA", "TTTTTTTT", band=8, **CONFIG_SW) == (0, 0, 0)
    assert O.sw_banded("", "ACGT", band=4) == (0, 0, 0)
    assert O.sw_banded("ACGT", "", band=4) == (0, 0, 0)
    assert O.sw_banded("", "", band=4) == (0, 0, 0)


def _kadane_first(vals):
    best, bi, cur = 0, 0, 0
    for i, v in enumerate(vals):
        cur = max(0, cur + v)
        if cur > best:
            best, bi = cur, i
    return best, bi


def test_sw_band0_is_kadane_on_main_diagonal():
    # band 0: up/left neighbours are outside the band, so e = f = open+extend
    # <= 0 and H(i,i) = max(0, H(i-1,i-1) + sub): a maximum-subarray scan.
    rng = random.Random(11)
    for _ in range(300):
        n = rng.randint(1, 60)
        q, r = rseq(rng, n), rseq(rng, rng.randint(1, 60))
        sc = rng.choice([DEFAULT_SW, CONFIG_SW])
        m = min(len(q), len(r))
        subs = [sc["match"] if q[i] == r[i] else sc["mismatch"] for i in range(m)]
        best, bi = _kadane_first(subs)
        exp = (best, bi, bi) if best > 0 else (0, 0, 0)
        assert O.sw_banded(q, r, band=0, **sc) == exp


@pytest.mark.parametrize("L,p", [(20, 0), (20, 19), (20, 10), (9, 4), (40, 3)])
def test_sw_single_substitution_closed_form(L, p):
    # One substitution in an identical pair: the best of the whole diagonal
    # (match*(L-1) + mismatch), the prefix (match*p) and the suffix; any gap
    # detour costs more than the mismatch under the default scores.
    rng = random.Random(L + 100 * p)
    s = rseq(rng, L)
    t = s[:p] + {"A": "C", "C": "G", "G": "T", "T": "A"}[s[p]] + s[p + 1:]
    sc = DEFAULT_SW
    whole = sc["match"] * (L - 1) + sc["mismatch"]
    pre, suf = sc["match"] * p, sc["match"] * (L - 1 - p)
    score, qe, re_ = O.sw_banded(s, t, band=4, **sc)
    assert score == max(whole, pre, suf)
    if whole > max(pre, suf):
        assert (qe, re_) == (L - 1, L - 1)


def test_sw_band_monotone_and_kadane_lower_bound():
    rng = random.Random(3)
    for _ in range(200):
        q = rseq(rng, rng.randint(1, 50))
        r = rseq(rng, rng.randint(1, 50))
        sc = rng.choice([DEFAULT_SW, CONFIG_SW])
        prev = -1
        for band in range(0, 12):
            s = O.sw_banded(q, r, band=band, **sc)[0]
            assert s >= prev
            prev = s
            for k in range(-band, band + 1):  # every band diagonal k = ri - qi
                subs = [sc["match"] if q[i] == r[i + k] else sc["mismatch"]
                        for i in range(len(q)) if 0 <= i + k < len(r)]
                assert s >= _kadane_first(subs)[0]


def test_sw_row_major_equals_antidiagonal_order():
    rng = random.Random(5)
    for _ in range(3000):
        q = rseq(rng, rng.randint(0, 40))
        r = rseq(rng, rng.randint(0, 40))
        band = rng.randint(0, 20)
        sc = rng.choice([DEFAULT_SW, CONFIG_SW, dict(match=1, mismatch=-1, gap_open=-1,
                                                     gap_extend=-1)])
        assert O.sw_banded(q, r, band=band, **sc) == O.sw_antidiag(q, r, band=band, **sc)


def _sw_topdown(q, r, band, match, mismatch, gap_open, gap_extend):
    """The align_cpu.mojo:229-309 recurrence evaluated top-down by recursion
    over matrix coordinates (no band-offset arrays, no row buffers)."""
    NEG = -1000000
    oe = gap_open + gap_extend

    def inband(i, j):
        return abs(j - i) <= band

    @functools.lru_cache(None)
    def H(i, j):
        if i < 0 or j < 0 or j >= len(r) or not inband(i, j):
            return 0
        sub = match if q[i] == r[j] else mismatch
        return max(0, H(i - 1, j - 1) + sub, E(i, j), H(i, j - 1) + oe)

    @functools.lru_cache(None)
    def E(i, j):
        if j < 0:
            return 0
        left = NEG if j - i == -band else (0 if j == 0 else E(i, j - 1))
        return max(left + gap_extend, H(i - 1, j) + oe)

    best = (0, 0, 0)
    for i in range(len(q)):
        for j in range(len(r)):
            if inband(i, j) and H(i, j) > best[0]:
                best = (H(i, j), i, j)
    return best


def test_sw_topdown_recursion_on_tiny_inputs():
    rng = random.Random(9)
    for _ in range(1500):
        q = rseq(rng, rng.randint(0, 7), "ACG")
        r = rseq(rng, rng.randint(0, 7), "ACG")
        band = rng.randint(0, 7)
        sc = rng.choice([DEFAULT_SW, CONFIG_SW, dict(match=3, mismatch=-1, gap_open=-1,
                                                     gap_extend=0)])
        assert O.sw_banded(q, r, band=band, **sc) == _sw_topdown(q, r, band, **sc)


# ------------------------------------------------------------ WFA

def test_wfa_reference_known_answers():
    # tests/test_alignment.mojo:87-138 with default_wfa_config (4, 6, 2, 200)
    cases = [("ACGTACGT", "ACGTACGT", 0), ("ACGT", "ACAT", 4), ("ACGT", "ACGGT", 8),
             ("ACGT", "ACGGGT", 10), ("ACGGT", "ACGT", 8), ("AAAA", "TTTT", 16)]
    for q, r, c in cases:
        assert O.wfa_affine(q, r) == c


def test_wfa_closed_forms():
    rng = random.Random(2)
    for L in (1, 5, 31, 32, 33, 100, 400):
        s = rseq(rng, L)
        assert O.wfa_affine(s, s, max_error=5) == 0
        assert O.wfa_affine("", s, max_error=6 + 2 * L) == 6 + 2 * L        # empty vs L
        assert O.wfa_affine(s, "", max_error=6 + 2 * L) == 6 + 2 * L
        assert O.wfa_affine("", s, max_error=6 + 2 * L - 1) == -1
    assert O.wfa_affine("", "", max_error=0) == 0
    # |qlen - rlen| > max_error: no diagonal reaches the target
    assert O.wfa_affine("ACGTACGTAC", "ACG", max_error=6) == -1
    # unrelated 100-bp sequences exceed max_error 50 (test_real_data.mojo:206-219)
    a, b = rseq(random.Random(1), 100), rseq(random.Random(2), 100)
    assert O.wfa_affine(a, b, max_error=50) == -1


def _enum_alignment_cost(q, r, x, o, e):
    """Minimum gap-affine cost over every alignment (exhaustive)."""
    best = [None]

    def go(i, j, prev, cost):
        if best[0] is not None and cost >= best[0]:
            return
        if i == len(q) and j == len(r):
            best[0] = cost
            return
        if i < len(q) and j < len(r):
            go(i + 1, j + 1, "M", cost + (0 if q[i] == r[j] else x))
        if i < len(q):
            go(i + 1, j, "I", cost + (e if prev == "I" else o + e))
        if j < len(r):
            go(i, j + 1, "D", cost + (e if prev == "D" else o + e))

    go(0, 0, "M", 0)
    return best[0]


def test_gotoh_equals_exhaustive_enumeration():
    rng = random.Random(4)
    for _ in range(400):
        q = rseq(rng, rng.randint(0, 5))
        r = rseq(rng, rng.randint(0, 5))
        x, o, e = rng.choice([(4, 6, 2), (1, 1, 1), (3, 0, 2), (2, 5, 1)])
        assert O.gotoh_global(q, r, x, o, e) == _enum_alignment_cost(q, r, x, o, e)


def test_wfa_equals_gotoh_random():
    rng = random.Random(6)
    for _ in range(3000):
        q = rseq(rng, rng.randint(0, 30))
        r = rseq(rng, rng.randint(0, 30)) if rng.random() < 0.5 else \
            "".join(c for c in q if rng.random() > 0.1)
        x, o, e = rng.choice([(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1)])
        me = rng.choice([0, 4, 12, 30, 80, 200])
        assert O.wfa_affine(q, r, x, o, e, me) == O.gotoh_global(q, r, x, o, e, me)


def test_wfa_equals_gotoh_exhaustive_small():
    seqs = ["".join(p) for n in range(0, 4) for p in itertools.product("ACG", repeat=n)]
    for q in seqs:
        for r in seqs:
            assert O.wfa_affine(q, r, 4, 6, 2, 40) == O.gotoh_global(q, r, 4, 6, 2, 40)


def test_wfa_equals_gotoh_long_diverged():
    rng = random.Random(8)
    for _ in range(20):
        t = rseq(rng, rng.randint(300, 900))
        q = []
        for ch in t:
            u = rng.random()
            if u < 0.02:
                q.append(rng.choice("ACGT"))
            elif u < 0.03:
                q.extend([ch] + [rng.choice("ACGT") for _ in range(rng.randint(1, 4))])
            elif u < 0.04:
                continue
            else:
                q.append(ch)
        q = "".join(q)
        g = O.gotoh_global(q, t)
        assert O.wfa_affine(q, t, max_error=g) == g
        assert O.wfa_affine(q, t, max_error=g - 1) == -1


# ------------------------------------------------------------ golden vectors

def test_golden_vectors_reproduce():
    with open(GOLDEN) as f:
        vec = json.load(f)
    for v in vec["sw_kat"]:
        s = O.sw_banded(v["q"], v["r"], band=v["band"], **v["scoring"])[0]
        if "score" in v:
            assert s == v["score"], v["src"]
        else:
            assert s >= v["score_at_least"], v["src"]
    for v in vec["wfa_kat"]:
        assert O.wfa_affine(v["q"], v["r"], v["x"], v["o"], v["e"], v["max_error"]) == v["cost"]
    for v in vec["sw_regress"]:
        assert O.sw_banded(v["q"], v["r"], band=v["band"], **v["scoring"]) == \
            (v["score"], v["q_end"], v["r_end"])
    for v in vec["wfa_regress"]:
        assert O.wfa_affine(v["q"], v["r"], v["x"], v["o"], v["e"], v["max_error"]) == v["cost"]


def test_batch_driver_matches_per_pair_calls():
    rng = random.Random(12)
    qs = [rseq(rng, rng.randint(0, 70)) for _ in range(50)]
    rs = [rseq(rng, rng.randint(0, 70)) for _ in range(50)]

    def cat(seqs):
        words, offs, lens, o = [], [], [], 0
        for s in seqs:
            w, _ = O.pack_ascii(s)
            words.append(w)
            offs.append(o)
            lens.append(len(s))
            o += w.size
        return np.concatenate(words), np.array(offs, np.int32), np.array(lens, np.int32)

    qp, qo, ql = cat(qs)
    rp, ro, rl = cat(rs)
    band = np.array([rng.randint(0, 20) for _ in qs], np.int32)
    s, qe, re_ = O.batch(0, qp, rp, qo, ro, ql, rl, band=band, threads=3, **CONFIG_SW)
    for i in range(50):
        assert (s[i], qe[i], re_[i]) == O.sw_banded(qs[i], rs[i], band=int(band[i]), **CONFIG_SW)
    me = np.array([rng.randint(0, 100) for _ in qs], np.int32)
    c, _, _ = O.batch(1, qp, rp, qo, ro, ql, rl, max_error=me, threads=2)
    for i in range(50):
        assert c[i] == O.wfa_affine(qs[i], rs[i], max_error=int(me[i]))


def test_sw_band_wider_than_both_sequences_is_clamped_exactly():
    # A band above max(qlen, rlen) adds no in-matrix cell and changes no
    # boundary value, so min(band, max(qlen, rlen)) gives identical results
    # (the GPU layouts rely on this to size LDS).
    rng = random.Random(13)
    for _ in range(2000):
        q = rseq(rng, rng.randint(0, 30))
        r = rseq(rng, rng.randint(0, 30))
        band = rng.randint(0, 80)
        sc = rng.choice([DEFAULT_SW, CONFIG_SW])
        eff = min(band, max(len(q), len(r)))
        assert O.sw_banded(q, r, band=band, **sc) == O.sw_banded(q, r, band=eff, **sc)


# ----------------------------------------------------------- WFA with CIGAR

def cigar_ops(cigar):
    """'3=1X2I' -> [(3, '='), (1, 'X'), (2, 'I')]"""
    import re
    return [(int(n), op) for n, op in re.findall(r"(\d+)([=XID])", cigar)]


def check_cigar(q, r, cigar, x=4, o=6, e=2):
    """Cost of the CIGAR (mismatch x, gap of l bases o + l e), and that it
    walks both sequences exactly: '=' on equal bases, 'X' on different."""
    i = j = cost = 0
    for n, op in cigar_ops(cigar):
        assert n > 0
        if op in "=X":
            for _ in range(n):
                assert (q[i] == r[j]) == (op == "="), (q, r, cigar)
                i += 1
                j += 1
            cost += x * n if op == "X" else 0
        elif op == "I":
            i += n
            cost += o + e * n
        else:
            j += n
            cost += o + e * n
    assert (i, j) == (len(q), len(r)), (q, r, cigar)
    return cost


def test_wfa_cigar_reference_known_answers():
    """test_alignment.mojo:201-279: exact CIGAR strings and scores."""
    assert O.wfa_affine_cigar("ACGTACGT", "ACGTACGT") == (0, "8=")
    assert O.wfa_affine_cigar("ACGT", "ACAT") == (4, "2=1X1=")
    assert O.wfa_affine_cigar("ACGGT", "ACGT") == (8, "3=1I1=")
    assert O.wfa_affine_cigar("ACGT", "ACGGT") == (8, "3=1D1=")
    assert O.wfa_affine_cigar("ACGT", "ACGGGT") == (10, "3=2D1=")
    for q, r in (("ACGT", "ACAT"), ("ACGGT", "ACGT"), ("ACGT", "ACGGT"), ("ACGT", "ACGGGT"),
                 ("AAAA", "TTTT"), ("ACGTACGT", "ACGTACGT")):
        c, cig = O.wfa_affine_cigar(q, r)
        assert c == O.wfa_affine(q, r)
        assert check_cigar(q, r, cig) == c


def test_wfa_cigar_edge_cases():
    assert O.wfa_affine_cigar("", "") == (0, "")
    assert O.wfa_affine_cigar("", "ACG") == (12, "3D")
    assert O.wfa_affine_cigar("ACGTA", "") == (16, "5I")
    assert O.wfa_affine_cigar("ACGT", "TTTT", max_error=11) == (-1, "")
    assert O.wfa_affine_cigar("A" * 40, "A" * 60, max_error=45) == (-1, "")


def test_wfa_cigar_cost_and_walk_random():
    """Random pairs (related and unrelated, several penalty sets): the cost
    equals wfa_affine_cpu's and the textbook Gotoh optimum, and the CIGAR
    reproduces that cost and walks both sequences exactly."""
    rng = random.Random(23)
    pens = [(4, 6, 2), (1, 0, 1), (3, 2, 1), (5, 1, 3), (2, 5, 1)]
    for _ in range(1500):
        q = rseq(rng, rng.randint(0, 60))
        r = "".join(c if rng.random() > 0.12 else rng.choice("ACGT") for c in q
                    if rng.random() > 0.06) if rng.random() < 0.7 else rseq(rng, rng.randint(0, 60))
        x, o, e = rng.choice(pens)
        c, cig = O.wfa_affine_cigar(q, r, x, o, e, max_error=400)
        assert c == O.wfa_affine(q, r, x, o, e, max_error=400)
        assert c == O.gotoh_global(q, r, x, o, e)
        assert check_cigar(q, r, cig, x, o, e) == c


def test_wfa_cigar_long_diverged_and_batch_driver():
    rng = random.Random(31)
    qs, rs, mes = [], [], []
    for _ in range(40):
        q = rseq(rng, rng.randint(200, 1200))
        r = "".join(c if rng.random() > 0.03 else rng.choice("ACGT") for c in q
                    if rng.random() > 0.01)
        c = O.wfa_affine(q, r, max_error=1000)
        assert c >= 0
        qs.append(q); rs.append(r); mes.append(c + rng.choice([0, 2, 30]))
    pq = [O.pack_ascii(s) for s in qs]
    pr = [O.pack_ascii(s) for s in rs]
    qoff = np.cumsum([0] + [w.size for w, _ in pq])[:-1].astype(np.int32)
    roff = np.cumsum([0] + [w.size for w, _ in pr])[:-1].astype(np.int32)
    cost, cigars = O.wfa_cigar_batch(np.concatenate([w for w, _ in pq]),
                                     np.concatenate([w for w, _ in pr]), qoff, roff,
                                     [len(s) for s in qs], [len(s) for s in rs],
                                     max_error=np.array(mes, np.int32), threads=4)
    for q, r, me, c, cig in zip(qs, rs, mes, cost, cigars):
        assert (c, cig) == O.wfa_affine_cigar(q, r, max_error=me)
        assert check_cigar(q, r, cig) == c


# -------------------------------------------------------------- banded NW

NW_NEG = -1000000


def test_nw_reference_known_answers():
    """test_alignment.mojo:61-80 (default_config, band 4)."""
    assert O.nw_banded("ACGT", "ACGT", band=4)[0] == 8
    assert O.nw_banded("AAAA", "TTTT", band=4)[0] == -12


@pytest.mark.parametrize("L,band", [(1, 0), (7, 0), (30, 3), (100, 32), (65, 64)])
def test_nw_closed_forms(L, band):
    rng = random.Random(L * 131 + band)
    q = rseq(rng, L)
    assert O.nw_banded(q, q, band=band) == (2 * L, L - 1, L - 1)
    p = rng.randrange(L)
    r = q[:p] + "ACGT"[("ACGT".index(q[p]) + 1) % 4] + q[p + 1:]
    # one substitution: match everywhere else unless two gaps (2 x -7) beat
    # the mismatch (-3 - 2 = -5 versus the match), which they never do
    assert O.nw_banded(q, r, band=band)[0] == 2 * (L - 1) - 3


def test_nw_band_too_narrow_and_empty():
    assert O.nw_banded("ACGTACGT", "ACG", band=4)[0] == NW_NEG   # |qlen - rlen| = 5 > 4
    assert O.nw_banded("", "", band=3) == (0, -1, -1)
    assert O.nw_banded("", "ACG", band=3) == (-5 + 3 * -2, -1, 2)   # leading reference gap
    assert O.nw_banded("", "ACGT", band=3)[0] == NW_NEG             # past the first row's band
    assert O.nw_banded("ACG", "", band=5)[0] == NW_NEG              # no leading query gap


def _nw_paths_best(q, r, band, match, mismatch, go, ge):
    """Every banded path from the origin, the reference's cost model: a
    start on row -1 at column c (value 0 for c = -1, go + (c + 1) ge for
    0 <= c < band, i.e. a leading reference gap), then diagonal steps
    (match / mismatch), down and right steps (go + ge each) through cells
    0 <= ri < rlen, |ri - qi| <= band; no cell of column -1 below row -1."""
    ql, rl = len(q), len(r)
    if ql == 0:
        if rl == 0:
            return 0
        return go + rl * ge if rl <= band else NW_NEG
    best = None

    @functools.lru_cache(maxsize=None)
    def walk(qi, ri):
        # best path value from cell (qi, ri) (already entered) to the end
        if (qi, ri) == (ql - 1, rl - 1):
            return 0
        out = []
        for dq, dr in ((1, 1), (1, 0), (0, 1)):
            a, b = qi + dq, ri + dr
            if a >= ql or b >= rl or b < 0 or abs(b - a) > band:
                continue
            step = (match if q[a] == r[b] else mismatch) if (dq, dr) == (1, 1) else go + ge
            rest = walk(a, b)
            if rest is not None:
                out.append(step + rest)
        return max(out) if out else None

    for c in range(-1, band):
        start = 0 if c == -1 else go + (c + 1) * ge
        for dq, dr in ((1, 1), (1, 0)):
            a, b = 0, c + dr
            if b < 0 or b >= rl or abs(b - a) > band:
                continue
            step = (match if q[0] == r[b] else mismatch) if dr == 1 else go + ge
            rest = walk(a, b)
            if rest is not None:
                v = start + step + rest
                best = v if best is None else max(best, v)
    return NW_NEG if best is None else best


def test_nw_equals_path_brute_force():
    rng = random.Random(77)
    scorings = [(2, -3, -5, -2), (1, -1, -1, -1), (3, 1, -2, 0), (2, -4, 0, -3)]
    for _ in range(2500):
        q = rseq(rng, rng.randint(0, 6))
        r = rseq(rng, rng.randint(0, 6))
        band = rng.randint(0, 6)
        sc = rng.choice(scorings)
        exp = _nw_paths_best(q, r, band, *sc)
        assert O.nw_banded(q, r, *sc, band=band)[0] == exp, (q, r, band, sc)
