"""CPU model of the wavefront-per-pair SW layout, checked against the oracle.

The layout (cgm_sw_wave_dev.h in the package's csrc/) gives one 64-lane
wavefront to one pair of effective band B <= 64 and sweeps the band's
anti-diagonals a = qi + ri.  Band diagonal d = ri - qi (|d| <= B) has a cell
on anti-diagonal a when d = a (mod 2), so lane l owns the two diagonals
de(l) = -B + 2l and do(l) = -B + 2l + 1 and computes one cell per step: on
its de diagonal when a = B (mod 2) ("even" steps), on its do diagonal
otherwise.  Every input of a cell is on the previous two anti-diagonals:
diag H(qi-1, ri-1) is the lane's own last value on that diagonal; up
H(qi-1, ri) and left H, E(qi, ri-1) are the values of diagonals d + 1 and
d - 1 from the previous step, the lane's own other diagonal or its
neighbour's (one lane down for de's left, one lane up for do's up).  The
band's outer edges read the reference's values (left of d = -B: H 0 and no
E; above d = B: H 0; with penalties <= 0 that missing E term never decides
a cell).  With B = 64 the diagonal d = 64 has no lane; its cell
is computed beside lane 63 (in a third register set of every lane).
Bases reach the lanes by the same shifts: on every step of a parity the
query bases move one lane up and the reference bases one lane down, with
one new base entering at each end.  Cells outside the matrix hold H = 0 and
E = 0, as the reference's skipped cells do.  Each diagonal keeps its first
strict maximum (rows ascend along a diagonal); the end cell is the maximum
of those with the smallest query row, then reference column, which is the
reference's first row-major maximum.  This restatement must equal
smith_waterman_banded (align_cpu.mojo:229-309) and copies with a planted
slip must not.
"""
import random

import pytest

from conftest import load_oracle

O = load_oracle()
NEG = -(1 << 40)


def sw_wave_model(q, r, band, match, mismatch, gap_open, gap_extend, lanes=64, mutant=None):
    """(score, query_end, ref_end) as the wave layout computes them, or None
    when the effective band exceeds the lanes."""
    ql, rl = len(q), len(r)
    B = min(band, max(ql, rl))
    L = lanes
    if B > L:
        return None
    oe = gap_open + gap_extend
    extra = B == L  # the even diagonal d = B has no lane

    def base(s, i):
        return s[i] if 0 <= i < len(s) else "A"

    He, Ee, Ho, Eo = [0] * L, [0] * L, [0] * L, [0] * L
    Hx = 0
    # per diagonal: first strict maximum and the anti-diagonal it was on
    bve, bae, bvo, bao = [0] * L, [0] * L, [0] * L, [0] * L
    bvx, bax = 0, 0
    a_first = (B & 1) - 1          # the first (odd) step
    # bases as they stand before the first step of each parity
    a_o, a_e = a_first - 2, a_first - 1
    qbo = [base(q, (a_o + B - 1) // 2 - l) for l in range(L)]
    rbo = [base(r, (a_o - B + 1) // 2 + l) for l in range(L)]
    qbe = [base(q, (a_e + B) // 2 - l) for l in range(L)]
    rbe = [base(r, (a_e - B) // 2 + l) for l in range(L)]

    def shift_up(v, new):      # lane l takes lane l - 1's value; lane 0 takes new
        return [new] + v[:-1]

    def shift_down(v, new):    # lane l takes lane l + 1's value; lane L - 1 takes new
        return v[1:] + [new]

    def cell(qi, ri, qb, rb, diag, up, hl, el):
        sub = match if qb == rb else mismatch
        e = max(el + gap_extend, up + oe)
        f = hl + oe
        return max(0, diag + sub, e, f), e

    a = a_first
    while a <= ql + rl:
        # odd step: lane l on do(l) = -B + 2l + 1
        if mutant == "base_backwards":
            qbo = shift_down(qbo, base(q, (a + B - 1) // 2 - (L - 1)))
        else:
            qbo = shift_up(qbo, base(q, (a + B - 1) // 2))
        rbo = shift_down(rbo, base(r, (a - B + 1) // 2 + L - 1))
        nHo, nEo = Ho[:], Eo[:]
        for l in range(L):
            qi, ri = (a + B - 1) // 2 - l, (a - B + 1) // 2 + l
            if mutant == "up_wrong_side":
                up = He[l - 1] if l > 0 else 0
            else:
                up = He[l + 1] if l + 1 < L else (Hx if extra else 0)
            if 0 <= qi < ql and 0 <= ri < rl and l <= B - 1:
                h, e = cell(qi, ri, qbo[l], rbo[l], Ho[l], up, He[l], Ee[l])
                if h > bvo[l] or (mutant == "last_max" and h == bvo[l] and h > 0):
                    bvo[l], bao[l] = h, a
            else:
                h, e = 0, 0
            nHo[l], nEo[l] = h, e
        Ho, Eo = nHo, nEo
        a += 1
        # even step: lane l on de(l) = -B + 2l, and d = B beside lane L - 1
        qbe = shift_up(qbe, base(q, (a + B) // 2))
        rbe = shift_down(rbe, base(r, (a - B) // 2 + L - 1))
        nHe, nEe = He[:], Ee[:]
        for l in range(L):
            qi, ri = (a + B) // 2 - l, (a - B) // 2 + l
            hl = Ho[l - 1] if l > 0 else 0
            el = Eo[l - 1] if l > 0 else NEG
            diag = Ho[l] if mutant == "diag_other" else He[l]
            if 0 <= qi < ql and 0 <= ri < rl and l <= B:
                h, e = cell(qi, ri, qbe[l], rbe[l], diag, Ho[l], hl, el)
                if h > bve[l] or (mutant == "last_max" and h == bve[l] and h > 0):
                    bve[l], bae[l] = h, a
            else:
                h, e = 0, 0
            nHe[l], nEe[l] = h, e
        if extra and mutant != "no_extra":
            qi, ri = (a - B) // 2, (a + B) // 2
            if 0 <= qi < ql and 0 <= ri < rl:
                h, _ = cell(qi, ri, base(q, qi), base(r, ri), Hx, 0, Ho[L - 1], Eo[L - 1])
                if h > bvx:
                    bvx, bax = h, a
            else:
                h = 0
            Hx = h
        He, Ee = nHe, nEe
        a += 1
    # merge: the largest maximum, then the smallest query row, then column
    cands = []
    for l in range(L):
        if bve[l] > 0:
            cands.append((bve[l], (bae[l] + B) // 2 - l, (bae[l] - B) // 2 + l))
        if bvo[l] > 0:
            cands.append((bvo[l], (bao[l] + B - 1) // 2 - l, (bao[l] - B + 1) // 2 + l))
    if bvx > 0:
        cands.append((bvx, (bax - B) // 2, (bax + B) // 2))
    if not cands:
        return 0, 0, 0
    best = max(c[0] for c in cands)
    _, bq, br = min((c for c in cands if c[0] == best), key=lambda c: (c[1], c[2]))
    return best, bq, br


SCORINGS = [(2, -3, -5, -2), (2, -4, -4, -2), (1, -1, -1, -1), (3, 0, -2, 0), (2, -1, 0, -1),
            (5, -4, 0, 0), (1, 0, 0, 0)]


def _pair(rng, maxlen, alpha="ACGT"):
    q = "".join(rng.choice(alpha) for _ in range(rng.randint(0, maxlen)))
    if rng.random() < 0.5:
        r = "".join(c if rng.random() > 0.1 else rng.choice(alpha) for c in q
                    if rng.random() > 0.05)
        if rng.random() < 0.3:
            r = "".join(rng.choice(alpha) for _ in range(rng.randint(0, 6))) + r
    else:
        r = "".join(rng.choice(alpha) for _ in range(rng.randint(0, maxlen)))
    return (q, r) if rng.random() < 0.5 else (r, q)


def test_wave_model_equals_oracle_few_lanes():
    """4 and 5 lanes: bands up to the lane count, so the extra diagonal,
    both band edges and the matrix corners are all exercised."""
    rng = random.Random(11)
    n = 0
    for _ in range(5000):
        lanes = rng.choice([4, 5])
        q, r = _pair(rng, 24, rng.choice(["AC", "ACGT"]))
        band = rng.randint(0, lanes)
        sc = rng.choice(SCORINGS)
        got = sw_wave_model(q, r, band, *sc, lanes=lanes)
        if got is None:
            continue
        n += 1
        assert got == O.sw_banded(q, r, *sc, band=band), (q, r, band, sc, lanes)
    assert n > 4000


def test_wave_model_equals_oracle_64_lanes():
    rng = random.Random(12)
    for i in range(150):
        q, r = _pair(rng, 160)
        band = rng.choice([0, 1, 7, 31, 32, 33, 63, 64, rng.randint(0, 64)])
        sc = rng.choice(SCORINGS)
        assert sw_wave_model(q, r, band, *sc) == O.sw_banded(q, r, *sc, band=band), \
            (q, r, band, sc)


@pytest.mark.parametrize("mutant", ["no_extra", "diag_other", "up_wrong_side", "last_max",
                                    "base_backwards"])
def test_planted_wave_mistakes_fail(mutant):
    rng = random.Random(13)
    bad = 0
    for _ in range(800):
        q, r = _pair(rng, 24, rng.choice(["AC", "ACGT"]))
        band = rng.choice([4, rng.randint(0, 4)])
        sc = rng.choice(SCORINGS)
        got = sw_wave_model(q, r, band, *sc, lanes=4, mutant=mutant)
        if got is not None and got != O.sw_banded(q, r, *sc, band=band):
            bad += 1
    assert bad > 0, mutant


def wave_feed_model(q, r, band, lanes=64):
    """The base feeds of the wavefront layout, restated register by register:
    yields, per step (anti-diagonals a and a + 1), the query base entering
    lane 0 on the odd and on the even step, the reference base entering lane
    lanes - 1, and (band = lanes) the d = band cell's query base (lane
    lanes - 1's even-step base before its shift) and reference base (lane
    lanes - 1 of the reference feed after it moved)."""
    L = lanes
    ql, rl = len(q), len(r)
    B = min(band, max(ql, rl))
    codes = {"A": 0, "C": 1, "G": 2, "T": 3}

    def b0(s, p):
        return codes[s[p]] if 0 <= p < len(s) else 0

    a_first = (B & 1) - 1
    a_end = (ql - 1) + min(rl - 1, ql - 1 + B)
    c1_0, c2_0 = (a_first + B - 1) // 2, (a_first - B + 1) // 2
    pq = [c1_0 + 1 + l for l in range(L)]
    pr = [c2_0 + (L - 1) + (L - 1) - l for l in range(L)]
    fq = [b0(q, p) for p in pq]
    fr = [b0(r, p) for p in pr]
    nq = [p + L for p in pq]  # positions whose words are in flight
    nr = [p + L for p in pr]
    qlast = [b0(q, c1_0)] * L
    qbe = [b0(q, (a_first - 1 + B) // 2 - l) for l in range(L)]
    fill = 0
    a = a_first
    while a <= a_end:
        q_odd = qlast[0]
        q_even = fq[0]
        r_new = fr[L - 1]
        qx = qbe[L - 1]
        qbe = [q_even] + qbe[:-1]
        qlast = fq[:]
        fq = fq[1:] + [0]       # lane l <- lane l + 1
        fr = [0] + fr[:-1]      # lane l <- lane l - 1
        fill += 1
        if fill == L:
            fill = 0
            pq = [p + L for p in pq]
            pr = [p + L for p in pr]
            assert pq == nq and pr == nr
            fq = [b0(q, p) for p in pq]
            fr = [b0(r, p) for p in pr]
            nq = [p + L for p in pq]
            nr = [p + L for p in pr]
        yield a, q_odd, q_even, r_new, qx, fr[L - 1]
        a += 2


def test_wave_feeds_deliver_the_direct_bases():
    """The feed registers hand each step exactly the bases the layout reads
    directly (query (a + B - 1) / 2 on the odd step, (a + 1 + B) / 2 on the
    even one, reference (a - B + 1) / 2 + 63 on both), across refills, and
    the d = 64 cell gets query (a + 1 - 64) / 2 and reference (a + 1 + 64) / 2."""
    rng = random.Random(14)
    codes = {"A": 0, "C": 1, "G": 2, "T": 3}
    refilled = 0
    for _ in range(60):
        q = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 400)))
        r = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 400)))
        band = rng.choice([0, 1, 17, 63, 64, rng.randint(0, 64)])
        B = min(band, max(len(q), len(r)))

        def b0(s, p):
            return codes[s[p]] if 0 <= p < len(s) else 0

        steps = 0
        for a, q_odd, q_even, r_new, qx, rx in wave_feed_model(q, r, band):
            steps += 1
            assert q_odd == b0(q, (a + B - 1) // 2), (q, r, band, a)
            assert q_even == b0(q, (a + 1 + B) // 2), (q, r, band, a)
            assert r_new == b0(r, (a - B + 1) // 2 + 63), (q, r, band, a)
            if B == 64:
                assert qx == b0(q, (a + 1 - 64) // 2), (q, r, a)
                assert rx == b0(r, (a + 1 + 64) // 2), (q, r, a)
        refilled += steps > 128
    assert refilled > 10  # cases that crossed at least two refills
