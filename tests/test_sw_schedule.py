"""CPU model of the GPU register-lane SW schedule, checked against the oracle.

The register lane kernels (csrc/cgm_sw_dev.h) sweep the full band frame of
every row without per-cell bounds tests: out-of-matrix cells only get a
forced mismatch; the row is kept as G = H + open + extend; the row maximum
uses the key (h << 8) | (64 - j mod 64) per 64-cell group and rows compare
with strict `>`.  With mismatch <= 0, extend <= 0 and open + extend <= 0 this
restatement must equal smith_waterman_banded (align_cpu.mojo:229-309).
"""
import random

import numpy as np

from conftest import load_oracle

O = load_oracle()
NEG = -1000000


def sw_register_schedule(q, r, band, match, mismatch, gap_open, gap_extend):
    ql, rl = len(q), len(r)
    b = min(band, max(ql, rl))
    W = 2 * b + 1
    oe = gap_open + gap_extend
    G = [oe] * W
    best, bq, br = 0, 0, 0
    for qi in range(ql):
        Gl, El = oe, NEG
        groups = {}
        for j in range(W):
            ri = qi - b + j
            eq = 0 <= ri < rl and q[qi] == r[ri]
            Gu = G[j + 1] if j + 1 < W else oe
            dterm = G[j] - oe + (match if eq else mismatch)
            e = max(El + gap_extend, Gu)
            h = max(0, dterm, e, Gl)
            key = (h << 8) | (64 - (j & 63))
            groups[j >> 6] = max(groups.get(j >> 6, -(1 << 62)), key)
            G[j] = h + oe
            Gl, El = G[j], e
        hrow, jrow = -(1 << 62), 0
        for g in sorted(groups):
            if (groups[g] >> 8) > hrow:
                hrow, jrow = groups[g] >> 8, 64 * g + 64 - (groups[g] & 255)
        if hrow > best:
            best, bq, br = hrow, qi, qi - b + jrow
    return best, bq, br


def test_register_schedule_equals_oracle():
    rng = random.Random(3)
    scorings = [(2, -3, -5, -2), (2, -4, -4, -2), (1, -1, -1, -1), (3, 0, -2, 0), (2, -1, 0, -1),
                (5, -4, 0, 0)]
    for _ in range(6000):
        q = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        if rng.random() < 0.5:
            r = "".join(c if rng.random() > 0.1 else rng.choice("ACGT") for c in q
                        if rng.random() > 0.05)
        else:
            r = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        band = rng.randint(0, 70)
        sc = rng.choice(scorings)
        assert sw_register_schedule(q, r, band, *sc) == O.sw_banded(q, r, *sc, band=band), \
            (q, r, band, sc)


def _w16(x):
    """int16 wrap-around (v_pk_add_u16 on two's-complement halves)."""
    x &= 0xFFFF
    return x - 0x10000 if x & 0x8000 else x


def _f16(x):
    """IEEE half precision rounding (round to nearest even)."""
    return float(np.float16(x))


def sw_dual_schedule(q, r, band, match, mismatch, gap_open, gap_extend, extra_rows=0,
                     enc="int16"):
    """The dual-row lane (csrc/cgm_sw_dev.h sw_dual_lane): halves [0, S) and
    [S, 2S) of the band frame, S = band + 1, the right half one row behind,
    every value in 16-bit lanes, cell 2S - 1 (past the frame) forced to
    H = 0, each half's first best row kept and searched for its first best
    cell.  enc "int16": two's-complement halves (v_pk_add_u16 wraps);
    "f16": IEEE half precision, every constant and every sum rounded to
    half (v_pk_add_f16).  `extra_rows` adds rows past the query, as a lane
    whose wavefront runs longer queries does."""
    ql, rl = len(q), len(r)
    B = min(band, max(ql, rl))
    S = B + 1
    if enc == "int16":
        cst = _w16

        def add(a, b):
            return _w16(a + b)
    else:
        cst = _f16

        def add(a, b):
            # the sum of two halves is exact in double; one rounding to half
            return _f16(a + b)
    oe = gap_open + gap_extend
    OE, EXT = cst(oe), cst(gap_extend)
    SUBM, SUBX = cst(match - oe), cst(mismatch - oe)
    NEG16 = cst(-16384)
    P = [[OE, OE] for _ in range(S)]  # [low half, high half]
    SV = [[OE, OE] for _ in range(S)]
    best = [0, 0]
    row = [0, 0]
    el_carry = NEG16
    p0_hi = OE
    rows = ql + extra_rows
    for t in range(rows + 1):
        rs = (t - B, t)
        ok = (t < ql, 1 <= t and t - 1 < ql)
        qb = (q[t] if t < ql else None, q[t - 1] if 1 <= t <= ql else None)
        Gl = [OE, P[S - 1][0]]
        El = [NEG16, el_carry]
        rowmax = [0, 0]
        for u in range(S):
            newP = [0, 0]
            for k in (0, 1):
                ri = rs[k] + u
                eq = ok[k] and 0 <= ri < rl and r[ri] == qb[k]
                if u + 1 < S:
                    Gu = P[u + 1][k]
                else:
                    Gu = p0_hi if k == 0 else OE
                dterm = add(P[u][k], SUBM if eq else SUBX)
                e = max(add(El[k], EXT), Gu)
                h = max(max(dterm, 0), max(e, Gl[k]))
                if u == S - 1 and k == 1:
                    h = 0
                rowmax[k] = max(rowmax[k], h)
                newP[k] = add(h, OE)
                Gl[k] = newP[k]
                El[k] = e
            P[u] = newP
            if u == 0:
                p0_hi = newP[1]
        el_carry = El[0]
        for k in (0, 1):
            if rowmax[k] > best[k]:
                best[k] = int(rowmax[k])
                row[k] = t - k
                for u in range(S):
                    SV[u][k] = P[u][k]
    cand = []
    for k in (0, 1):
        if best[k] == 0:
            cand.append((0, 0, 0))
            continue
        u = next(u for u in range(S) if int(SV[u][k]) - oe == best[k])
        cand.append((best[k], row[k], row[k] - B + k * S + u))
    lo, hi = cand
    if hi[0] > lo[0] or (hi[0] == lo[0] and hi[0] > 0 and hi[1] < lo[1]):
        return hi
    return lo


def test_dual_schedule_equals_oracle():
    rng = random.Random(5)
    scorings = [(2, -3, -5, -2), (2, -4, -4, -2), (1, -1, -1, -1), (3, 0, -2, 0), (2, -1, 0, -1),
                (5, -4, 0, 0), (1, 0, 0, 0)]
    for _ in range(5000):
        q = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        if rng.random() < 0.5:
            r = "".join(c if rng.random() > 0.1 else rng.choice("ACGT") for c in q
                        if rng.random() > 0.05)
        else:
            r = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 40)))
        band = rng.randint(0, 70)
        sc = rng.choice(scorings)
        extra = rng.choice([0, 0, 1, 7])
        want = O.sw_banded(q, r, *sc, band=band)
        for enc in ("int16", "f16"):
            got = sw_dual_schedule(q, r, band, *sc, extra_rows=extra, enc=enc)
            assert got == want, (q, r, band, sc, enc, got, want)


def _f16_unit(match, mismatch, gap_open, gap_extend):
    """Largest power of two <= 8 dividing every score (sw_dual_lane's t)."""
    m = abs(match) | abs(mismatch) | abs(gap_open) | abs(gap_extend)
    return min(m & -m, 8) if m else 8


def _mutate(rng, x, subs, indels):
    y = list(x)
    for _ in range(subs):
        p = rng.randrange(len(y))
        y[p] = rng.choice("ACGT".replace(y[p], ""))
    for _ in range(indels):
        p = rng.randrange(len(y))
        if rng.random() < 0.5:
            del y[p]
        else:
            y.insert(p, rng.choice("ACGT"))
    return "".join(y)


def test_dual_f16_exact_up_to_the_range_limit():
    """Half-precision lanes at the longest pairs they admit: match *
    min(qlen, rlen) + match = 2048 t (t = 1 and t = 2), identical and
    mutated, equal the oracle, so every value of the recurrence is a
    representable multiple of t there."""
    rng = random.Random(11)
    for sc in ((3, -2, -5, -1), (2, -4, -4, -2), (1, -1, -1, -1)):
        t = _f16_unit(*sc)
        L = (2048 * t - sc[0]) // sc[0]
        assert sc[0] * L + sc[0] <= 2048 * t < sc[0] * (L + 1) + sc[0]
        a = "".join(rng.choice("ACGT") for _ in range(L))
        for q, r, band in ((a, a, 2), (_mutate(rng, a, 20, 4), a, 5), (a, _mutate(rng, a, 9, 2), 3)):
            want = O.sw_banded(q, r, *sc, band=band)
            assert sw_dual_schedule(q, r, band, *sc, enc="f16") == want, (sc, band)
        assert O.sw_banded(a, a, *sc, band=2)[0] == sc[0] * L


def test_dual_f16_is_wrong_past_the_range_limit():
    """The bound is needed: with t = 1 an identical pair one base past the
    limit reaches an odd score above 2048, which half precision cannot hold,
    and the half-precision schedule then differs from the oracle (the GPU
    runs such a pass on the int16 lanes)."""
    sc = (3, -2, -5, -1)
    L = 683  # 3 * 683 = 2049
    a = "".join(random.Random(3).choice("ACGT") for _ in range(L))
    want = O.sw_banded(a, a, *sc, band=1)
    assert want[0] == 2049
    assert sw_dual_schedule(a, a, 1, *sc, enc="f16") != want
    assert sw_dual_schedule(a, a, 1, *sc, enc="int16") == want
