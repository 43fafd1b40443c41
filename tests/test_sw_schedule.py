"""CPU model of the GPU register-lane SW schedule, checked against the oracle.

The register lane kernels (csrc/cgm_sw_dev.h) sweep the full band frame of
every row without per-cell bounds tests: out-of-matrix cells only get a
forced mismatch; the row is kept as G = H + open + extend; the row maximum
uses the key (h << 8) | (64 - j mod 64) per 64-cell group and rows compare
with strict `>`.  With mismatch <= 0, extend <= 0 and open + extend <= 0 this
restatement must equal smith_waterman_banded (align_cpu.mojo:229-309).
"""
import random

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
