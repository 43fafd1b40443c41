"""CPU model of the register-resident WFA layout, checked against the oracle.

The layout (cgm_wfa_reg.h in the package's csrc/) gives every lane one
diagonal per block: block gb holds the 64 diagonals kbase + 64 gb ..
kbase + 64 gb + 63 of the pair's alive span [kbase, kmax], and keeps each
diagonal's score history in registers as rings of NI = e / g + 1 I and D
slots and NM M slots (max(x, o + e) / g + 1 rounded up to a multiple of NI),
read without existence tests (a slot not yet written holds the invalid
value).  Neighbour diagonals k - 1 and k + 1 come from the adjacent lanes
(wave shifts), across block edges from the adjacent block's edge values of
the step, which every block computes from its older slots and publishes
before the step.  Cells that max_error rules out (e |k - t| > max_error - s,
t = qlen - rlen) are held invalid besides the reference's own range
(|k| <= R(s), -rlen <= k <= qlen): per score the alive diagonals form one
range, lanes outside it are invalid, and a block wholly outside it skips the
step after it has written the invalid value for NM scores.  This
restatement of that data flow must equal wfa_affine_cpu
(align_cpu.mojo:75-226) on every input, and copies with a planted slip must
not.  span_bisect restates the kernel's bisection for the alive span
(reg_span_side) and must equal the scan over all scores.
"""
import math
import random

import pytest

from conftest import load_oracle

O = load_oracle()

NEG = -(1 << 30)
LANES = 64


def span(ql, rl, x, o, e, me):
    """Diagonals any score can hold alive: k in [lo(s), hi(s)] with
    lo(s) = max(-rlen, -R(s), t - P(s)), hi(s) = min(qlen, R(s), t + P(s)),
    R(s) the reach, P(s) = (me - s) // e the gap budget left, t = qlen - rlen;
    (lo, hi) over all scores s = 0, g, ... <= me."""
    g = math.gcd(math.gcd(x, o), e)
    t = ql - rl

    def R(s):
        return (s - o) // e if s >= o + e else 0
    lo, hi = 0, 0
    for s in range(0, me + 1, g):
        P = (me - s) // e
        lo = min(lo, max(-rl, -R(s), t - P))
        hi = max(hi, min(ql, R(s), t + P))
    return lo, hi


def span_bisect(ql, rl, x, o, e, me):
    """span() as the kernel computes it: G(u) = max over s of
    min(R(s), u + P(s)) at the crossing of the rising R and the falling
    u + P, found by bisection on s / g; hi = min(qlen, G(t)),
    lo = max(-rlen, -G(-t)).  Defined for e |t| <= me."""
    g = math.gcd(math.gcd(x, o), e)
    t = ql - rl

    def R(s):
        return (s - o) // e if s >= o + e else 0

    def side(u):
        def H(s):
            return u + (me - s) // e
        a, b = 0, me // g
        while a < b:
            m = (a + b + 1) >> 1
            if R(m * g) <= H(m * g):
                a = m
            else:
                b = m - 1
        v = R(a * g)
        return max(v, H((a + 1) * g)) if a < me // g else v
    return max(-rl, -side(-t)), min(ql, side(t))


def wfa_reg_model(q, r, x, o, e, me, NB, NW, lanes=LANES, prune=True, lag=None):
    """Cost or -1 as the register layout computes it, or None when the
    pair's diagonals do not fit NB x NW blocks.  prune: cells whose
    remaining gap budget cannot reach the target diagonal (|k - t| >
    (me - s) // e) are held invalid.  lag: scores a block that left the
    alive range keeps running with every lane invalid (NM by default)."""
    ql, rl = len(q), len(r)
    if ql == 0 and rl == 0:
        return 0
    g = math.gcd(math.gcd(x, o), e)
    DX, DO, DE = x // g, (o + e) // g, e // g
    NI = DE + 1
    NM = -(-(max(DX, DO) + 1) // NI) * NI   # rounded up to a multiple of NI

    def R(s):
        return (s - o) // e if s >= o + e else 0
    target = ql - rl
    if prune and abs(target) * e > me:
        return -1  # the length difference alone costs more than me
    if prune:
        kbase, kmax = span(ql, rl, x, o, e, me)
    else:
        K = min(me, max(ql, rl), R(me))
        kbase, kmax = -min(rl, K), min(ql, K)
    nblk = (kmax - kbase + 1 + lanes - 1) // lanes
    if nblk > NB * NW:
        return None

    def ext(row, k):
        col = row - k
        while row < ql and 0 <= col < rl and q[row] == r[col]:
            row += 1
            col += 1
        return row

    # registers: per global block, per lane, ring slots
    M = [[[NEG] * NM for _ in range(lanes)] for _ in range(nblk)]
    I = [[[NEG] * NI for _ in range(lanes)] for _ in range(nblk)]
    D = [[[NEG] * NI for _ in range(lanes)] for _ in range(nblk)]
    b0, l0 = divmod(0 - kbase, lanes)
    M[b0][l0][0] = ext(0, 0)
    if target == 0 and M[b0][l0][0] >= ql:
        return 0
    # edge buffers (double, by step parity): ER = max(M_o, I_e) of lane 63,
    # EL = max(M_o, D_e) of lane 0, for the step they are published for
    ER = [[NEG] * nblk for _ in range(2)]
    EL = [[NEG] * nblk for _ in range(2)]
    # a block runs at score s when it may hold a cell alive at one of the
    # scores s - NM g .. s (a superset: below max(-rlen, -R(s),
    # t - P(s - NM g)) and above min(qlen, R(s), t + P(s - NM g)) no such
    # cell exists), every lane outside the alive range of s held invalid; so
    # a block that leaves the range still writes the invalid value for NM
    # scores, and its published edges come from its older slots until then
    lag = NM if lag is None else lag

    def alive_range(s, Rs):
        lo, hi = max(-rl, -Rs), min(ql, Rs)
        if prune:
            P = (me - s) // e
            lo, hi = max(lo, target - P), min(hi, target + P)
        return lo, hi

    def run_blocks(s, Rs):
        if not prune:
            lo, hi = alive_range(s, Rs)
        else:
            Pl = (me - max(s - lag * g, 0)) // e
            lo, hi = max(-rl, -Rs, target - Pl), min(ql, Rs, target + Pl)
        return (lo - kbase) // lanes, min((hi - kbase) // lanes, nblk - 1)

    def publish(gb, si_next):
        # a block's edge values for the next step, from its slots of the
        # scores that step reads (written before the step, so after this one)
        mo, ie = (si_next - DO) % NM, (si_next - DE) % NI
        ER[si_next & 1][gb] = max(M[gb][lanes - 1][mo], I[gb][lanes - 1][ie])
        EL[si_next & 1][gb] = max(M[gb][0][mo], D[gb][0][ie])

    publish(b0, 1)   # after score 0: the block holding diagonal 0
    Rs = 0
    si = 0
    while (si + 1) * g <= me:
        si += 1
        s = si * g
        while (Rs + 1) * e + o <= s:
            Rs += 1
        ms, is_ = si % NM, si % NI
        mx, mo, ie = (si - DX) % NM, (si - DO) % NM, (si - DE) % NI
        found = False
        buf = si & 1
        lo_s, hi_s = alive_range(s, Rs)
        blo, bhi = run_blocks(s, Rs)
        for gb in range(nblk):
            if gb < blo or gb > bhi:
                continue
            ti = [max(M[gb][l][mo], I[gb][l][ie]) for l in range(lanes)]
            td = [max(M[gb][l][mo], D[gb][l][ie]) for l in range(lanes)]
            ti_in = ER[buf][gb - 1] if gb > 0 else NEG
            td_in = EL[buf][gb + 1] if gb + 1 < nblk else NEG
            for l in range(lanes):
                k = kbase + lanes * gb + l
                iv = (ti[l - 1] if l > 0 else ti_in) + 1
                dv = td[l + 1] if l < lanes - 1 else td_in
                mv = max(M[gb][l][mx] + 1, iv, dv)
                if k < lo_s or k > hi_s:  # held invalid
                    iv = dv = mv = NEG
                mr = ext(mv, k) if mv >= 0 else NEG
                I[gb][l][is_] = iv
                D[gb][l][is_] = dv
                M[gb][l][ms] = mr
                if k == target and mr >= ql:
                    found = True
            publish(gb, si + 1)
        if found:
            return s
    return -1


def _pairs(rng, n, maxlen):
    for _ in range(n):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, maxlen)))
        if rng.random() < 0.5:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, maxlen)))
        else:
            b = "".join(c if rng.random() > 0.15 else rng.choice("ACGT")
                        for c in a if rng.random() > 0.1)
            if rng.random() < 0.3:
                b = b + "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 12)))
        yield a, b


PENS = [(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3), (8, 12, 4)]


def test_register_layout_equals_oracle_small_blocks():
    """Blocks of 4 lanes, so that pairs of a few dozen bases span many
    blocks, edges and the last block's cap."""
    rng = random.Random(7)
    n = 0
    for a, b in _pairs(rng, 2500, 30):
        x, o, e = rng.choice(PENS)
        me = rng.choice([0, 1, 5, 20, 60, 200])
        got = wfa_reg_model(a, b, x, o, e, me, NB=8, NW=4, lanes=4)
        if got is None:
            continue
        n += 1
        assert got == O.wfa_affine(a, b, x=x, o=o, e=e, max_error=me), (a, b, x, o, e, me)
    assert n > 2000


def test_register_layout_equals_oracle_full_lanes():
    rng = random.Random(8)
    for a, b in _pairs(rng, 300, 150):
        x, o, e = rng.choice(PENS)
        me = rng.choice([30, 100, 400])
        got = wfa_reg_model(a, b, x, o, e, me, NB=2, NW=2)
        if got is not None:
            assert got == O.wfa_affine(a, b, x=x, o=o, e=e, max_error=me), (a, b, x, o, e, me)


def test_pruned_layout_with_tight_max_error_equals_oracle():
    """max_error at the cost, just above and just below it (the pruning
    then removes most cells), over several penalty sets."""
    rng = random.Random(12)
    n = 0
    for a, b in _pairs(rng, 700, 40):
        x, o, e = rng.choice(PENS)
        c = O.wfa_affine(a, b, x=x, o=o, e=e, max_error=400)
        if c < 0:
            continue
        for me in (max(c - 1, 0), c, c + 1, c + 7):
            got = wfa_reg_model(a, b, x, o, e, me, NB=8, NW=4, lanes=4)
            assert got is not None
            assert got == O.wfa_affine(a, b, x=x, o=o, e=e, max_error=me), (a, b, x, o, e, me)
            n += 1
    assert n > 1500


def test_span_bisection_equals_scan():
    rng = random.Random(21)
    n = 0
    for _ in range(20000):
        x, o, e = rng.choice(PENS)
        ql, rl = rng.randint(0, 3000), rng.randint(0, 3000)
        me = rng.choice([rng.randint(0, 40), rng.randint(0, 4000)])
        if abs(ql - rl) * e > me:
            continue
        n += 1
        assert span_bisect(ql, rl, x, o, e, me) == span(ql, rl, x, o, e, me), (ql, rl, x, o, e, me)
    assert n > 4000


def test_pruning_shrinks_the_span():
    # a 2 kb pair at max_error 600 (x=4, o=6, e=2): reach alone allows
    # diagonals -297..297, the gap budget keeps about half
    lo, hi = span(2000, 2000, 4, 6, 2, 600)
    assert -160 <= lo <= -140 and 140 <= hi <= 160


@pytest.mark.parametrize("mistake", ["edge_shift", "ring_back", "ghost_zero", "late_skip",
                                     "over_prune", "no_flush", "short_flush"])
def test_planted_layout_mistakes_fail(mistake):
    """Copies of the model with one slip each must differ from the oracle."""
    rng = random.Random(9)
    pairs = list(_pairs(rng, 600, 30))

    def broken(q, r, x, o, e, me):
        g = math.gcd(math.gcd(x, o), e)
        if mistake == "ring_back":      # read M[s - x] one slot too far back
            x2 = x + g
            return wfa_reg_model(q, r, x2, o, e, me, 8, 4, lanes=4) if x2 else None
        if mistake == "edge_shift":     # neighbour across a block edge from the wrong side
            return _model_variant(q, r, x, o, e, me, swap_edges=True)
        if mistake == "ghost_zero":     # the missing neighbour of the outer blocks read as row 0
            return _model_variant(q, r, x, o, e, me, ghost=0)
        if mistake == "no_flush":       # a block leaving the alive range skips at once
            return wfa_reg_model(q, r, x, o, e, me, 8, 4, lanes=4, lag=0)
        if mistake == "short_flush":    # ... or after one more score
            return wfa_reg_model(q, r, x, o, e, me, 8, 4, lanes=4, lag=1)
        if mistake == "over_prune":     # the gap budget counted with one extension too few
            return wfa_reg_model(q, r, x, o, e, max(me - e, 0), 8, 4, lanes=4) if me >= e else None
        return _model_variant(q, r, x, o, e, me, skip_margin=1)  # late_skip
    bad = 0
    for a, b in pairs:
        x, o, e = rng.choice(PENS)
        me = rng.choice([5, 20, 60])
        if rng.random() < 0.5:   # max_error at the cost: the pruning is tight
            me = max(0, O.wfa_affine(a, b, x=x, o=o, e=e, max_error=400)) + rng.choice([0, 1])
        got = broken(a, b, x, o, e, me)
        if got is not None and got != O.wfa_affine(a, b, x=x, o=o, e=e, max_error=me):
            bad += 1
    assert bad > 0, mistake


def _model_variant(q, r, x, o, e, me, swap_edges=False, ghost=NEG, skip_margin=0, lanes=4,
                   NBW=32):
    """wfa_reg_model with a planted slip (used only by the test above)."""
    ql, rl = len(q), len(r)
    if ql == 0 and rl == 0:
        return 0
    g = math.gcd(math.gcd(x, o), e)
    DX, DO, DE = x // g, (o + e) // g, e // g
    NM, NI = max(DX, DO) + 1, DE + 1

    def R(s):
        return (s - o) // e if s >= o + e else 0
    K = min(me, max(ql, rl), R(me))
    kbase, kmax = -min(rl, K), min(ql, K)
    nblk = (kmax - kbase + 1 + lanes - 1) // lanes
    if nblk > NBW:
        return None
    target = ql - rl

    def ext(row, k):
        col = row - k
        while row < ql and 0 <= col < rl and q[row] == r[col]:
            row += 1
            col += 1
        return row
    M = [[[NEG] * NM for _ in range(lanes)] for _ in range(nblk)]
    I = [[[NEG] * NI for _ in range(lanes)] for _ in range(nblk)]
    D = [[[NEG] * NI for _ in range(lanes)] for _ in range(nblk)]
    b0, l0 = divmod(-kbase, lanes)
    M[b0][l0][0] = ext(0, 0)
    if target == 0 and M[b0][l0][0] >= ql:
        return 0

    def active(gb, Rs):
        k0 = kbase + lanes * gb
        return k0 <= Rs - skip_margin and k0 + lanes - 1 >= -Rs + skip_margin
    Rs, si = 0, 0
    while (si + 1) * g <= me:
        si += 1
        s = si * g
        while (Rs + 1) * e + o <= s:
            Rs += 1
        ms, is_ = si % NM, si % NI
        mx, mo, ie = (si - DX) % NM, (si - DO) % NM, (si - DE) % NI
        ti_all = [[max(M[gb][l][mo], I[gb][l][ie]) for l in range(lanes)] for gb in range(nblk)]
        td_all = [[max(M[gb][l][mo], D[gb][l][ie]) for l in range(lanes)] for gb in range(nblk)]
        found = False
        for gb in range(nblk):
            if not active(gb, Rs):
                continue
            ti, td = ti_all[gb], td_all[gb]
            if swap_edges:
                ti_in = td_all[gb - 1][0] if gb > 0 else NEG
                td_in = ti_all[gb + 1][lanes - 1] if gb + 1 < nblk else NEG
            else:
                ti_in = ti_all[gb - 1][lanes - 1] if gb > 0 else ghost
                td_in = td_all[gb + 1][0] if gb + 1 < nblk else ghost
            for l in range(lanes):
                k = kbase + lanes * gb + l
                iv = (ti[l - 1] if l > 0 else ti_in) + 1
                dv = td[l + 1] if l < lanes - 1 else td_in
                mv = max(M[gb][l][mx] + 1, iv, dv)
                mr = ext(mv, k) if mv >= 0 else mv
                if k > kmax:
                    iv, dv, mr = min(iv, NEG), min(dv, NEG), min(mr, NEG)
                I[gb][l][is_] = iv
                D[gb][l][is_] = dv
                M[gb][l][ms] = mr
                if k == target and mr >= ql:
                    found = True
        if found:
            return s
    return -1
