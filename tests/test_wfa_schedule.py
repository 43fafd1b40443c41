"""CPU model of the GPU WFA wave schedule, checked against the oracle.

The wave kernel (csrc/cgm_wfa.hip) skips scores that are not multiples of
g = gcd(x, o, e), visits only diagonals |k| <= (s - o) / e, fuses NEXT with
EXTEND per diagonal and never clears score slots.  This restatement of that
schedule in Python must equal wfa_affine_cpu (align_cpu.mojo:75-226) on every
input: it is what makes those reductions exact.
"""
import math
import random

from conftest import load_oracle

O = load_oracle()

NEG = -(1 << 30)

def wfa_sched(q, r, x, o, e, me):
    ql, rl = len(q), len(r)
    if ql == 0 and rl == 0: return 0
    target = ql - rl
    g = math.gcd(math.gcd(x, o), e)
    R = lambda s: (s - o) // e if s >= o + e else 0
    if abs(target) > min(me, R(me)) and target != 0: return -1
    K = min(me, max(ql, rl), R(me))
    W = 2 * K + 1
    nM = max(x, o + e) // g + 1
    nI = e // g + 1
    M = [[NEG] * W for _ in range(nM)]
    I = [[NEG] * W for _ in range(nI)]
    D = [[NEG] * W for _ in range(nI)]
    def ext(row, k):
        col = row - k
        while row < ql and 0 <= col < rl and q[row] == r[col]:
            row += 1; col += 1
        return row
    M[0][K] = ext(0, 0)
    if target == 0 and M[0][K] >= ql: return 0
    s = g
    while s <= me:
        si = s // g
        ms, is_ = si % nM, si % nI
        Rs = R(s)
        lo = max(-rl, -me, -Rs); hi = min(ql, me, Rs)
        mx = ((s - x) // g) % nM if s >= x else -1
        mo = ((s - o - e) // g) % nM if s >= o + e else -1
        ie = ((s - e) // g) % nI if s >= e else -1
        found = False
        for k in range(lo, hi + 1):
            di = k + K
            iv = dv = mv = NEG
            if mo >= 0:
                if di > 0 and M[mo][di - 1] >= 0: iv = max(iv, M[mo][di - 1] + 1)
                if di < W - 1 and M[mo][di + 1] >= 0: dv = max(dv, M[mo][di + 1])
            if ie >= 0:
                if di > 0 and I[ie][di - 1] >= 0: iv = max(iv, I[ie][di - 1] + 1)
                if di < W - 1 and D[ie][di + 1] >= 0: dv = max(dv, D[ie][di + 1])
            if mx >= 0 and M[mx][di] >= 0: mv = M[mx][di] + 1
            mv = max(mv, iv, dv)
            I[is_][di] = iv; D[is_][di] = dv
            if mv >= 0: mv = ext(mv, k)
            M[ms][di] = mv
            if k == target and mv >= ql: found = True
        if found: return s
        s += g
    return -1



def test_wave_schedule_equals_oracle():
    rng = random.Random(1)
    pens = [(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3)]
    for _ in range(4000):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 25)))
        if rng.random() < 0.5:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 25)))
        else:
            b = "".join(c if rng.random() > 0.15 else rng.choice("ACGT")
                        for c in a if rng.random() > 0.1)
        x, o, e = rng.choice(pens)
        me = rng.choice([0, 2, 5, 12, 30, 80, 200])
        assert wfa_sched(a, b, x, o, e, me) == O.wfa_affine(a, b, x, o, e, me), (a, b, x, o, e, me)
