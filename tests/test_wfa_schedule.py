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


def wfa_tiled(q, r, x, o, e, me, TB, TW):
    """The trapezoid-tiled schedule of the team kernel for pairs whose
    wavefronts exceed LDS: blocks of up to TB score steps; per block the
    widest range is cut into tiles of TW diagonals; a tile copies every slot
    of its diagonals plus TB on each side from slab set `cur` (entries
    outside [-K, K] read as invalid), runs the block's steps on a range that
    narrows by one diagonal per step on each side, returns at the first step
    that reaches the target, and writes its own diagonals of every slot to
    slab set `cur ^ 1`."""
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
    slots = nM + 2 * nI
    slab = [[[NEG] * W for _ in range(slots)] for _ in range(2)]
    def ext(row, k):
        col = row - k
        while row < ql and 0 <= col < rl and q[row] == r[col]:
            row += 1; col += 1
        return row
    slab[0][0][K] = ext(0, 0)
    if target == 0 and slab[0][0][K] >= ql: return 0
    LW = TW + 2 * TB
    cur, s0 = 0, 0
    while s0 + g <= me:
        nsteps = min(TB, (me - s0) // g)
        s_end = s0 + nsteps * g
        Re = R(s_end)
        lo_e, hi_e = max(-rl, -me, -Re), min(ql, me, Re)
        src, dst = slab[cur], slab[cur ^ 1]
        for a in range(lo_e, hi_e + 1, TW):
            b = min(a + TW - 1, hi_e)
            base = a - TB
            loc = [[src[sl][k + K] if -K <= k <= K else NEG for k in range(base, base + LW)]
                   for sl in range(slots)]
            M, I, D = loc[:nM], loc[nM:nM + nI], loc[nM + nI:]
            for t in range(1, nsteps + 1):
                s = s0 + t * g
                si = s // g
                ms, is_ = si % nM, si % nI
                Rs = R(s)
                shrink = nsteps - t
                klo = max(-rl, -me, -Rs, a - shrink)
                khi = min(ql, me, Rs, b + shrink)
                mx = ((s - x) // g) % nM if s >= x else -1
                mo = ((s - o - e) // g) % nM if s >= o + e else -1
                ie = ((s - e) // g) % nI if s >= e else -1
                found = False
                for k in range(klo, khi + 1):
                    di = k - base
                    assert 1 <= di <= LW - 2
                    A = M[mo][di - 1] if mo >= 0 else NEG
                    Bv = M[mo][di + 1] if mo >= 0 else NEG
                    C = I[ie][di - 1] if ie >= 0 else NEG
                    Dv = D[ie][di + 1] if ie >= 0 else NEG
                    m = M[mx][di] if mx >= 0 else NEG
                    iv = max(A, C) + 1
                    dv = max(Bv, Dv)
                    mv = max(m + 1, iv, dv)
                    I[is_][di] = iv if iv >= 0 else NEG
                    D[is_][di] = dv if dv >= 0 else NEG
                    if mv >= 0:
                        mv = ext(mv, k)
                        found |= (k == target and mv >= ql)
                    M[ms][di] = mv if mv >= 0 else NEG
                if found: return s
            for sl in range(slots):
                for k in range(a, b + 1):
                    dst[sl][k + K] = loc[sl][k - base]
        cur ^= 1
        s0 = s_end
    return -1


def test_tiled_schedule_equals_oracle():
    rng = random.Random(2)
    pens = [(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3)]
    for _ in range(2500):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        if rng.random() < 0.5:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        else:
            b = "".join(c if rng.random() > 0.15 else rng.choice("ACGT")
                        for c in a if rng.random() > 0.1)
        x, o, e = rng.choice(pens)
        me = rng.choice([0, 2, 5, 12, 30, 80, 200])
        TB, TW = rng.randint(1, 6), rng.randint(1, 9)
        assert wfa_tiled(a, b, x, o, e, me, TB, TW) == O.wfa_affine(a, b, x, o, e, me), \
            (a, b, x, o, e, me, TB, TW)


def test_incremental_step_state_equals_division_form():
    """The kernels advance slot indices and the reach per score step with no
    division (WfaSteps in csrc/cgm_wfa_dev.h); restated here, they must equal
    si % nM, si % nI, ((s - x) / g) % nM, ((s - o - e) / g) % nM,
    ((s - e) / g) % nI and R(s) at every score."""
    for x, o, e in [(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3),
                    (8, 12, 4), (4, 0, 2), (7, 5, 7)]:
        g = math.gcd(math.gcd(x, o), e)
        nM, nI = max(x, o + e) // g + 1, e // g + 1
        dx, dO, de = x // g, (o + e) // g, e // g
        back = lambda slot, d, n: slot - d if slot >= d else slot - d + n  # noqa: E731
        si = ms = is_ = R = 0
        for s in range(g, 2000, g):
            si += 1
            ms = 0 if ms + 1 == nM else ms + 1
            is_ = 0 if is_ + 1 == nI else is_ + 1
            while (R + 1) * e + o <= si * g:
                R += 1
            assert si == s // g and ms == si % nM and is_ == si % nI
            assert R == ((s - o) // e if s >= o + e else 0)
            assert (back(ms, dx, nM) if si >= dx else -1) == (((s - x) // g) % nM if s >= x else -1)
            assert (back(ms, dO, nM) if si >= dO else -1) == \
                (((s - o - e) // g) % nM if s >= o + e else -1)
            assert (back(is_, de, nI) if si >= de else -1) == (((s - e) // g) % nI if s >= e else -1)


NEG16 = -32768
FAST_SLACK = 512  # bytes of slack after the rows (kFastSlack)


def _bitrev32(v):
    return int(format(v, "032b")[::-1], 2)


def lsb_words(seq):
    """LDS staging of the branch-free team layout (wfa_pair_team_fast): each
    32-bit half of the oracle's packed MSB-first words bit-reversed, high
    half first, then four zero words."""
    w, _ = O.pack_ascii(seq)
    out = []
    for v in w.tolist():
        out += [_bitrev32(v >> 32), _bitrev32(v & 0xFFFFFFFF)]
    return out + [0, 0, 0, 0]


def _alignbit(a, b, sh):
    return (((a << 32) | b) >> (sh & 31)) & 0xFFFFFFFF


def window_lsb(s, pos):
    p, sh = pos >> 4, 2 * (pos & 15)
    assert 0 <= p and p + 2 < len(s), (pos, len(s))
    return _alignbit(s[p + 1], s[p], sh), _alignbit(s[p + 2], s[p + 1], sh)


def _lowest_bit(x):
    """v_ffbl_b32: index of the lowest set bit, 0xFFFFFFFF for 0."""
    return ((x & -x).bit_length() - 1) if x else 0xFFFFFFFF


def match32(a, b):
    """min(ffbl(x0), 32 + min(ffbl(x1), 32)) >> 1, as the kernels count it."""
    t1 = min(_lowest_bit(a[1] ^ b[1]), 32)
    return min(_lowest_bit(a[0] ^ b[0]), 32 + t1) >> 1


def extend_lsb(qs, ql, rs, rl, row, col):
    while row < ql and col < rl:
        n = match32(window_lsb(qs, row), window_lsb(rs, col))
        lim = min(ql - row, rl - col)
        if n >= lim:
            return row + lim
        row += n
        col += n
        if n < 32:
            break
    return row


def wfa_team_fast(q, r, x, o, e, me, U=4, NT=256):
    """wfa_pair_team_fast restated on flat arrays with its own index
    arithmetic: guard entries around every row, I and D of a diagonal side by
    side, ring slots read before their score exists (no existence test),
    wavefront-wide groups of which lanes past `hi` read whatever their
    addresses hold and write nothing, LSB-first 32-base windows with clamped
    positions.  Every read must stay inside the layout, every stored D must
    be valid or exactly the invalid value."""
    ql, rl = len(q), len(r)
    if ql == 0 and rl == 0: return 0
    target = ql - rl
    g = math.gcd(math.gcd(x, o), e)
    R = lambda s: (s - o) // e if s >= o + e else 0
    K = min(me, max(ql, rl), R(me))
    W = 2 * K + 1
    if target != 0 and abs(target) > K: return -1
    nM, nI = max(x, o + e) // g + 1, e // g + 1
    dx, dO, de = x // g, (o + e) // g, e // g
    qs, rs = lsb_words(q), lsb_words(r)
    WP = W + 2
    m_ents = (nM * WP + 1) & ~1
    size = m_ents + 2 * nI * WP + FAST_SLACK // 2
    L = [NEG16] * size
    IDb = m_ents
    L[K + 1] = extend_lsb(qs, ql, rs, rl, 0, 0)
    if target == 0 and L[K + 1] >= ql: return 0
    back = lambda slot, d, n: slot - d if slot >= d else slot - d + n  # noqa: E731
    si = ms = is_ = Rr = 0
    for s in range(g, me + 1, g):
        si += 1
        ms = 0 if ms + 1 == nM else ms + 1
        is_ = 0 if is_ + 1 == nI else is_ + 1
        while (Rr + 1) * e + o <= s:
            Rr += 1
        lo, hi = max(-rl, -me, -Rr), min(ql, me, Rr)
        Mo, Mx = back(ms, dO, nM) * WP, back(ms, dx, nM) * WP
        IDe = IDb + back(is_, de, nI) * 2 * WP
        Ms, IDs = ms * WP, IDb + is_ * 2 * WP
        found = False
        for wave0 in range(0, NT, 64):
            kb = lo + wave0
            while kb <= hi:
                ng = min(U, (hi - kb) // NT + 1)
                mv, mr = {}, {}
                for u in range(ng):
                    for lane in range(64):
                        di = kb + lane + K + u * NT
                        idx = [Mo + di, Mo + di + 2, IDe + 2 * di, IDe + 2 * di + 5, Mx + di + 1]
                        assert all(0 <= i < size for i in idx), (di, size)
                        a, b, c, d, m = (L[i] for i in idx)
                        iv, dv = max(a, c) + 1, max(b, d)
                        mv[u, lane] = max(m + 1, iv, dv)
                        if di - K <= hi:
                            assert dv >= 0 or dv == NEG16
                            L[IDs + 2 * di + 2] = NEG16 if iv < 0 else iv
                            L[IDs + 2 * di + 3] = dv
                for u in range(ng):
                    for lane in range(64):
                        k = kb + lane + u * NT
                        row = mv[u, lane]
                        col = row - k
                        wq = window_lsb(qs, min(max(row, 0), ql))
                        wr = window_lsb(rs, min(max(col, 0), rl))
                        lim = min(ql - row, rl - col)
                        n = min(match32(wq, wr), max(lim, 0))
                        mr[u, lane] = row + n
                        if row >= 0 and n == 32 and lim > 32 and k <= hi:
                            mr[u, lane] = extend_lsb(qs, ql, rs, rl, row + 32, col + 32)
                for u in range(ng):
                    for lane in range(64):
                        k = kb + lane + u * NT
                        if k <= hi:
                            row = mv[u, lane]
                            found |= k == target and row >= 0 and mr[u, lane] >= ql
                            v = NEG16 if row < 0 else mr[u, lane]
                            assert -32768 <= v <= 32767
                            L[Ms + k + K + 1] = v
                kb += U * NT
        if found: return s
    return -1


def test_lsb_windows_equal_base_compare():
    rng = random.Random(5)
    for _ in range(300):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 90)))
        b = "".join(c if rng.random() > 0.05 else rng.choice("ACGT") for c in a)
        qs, rs = lsb_words(a), lsb_words(b)
        for i in range(len(a)):
            j = rng.randint(0, len(b) - 1)
            n = match32(window_lsb(qs, i), window_lsb(rs, j))
            t = 0
            while t < 32 and i + t < len(a) and j + t < len(b) and a[i + t] == b[j + t]:
                t += 1
            lim = min(len(a) - i, len(b) - j, 32)
            # bases past an end are zero bits (code of A): only the capped
            # count has to agree
            assert min(n, lim) == t, (a, b, i, j)


def test_fast_team_layout_equals_oracle():
    rng = random.Random(6)
    pens = [(4, 6, 2), (1, 1, 1), (3, 2, 1), (5, 0, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3)]
    for it in range(500):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 120)))
        if rng.random() < 0.3:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 60)))
        else:
            div = rng.choice([0.0, 0.01, 0.05, 0.15])
            b = "".join(c if rng.random() > div else rng.choice("ACGT")
                        for c in a if rng.random() > div / 2)
        x, o, e = rng.choice(pens)
        me = rng.choice([0, 2, 5, 12, 30, 80, 200])
        U, NT = (4, 256) if it % 2 else (2, 64)
        assert wfa_team_fast(a, b, x, o, e, me, U, NT) == O.wfa_affine(a, b, x, o, e, me), \
            (a, b, x, o, e, me, U, NT)


def wfa_cigar_sched(q, r, x, o, e, me):
    """The traceback schedule of the CIGAR kernel: the cost first (any exact
    schedule), then a history of every score step 0..cost/g, each row the
    diagonals -K..K with K = min(me, max(qlen, rlen), R(cost)) and only
    |k| <= R(s) computed, then the reference's backtrace (align_cpu.mojo:
    510-600) with M before extension recomputed from the stored rows.
    Returns (cost, CIGAR string)."""
    ql, rl = len(q), len(r)
    cost = wfa_sched(q, r, x, o, e, me)
    if cost < 0 or (ql == 0 and rl == 0):
        return cost, ""
    g = math.gcd(math.gcd(x, o), e)
    R = lambda s: (s - o) // e if s >= o + e else 0
    K = min(me, max(ql, rl), R(cost))
    W = 2 * K + 1
    rows = cost // g + 1
    M = [[NEG] * W for _ in range(rows)]
    I = [[NEG] * W for _ in range(rows)]
    D = [[NEG] * W for _ in range(rows)]
    def ext(row, k):
        col = row - k
        while row < ql and 0 <= col < rl and q[row] == r[col]:
            row += 1; col += 1
        return row
    M[0][K] = ext(0, 0)
    dx, dO, de = x // g, (o + e) // g, e // g
    for si in range(1, rows):
        s = si * g
        lo, hi = max(-rl, -me, -R(s)), min(ql, me, R(s))
        for k in range(max(lo, -K), min(hi, K) + 1):
            di = k + K
            a = M[si - dO][di - 1] if si >= dO and di > 0 else NEG
            b = M[si - dO][di + 1] if si >= dO and di < W - 1 else NEG
            c = I[si - de][di - 1] if si >= de and di > 0 else NEG
            d = D[si - de][di + 1] if si >= de and di < W - 1 else NEG
            m = M[si - dx][di] if si >= dx else NEG
            iv = max(a + 1 if a >= 0 else NEG, c + 1 if c >= 0 else NEG)
            dv = max(b, d)
            mv = max(m + 1 if m >= 0 else NEG, iv, dv)
            I[si][di], D[si][di] = iv, dv
            M[si][di] = ext(mv, k) if mv >= 0 else NEG
    get = lambda A, si, k: A[si][k + K] if -K <= k <= K else NEG
    def m_pre(si, k):
        if si == 0:
            return 0 if k == 0 else NEG
        v = NEG
        if si >= dx and get(M, si - dx, k) >= 0:
            v = max(v, get(M, si - dx, k) + 1)
        for A in (I, D):
            if get(A, si, k) >= 0:
                v = max(v, get(A, si, k))
        return v
    ops = []
    si, d, mat = cost // g, ql - rl, 0
    row = get(M, si, d)
    while True:
        if mat == 0:
            pre = m_pre(si, d)
            ops += ["="] * (row - pre)
            row = pre
            if si == 0 and d == 0:
                break
            pm = get(M, si - dx, d) if si >= dx else NEG
            if pm >= 0 and pm + 1 == row:
                ops.append("X"); si -= dx; row = pm; continue
            if get(I, si, d) >= 0 and get(I, si, d) == row:
                mat = 1; continue
            if get(D, si, d) >= 0 and get(D, si, d) == row:
                mat = 2; continue
            break
        elif mat == 1:
            ops.append("I")
            mv = get(M, si - dO, d - 1) if si >= dO else NEG
            if mv >= 0 and mv + 1 == row:
                si -= dO; d -= 1; row = mv; mat = 0; continue
            iv = get(I, si - de, d - 1) if si >= de else NEG
            if iv >= 0 and iv + 1 == row:
                si -= de; d -= 1; row = iv; continue
            break
        else:
            ops.append("D")
            mv = get(M, si - dO, d + 1) if si >= dO else NEG
            if mv >= 0 and mv == row:
                si -= dO; d += 1; row = mv; mat = 0; continue
            dv = get(D, si - de, d + 1) if si >= de else NEG
            if dv >= 0 and dv == row:
                si -= de; d += 1; row = dv; continue
            break
    out, i = "", len(ops) - 1
    while i >= 0:
        n = 1
        while i - n >= 0 and ops[i - n] == ops[i]:
            n += 1
        out += f"{n}{ops[i]}"
        i -= n
    return cost, out


def test_cigar_schedule_equals_oracle():
    rng = random.Random(8)
    pens = [(4, 6, 2), (1, 0, 1), (3, 2, 1), (5, 1, 3), (2, 7, 1), (6, 4, 2), (9, 3, 3)]
    for _ in range(2500):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        if rng.random() < 0.4:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        else:
            b = "".join(c if rng.random() > 0.15 else rng.choice("ACGT")
                        for c in a if rng.random() > 0.1)
        x, o, e = rng.choice(pens)
        me = rng.choice([0, 3, 8, 20, 60, 150, 400])
        assert wfa_cigar_sched(a, b, x, o, e, me) == O.wfa_affine_cigar(a, b, x, o, e, me), \
            (a, b, x, o, e, me)
