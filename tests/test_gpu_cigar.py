"""GPU parity of gap-affine WFA with traceback (cgm_wfa_cigar_*) against the
oracle's restatement of wfa_affine_cigar_cpu (align_cpu.mojo:380-623):
identical costs and CIGAR strings on known answers, random pairs and a
sample of the 1-10 kb workload; on the whole 20,000-pair workload every
CIGAR has the pair's cost and walks both sequences exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PEN = dict(x=4, o=6, e=2)


def packed_batch(orc, seqs):
    words, offs = [], []
    nw = 0
    for s in seqs:
        w, _ = orc.pack_ascii(s)
        offs.append(nw)
        nw += w.size
        words.append(w)
    packed = np.concatenate(words) if words else np.zeros(0, np.uint64)
    return packed, np.array(offs, np.int32), np.array([len(s) for s in seqs], np.int32)


def gpu_cigars(cgm, orc, qs, rs, max_error, pen=PEN):
    qp, qo, ql = packed_batch(orc, qs)
    rp, ro, rl = packed_batch(orc, rs)
    cfg = cgm.WFAConfig(pen["x"], pen["o"], pen["e"], int(np.max(max_error)))
    return cgm.wfa_affine_cigar_batch(qp, qo, ql, rp, ro, rl, cfg,
                                      max_error=np.asarray(max_error, np.int32))


def test_cigar_reference_known_answers(cgm, dev):
    """test_alignment.mojo:201-279 through the drop-in signature."""
    cfg = cgm.default_wfa_config()
    cases = [("ACGTACGT", "ACGTACGT", 0, "8="), ("ACGT", "ACAT", 4, "2=1X1="),
             ("ACGGT", "ACGT", 8, "3=1I1="), ("ACGT", "ACGGT", 8, "3=1D1="),
             ("ACGT", "ACGGGT", 10, "3=2D1="), ("AAAA", "TTTT", 16, "4X")]
    for q, r, cost, cigar in cases:
        bq, br = cgm.SequenceBatch(), cgm.SequenceBatch()
        bq.add_sequence(q)
        br.add_sequence(r)
        res = cgm.wfa_affine_cigar_cpu(cgm.get_view(bq, 0), cgm.get_view(br, 0), cfg)
        assert (res.score, res.cigar) == (cost, cigar), (q, r)


def test_cigar_random_pairs_equal_oracle(cgm, dev, orc):
    rng = np.random.default_rng(41)
    for pen in (PEN, dict(x=1, o=0, e=1), dict(x=3, o=2, e=1), dict(x=5, o=1, e=3)):
        qs, rs = [], []
        for _ in range(400):
            q = "".join(rng.choice(list("ACGT"), int(rng.integers(0, 90))))
            if rng.random() < 0.7:
                r = "".join(c if rng.random() > 0.1 else str(rng.choice(list("ACGT")))
                            for c in q if rng.random() > 0.05)
            else:
                r = "".join(rng.choice(list("ACGT"), int(rng.integers(0, 90))))
            qs.append(q)
            rs.append(r)
        me = rng.integers(0, 300, len(qs)).astype(np.int32)
        print(f"[case] cigar random pairs {pen}", flush=True)
        cost, cigars = gpu_cigars(cgm, orc, qs, rs, me, pen)
        for q, r, m, c, cig in zip(qs, rs, me, cost, cigars):
            assert (int(c), cig) == orc.wfa_affine_cigar(q, r, max_error=int(m), **pen), (q, r, m)


def test_cigar_edge_cases(cgm, dev, orc):
    qs = ["", "", "ACGTA", "ACGT", "A" * 40, "ACGT" * 30, "N" * 10]
    rs = ["", "ACG", "", "TTTT", "A" * 60, "ACGT" * 30, "A" * 10]
    me = np.array([0, 50, 50, 11, 45, 5, 10], np.int32)
    print("[case] cigar edge cases", flush=True)
    cost, cigars = gpu_cigars(cgm, orc, qs, rs, me)
    for q, r, m, c, cig in zip(qs, rs, me, cost, cigars):
        assert (int(c), cig) == orc.wfa_affine_cigar(q, r, max_error=int(m)), (q, r, m)


def test_cigar_int32_history_rows(cgm, dev, orc):
    """Rows past int16 (query of 32,500 bases): the int32 history path."""
    rng = np.random.default_rng(3)
    q = "".join(rng.choice(list("ACGT"), 32500))
    r = list(q)
    for p in rng.choice(len(r), 150, replace=False):
        r[p] = "ACGT"[("ACGT".index(r[p]) + 1) % 4]
    r = "".join(r)
    exp = orc.wfa_affine_cigar(q, r, max_error=620)
    assert exp[0] == 600
    print("[case] cigar int32 rows", flush=True)
    cost, cigars = gpu_cigars(cgm, orc, [q], [r], np.array([620], np.int32))
    assert (int(cost[0]), cigars[0]) == exp


def test_cigar_wfa_20k_sample_equals_oracle(cgm, dev, orc, gen_mod):
    w = gen_mod.generate(gen_mod.WFA_1_10KB, 2000)
    idx = np.nonzero(w.max_error <= 1500)[0]
    idx = idx[np.linspace(0, idx.size - 1, 120).astype(int)]
    s = w.subset(idx)
    cfg = cgm.WFAConfig(4, 6, 2, int(s.max_error.max()))
    print(f"[case] cigar wfa 20k x 1-10 kb sample {s.n_pairs} pairs", flush=True)
    cost, cigars = cgm.wfa_affine_cigar_batch(s.q_packed, s.q_off, s.q_len, s.r_packed, s.r_off,
                                              s.r_len, cfg, max_error=s.max_error)
    e_cost, e_cig = orc.wfa_cigar_batch(s.q_packed, s.r_packed, s.q_off, s.r_off, s.q_len,
                                        s.r_len, max_error=s.max_error, threads=8)
    assert np.array_equal(cost, e_cost)
    assert cigars == e_cig


def walk_cigar(qc, rc, cigar_runs, x=4, o=6, e=2):
    """Cost of run-length operations, checking '=' / 'X' against the bases
    and that both sequences are consumed exactly."""
    cnt = cigar_runs >> 2
    op = cigar_runs & 3
    adv_q = np.where(op != 3, cnt, 0)
    adv_r = np.where(op != 2, cnt, 0)
    assert adv_q.sum() == qc.size and adv_r.sum() == rc.size
    qstart = np.concatenate([[0], np.cumsum(adv_q)[:-1]])
    rstart = np.concatenate([[0], np.cumsum(adv_r)[:-1]])
    diag = op <= 1
    n = cnt[diag]
    qi = np.repeat(qstart[diag], n) + (np.arange(n.sum()) - np.repeat(np.cumsum(n) - n, n))
    ri = np.repeat(rstart[diag], n) + (np.arange(n.sum()) - np.repeat(np.cumsum(n) - n, n))
    eq = np.repeat(op[diag] == 0, n)
    assert np.array_equal(qc[qi] == rc[ri], eq)
    return int(x * cnt[op == 1].sum() + ((op >= 2) * (o + e * cnt)).sum())


def codes(packed, off, length):
    i = np.arange(int(length))
    w = packed[int(off) + i // 32]
    return ((w >> (62 - 2 * (i % 32)).astype(np.uint64)) & np.uint64(3)).astype(np.int8)


def test_cigar_wfa_20k_full_size_properties(cgm, dev, gen_mod):
    """All 20,000 pairs of 1-10 kb: every CIGAR's cost is the pair's WFA cost
    and its operations walk both sequences exactly."""
    w = gen_mod.generate(gen_mod.WFA_1_10KB, 20_000)
    cfg = cgm.WFAConfig(4, 6, 2, int(w.max_error.max()))
    print("[case] cigar wfa 20k x 1-10 kb full 20000 pairs", flush=True)
    up = {k: dev.upload(getattr(w, k)) for k in ("q_packed", "r_packed", "q_len", "r_len",
                                                  "q_off", "r_off", "max_error")}
    n = w.n_pairs
    cap = cgm.cigar_capacity(w.max_error, cfg)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(cap)
    d_off = dev.upload(off)
    d_runs = cgm.DeviceArray(int(off[-1]), np.uint32)
    d_cost, d_len = cgm.DeviceArray(n, np.int32), cgm.DeviceArray(n, np.int32)
    cgm._lib.call("cgm_wfa_cigar_launch", up["q_packed"].dptr, up["r_packed"].dptr,
                  up["q_len"].dptr, up["r_len"].dptr, up["q_off"].dptr, up["r_off"].dptr, n,
                  w.q_packed.size, w.r_packed.size, up["max_error"].dptr, 0, 4, 6, 2,
                  d_cost.dptr, d_runs.dptr, d_off.dptr, d_len.dptr, dev.stream)
    dev.synchronize()
    cost = dev.download(d_cost)
    lens = dev.download(d_len)
    runs = dev.download(d_runs)
    plain = cgm.wfa_affine_batch(w.q_packed, w.q_off, w.q_len, w.r_packed, w.r_off, w.r_len, cfg,
                                 max_error=w.max_error, layout=1)
    assert np.array_equal(cost, plain)
    assert np.all(cost >= 0) and np.all(lens > 0)
    for i in range(n):
        r_ = runs[off[i]:off[i] + lens[i]].astype(np.int64)
        c = walk_cigar(codes(w.q_packed, w.q_off[i], w.q_len[i]),
                       codes(w.r_packed, w.r_off[i], w.r_len[i]), r_)
        assert c == cost[i], i


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen
