"""GPU parity of the gap-affine WFA kernels against the CPU oracle
(bit-exact cost or -1), through the C ABI."""
import json
import os

import numpy as np
import pytest

from gpu_helpers import Uploaded, assert_same, wfa_gpu, wfa_oracle, workload_from_strings

pytestmark = pytest.mark.gpu

LAYOUTS = {"plain": 1, "wave": 3, "auto": 0}
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "alignment_vectors.json")


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


def test_wfa_reference_known_answers(cgm, dev):
    cfg = cgm.default_wfa_config()
    cases = [("ACGTACGT", "ACGTACGT", 0), ("ACGT", "ACAT", 4), ("ACGT", "ACGGT", 8),
             ("ACGT", "ACGGGT", 10), ("ACGGT", "ACGT", 8), ("AAAA", "TTTT", 16)]
    for q, r, c in cases:
        b = cgm.SequenceBatch()
        b.add_sequence(q)
        b.add_sequence(r)
        assert cgm.wfa_affine_cpu(cgm.get_view(b, 0), cgm.get_view(b, 1), cfg) == c


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_wfa_golden_vectors(cgm, dev, layout):
    vec = json.load(open(GOLDEN))["wfa_regress"]
    pens = sorted({(v["x"], v["o"], v["e"]) for v in vec})
    for x, o, e in pens:
        vs = [v for v in vec if (v["x"], v["o"], v["e"]) == (x, o, e)]
        w = workload_from_strings([(v["q"], v["r"]) for v in vs],
                                  max_error=[v["max_error"] for v in vs])
        got = wfa_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout], pen=dict(x=x, o=o, e=e))
        assert_same(got, np.array([v["cost"] for v in vs]), f"golden {layout} {(x, o, e)}")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_wfa_20k_sample_parity(cgm, dev, orc, gen_mod, layout):
    w = gen_mod.generate(gen_mod.WFA_1_10KB, 20000)
    idx = np.random.default_rng(3).choice(w.n_pairs, 160, replace=False)
    sub = w.subset(idx)
    print(f"[case] wfa 1-10 kb sample 160 pairs layout={layout}", flush=True)
    got = wfa_gpu(cgm, dev, Uploaded(dev, sub), LAYOUTS[layout])
    exp = wfa_oracle(orc, sub)
    assert (exp >= 0).all()
    assert_same(got, exp, f"wfa 1-10 kb {layout}")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_wfa_edge_cases(cgm, dev, orc, layout):
    s = "ACGTTGCAACGGTCAGTCCGATAGCTAGGCTA" * 4
    cases = [("", "", 0), ("", "ACGT", 14), ("ACGT", "", 14), ("ACGT", "", 13), (s, s, 0),
             (s, s[:-3], 12), (s, s[:-3], 11), (s[:40], s, 3), (s, s[::-1], 400),
             (s, s[::-1], 20), ("NNNN" + s, "AAAA" + s, 10), (s[:33], s[1:34], 40),
             (s[:31] + "G" + s[32:], s, 4), ("A" * 70, "A" * 64 + "C" * 6, 100)]
    w = workload_from_strings([(q, r) for q, r, _ in cases], max_error=[m for _, _, m in cases])
    got = wfa_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    exp = wfa_oracle(orc, w)
    assert (exp == -1).any() and (exp > 0).any()
    assert_same(got, exp, f"edge {layout}")


def test_wfa_scalar_max_error_and_bad_pairs(cgm, dev, orc):
    s = "ACGTTGCAACGGTCAGTCCGATAGCTAGGCTA" * 3
    w = workload_from_strings([(s, s[1:]), (s, s[:50]), (s, s)])
    w.r_off[2] = 99_999
    got = wfa_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS["auto"], me_scalar=30)
    exp = [orc.wfa_affine(s, s[1:], max_error=30), orc.wfa_affine(s, s[:50], max_error=30)]
    assert got.tolist() == exp + [-(2 ** 31)]


def test_wfa_is_deterministic(cgm, dev, gen_mod):
    w = gen_mod.generate(gen_mod.WFA_1_10KB, 400, seed=4)
    up = Uploaded(dev, w)
    assert_same(wfa_gpu(cgm, dev, up, 0), wfa_gpu(cgm, dev, up, 0), "run twice")


def big_wfa_pairs(cgm, gen_mod, orc):
    """Mixed-batch WFA pairs whose wavefronts exceed LDS (max_error 2.6k-13k,
    5-20 kb), the same pairs with max_error just below their cost, and a
    34 kb pair whose rows do not fit int16."""
    w = gen_mod.generate(gen_mod.MIXED, 20000, seed=11)
    wf = np.nonzero(w.kind == 1)[0]
    me = w.max_error[wf]
    idx = np.concatenate([wf[(me >= lo) & (me < hi)][:n]
                          for lo, hi, n in ((2600, 4000, 6), (4000, 7000, 4), (7000, 13000, 1))])
    big = w.subset(idx)
    below = w.subset(idx)
    below.max_error = np.maximum(wfa_oracle(orc, big) - 2, 0).astype(np.int32)
    rng = np.random.default_rng(12)
    t = "".join(rng.choice(list("ACGT"), 34000))
    qs = list(t)
    for pos in rng.choice(34000, 40, replace=False):
        qs[pos] = "ACGT"[("ACGT".index(qs[pos]) + 1) % 4]
    huge = workload_from_strings([("".join(qs[:20000] + qs[20005:]), t)], max_error=[400])
    huge.kind[:] = 1
    return gen_mod.concat([big, below, huge])


def test_wfa_large_pairs_tiled_and_slab(cgm, dev, orc, gen_mod):
    w = big_wfa_pairs(cgm, gen_mod, orc)
    print(f"[case] large WFA pairs n={w.n_pairs} max_error max={w.max_error.max()}", flush=True)
    exp = wfa_oracle(orc, w)
    assert (exp >= 0).sum() > 10 and (exp == -1).sum() > 5
    up = Uploaded(dev, w)
    for layout in ("wave", "auto"):
        print(f"[case] large WFA layout={layout}", flush=True)
        assert_same(wfa_gpu(cgm, dev, up, LAYOUTS[layout]), exp, f"large {layout}")
    kind = dev.upload(w.kind)
    cfg = cgm.AlignConfig(2, -4, -4, -2, 0)
    for lds in (24 * 1024, 12 * 1024, 0):
        print(f"[case] large WFA mixed plan lds_bytes={lds}", flush=True)
        plan = cgm.MixedPlan(w.n_pairs, int(w.max_error.max()), lds_bytes=lds)
        out = cgm.SwResult(w.n_pairs)
        plan.align(dev, (up.q, up.r, up.ql, up.rl, up.qo, up.ro), w.n_pairs, kind, up.band,
                   up.me, cfg, out)
        dev.synchronize()
        assert_same(dev.download(out.scores, w.n_pairs), exp, f"large mixed lds={lds}")
        plan.close()


def _layout_class_pairs(seed):
    """Pairs whose max_error spreads them over every layout class: register
    layouts of 128 to 8192 diagonals, the team layouts beyond."""
    rng = np.random.default_rng(seed)
    pairs, mes = [], []
    for i, (L, me) in enumerate([(L, me) for L in (60, 400, 1500, 4000, 9000)
                                 for me in (8, 40, 120, 250, 600, 1500, 4000, 9000)] * 3):
        a = "".join(rng.choice(list("ACGT"), L))
        b = "".join(c if rng.random() > 0.03 else rng.choice(list("ACGT")) for c in a
                    if rng.random() > 0.01)
        if i % 3 == 1:
            b = b[: len(b) * 2 // 3]
        pairs.append((a, b) if i % 2 else (b, a))
        mes.append(me)
    return pairs, mes


@pytest.mark.parametrize("pen", [dict(x=4, o=6, e=2), dict(x=8, o=12, e=4), dict(x=3, o=5, e=2)])
def test_wfa_every_layout_class(cgm, dev, orc, pen):
    """The automatic layout (register classes where the penalties are
    2 : 4 : 1 multiples, team layouts otherwise) and the team-only layout
    equal the oracle on pairs spread over every class."""
    pairs, mes = _layout_class_pairs(11)
    w = workload_from_strings(pairs, max_error=mes)
    up = Uploaded(dev, w)
    exp = wfa_oracle(orc, w, pen=pen)
    for layout in ("auto", "wave"):
        print(f"[case] wfa every layout class {pen} layout={layout}", flush=True)
        assert_same(wfa_gpu(cgm, dev, up, LAYOUTS[layout], pen=pen), exp,
                    f"wfa classes {pen} {layout}")


@pytest.mark.parametrize("pen", [dict(x=4, o=6, e=2), dict(x=8, o=12, e=4)])
def test_wfa_max_error_at_the_cost(cgm, dev, orc, pen):
    """max_error one below, at and just above each pair's cost, where the
    register layout's max_error pruning removes most cells and whole
    64-diagonal blocks leave the alive range; also length differences whose
    gap cost alone exceeds max_error."""
    pairs, _ = _layout_class_pairs(17)
    w = workload_from_strings(pairs, max_error=[20000] * len(pairs))
    cost = wfa_oracle(orc, w, pen=pen)
    assert (cost >= 0).all()
    for delta in (-1, 0, 1, pen["e"] + 1):
        mes = [max(int(c) + delta, 0) for c in cost]
        wt = workload_from_strings(pairs, max_error=mes)
        print(f"[case] wfa max_error = cost + {delta} {pen}", flush=True)
        got = wfa_gpu(cgm, dev, Uploaded(dev, wt), LAYOUTS["auto"], pen=pen)
        exp = wfa_oracle(orc, wt, pen=pen)
        assert_same(got, exp, f"wfa max_error = cost + {delta} {pen}")
        assert ((got == cost) == ((delta >= 0) | (cost == 0))).all()
