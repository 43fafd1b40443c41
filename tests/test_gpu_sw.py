"""GPU parity of the banded SW kernels against the CPU oracle (bit-exact on
score, query_end and ref_end), through the C ABI."""
import json
import os

import numpy as np
import pytest

from gpu_helpers import (CONFIG_SW, DEFAULT_SW, Uploaded, assert_same, sw_gpu, sw_oracle,
                         workload_from_strings)

pytestmark = pytest.mark.gpu

LAYOUTS = {"plain": 1, "lane": 2, "auto": 0, "wave": 3}
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "alignment_vectors.json")


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


def test_sw_reference_known_answers(cgm, dev):
    cfg = cgm.default_config()
    b = cgm.SequenceBatch()
    for s in ("ACGTACGT", "ACGAACGT", "AAAA", "TTTT"):
        b.add_sequence(s)
    cfg.band_width = 8
    r = cgm.smith_waterman_banded(cgm.get_view(b, 0), cgm.get_view(b, 0), cfg)
    assert (r.score, r.query_start, r.ref_start, r.query_end, r.ref_end, r.cigar) == \
        (16, 0, 0, 7, 7, [])
    cfg.band_width = 4
    r = cgm.smith_waterman_banded(cgm.get_view(b, 0), cgm.get_view(b, 1), cfg)
    assert r.score > 0 and r.score >= 8
    r = cgm.smith_waterman_banded(cgm.get_view(b, 2), cgm.get_view(b, 3), cfg)
    assert r.score == 0


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_golden_vectors(cgm, dev, layout):
    vec = json.load(open(GOLDEN))["sw_regress"]
    for sc in (DEFAULT_SW, CONFIG_SW):
        vs = [v for v in vec if v["scoring"] == sc]
        w = workload_from_strings([(v["q"], v["r"]) for v in vs], band=[v["band"] for v in vs])
        got = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout], scoring=sc)
        exp = (np.array([v["score"] for v in vs]), np.array([v["q_end"] for v in vs]),
               np.array([v["r_end"] for v in vs]))
        assert_same(got, exp, f"golden {layout} {sc}")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_1k_150bp_parity(cgm, dev, orc, gen_mod, layout):
    w = gen_mod.generate(gen_mod.SW_150BP, 1000, seed=101)
    print(f"[case] sw 1000 pairs x 150 bp layout={layout}", flush=True)
    got = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    assert_same(got, sw_oracle(orc, w), f"sw 1k x 150 bp {layout}")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_200k_sample_parity(cgm, dev, orc, gen_mod, layout):
    w = gen_mod.generate(gen_mod.SW_100_300BP, 4000, seed=202)
    print(f"[case] sw 4000 pairs x 100-300 bp layout={layout}", flush=True)
    got = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    assert_same(got, sw_oracle(orc, w), f"sw 100-300 bp sample {layout}")


def test_sw_200k_full_size(cgm, dev, orc, gen_mod):
    """200,000 pairs: every layout agrees on the whole batch, and a random
    sample of 2,000 pairs equals the oracle."""
    w = gen_mod.generate(gen_mod.SW_100_300BP, 200_000)
    up = Uploaded(dev, w)
    print("[case] sw full 200000 pairs x 100-300 bp", flush=True)
    ref = sw_gpu(cgm, dev, up, LAYOUTS["plain"])
    for name in ("lane", "auto", "wave"):
        assert_same(sw_gpu(cgm, dev, up, LAYOUTS[name]), ref, f"sw 200k full {name} vs plain")
    idx = np.random.default_rng(0).choice(w.n_pairs, 2000, replace=False)
    sub = w.subset(idx)
    assert_same(tuple(a[idx] for a in ref), sw_oracle(orc, sub), "sw 200k full sample")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_edge_cases(cgm, dev, orc, layout):
    pairs, bands = [], []
    s = "ACGTTGCAACGGTCAGTCCGATAGCTAGGCTA" * 5
    for q, r, b in [("", "ACGT", 4), ("ACGT", "", 4), ("", "", 0), ("A", "A", 0), ("A", "C", 3),
                    (s, s, 0), (s, s, 1), (s[:31], s[:33], 2), (s[:64], s[:65], 64),
                    (s[:65], s[:63], 100), (s[:10], s, 5), (s, s[:10], 5), (s, s[40:], 3),
                    (s[40:], s, 50), ("NNNNACGT", "AAAAACGT", 4), ("ACGTN", "ACGTA", 2),
                    (s, s[::-1], 32), (s[:100], s[50:150], 64), (s, s, 1000)]:
        pairs.append((q, r))
        bands.append(b)
    w = workload_from_strings(pairs, band=bands)
    got = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    assert_same(got, sw_oracle(orc, w), f"edge {layout}")


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_long_identical_pair_exceeds_int16(cgm, dev, layout):
    """A 20 kb identical pair scores 40,000 (above int16) at the last cell."""
    rng = np.random.default_rng(1)
    s = "".join(rng.choice(list("ACGT"), 20000))
    w = workload_from_strings([(s, s), (s[:17000], s[:17000])], band=[16, 3])
    sc, qe, re_ = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    assert sc.tolist() == [40000, 34000]
    assert qe.tolist() == [19999, 16999] and re_.tolist() == [19999, 16999]


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_bad_pairs_are_flagged_not_read(cgm, dev, layout):
    w = workload_from_strings([("ACGT" * 10, "ACGT" * 10)] * 4, band=[4, -1, 4, 4])
    w.q_off[2] = 10_000       # outside the query buffer
    w.r_len[3] = 1_000_000    # runs past the reference buffer
    sc, qe, re_ = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    bad = -(2 ** 31)
    assert sc.tolist() == [80, bad, bad, bad]
    assert qe.tolist()[1:] == [-1, -1, -1]


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_band_limit(cgm, dev, orc, layout):
    """Effective band 1600 (the one-workgroup-per-pair LDS limit) is computed
    and equals the oracle; 1601 is flagged CGM_BAD_PAIR; a band above both
    lengths is clamped to the longer length first (5000 -> 1500)."""
    rng = np.random.default_rng(23)
    r = "".join(rng.choice(list("ACGT"), 2000))
    q = list(r)
    for p in rng.choice(2000, 40, replace=False):
        q[p] = "ACGT"[("ACGT".index(q[p]) + 1) % 4]
    q = "".join(q[:700] + q[712:1300]) + "GATTACA" + "".join(q[1300:])
    pairs = [(q, r), (q, r), (q[:1500], r[:1400])]
    w = workload_from_strings(pairs, band=[1600, 1601, 5000])
    print(f"[case] sw band limit layout={layout} lengths {w.q_len.tolist()}/{w.r_len.tolist()}",
          flush=True)
    sc, qe, re_ = sw_gpu(cgm, dev, Uploaded(dev, w), LAYOUTS[layout])
    ref = sw_oracle(orc, w)
    for i in (0, 2):
        assert (sc[i], qe[i], re_[i]) == (ref[0][i], ref[1][i], ref[2][i]), i
    assert sc[0] > 3000
    assert (sc[1], qe[1], re_[1]) == (-(2 ** 31), -1, -1)


def test_sw_is_deterministic(cgm, dev, gen_mod):
    w = gen_mod.generate(gen_mod.SW_100_300BP, 20000, seed=9)
    up = Uploaded(dev, w)
    a = sw_gpu(cgm, dev, up, LAYOUTS["auto"])
    b = sw_gpu(cgm, dev, up, LAYOUTS["auto"])
    assert_same(a, b, "run twice")


def plan_sw(cgm, dev, up, plan, scoring):
    cfg = cgm.AlignConfig(scoring["match"], scoring["mismatch"], scoring["gap_open"],
                          scoring["gap_extend"], 0)
    out = cgm.SwResult(up.n)
    plan.align(dev, (up.q, up.r, up.ql, up.rl, up.qo, up.ro), up.n, None, up.band, None, cfg, out)
    dev.synchronize()
    return tuple(dev.download(a, up.n) for a in (out.scores, out.query_end, out.ref_end))


@pytest.mark.parametrize("dual", [1, 0])
def test_sw_register_lane_forms_equal_oracle(cgm, dev, orc, gen_mod, dual):
    """The int32 and the dual-row register lanes (tuning sw_dual: half
    precision for these lengths and scorings) on random pairs over several
    scorings and every band 0-64."""
    rng = np.random.default_rng(31)
    for k, sc in enumerate((CONFIG_SW, DEFAULT_SW, dict(match=1, mismatch=-1, gap_open=-1,
                                                        gap_extend=-1),
                            dict(match=3, mismatch=0, gap_open=-2, gap_extend=0),
                            dict(match=5, mismatch=-4, gap_open=0, gap_extend=0))):
        w = gen_mod.generate(gen_mod.SW_100_300BP, 3000, seed=40 + k)
        w.band[:] = rng.integers(0, 65, w.n_pairs)
        up = Uploaded(dev, w)
        plan = cgm.MixedPlan(w.n_pairs, 0)
        plan.set_tuning(sw_dual=dual)
        print(f"[case] register lanes dual={dual} scoring {sc}", flush=True)
        assert_same(plan_sw(cgm, dev, up, plan, sc), sw_oracle(orc, w, sc), f"dual={dual} {sc}")
        plan.close()


@pytest.mark.parametrize("split", [0, 30, 55, 100])
def test_sw_unit_queue_split_equals_oracle(cgm, dev, orc, gen_mod, split):
    """Tuning sw_split cuts the longest-first 64-pair units into a long and
    a short queue (0: one queue taken by every wavefront; 100: only the
    first wavefront of each SIMD takes units); every unit is aligned once
    and the results equal the oracle."""
    w = gen_mod.generate(gen_mod.SW_100_300BP, 20000, seed=77)
    up = Uploaded(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, 0)
    plan.set_tuning(sw_split=split)
    print(f"[case] unit queue split={split}", flush=True)
    assert_same(plan_sw(cgm, dev, up, plan, CONFIG_SW), sw_oracle(orc, w, CONFIG_SW),
                f"split={split}")
    plan.close()


def test_sw_dual_lanes_int16_boundary(cgm, dev, orc):
    """match * min(qlen, rlen) + match at 32766 (dual-row lanes) and 32768
    (int32 lanes), alone and together in one wavefront (which then runs the
    int32 lanes); every result equals the oracle and the identical pairs
    score 2 * length."""
    rng = np.random.default_rng(8)
    a = "".join(rng.choice(list("ACGT"), 16383))
    mut = list(a)
    for p in rng.choice(len(mut), 40, replace=False):
        mut[p] = "ACGT"[("ACGT".index(mut[p]) + 1) % 4]
    mut = "".join(mut)
    pairs = [(a[:16382], a[:16382]), (a, a), (a[:16382], mut[:16382]), (mut, a),
             (a[:300], a[:300])]
    w = workload_from_strings(pairs, band=[8, 9, 12, 12, 64])
    up = Uploaded(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, 0)
    plan.set_tuning(sw_dual=1)
    print("[case] dual int16 boundary", flush=True)
    got = plan_sw(cgm, dev, up, plan, CONFIG_SW)
    assert_same(got, sw_oracle(orc, w, CONFIG_SW), "int16 boundary")
    scores, q_end = got[0], got[1]
    assert scores[0] == 32764 and q_end[0] == 16381 and scores[1] == 32766
    plan.close()


@pytest.mark.parametrize("scoring,t", [
    (dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2), 2),
    (dict(match=3, mismatch=-2, gap_open=-5, gap_extend=-1), 1),
    (dict(match=8, mismatch=-16, gap_open=-24, gap_extend=-8), 8),
])
def test_sw_dual_lanes_f16_boundary(cgm, dev, orc, scoring, t):
    """The half-precision dual-row lanes hold exact integers while
    match * min(qlen, rlen) + match <= 2048 t (t: the largest power of two
    <= 8 dividing every score); one base more runs the int16 lanes.  Pairs at
    both sides of that length, identical and mutated, alone and sharing a
    wavefront (which then runs int16), and random pairs of 100-700 bp: every
    result equals the oracle, identical pairs score match * length."""
    rng = np.random.default_rng(9 + t)
    m = scoring["match"]
    L = (2048 * t - m) // m  # longest length inside the half-precision range
    a = "".join(rng.choice(list("ACGT"), L + 1))

    def mutate(x, k):
        y = list(x)
        for p in rng.choice(len(y), k, replace=False):
            y[p] = "ACGT"[("ACGT".index(y[p]) + 1 + int(rng.integers(3))) % 4]
        if len(y) > 10:
            del y[len(y) // 3]
            y.insert(2 * len(y) // 3, "A")
        return "".join(y)

    pairs = [(a[:L], a[:L]), (a, a), (a[:L], mutate(a[:L], 12)), (mutate(a, 12), a),
             (a[:L], a[:L]), (a, a)]
    bands = [12, 12, 40, 40, 64, 33]
    for _ in range(120):
        n = int(rng.integers(100, 700))
        x = "".join(rng.choice(list("ACGT"), n))
        pairs.append((mutate(x, int(rng.integers(0, n // 20 + 1))), x))
        bands.append(int(rng.integers(0, 65)))
    w = workload_from_strings(pairs, band=bands)
    up = Uploaded(dev, w)
    plan = cgm.MixedPlan(w.n_pairs, 0)
    plan.set_tuning(sw_dual=1)
    print(f"[case] dual f16 boundary t={t} L={L}", flush=True)
    got = plan_sw(cgm, dev, up, plan, scoring)
    assert_same(got, sw_oracle(orc, w, scoring), f"f16 boundary t={t}")
    assert got[0][0] == m * L and got[1][0] == L - 1 and got[0][1] == m * (L + 1)
    assert m * L + m <= 2048 * t < m * (L + 1) + m
    plan.close()


def test_sw_wave_long_pairs(cgm, dev, orc):
    """Reads of 300 bp to 6 kb with planted substitutions, indels and
    flanks, at bands 0-64 (64: the extra diagonal) and bands beyond the
    layout's limit (placed as the automatic layout places them)."""
    import random
    rng = random.Random(31)
    pairs, bands = [], []
    for i in range(240):
        L = rng.randint(300, 6000)
        a = "".join(rng.choice("ACGT") for _ in range(L))
        b = []
        for c in a:
            u = rng.random()
            if u < 0.02:
                b.append(rng.choice("ACGT"))
            elif u < 0.025:
                continue
            elif u < 0.03:
                b.append(c + rng.choice("ACGT"))
            else:
                b.append(c)
        b = "".join(b)
        if i % 5 == 1:
            b = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 80))) + b
        if i % 7 == 2:
            b = b[: len(b) * 3 // 4]
        pairs.append((a, b) if i % 2 else (b, a))
        bands.append(rng.choice([0, 1, 2, 31, 32, 33, 63, 64, 64, 70, rng.randint(0, 64)]))
    w = workload_from_strings(pairs, band=bands)
    up = Uploaded(dev, w)
    for sc in (CONFIG_SW, DEFAULT_SW):
        print(f"[case] sw wave long pairs scoring={sc}", flush=True)
        assert_same(sw_gpu(cgm, dev, up, LAYOUTS["wave"], scoring=sc), sw_oracle(orc, w, sc),
                    f"sw wave long pairs {sc}")


def test_sequence_batch_forms_ignore_nmask(cgm, dev, orc, gen_mod):
    """cgm_sw_banded_batch / cgm_wfa_affine_batch take each side's SoA with
    its NMask words; the masks are never read, so the real masks, inverted
    masks and NULL give the oracle's results (N scored as the A it is
    packed as, as the reference scores it)."""
    from gpu_helpers import WFA_PEN, wfa_oracle
    w = gen_mod.generate(gen_mod.SW_150BP, 2000, seed=131)
    assert w.q_nmask.any()  # the workload holds N positions
    w3 = gen_mod.generate(gen_mod.WFA_1_10KB, 64, seed=132)
    exp_sw = sw_oracle(orc, w)
    exp_wfa = wfa_oracle(orc, w3)
    for variant in ("masks", "inverted", "null"):
        for wl, kind in ((w, "sw"), (w3, "wfa")):
            up = Uploaded(dev, wl)
            if variant == "null":
                qn = rn = None
            else:
                qm, rm = wl.q_nmask, wl.r_nmask
                if variant == "inverted":
                    qm, rm = ~qm, ~rm
                qn, rn = dev.upload(qm), dev.upload(rm)
            qnp = qn.dptr if qn is not None else None
            rnp = rn.dptr if rn is not None else None
            if kind == "sw":
                out = [cgm.DeviceArray(wl.n_pairs, np.int32) for _ in range(3)]
                cgm._lib.call("cgm_sw_banded_batch", up.q.dptr, qnp, up.qo.dptr, up.ql.dptr,
                              up.q_words, up.r.dptr, rnp, up.ro.dptr, up.rl.dptr, up.r_words,
                              wl.n_pairs, up.band.dptr, 0, CONFIG_SW["match"],
                              CONFIG_SW["mismatch"], CONFIG_SW["gap_open"],
                              CONFIG_SW["gap_extend"], out[0].dptr, out[1].dptr, out[2].dptr,
                              dev.stream)
                dev.synchronize()
                got = tuple(dev.download(o, wl.n_pairs) for o in out)
                assert_same(got, exp_sw, f"cgm_sw_banded_batch nmask={variant}")
            else:
                out = cgm.DeviceArray(wl.n_pairs, np.int32)
                cgm._lib.call("cgm_wfa_affine_batch", up.q.dptr, qnp, up.qo.dptr, up.ql.dptr,
                              up.q_words, up.r.dptr, rnp, up.ro.dptr, up.rl.dptr, up.r_words,
                              wl.n_pairs, up.me.dptr, 0, WFA_PEN["x"], WFA_PEN["o"],
                              WFA_PEN["e"], out.dptr, dev.stream)
                dev.synchronize()
                assert_same(dev.download(out, wl.n_pairs), exp_wfa,
                            f"cgm_wfa_affine_batch nmask={variant}")
