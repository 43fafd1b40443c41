"""GPU parity of banded Needleman-Wunsch (cgm_nw_banded_*) against the
oracle's restatement of needleman_wunsch_banded (align_cpu.mojo:312-377):
known answers, random pairs over several scorings and bands (register lanes
up to band 64, one workgroup per pair above), edge cases, and the batched
workload's 200,000 pairs with both layouts equal on every pair."""
import numpy as np
import pytest

from gpu_helpers import CONFIG_SW, DEFAULT_SW

pytestmark = pytest.mark.gpu


def cfg_of(cgm, sc, band=0):
    return cgm.AlignConfig(sc["match"], sc["mismatch"], sc["gap_open"], sc["gap_extend"], band)


def nw_oracle(orc, w, sc, band=None, threads=8):
    return orc.batch(2, w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                     band=w.band if band is None else band, threads=threads, **sc)


def assert_same(got, exp, what):
    for g, e, name in zip(got, exp, ("score", "query_end", "ref_end")):
        bad = np.nonzero(g != e)[0]
        assert bad.size == 0, f"{what}: {name} differs at {bad[:8].tolist()} " \
                              f"(got {g[bad[:4]].tolist()}, want {e[bad[:4]].tolist()})"


def test_nw_reference_known_answers(cgm, dev):
    """test_alignment.mojo:61-80 through the drop-in signature."""
    cfg = cgm.default_config()
    cfg.band_width = 4
    for q, r, want in (("ACGT", "ACGT", 8), ("AAAA", "TTTT", -12)):
        bq, br = cgm.SequenceBatch(), cgm.SequenceBatch()
        bq.add_sequence(q)
        br.add_sequence(r)
        res = cgm.needleman_wunsch_banded(cgm.get_view(bq, 0), cgm.get_view(br, 0), cfg)
        assert (res.score, res.query_end, res.ref_end) == (want, 3, 3)


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


def random_pairs(gen_mod, n, seed, band_lo, band_hi):
    w = gen_mod.generate(gen_mod.SW_100_300BP, n, seed=seed)
    rng = np.random.default_rng(seed)
    w.band[:] = rng.integers(band_lo, band_hi + 1, n)
    return w


@pytest.mark.parametrize("layout", ["auto", "plain"])
def test_nw_random_pairs_equal_oracle(cgm, dev, orc, gen_mod, layout):
    lay = {"auto": 0, "plain": 1}[layout]
    for k, sc in enumerate((CONFIG_SW, DEFAULT_SW, dict(match=1, mismatch=-1, gap_open=-1,
                                                        gap_extend=-1),
                            dict(match=3, mismatch=1, gap_open=-2, gap_extend=0))):
        w = random_pairs(gen_mod, 1500, 11 + k, 0, 100)
        print(f"[case] nw random {layout} scoring {sc}", flush=True)
        got = cgm.nw_banded_batch(w.q_packed, w.q_off, w.q_len, w.r_packed, w.r_off, w.r_len,
                                  cfg_of(cgm, sc), band=w.band, layout=lay)
        assert_same(got, nw_oracle(orc, w, sc), f"nw {layout} {sc}")


def test_nw_edge_cases(cgm, dev, orc):
    qs = ["", "", "ACG", "ACGTACGT", "ACGT" * 40, "A" * 70, "ACGTN" * 20, "T" * 300, "GA"]
    rs = ["", "ACG", "", "ACG", "ACGT" * 40, "A" * 3, "ACGTA" * 20, "T" * 290, "GATTACA" * 50]
    bands = [3, 3, 5, 4, 0, 80, 1, 200, 2000]
    pq = [orc.pack_ascii(s) for s in qs]
    pr = [orc.pack_ascii(s) for s in rs]
    qp = np.concatenate([w for w, _ in pq])
    rp = np.concatenate([w for w, _ in pr])
    qo = np.cumsum([0] + [w.size for w, _ in pq])[:-1]
    ro = np.cumsum([0] + [w.size for w, _ in pr])[:-1]
    ql = [len(s) for s in qs]
    rl = [len(s) for s in rs]
    sc = DEFAULT_SW
    exp = orc.batch(2, qp, rp, qo, ro, ql, rl, band=bands, **sc)
    for lay in (0, 1):
        print(f"[case] nw edge cases layout {lay}", flush=True)
        got = cgm.nw_banded_batch(qp, qo, ql, rp, ro, rl, cfg_of(cgm, sc), band=bands, layout=lay)
        assert_same(got, exp, f"nw edge layout {lay}")


def test_nw_bad_pairs_flagged(cgm, dev):
    """Offsets past the buffer and a negative band are reported, not read."""
    qp = np.zeros(4, np.uint64)
    rp = np.zeros(4, np.uint64)
    got = cgm.nw_banded_batch(qp, [0, 3, 0], [32, 64, 10], rp, [0, 0, 0], [32, 10, 10],
                              cfg_of(cgm, DEFAULT_SW), band=[4, 4, -1])
    assert got[0][0] == 64 and got[0][1] == cgm._lib.CGM_BAD_PAIR and \
        got[0][2] == cgm._lib.CGM_BAD_PAIR
    assert got[1][1] == -1 and got[2][2] == -1


def test_nw_batched_workload_full_size(cgm, dev, orc, gen_mod):
    """200,000 pairs (100-300 bp, band 16-64): both layouts agree on every
    pair and a 3,000-pair sample equals the oracle."""
    w = gen_mod.generate(gen_mod.SW_100_300BP, 200_000)
    sc = CONFIG_SW
    print("[case] nw full 200000 pairs", flush=True)
    auto = cgm.nw_banded_batch(w.q_packed, w.q_off, w.q_len, w.r_packed, w.r_off, w.r_len,
                               cfg_of(cgm, sc), band=w.band, layout=0)
    plain = cgm.nw_banded_batch(w.q_packed, w.q_off, w.q_len, w.r_packed, w.r_off, w.r_len,
                                cfg_of(cgm, sc), band=w.band, layout=1)
    assert_same(auto, plain, "nw auto vs plain")
    idx = np.random.default_rng(2).choice(w.n_pairs, 3000, replace=False)
    s = w.subset(idx)
    exp = nw_oracle(orc, s, sc)
    assert_same(tuple(a[idx] for a in auto), exp, "nw sample vs oracle")
