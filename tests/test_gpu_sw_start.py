"""GPU parity of SW start coordinates (cgm_sw_banded_start_launch) with the
oracle's or_sw_banded_start, through the C ABI: score and ends equal the
plain SW entry point's, starts equal the oracle's, on every layout."""
import random

import numpy as np
import pytest

from gpu_helpers import CONFIG_SW, DEFAULT_SW, Uploaded, assert_same, sw_gpu, workload_from_strings

pytestmark = pytest.mark.gpu

LAYOUTS = {"plain": 1, "lane": 2, "auto": 0}


@pytest.fixture(scope="module")
def gen_mod(cgm):
    from cgm_amd import gen
    return gen


def start_gpu(cgm, dev, up, layout, scoring=CONFIG_SW):
    cfg = cgm.AlignConfig(scoring["match"], scoring["mismatch"], scoring["gap_open"],
                          scoring["gap_extend"], 0)
    out = cgm.launch_banded_sw_start(dev, up.q, up.r, up.ql, up.rl, up.qo, up.ro, up.n,
                                     up.q_words, band=0, cfg=cfg, band_buf=up.band,
                                     r_total_words=up.r_words, layout=layout)
    dev.synchronize()
    return tuple(dev.download(a, up.n) for a in (out.scores, out.query_start, out.ref_start,
                                                  out.query_end, out.ref_end))


def start_oracle(orc, w, scoring=CONFIG_SW):
    return orc.sw_start_batch(w.q_packed, w.r_packed, w.q_off, w.r_off, w.q_len, w.r_len,
                              band=w.band, threads=8, **scoring)


def rseq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(alphabet) for _ in range(n))


def test_sw_start_known_answers(cgm, dev):
    b = cgm.SequenceBatch()
    for s in ("AAGGTTA", "CGGTTCC", "ACGTACGT"):
        b.add_sequence(s)
    cfg = cgm.AlignConfig(2, -4, -4, -2, 3)
    r = cgm.smith_waterman_banded(cgm.get_view(b, 0), cgm.get_view(b, 1), cfg, with_start=True)
    assert (r.score, r.query_start, r.ref_start, r.query_end, r.ref_end) == (8, 2, 1, 5, 4)
    r0 = cgm.smith_waterman_banded(cgm.get_view(b, 0), cgm.get_view(b, 1), cfg)
    assert (r0.query_start, r0.ref_start) == (0, 0)   # the reference's zeros by default
    cfg.band_width = 8
    r = cgm.smith_waterman_banded(cgm.get_view(b, 2), cgm.get_view(b, 2), cfg, with_start=True)
    assert (r.score, r.query_start, r.ref_start, r.query_end, r.ref_end) == (16, 0, 0, 7, 7)


@pytest.mark.parametrize("layout", list(LAYOUTS))
def test_sw_start_1k_150bp(cgm, dev, orc, gen_mod, layout):
    w = gen_mod.generate(gen_mod.SW_150BP, 1000, seed=303)
    print(f"[case] sw start 1000 x 150 bp layout={layout}", flush=True)
    up = Uploaded(dev, w)
    got = start_gpu(cgm, dev, up, LAYOUTS[layout])
    assert_same(got, start_oracle(orc, w), f"sw start 1k x 150 bp {layout}")
    assert_same(got[0::3] + got[4:], sw_gpu(cgm, dev, up, LAYOUTS[layout]), "same score/ends")


@pytest.mark.parametrize("scoring", [CONFIG_SW, DEFAULT_SW,
                                     dict(match=1, mismatch=0, gap_open=0, gap_extend=0),
                                     dict(match=3, mismatch=-1, gap_open=-1, gap_extend=-1)])
def test_sw_start_100_300bp_sample(cgm, dev, orc, gen_mod, scoring):
    w = gen_mod.generate(gen_mod.SW_100_300BP, 3000, seed=404)
    print(f"[case] sw start 3000 x 100-300 bp scoring={scoring}", flush=True)
    got = start_gpu(cgm, dev, Uploaded(dev, w), 0, scoring)
    assert_same(got, start_oracle(orc, w, scoring), f"sw start 100-300 bp {scoring}")


def test_sw_start_random_small_every_band(cgm, dev, orc):
    """Short random and planted-gap pairs over bands 0-100 (lane buckets and
    the workgroup layout), zero-sum scorings with many ties."""
    rng = random.Random(5)
    pairs, bands = [], []
    for i in range(3000):
        alpha = rng.choice(["AC", "ACG", "ACGT"])
        if i % 2:
            s = rseq(rng, rng.randint(5, 120), alpha)
            p, l = rng.randint(0, len(s)), rng.randint(1, 6)
            t = s[:p] + rseq(rng, l, alpha) + s[p:]
            pairs.append((s, t) if i % 4 == 1 else (t, s))
        else:
            pairs.append((rseq(rng, rng.randint(0, 90), alpha), rseq(rng, rng.randint(0, 90), alpha)))
        bands.append(rng.randint(0, 100))
    w = workload_from_strings(pairs, band=bands)
    up = Uploaded(dev, w)
    for sc in (CONFIG_SW, dict(match=2, mismatch=-2, gap_open=-2, gap_extend=-2)):
        print(f"[case] sw start 3000 random small pairs scoring={sc}", flush=True)
        got = start_gpu(cgm, dev, up, 0, sc)
        assert_same(got, start_oracle(orc, w, sc), f"sw start random small {sc}")


def test_sw_start_planted_kmer_closed_form(cgm, dev):
    """An identical k-mer over {G,T} between flanks over {A} and {C}: the
    alignment is the k-mer, start and end exactly known."""
    rng = random.Random(9)
    pairs, bands, exp = [], [], []
    for _ in range(500):
        k, x, u = rng.randint(1, 40), rng.randint(0, 60), rng.randint(0, 60)
        km = rseq(rng, k, "GT")
        pairs.append(("A" * x + km + "A" * rng.randint(0, 30), "C" * u + km + "C" * rng.randint(0, 30)))
        bands.append(abs(x - u) + rng.randint(0, 70))
        exp.append((2 * k, x, u, x + k - 1, u + k - 1))
    w = workload_from_strings(pairs, band=bands)
    got = start_gpu(cgm, dev, Uploaded(dev, w), 0)
    exp = tuple(np.array([e[i] for e in exp], np.int32) for i in range(5))
    assert_same(got, exp, "planted k-mer starts")


def test_sw_start_long_and_wide_pairs(cgm, dev, orc):
    """Pairs whose packed values need int64 (20 kb identical, band 64; 2 kb,
    band 128 and 700) and a bad pair: starts -1 for the bad pair, (0, 0) for
    score 0."""
    rng = random.Random(12)
    long_s = rseq(rng, 20000)
    mid = rseq(rng, 2000)
    mid_q = mid[:900] + mid[905:1500] + rseq(rng, 3) + mid[1500:]
    pairs = [(long_s, long_s), (mid_q, mid), (mid, mid_q), ("AAAA", "TTTT"), ("", "ACGT")]
    bands = [64, 128, 700, 4, 4]
    w = workload_from_strings(pairs, band=bands)
    got = start_gpu(cgm, dev, Uploaded(dev, w), 0)
    assert_same(got, start_oracle(orc, w), "long / wide starts")
    assert (int(got[0][0]), int(got[1][0]), int(got[2][0])) == (40000, 0, 0)
    assert [int(got[1][i]) for i in (3, 4)] == [0, 0]
    w.q_len[4] = 10 ** 6   # beyond its buffer: refused, nothing read
    got = start_gpu(cgm, dev, Uploaded(dev, w), 0)
    assert (int(got[0][4]), int(got[1][4]), int(got[2][4])) == (-(2 ** 31), -1, -1)


def test_sw_start_deterministic(cgm, dev, gen_mod):
    w = gen_mod.generate(gen_mod.SW_100_300BP, 20000, seed=505)
    up = Uploaded(dev, w)
    a = start_gpu(cgm, dev, up, 0)
    b = start_gpu(cgm, dev, up, 0)
    assert_same(a, b, "sw start run twice")
