"""Writes tests/golden/alignment_vectors.json.

Contents:
  * reference known answers, copied as data from the reference tests
    (tests/test_alignment.mojo:24-58 for SW, :87-138 for WFA) with the
    expected values those tests assert;
  * regression vectors: seeded random small pairs (ASCII, with N) and the
    outputs of the CPU oracle (oracle/cgm_oracle.c) on them.  They pin the
    GPU kernels to the oracle and guard the oracle against unintended
    changes; the oracle itself is pinned by tests/test_oracle.py.
Run: python tests/golden/make_golden.py
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

DEFAULT_SW = dict(match=2, mismatch=-3, gap_open=-5, gap_extend=-2)   # align_cpu.mojo:46-53
CONFIG_SW = dict(match=2, mismatch=-4, gap_open=-4, gap_extend=-2)    # benchmark workloads


def rand_seq(rng, n, p_n=0.03):
    return "".join("N" if rng.random() < p_n else rng.choice("ACGT") for _ in range(n))


def mutate(rng, s, d):
    out = []
    for ch in s:
        u = rng.random()
        if u < d * 0.6:
            out.append(rng.choice("ACGT"))
        elif u < d * 0.8:
            out.append(ch)
            out.append(rng.choice("ACGT"))
        elif u < d:
            continue
        else:
            out.append(ch)
    return "".join(out)


def main():
    rng = random.Random(0x5EED)
    vec = {"sw_kat": [], "wfa_kat": [], "sw_regress": [], "wfa_regress": []}
    # test_alignment.mojo:24-58 (default_config with band overrides)
    vec["sw_kat"] = [
        dict(q="ACGTACGT", r="ACGTACGT", band=8, scoring=DEFAULT_SW, score=16,
             src="tests/test_alignment.mojo:24-32"),
        dict(q="ACGTACGT", r="ACGAACGT", band=4, scoring=DEFAULT_SW, score_at_least=8,
             src="tests/test_alignment.mojo:35-47"),
        dict(q="AAAA", r="TTTT", band=4, scoring=DEFAULT_SW, score=0,
             src="tests/test_alignment.mojo:50-58"),
    ]
    # test_alignment.mojo:87-138 (default_wfa_config: x=4 o=6 e=2 max_error=200)
    for q, r, c, src in [("ACGTACGT", "ACGTACGT", 0, ":87-93"), ("ACGT", "ACAT", 4, ":96-102"),
                         ("ACGT", "ACGGT", 8, ":105-111"), ("ACGT", "ACGGGT", 10, ":114-120"),
                         ("ACGGT", "ACGT", 8, ":123-129"), ("AAAA", "TTTT", 16, ":132-138")]:
        vec["wfa_kat"].append(dict(q=q, r=r, x=4, o=6, e=2, max_error=200, cost=c,
                                   src="tests/test_alignment.mojo" + src))
    for i in range(300):
        n1 = rng.choice([0, 1, 2, 5, 17, 31, 32, 33, 64, 65, 100])
        q = rand_seq(rng, n1)
        r = mutate(rng, q, rng.choice([0.0, 0.05, 0.3])) if rng.random() < 0.7 else \
            rand_seq(rng, rng.randint(0, 100))
        band = rng.choice([0, 1, 3, 8, 16, 32, 64, 100])
        sc = DEFAULT_SW if i % 2 else CONFIG_SW
        s, qe, re_ = oracle.sw_banded(q, r, band=band, **sc)
        vec["sw_regress"].append(dict(q=q, r=r, band=band, scoring=sc, score=s, q_end=qe,
                                      r_end=re_))
    for i in range(300):
        n1 = rng.choice([0, 1, 2, 5, 17, 31, 32, 33, 64, 65, 150])
        q = rand_seq(rng, n1)
        r = mutate(rng, q, rng.choice([0.0, 0.02, 0.1, 0.3])) if rng.random() < 0.8 else \
            rand_seq(rng, rng.randint(0, 60))
        me = rng.choice([0, 3, 10, 40, 200])
        x, o, e = (4, 6, 2) if i % 3 else rng.choice([(1, 1, 1), (3, 2, 1), (5, 0, 3)])
        c = oracle.wfa_affine(q, r, x=x, o=o, e=e, max_error=me)
        vec["wfa_regress"].append(dict(q=q, r=r, x=x, o=o, e=e, max_error=me, cost=c))
    with open(os.path.join(HERE, "alignment_vectors.json"), "w") as f:
        json.dump(vec, f, indent=0, separators=(",", ":"))
    print({k: len(v) for k, v in vec.items()})


if __name__ == "__main__":
    main()
