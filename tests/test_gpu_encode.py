"""GPU parity of the device ingest (cgm_encode_ascii_*): ASCII -> 2-bit packed
words + NMask + word offsets, against the oracle's restatement of
SequenceBatch.add_sequence (sequence.mojo:38-66, dna.mojo:26-40), and a
decode/encode round trip of whole generated batches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_batch(orc, seqs):
    """(packed, nmask, word offsets, lengths) as appending seqs one by one."""
    packed, masks, offs, lens = [], [], [], []
    nw = 0
    for s in seqs:
        w, m = orc.pack_ascii(s)
        offs.append(nw)
        lens.append(len(s))
        packed.append(w)
        masks.append(m)
        nw += w.size
    cat = lambda a: np.concatenate(a) if a else np.zeros(0, np.uint64)  # noqa: E731
    return cat(packed), cat(masks), np.array(offs, np.int32), np.array(lens, np.int32)


def check_against_oracle(cgm, orc, seqs, what):
    got = cgm.encode_batch_gpu(seqs)
    exp = oracle_batch(orc, seqs)
    for g, e, name in zip((got.packed, got.n_masks, got.offsets, got.lengths), exp,
                          ("packed", "nmask", "offsets", "lengths")):
        assert g.shape == e.shape and np.array_equal(g, e), f"{what}: {name} differs"


def test_encode_known_cases(cgm, dev, orc):
    seqs = ["", "A", "ACGT", "acgt", "ACGTN", "NNNN", "A" * 31, "C" * 32, "G" * 33, "T" * 64,
            "ACGTRYKM-.*nx" * 5, "acgtACGT" * 20 + "N", "T" * 65, "", "GATTACA" * 143]
    print(f"[case] encode known cases ({len(seqs)} sequences)", flush=True)
    check_against_oracle(cgm, orc, seqs, "known cases")


def test_encode_random_bytes_and_lengths(cgm, dev, orc):
    rng = np.random.default_rng(5)
    alphabet = np.frombuffer(b"ACGTacgtNnRY-", dtype=np.uint8)
    seqs = []
    for i in range(3000):
        n = int(rng.integers(0, 400))
        if i % 3 == 0:
            seqs.append(rng.integers(0, 256, n, dtype=np.uint8).tobytes())  # any byte value
        else:
            seqs.append(rng.choice(alphabet, n).tobytes())
    print(f"[case] encode random bytes ({len(seqs)} sequences)", flush=True)
    check_against_oracle(cgm, orc, seqs, "random bytes")


def test_encode_launch_offsets_bad_entries_and_capacity(cgm, dev, orc):
    """Device-pointer form: arbitrary (overlapping, unaligned) byte starts,
    out-of-buffer entries flagged with length -1, and a words_cap below the
    batch's need writes nothing past the cap."""
    rng = np.random.default_rng(9)
    buf = rng.choice(np.frombuffer(b"ACGTNacgt", dtype=np.uint8), 5000).astype(np.uint8)
    starts = rng.integers(0, 4000, 400).astype(np.int64)
    lens = rng.integers(0, 1000, 400).astype(np.int32)
    lens = np.minimum(lens, 5000 - starts).astype(np.int32)
    starts[7], lens[7] = 4990, 20      # runs past the end of the buffer
    starts[11], lens[11] = -3, 5       # negative start
    exp_lens = lens.copy()
    exp_lens[[7, 11]] = -1
    seqs = [buf[s:s + l].tobytes() if k not in (7, 11) else b""
            for k, (s, l) in enumerate(zip(starts, lens))]
    e_packed, e_mask, e_off, _ = oracle_batch(orc, seqs)
    total = int(e_packed.size)
    d_buf, d_st, d_ln = dev.upload(buf), dev.upload(starts), dev.upload(lens)
    cap = total + 64
    sentinel = np.uint64(0xDEADBEEFCAFEF00D)
    d_packed = dev.upload(np.full(cap, sentinel, np.uint64))
    d_mask = dev.upload(np.full(cap, sentinel, np.uint64))
    d_off, d_len = dev.alloc_int32(len(lens)), dev.alloc_int32(len(lens))
    d_tot = dev.upload(np.zeros(1, np.int64))
    print("[case] encode launch, full capacity", flush=True)
    cgm.encode_ascii_launch(dev, d_buf, d_st, d_ln, len(lens), cap, d_packed, d_mask, d_off,
                            d_len, d_tot)
    dev.synchronize()
    assert int(dev.download(d_tot)[0]) == total
    assert np.array_equal(dev.download(d_len), exp_lens)
    assert np.array_equal(dev.download(d_off), e_off)
    p = dev.download(d_packed)
    m = dev.download(d_mask)
    assert np.array_equal(p[:total], e_packed) and np.array_equal(m[:total], e_mask)
    assert np.all(p[total:] == sentinel) and np.all(m[total:] == sentinel)
    # a cap of half the batch: the first half is written, nothing after it
    half = total // 2
    d_packed2 = dev.upload(np.full(cap, sentinel, np.uint64))
    print("[case] encode launch, half capacity", flush=True)
    cgm.encode_ascii_launch(dev, d_buf, d_st, d_ln, len(lens), half, d_packed2, None, d_off,
                            d_len, d_tot)
    dev.synchronize()
    assert int(dev.download(d_tot)[0]) == total
    p2 = dev.download(d_packed2)
    assert np.array_equal(p2[:half], e_packed[:half]) and np.all(p2[half:] == sentinel)


@pytest.mark.parametrize("workload,n", [("SW_100_300BP", 200_000), ("SERVING", 100_000)])
def test_encode_round_trip_full_batches(cgm, dev, workload, n):
    """Decoding a whole generated batch to ASCII and encoding it on the GPU
    gives back the generator's words, masks and offsets (queries and
    references)."""
    from cgm_amd import gen
    w = gen.generate(getattr(gen, workload), n)
    for side, packed, mask, offs, lens in (("query", w.q_packed, w.q_nmask, w.q_off, w.q_len),
                                           ("reference", w.r_packed, w.r_nmask, w.r_off,
                                            w.r_len)):
        seqs = cgm.decode_ascii(packed, mask, offs, lens)
        print(f"[case] encode round trip {workload} {side} ({len(seqs)} sequences)",
              flush=True)
        got = cgm.encode_batch_gpu(seqs)
        assert np.array_equal(got.offsets, offs), side
        assert np.array_equal(got.lengths, lens), side
        assert np.array_equal(got.packed, packed), side
        assert np.array_equal(got.n_masks, mask), side
