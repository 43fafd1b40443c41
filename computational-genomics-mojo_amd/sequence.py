"""Host-side SequenceBatch: 2-bit packed structure-of-arrays DNA batches.

Mirrors genomics/core/sequence.mojo:23-137 and genomics/core/dna.mojo:12-141
of the reference: 32 bases per uint64 word, base i of a word at bits
(63-2i)..(62-2i) (MSB first), A=0 C=1 G=2 T=3, any other byte stored as A and
flagged at bit i of the word's NMask (LSB = base 0); every sequence starts on
a new word; offsets are word offsets, lengths are in bases.
"""
from __future__ import annotations

import numpy as np

BASE_BITS = 2           # dna.mojo:12
BASES_PER_WORD = 32     # dna.mojo:13


class Base:
    """dna.mojo:16-22"""
    A = 0
    C = 1
    G = 2
    T = 3
    N = 0xFF


_CODE = np.full(256, 0xFF, dtype=np.uint8)
for _ch, _v in ((b"A", 0), (b"a", 0), (b"C", 1), (b"c", 1), (b"G", 2), (b"g", 2),
                (b"T", 3), (b"t", 3)):
    _CODE[_ch[0]] = _v
_SHIFTS = (62 - 2 * np.arange(32, dtype=np.uint64)).astype(np.uint64)
_BITS = (np.uint64(1) << np.arange(32, dtype=np.uint64)).astype(np.uint64)


def ascii_to_2bit(c: int) -> int:
    """dna.mojo:26-40"""
    return int(_CODE[c & 0xFF])


def base_to_ascii(b: int) -> int:
    """dna.mojo:43-52"""
    return (65, 67, 71, 84)[b] if 0 <= b < 3 else 84


def get_base_from_word(word: int, pos: int) -> int:
    """dna.mojo:97-100"""
    return (int(word) >> (62 - pos * BASE_BITS)) & 3


def words_for_length(length: int) -> int:
    return (length + BASES_PER_WORD - 1) // BASES_PER_WORD


def pack_codes(codes: np.ndarray):
    """Packs an array of codes (0..3, 0xFF = N) into (words, nmask)."""
    n = int(codes.size)
    nw = words_for_length(n)
    padded = np.zeros(nw * 32, dtype=np.uint8)
    padded[:n] = codes
    is_n = padded == 0xFF
    padded[is_n] = 0
    blocks = padded.reshape(nw, 32).astype(np.uint64)
    words = np.bitwise_or.reduce(blocks << _SHIFTS, axis=1) if nw else np.zeros(0, np.uint64)
    nmask = (np.bitwise_or.reduce(np.where(is_n.reshape(nw, 32), _BITS, np.uint64(0)), axis=1)
             if nw else np.zeros(0, np.uint64))
    return words.astype(np.uint64), nmask.astype(np.uint64)


def encode_ascii(seq) -> np.ndarray:
    b = seq.encode() if isinstance(seq, str) else bytes(seq)
    return _CODE[np.frombuffer(b, dtype=np.uint8)] if b else np.zeros(0, np.uint8)


class SequenceView:
    """sequence.mojo:78-120 -- one sequence's words and NMask words."""

    def __init__(self, packed: np.ndarray, n_masks: np.ndarray, length: int):
        self.packed = packed
        self.n_masks = n_masks
        self.length = int(length)
        self.word_count = words_for_length(self.length)

    def get_base(self, pos: int) -> int:
        return get_base_from_word(self.packed[pos // BASES_PER_WORD], pos % BASES_PER_WORD)

    def is_n(self, pos: int) -> bool:
        return bool((int(self.n_masks[pos // BASES_PER_WORD]) >> (pos % BASES_PER_WORD)) & 1)


class SequenceBatch:
    """sequence.mojo:23-75 -- packed, n_masks, offsets, lengths, count."""

    def __init__(self):
        self._packed = []
        self._masks = []
        self._offsets = []
        self._lengths = []
        self._words = 0
        self._cache = None

    @classmethod
    def from_arrays(cls, packed, n_masks, offsets, lengths):
        b = cls()
        b._cache = (np.ascontiguousarray(packed, np.uint64),
                    np.ascontiguousarray(n_masks, np.uint64) if n_masks is not None
                    else np.zeros(len(packed), np.uint64),
                    np.ascontiguousarray(offsets, np.int32),
                    np.ascontiguousarray(lengths, np.int32))
        b._words = int(len(packed))
        b._offsets = list(b._cache[2])
        b._lengths = list(b._cache[3])
        b._packed = [b._cache[0]]
        b._masks = [b._cache[1]]
        return b

    def add_sequence(self, src, length: int | None = None) -> None:
        """sequence.mojo:38-66 -- encode and append an ASCII sequence."""
        codes = encode_ascii(src)
        if length is not None:
            codes = codes[:length]
        w, m = pack_codes(codes)
        self._offsets.append(self._words)
        self._lengths.append(int(codes.size))
        self._packed.append(w)
        self._masks.append(m)
        self._words += int(w.size)
        self._cache = None

    @property
    def count(self) -> int:
        return len(self._lengths)

    def _arrays(self):
        if self._cache is None:
            packed = np.concatenate(self._packed) if self._packed else np.zeros(0, np.uint64)
            masks = np.concatenate(self._masks) if self._masks else np.zeros(0, np.uint64)
            self._cache = (packed.astype(np.uint64), masks.astype(np.uint64),
                           np.asarray(self._offsets, np.int32), np.asarray(self._lengths, np.int32))
            self._packed, self._masks = [self._cache[0]], [self._cache[1]]
        return self._cache

    @property
    def packed(self) -> np.ndarray:
        return self._arrays()[0]

    @property
    def n_masks(self) -> np.ndarray:
        return self._arrays()[1]

    @property
    def offsets(self) -> np.ndarray:
        return self._arrays()[2]

    @property
    def lengths(self) -> np.ndarray:
        return self._arrays()[3]

    def total_bases(self) -> int:
        return int(np.sum(self.lengths, dtype=np.int64))

    def words_for(self, idx: int) -> int:
        return words_for_length(int(self.lengths[idx]))


def concat_ascii(seqs):
    """Sequences (str or bytes) back to back as one uint8 buffer, with the
    int64 byte start and int32 length of each: the input of the device
    encoder (cgm_encode_ascii_*)."""
    parts = [q.encode() if isinstance(q, str) else bytes(q) for q in seqs]
    lens = np.fromiter((len(p) for p in parts), dtype=np.int64, count=len(parts))
    starts = np.zeros(len(parts), np.int64)
    if len(parts) > 1:
        starts[1:] = np.cumsum(lens)[:-1]
    buf = np.frombuffer(b"".join(parts), dtype=np.uint8).copy()
    return buf, starts, lens.astype(np.int32)


def decode_ascii(packed, nmask, offsets, lengths) -> list:
    """Sequences of a packed batch back to ASCII: bases as A/C/G/T
    (base_to_ascii, dna.mojo:43-52), NMask-flagged positions as 'N'."""
    out = []
    lut = np.frombuffer(b"ACGT", dtype=np.uint8)
    for o, n in zip(offsets, lengths):
        o, n = int(o), int(n)
        nw = words_for_length(n)
        w = np.asarray(packed[o:o + nw], np.uint64)
        m = np.asarray(nmask[o:o + nw], np.uint64) if nmask is not None else np.zeros(nw, np.uint64)
        i = np.arange(n)
        codes = (w[i // 32] >> (62 - 2 * (i % 32)).astype(np.uint64)) & np.uint64(3)
        flags = (m[i // 32] >> (i % 32).astype(np.uint64)) & np.uint64(1)
        s = lut[codes.astype(np.int64)]
        s[flags.astype(bool)] = ord("N")
        out.append(s.tobytes())
    return out


def encode_batch_gpu(seqs) -> SequenceBatch:
    """SequenceBatch of `seqs` encoded on the GPU (cgm_encode_ascii_host):
    the batch add_sequence (sequence.mojo:38-66) would build by appending
    them one by one."""
    from . import _lib
    buf, starts, lens = concat_ascii(seqs)
    n = len(lens)
    cap = n + int(buf.size) // BASES_PER_WORD + 1
    packed = np.zeros(cap, np.uint64)
    masks = np.zeros(cap, np.uint64)
    offs = np.zeros(n, np.int32)
    lout = np.zeros(n, np.int32)
    import ctypes
    total = ctypes.c_int64(0)
    _lib.call("cgm_encode_ascii_host", _lib.ptr(buf if buf.size else np.zeros(1, np.uint8)),
              int(buf.size), _lib.ptr(starts), _lib.ptr(lens), n, _lib.ptr(packed),
              _lib.ptr(masks), cap, _lib.ptr(offs), _lib.ptr(lout), ctypes.byref(total))
    nw = int(total.value)
    return SequenceBatch.from_arrays(packed[:nw], masks[:nw], offs, lout)


def encode_ascii_launch(device, ascii, seq_start, seq_len, n_seqs: int, words_cap: int,
                        packed, nmask, word_off, lengths, total_words=None) -> None:
    """Device-pointer form (cgm_encode_ascii_launch) on the device's stream;
    every array is a DeviceArray (nmask / total_words may be None)."""
    from . import _lib
    _lib.call("cgm_encode_ascii_launch", _lib.ptr(ascii), int(ascii.count), _lib.ptr(seq_start),
              _lib.ptr(seq_len), int(n_seqs), _lib.ptr(packed), _lib.ptr(nmask), int(words_cap),
              _lib.ptr(word_off), _lib.ptr(lengths), _lib.ptr(total_words), device.stream)


def get_view(batch: SequenceBatch, idx: int) -> SequenceView:
    """sequence.mojo:123-137 -- copy sequence idx into a SequenceView."""
    off = int(batch.offsets[idx])
    n = batch.words_for(idx)
    return SequenceView(batch.packed[off:off + n].copy(), batch.n_masks[off:off + n].copy(),
                        int(batch.lengths[idx]))
