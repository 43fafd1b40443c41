// cgm_encode.hip -- ASCII -> 2-bit packed SequenceBatch encoding on the
// device (the ingest side of the alignment path).
//
// Semantics of SequenceBatch.add_sequence (genomics/core/sequence.mojo:38-66)
// with the code table of ascii_to_2bit (genomics/core/dna.mojo:26-40): each
// sequence starts on a new 64-bit word, base i of a word sits at bits
// (63-2i)..(62-2i), A/a=0 C/c=1 G/g=2 T/t=3, any other byte is stored as 0
// and flagged at bit i of the word's NMask (LSB = base 0); bits past the
// sequence's end are 0 in both.  Word offsets are the running sum of
// ceil(len / 32), as the reference's append order gives.
//
// Pipeline (one stream, stream-ordered workspace):
//  1. per sequence: validate (start, len) against the byte buffer, words;
//  2. exclusive scan of the word counts (hipCUB) -> word offsets;
//  3. per sequence: int32 offsets/lengths out, and the first sequence of
//     every 256-word block (so the packing kernel needs no global search);
//  4. one thread per output word: find its sequence by a short binary
//     search inside its block's sequence range, load the 32 bytes with
//     aligned dword loads + alignbyte, and encode four bytes per SWAR step.
// The packing kernel reads each input byte once and writes 16 bytes per 32
// bases, so it is bound by HBM bandwidth.

#include <stdint.h>

#include <vector>

#include <hipcub/hipcub.hpp>

#include "cgm_internal.h"

namespace cgm {

namespace {

constexpr int kEncBlock = 256;  // words per packing workgroup

__global__ void enc_count_kernel(const int64_t *seq_start, const int32_t *seq_len, int64_t n,
                                 int64_t ascii_bytes, int64_t *nwords, int32_t *lengths) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t s = seq_start[i];
    const int64_t l = seq_len[i];
    const bool ok = s >= 0 && l >= 0 && s <= ascii_bytes && l <= ascii_bytes - s;
    nwords[i] = ok ? (l + 31) >> 5 : 0;
    lengths[i] = ok ? (int32_t)l : -1;
}

// Offsets out (int32, as the reference's device upload narrows them), the
// total word count, and block_first[b] = the sequence holding word 256 b
// for the nblocks blocks the packing kernel can reach.
__global__ void enc_offsets_kernel(const int64_t *off64, const int64_t *nwords, int64_t n,
                                   int32_t *word_off, int32_t *block_first, int64_t nblocks,
                                   int64_t *total) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = off64[i], nw = nwords[i];
    word_off[i] = (int32_t)o;
    if (i == n - 1) *total = o + nw;
    for (int64_t b = (o + kEncBlock - 1) / kEncBlock; b * kEncBlock < o + nw && b < nblocks; ++b)
        block_first[b] = (int32_t)i;
}

// Four ASCII bytes (byte 0 = first base) -> 8 packed bits (first base in the
// two most significant bits) and a 4-bit NMask nibble (first base in bit
// 0).  `present` has 0xFF in each byte that belongs to the sequence.
__device__ __forceinline__ void encode4(uint32_t v, uint32_t present, uint32_t &chunk,
                                        uint32_t &nib) {
    const uint32_t u = v & 0xDFDFDFDFu;                    // fold lower case
    const uint32_t x = (u >> 1) & 0x03030303u;             // A0 C1 G3 T2
    const uint32_t code = x ^ ((x >> 1) & 0x01010101u);    // A0 C1 G2 T3
    const uint32_t letter = __builtin_amdgcn_perm(0u, 0x54474341u, code);  // 'A' 'C' 'G' 'T'
    const uint32_t t = u ^ letter;                         // 0 exactly for A/C/G/T bytes
    const uint32_t bad = ((((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t) & 0x80808080u) & present;
    const uint32_t bad01 = bad >> 7;                       // 1 per non-ACGT byte
    const uint32_t keep = present & ~(bad01 * 0xFFu);
    const uint32_t c = code & keep;
    // bytes b0..b3 (2 bits each) -> b0 b1 b2 b3 from bit 31 down; no carries
    chunk = (c * 0x40100401u) >> 24;
    // flags of bytes 0..3 -> bits 0..3; no carries
    nib = ((bad01 * 0x00204081u) >> 21) & 0xFu;
}

__global__ __launch_bounds__(kEncBlock) void enc_pack_kernel(
    const uint8_t *ascii, int64_t ascii_bytes, const int64_t *seq_start, const int32_t *lengths,
    const int64_t *off64, int64_t n, const int32_t *block_first, int64_t nblocks_cap,
    const int64_t *total, int64_t words_cap, uint64_t *packed, uint64_t *nmask) {
    const int64_t w = blockIdx.x * (int64_t)kEncBlock + threadIdx.x;
    const int64_t tot = *total;
    if (w >= tot || w >= words_cap) return;
    const int64_t b = w / kEncBlock;
    // the covering sequence is the last one whose offset is <= w, between
    // this block's first sequence and the next block's
    int64_t lo = block_first[b];
    int64_t hi = (b + 1 < nblocks_cap && (b + 1) * kEncBlock < tot) ? block_first[b + 1] : n - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (off64[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    const int64_t s = lo;
    const int64_t k = w - off64[s];
    const int64_t pos = seq_start[s] + 32 * k;
    const int cnt = (int)min((int64_t)32, (int64_t)lengths[s] - 32 * k);
    // dwords covering [pos, pos + cnt), 4-byte aligned in the address
    // space; a dword that is not wholly inside the buffer is read byte by
    // byte, and bytes outside it read as 0 (they are masked out anyway)
    const uintptr_t lo_addr = (uintptr_t)ascii, hi_addr = lo_addr + (uintptr_t)ascii_bytes;
    const uintptr_t start = lo_addr + (uintptr_t)pos;
    const uintptr_t base = start & ~(uintptr_t)3;
    const int sh = (int)(start & 3);
    uint32_t raw[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const uintptr_t a = base + 4 * j;
        if (a >= lo_addr && a + 4 <= hi_addr) {
            raw[j] = *(const uint32_t *)a;
        } else {
            uint32_t v = 0;
            for (int t = 0; t < 4; ++t)
                if (a + t >= lo_addr && a + t < hi_addr) v |= (uint32_t)(*(const uint8_t *)(a + t)) << (8 * t);
            raw[j] = v;
        }
    }
    uint32_t whi = 0, wlo = 0, m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t v = __builtin_amdgcn_alignbyte(raw[j + 1], raw[j], sh);
        const int left = cnt - 4 * j;  // bytes of this dword inside the sequence
        const uint32_t present = left >= 4 ? 0xFFFFFFFFu
                                 : left <= 0 ? 0u
                                             : (0xFFFFFFFFu >> (8 * (4 - left)));
        uint32_t chunk, nib;
        encode4(v, present, chunk, nib);
        if (j < 4) whi |= chunk << (24 - 8 * j);
        else wlo |= chunk << (24 - 8 * (j - 4));
        m |= nib << (4 * j);
    }
    packed[w] = ((uint64_t)whi << 32) | wlo;
    if (nmask) nmask[w] = m;
}

}  // namespace

}  // namespace cgm

using namespace cgm;

extern "C" {

int cgm_encode_ascii_launch(const uint8_t *ascii, int64_t ascii_bytes, const int64_t *seq_start,
                            const int32_t *seq_len, int64_t n_seqs, uint64_t *packed,
                            uint64_t *nmask, int64_t words_cap, int32_t *word_off,
                            int32_t *lengths, int64_t *total_words, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (n_seqs < 0 || n_seqs > (int64_t(1) << 31) - 1)
        return set_error(CGM_ERR_INVALID, "encode: n_seqs out of range");
    if (ascii_bytes < 0 || words_cap < 0) return set_error(CGM_ERR_INVALID, "encode: negative size");
    if (words_cap > (int64_t(1) << 31) - 1)
        return set_error(CGM_ERR_INVALID, "encode: words_cap must fit int32 word offsets");
    if (n_seqs == 0) {
        if (total_words) CGM_HIP(hipMemsetAsync(total_words, 0, 8, stream));
        return CGM_OK;
    }
    if (!seq_start || !seq_len || !word_off || !lengths || (!ascii && ascii_bytes > 0) ||
        (!packed && words_cap > 0))
        return set_error(CGM_ERR_INVALID, "encode: null buffer");
    const int64_t n = n_seqs;
    const int64_t nblocks = (words_cap + kEncBlock - 1) / kEncBlock + 1;
    size_t scan_bytes = 0;
    CGM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const int64_t *)nullptr,
                                             (int64_t *)nullptr, (int)n, stream));
    // stream-ordered workspace: word counts, offsets, block table, total, scan
    const size_t a8 = ((size_t)n * 8 + 255) & ~(size_t)255;
    const size_t ab = ((size_t)nblocks * 4 + 255) & ~(size_t)255;
    const size_t ws_bytes = 2 * a8 + ab + 256 + scan_bytes;
    void *ws = nullptr;
    CGM_HIP(hipMallocAsync(&ws, ws_bytes, stream));
    uint8_t *p = (uint8_t *)ws;
    int64_t *nwords = (int64_t *)p;
    int64_t *off64 = (int64_t *)(p + a8);
    int32_t *block_first = (int32_t *)(p + 2 * a8);
    int64_t *total = (int64_t *)(p + 2 * a8 + ab);
    void *scan_tmp = p + 2 * a8 + ab + 256;
    const unsigned g = (unsigned)((n + 255) / 256);
    int rc = CGM_OK;
    hipLaunchKernelGGL(enc_count_kernel, dim3(g), dim3(256), 0, stream, seq_start, seq_len, n,
                       ascii_bytes, nwords, lengths);
    size_t sb = scan_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, sb, nwords, off64, (int)n, stream);
    if (e != hipSuccess) rc = hip_error(e, "encode: scan");
    if (!rc) {
        hipLaunchKernelGGL(enc_offsets_kernel, dim3(g), dim3(256), 0, stream, off64, nwords, n,
                           word_off, block_first, nblocks, total);
        if (words_cap > 0)
            hipLaunchKernelGGL(enc_pack_kernel, dim3((unsigned)((words_cap + kEncBlock - 1) / kEncBlock)),
                               dim3(kEncBlock), 0, stream, ascii, ascii_bytes, seq_start, lengths,
                               off64, n, block_first, nblocks, total, words_cap, packed, nmask);
        if (total_words) {
            e = hipMemcpyAsync(total_words, total, 8, hipMemcpyDeviceToDevice, stream);
            if (e != hipSuccess) rc = hip_error(e, "encode: total");
        }
        e = hipGetLastError();
        if (!rc && e != hipSuccess) rc = hip_error(e, "encode: launch");
    }
    e = hipFreeAsync(ws, stream);
    if (!rc && e != hipSuccess) rc = hip_error(e, "encode: workspace free");
    return rc;
}

int cgm_encode_ascii_host(const uint8_t *ascii, int64_t ascii_bytes, const int64_t *seq_start,
                          const int32_t *seq_len, int64_t n_seqs, uint64_t *packed,
                          uint64_t *nmask, int64_t words_cap, int32_t *word_off, int32_t *lengths,
                          int64_t *total_words) {
    if (n_seqs < 0 || ascii_bytes < 0 || words_cap < 0)
        return set_error(CGM_ERR_INVALID, "encode: negative size");
    if (n_seqs == 0) {
        if (total_words) *total_words = 0;
        return CGM_OK;
    }
    hipStream_t s = nullptr;
    CGM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        std::vector<void *> bufs;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            for (void *b : bufs) (void)hipFree(b);
            (void)hipStreamDestroy(s);
        }
    } guard{s, {}};
    auto dalloc = [&](size_t bytes, void **d) -> int {
        CGM_HIP(hipMalloc(d, bytes > 0 ? bytes : 8));
        guard.bufs.push_back(*d);
        return CGM_OK;
    };
    void *d_ascii, *d_start, *d_len, *d_packed, *d_mask = nullptr, *d_off, *d_lens, *d_total;
    int rc = CGM_OK;
    if ((rc = dalloc((size_t)ascii_bytes, &d_ascii)) || (rc = dalloc((size_t)n_seqs * 8, &d_start)) ||
        (rc = dalloc((size_t)n_seqs * 4, &d_len)) || (rc = dalloc((size_t)words_cap * 8, &d_packed)) ||
        (nmask && (rc = dalloc((size_t)words_cap * 8, &d_mask))) ||
        (rc = dalloc((size_t)n_seqs * 4, &d_off)) || (rc = dalloc((size_t)n_seqs * 4, &d_lens)) ||
        (rc = dalloc(8, &d_total)))
        return rc;
    if (ascii_bytes > 0)
        CGM_HIP(hipMemcpyAsync(d_ascii, ascii, (size_t)ascii_bytes, hipMemcpyHostToDevice, s));
    CGM_HIP(hipMemcpyAsync(d_start, seq_start, (size_t)n_seqs * 8, hipMemcpyHostToDevice, s));
    CGM_HIP(hipMemcpyAsync(d_len, seq_len, (size_t)n_seqs * 4, hipMemcpyHostToDevice, s));
    if ((rc = cgm_encode_ascii_launch((const uint8_t *)d_ascii, ascii_bytes,
                                      (const int64_t *)d_start, (const int32_t *)d_len, n_seqs,
                                      (uint64_t *)d_packed, (uint64_t *)d_mask, words_cap,
                                      (int32_t *)d_off, (int32_t *)d_lens, (int64_t *)d_total, s)))
        return rc;
    int64_t tot = 0;
    CGM_HIP(hipMemcpyAsync(&tot, d_total, 8, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipStreamSynchronize(s));
    const int64_t nw = tot < words_cap ? tot : words_cap;
    if (nw > 0) {
        CGM_HIP(hipMemcpyAsync(packed, d_packed, (size_t)nw * 8, hipMemcpyDeviceToHost, s));
        if (nmask) CGM_HIP(hipMemcpyAsync(nmask, d_mask, (size_t)nw * 8, hipMemcpyDeviceToHost, s));
    }
    CGM_HIP(hipMemcpyAsync(word_off, d_off, (size_t)n_seqs * 4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipMemcpyAsync(lengths, d_lens, (size_t)n_seqs * 4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipStreamSynchronize(s));
    if (total_words) *total_words = tot;
    return CGM_OK;
}

}  // extern "C"
