// cgm_wfa.hip -- gap-affine WFA kernels for gfx950.
//
// Semantics are those of wfa_affine_cpu (reference genomics/cpu/
// align_cpu.mojo:75-226): diagonals k = row - col, furthest-reaching query
// row per diagonal for the M / I / D wavefronts,
//     I[s,k] = max(M[s-o-e,k-1] + 1, I[s-e,k-1] + 1)
//     D[s,k] = max(M[s-o-e,k+1],     D[s-e,k+1])
//     M[s,k] = max(M[s-x,k] + 1, I[s,k], D[s,k])
// (missing operands skipped), a greedy extension along each diagonal after
// every score, diagonals limited to [max(-rlen,-s), min(qlen,s)] widened to
// the target diagonal qlen - rlen and clamped to +-max_error, and the result
// is the first s with M[s, qlen-rlen] >= qlen, or -1 past max_error.

#include <algorithm>

#include "cgm_wfa_reg.h"

namespace cgm {

namespace {

using namespace wfadev;

constexpr int kWfaNeg = -1000000;  // align_cpu.mojo:107

// ---------------------------------------------------------------------------
// Plain layout: one 128-thread workgroup per pair (grid-stride over pairs),
// circular M/I/D buffers of max(x,o+e)+1 and e+1 score slots over all
// 2*max_error+1 diagonals, every slot cleared per score, NEXT over the
// diagonal range, a barrier, then a base-by-base EXTEND, a barrier and the
// termination check -- the structure of the reference's own GPU kernel
// (genomics/gpu/align_gpu.mojo:261-453).  Wavefronts live in a global
// scratch slab per workgroup, so max_error is not limited by LDS.
// ---------------------------------------------------------------------------
constexpr int kPlainThreads = 128;

__global__ __launch_bounds__(kPlainThreads) void wfa_plain_kernel(Batch b, const int32_t *me_arr,
                                                                  int32_t me_scalar, int32_t me_cap,
                                                                  WfaPen p, int32_t *out,
                                                                  int32_t *scratch, int64_t slab,
                                                                  const int32_t *kind,
                                                                  int32_t *q_end, int32_t *r_end) {
    __shared__ int found;
    int32_t *M = scratch + (int64_t)blockIdx.x * slab;
    const int tid = threadIdx.x;
    const int m_hist = max(p.x, p.o + p.e) + 1;
    const int ie_hist = p.e + 1;
    for (int64_t pair = blockIdx.x; pair < b.n_pairs; pair += gridDim.x) {
        if (kind && kind[pair] != 1) continue;
        if (tid == 0 && q_end) q_end[pair] = -1;
        if (tid == 0 && r_end) r_end[pair] = -1;
        const uint64_t *q, *r;
        int qlen, rlen;
        const bool ok = load_pair(b, pair, q, r, qlen, rlen);
        const int me = me_arr ? me_arr[pair] : me_scalar;
        if (!ok || me < 0 || me > me_cap) {
            if (tid == 0) out[pair] = CGM_BAD_PAIR;
            continue;
        }
        if (qlen == 0 && rlen == 0) {
            if (tid == 0) out[pair] = 0;
            continue;
        }
        const int nd = 2 * me + 1;
        int32_t *I = M + (int64_t)m_hist * nd;
        int32_t *D = I + (int64_t)ie_hist * nd;
        for (int j = tid; j < m_hist * nd; j += kPlainThreads) M[j] = kWfaNeg;
        for (int j = tid; j < ie_hist * nd; j += kPlainThreads) {
            I[j] = kWfaNeg;
            D[j] = kWfaNeg;
        }
        __syncthreads();
        const int target = qlen - rlen;
        if (tid == 0) {
            int row = 0;
            while (row < qlen && row < rlen && base_at(q, row) == base_at(r, row)) ++row;
            M[me] = row;
            found = (target == 0 && row >= qlen) ? 1 : 0;
        }
        __syncthreads();
        int result = found ? 0 : -1;
        for (int s = 1; s <= me && result < 0; ++s) {
            const int ms = s % m_hist, is = s % ie_hist;
            for (int j = tid; j < nd; j += kPlainThreads) {
                M[ms * nd + j] = kWfaNeg;
                I[is * nd + j] = kWfaNeg;
                D[is * nd + j] = kWfaNeg;
            }
            __syncthreads();
            int lo = max(-rlen, -s), hi = min(qlen, s);
            if (target < lo || target > hi) {
                lo = min(lo, target);
                hi = max(hi, target);
            }
            lo = max(lo, -me);
            hi = min(hi, me);
            const int m_oe = (s >= p.o + p.e) ? ((s - p.o - p.e) % m_hist) * nd : -1;
            const int i_e = (s >= p.e) ? ((s - p.e) % ie_hist) * nd : -1;
            const int m_x = (s >= p.x) ? ((s - p.x) % m_hist) * nd : -1;
            for (int d = lo + tid; d <= hi; d += kPlainThreads) {
                const int di = d + me;
                int iv = kWfaNeg, dv = kWfaNeg, mv = kWfaNeg;
                if (di > 0) {
                    if (m_oe >= 0) {
                        int v = M[m_oe + di - 1];
                        if (v != kWfaNeg) iv = max(iv, v + 1);
                    }
                    if (i_e >= 0) {
                        int v = I[i_e + di - 1];
                        if (v != kWfaNeg) iv = max(iv, v + 1);
                    }
                }
                if (di < nd - 1) {
                    if (m_oe >= 0) {
                        int v = M[m_oe + di + 1];
                        if (v != kWfaNeg) dv = max(dv, v);
                    }
                    if (i_e >= 0) {
                        int v = D[i_e + di + 1];
                        if (v != kWfaNeg) dv = max(dv, v);
                    }
                }
                if (m_x >= 0) {
                    int v = M[m_x + di];
                    if (v != kWfaNeg) mv = max(mv, v + 1);
                }
                if (iv != kWfaNeg) mv = max(mv, iv);
                if (dv != kWfaNeg) mv = max(mv, dv);
                I[is * nd + di] = iv;
                D[is * nd + di] = dv;
                M[ms * nd + di] = mv;
            }
            __syncthreads();
            for (int d = lo + tid; d <= hi; d += kPlainThreads) {
                const int di = d + me;
                int row = M[ms * nd + di];
                if (row == kWfaNeg || row < 0) continue;
                int col = row - d;
                while (row < qlen && col >= 0 && col < rlen && base_at(q, row) == base_at(r, col)) {
                    ++row;
                    ++col;
                }
                M[ms * nd + di] = row;
                if (d == target && row >= qlen) found = 1;
            }
            __syncthreads();
            if (target >= lo && target <= hi && found) result = s;
            __syncthreads();
        }
        if (tid == 0) out[pair] = result;
        __syncthreads();
    }
}

constexpr int kTeamThreads = 256;     // workgroups of the small-LDS pairs
constexpr int kTeamThreadsBig = 512;  // workgroups of the other pairs
constexpr int kTeamUnroll = 4;        // diagonal groups per wavefront and pass of a score

// Persistent workgroups take items of `order` from the atomic counter
// `queue`: all items (part 0), the items before *split (part 1) or those
// from *split on (part 2).
// Both launches are LDS-bound at 6 (512 threads, 3 workgroups per CU) and 8
// (256 threads, 8 per CU) wavefronts per SIMD; the register budget is held
// to that occupancy (the tiled path for the rare pairs beyond LDS would
// otherwise raise the whole kernel to ~90 VGPRs, 5 wavefronts per SIMD).
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT == 512 ? 6 : 8)))
void wfa_wave_kernel(
    Batch b, const int32_t *me_arr, int32_t me_scalar, int32_t me_cap, WfaPen p, WfaGeom gm,
    int32_t *out, const int32_t *order, const int32_t *range, int32_t *queue,
    int32_t lds_bytes, int32_t *slab, int64_t slab_words) {
    extern __shared__ uint8_t lds_raw[];
    __shared__ int next_item;
    const int64_t first = range[0];
    const int64_t last = range[1];
    int32_t *my_slab = slab ? slab + (int64_t)blockIdx.x * slab_words : nullptr;
    for (;;) {
        if (threadIdx.x == 0) next_item = atomicAdd(queue, 1);
        __syncthreads();
        const int64_t item = first + next_item;
        __syncthreads();
        if (item >= last) break;
        const int64_t pair = order ? order[item] : item;
        const uint64_t *q, *r;
        int qlen, rlen;
        const bool ok = load_pair(b, pair, q, r, qlen, rlen);
        const int me = me_arr ? me_arr[pair] : me_scalar;
        const int res = (!ok || me < 0 || me > me_cap)
                            ? CGM_BAD_PAIR
                            : wfa_pair_block<kTeamUnroll, NT>(q, qlen, r, rlen, me, p, gm, lds_raw,
                                                              lds_bytes, my_slab, slab_words);
        if (threadIdx.x == 0) out[pair] = res;
        __syncthreads();
    }
}

// Register layout: one pair per workgroup of NW wavefronts (NB blocks of
// 64 diagonals each), pairs order[range[0] .. range[1]) from an atomic queue.
template <int NB, int NW, int WPE>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void wfa_reg_kernel(Batch b, const int32_t *me_arr, int32_t me_scalar, WfaPen p, int32_t g,
                    int32_t *out, const int32_t *order, const int32_t *range, int32_t *queue) {
    extern __shared__ uint8_t lds_raw[];
    __shared__ int next_item;
    const int64_t first = range[0];
    const int64_t last = range[1];
    for (;;) {
        if (threadIdx.x == 0) next_item = atomicAdd(queue, 1);
        __syncthreads();
        const int64_t item = first + next_item;
        __syncthreads();
        if (item >= last) break;
        const int64_t pair = order[item];
        const uint64_t *q, *r;
        int qlen, rlen;
        load_pair(b, pair, q, r, qlen, rlen);  // checked by the classifier
        const int me = me_arr ? me_arr[pair] : me_scalar;
        const int res = wfareg::wfa_reg_pair<NB, NW, wfareg::kRegDX, wfareg::kRegDO,
                                             wfareg::kRegDE>(q, qlen, r, rlen, me, p, g, lds_raw);
        if (threadIdx.x == 0) out[pair] = res;
        __syncthreads();
    }
}

// Layout class and longest-first key of every pair (class first); bad
// pairs get CGM_BAD_PAIR here and a key that sorts them last, outside every
// class range.  counts[c]: pairs of class c.
__global__ void wfa_class_kernel(Batch b, const int32_t *me_arr, int32_t me_scalar,
                                 int32_t me_cap, WfaPen p, WfaGeom gm, int32_t reg_pen,
                                 int32_t *out, uint32_t *keys, int32_t *idx, int32_t *counts) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool in = i < b.n_pairs;
    int cls = -1;
    uint32_t key = 0;
    if (in) {
        const uint64_t *q, *r;
        int ql, rl;
        const int me = me_arr ? me_arr[i] : me_scalar;
        if (load_pair(b, i, q, r, ql, rl) && me >= 0 && me <= me_cap) {
            cls = wfareg::wfa_pair_class(ql, rl, me, p, gm, reg_pen != 0);
            key = wfareg::wfa_class_key(cls, wfa_work_key(me, p, gm));
        } else {
            out[i] = CGM_BAD_PAIR;
        }
        keys[i] = key;
        idx[i] = (int32_t)i;
    }
    const int lane = threadIdx.x & 63;
    for (int c = 0; c < wfareg::kWfaClasses; ++c) {
        const uint64_t m = __ballot(cls == c);
        if (m && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&counts[c], __popcll(m));
    }
}

// ranges[2c], ranges[2c + 1]: class c's part of the sorted order, which
// starts at `start` (a device value, or 0 when null)
__global__ void wfa_class_ranges_kernel(const int32_t *counts, const int32_t *start,
                                        int32_t *ranges) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t s = start ? *start : 0;
    for (int c = 0; c < wfareg::kWfaClasses; ++c) {
        ranges[2 * c] = s;
        s += counts[c];
        ranges[2 * c + 1] = s;
    }
}

__global__ void zero_i32_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

}  // namespace

int launch_wfa_plain(const Batch &b, const int32_t *max_error, int32_t max_error_scalar, WfaPen p,
                     int32_t *out, hipStream_t stream) {
    return launch_wfa_plain_kind(b, max_error, max_error_scalar, p, out, nullptr, nullptr, nullptr,
                                 stream);
}

int launch_wfa_plain_kind(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                          WfaPen p, int32_t *out, const int32_t *kind, int32_t *q_end,
                          int32_t *r_end, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_wfa_pen(p)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int32_t me_max = max_error_scalar;
    if (max_error) {
        if (int rc = device_max_i32(max_error, b.n_pairs, stream, &me_max)) return rc;
    }
    if (me_max < 0) me_max = 0;
    if (me_max > kMaxWfaError) me_max = kMaxWfaError;
    const int m_hist = (p.x > p.o + p.e ? p.x : p.o + p.e) + 1;
    const int ie_hist = p.e + 1;
    const int64_t slab = (int64_t)(m_hist + 2 * ie_hist) * (2 * (int64_t)me_max + 1);
    int64_t grid = b.n_pairs < 2048 ? b.n_pairs : 2048;
    // Keep the scratch slabs within 4 GiB.
    while (grid > 1 && grid * slab * (int64_t)sizeof(int32_t) > (int64_t(4) << 30)) grid /= 2;
    int32_t *scratch = nullptr;
    CGM_HIP(hipMallocAsync((void **)&scratch, (size_t)(grid * slab * sizeof(int32_t)), stream));
    hipLaunchKernelGGL(wfa_plain_kernel, dim3((unsigned)grid), dim3(kPlainThreads), 0, stream, b,
                       max_error, max_error_scalar, me_max, p, out, scratch, slab, kind, q_end,
                       r_end);
    hipError_t le = hipGetLastError();
    CGM_HIP(hipFreeAsync(scratch, stream));
    CGM_HIP(le);
    return CGM_OK;
}

}  // namespace cgm

namespace cgm {

int wfa_geometry(const WfaPen &p, int *g, int *nM, int *nI) {
    auto gcd = [](int a, int b) {
        while (b) {
            int t = a % b;
            a = b;
            b = t;
        }
        return a;
    };
    *g = gcd(gcd(p.x, p.o), p.e);
    *nM = (p.x > p.o + p.e ? p.x : p.o + p.e) / *g + 1;
    *nI = p.e / *g + 1;
    return CGM_OK;
}

bool wfa_reg_penalties(const WfaPen &p) {
    int g = 0, nM = 0, nI = 0;
    wfa_geometry(p, &g, &nM, &nI);
    return p.x == wfareg::kRegDX * g && p.o + p.e == wfareg::kRegDO * g &&
           p.e == wfareg::kRegDE * g;
}

int launch_wfa_classify(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                        int32_t me_cap, WfaPen p, bool reg, int32_t *out, uint32_t *keys,
                        int32_t *idx, int32_t *counts, hipStream_t stream) {
    WfaGeom gm;
    wfa_geometry(p, &gm.g, &gm.nM, &gm.nI);
    hipLaunchKernelGGL(wfa_class_kernel, dim3((unsigned)((b.n_pairs + 255) / 256)), dim3(256), 0,
                       stream, b, max_error, max_error_scalar, me_cap, p, gm,
                       (int32_t)(reg && wfa_reg_penalties(p)), out, keys, idx, counts);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_wfa_class_ranges(const int32_t *counts, const int32_t *start, int32_t *ranges,
                            hipStream_t stream) {
    hipLaunchKernelGGL(wfa_class_ranges_kernel, dim3(1), dim3(64), 0, stream, counts, start,
                       ranges);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_wfa_classes(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                       int32_t me_cap, WfaPen p, bool reg, int32_t *out, const int32_t *order,
                       const int32_t *ranges, int32_t *queues, int32_t *slab, int64_t slab_words,
                       int grid_big, hipStream_t s_long, hipStream_t s_short) {
    using namespace wfareg;
    WfaGeom gm;
    wfa_geometry(p, &gm.g, &gm.nM, &gm.nI);
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    if (int rc = set_kernel_lds((const void *)wfa_wave_kernel<kTeamThreadsBig>, kWfaLdsBytes))
        return rc;
    if (int rc = set_kernel_lds((const void *)wfa_wave_kernel<kTeamThreads>, kWfaSmallLdsBytes))
        return rc;
    // the longest pairs (team layout beyond the register classes, then the
    // widest register classes) on one stream, the rest on the other, so that
    // the two halves fill each other's tails
    if (grid_big > 0)
        hipLaunchKernelGGL(wfa_wave_kernel<kTeamThreadsBig>, dim3((unsigned)grid_big),
                           dim3(kTeamThreadsBig), (size_t)kWfaLdsBytes, s_long, b, max_error,
                           max_error_scalar, me_cap, p, gm, out, order, ranges + 2 * kWfaTeamBig,
                           queues + kWfaTeamBig, kWfaLdsBytes, slab, slab_words);
#define CGM_REG_CLASS(C, STREAM)                                                                 \
    do {                                                                                         \
        constexpr int NB = reg_class_nb(C), NW = reg_class_nw(C), WPE = reg_class_wpe(C);       \
        constexpr int LDS = reg_class_lds(C);                                                    \
        const void *fn = (const void *)wfa_reg_kernel<NB, NW, WPE>;                             \
        if (int rc = set_kernel_lds(fn, LDS)) return rc;                                         \
        const int per_cu = std::min(4 * WPE / NW, 160 * 1024 / LDS);                            \
        hipLaunchKernelGGL((wfa_reg_kernel<NB, NW, WPE>), dim3((unsigned)(cus * per_cu)),       \
                           dim3(NW * 64), (size_t)LDS, STREAM, b, max_error, max_error_scalar, p, \
                           gm.g, out, order, ranges + 2 * (C), queues + (C));                    \
    } while (0)
    if (reg && wfa_reg_penalties(p)) {
        CGM_REG_CLASS(6, s_long);
        CGM_REG_CLASS(5, s_long);
        CGM_REG_CLASS(4, s_long);
        CGM_REG_CLASS(3, s_short);
        CGM_REG_CLASS(2, s_short);
        CGM_REG_CLASS(1, s_short);
        CGM_REG_CLASS(0, s_short);
    }
#undef CGM_REG_CLASS
    hipLaunchKernelGGL(wfa_wave_kernel<kTeamThreads>, dim3((unsigned)(cus * kWfaSmallBlocksPerCu)),
                       dim3(kTeamThreads), (size_t)kWfaSmallLdsBytes, s_short, b, max_error,
                       max_error_scalar, me_cap, p, gm, out, order, ranges + 2 * kWfaTeamSmall,
                       queues + kWfaTeamSmall, kWfaSmallLdsBytes, nullptr, (int64_t)0);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int wfa_slab_words(int32_t me_cap, const WfaPen &p) {
    int g = 0, nM = 0, nI = 0;
    wfa_geometry(p, &g, &nM, &nI);
    const int K = me_cap < reach_host(me_cap, p) ? me_cap : reach_host(me_cap, p);
    return (nM + 2 * nI) * (2 * K + 1);
}

int launch_wfa_wave(const Batch &b, const int32_t *max_error, int32_t max_error_scalar, WfaPen p,
                    int32_t *out, bool reg, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_wfa_pen(p)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int32_t me_max = max_error_scalar;
    if (max_error) {
        if (int rc = device_max_i32(max_error, b.n_pairs, stream, &me_max)) return rc;
    }
    if (me_max < 0) me_max = 0;
    if (me_max > kMaxWfaError) me_max = kMaxWfaError;
    const int64_t slab_words = wfa_slab_words(me_max, p);
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    int64_t grid = (int64_t)cus * kWfaBlocksPerCu;
    // keep the HBM slabs within 4 GiB
    while (grid > 1 && grid * slab_words * 4 > (int64_t(4) << 30)) grid /= 2;
    // small: queue counters [0, C), class counts [C, 2C), class ranges [2C, 4C)
    constexpr int C = wfareg::kWfaClasses;
    int32_t *slab = nullptr, *small = nullptr, *order = nullptr, *idx = nullptr;
    uint32_t *keys = nullptr;
    CGM_HIP(hipMallocAsync((void **)&small, 4 * C * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&slab, (size_t)(grid * slab_words * 4), stream));
    CGM_HIP(hipMallocAsync((void **)&order, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&keys, (size_t)b.n_pairs * 4, stream));
    CGM_HIP(hipMallocAsync((void **)&idx, (size_t)b.n_pairs * 4, stream));
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, stream, small, 4 * C);
    // every class longest first (a late long pair would leave the GPU idle)
    int rc = launch_wfa_classify(b, max_error, max_error_scalar, me_max, p, reg, out, keys, idx,
                                 small + C, stream);
    if (!rc) rc = sort_desc_by_key(keys, idx, order, b.n_pairs, stream);
    if (!rc) rc = launch_wfa_class_ranges(small + C, nullptr, small + 2 * C, stream);
    hipStream_t aux[2];
    if (!rc) rc = fork_streams(stream, 2, aux);
    if (!rc) {
        rc = launch_wfa_classes(b, max_error, max_error_scalar, me_max, p, reg, out, order,
                                small + 2 * C, small, slab, slab_words, (int)grid, aux[0], aux[1]);
        int rc2 = join_streams(stream, 2, aux);
        if (!rc) rc = rc2;
    }
    hipError_t f1 = hipFreeAsync(slab, stream), f2 = hipFreeAsync(small, stream);
    hipError_t f3 = hipFreeAsync(order, stream), f4 = hipFreeAsync(keys, stream);
    hipError_t f5 = hipFreeAsync(idx, stream);
    if (rc) return rc;
    CGM_HIP(f1);
    CGM_HIP(f2);
    CGM_HIP(f3);
    CGM_HIP(f4);
    CGM_HIP(f5);
    return CGM_OK;
}

}  // namespace cgm
