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

#include "cgm_wfa_dev.h"

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

// WFA work key for the longest-first order: sum over scores of the
// diagonal count, from max_error (an upper bound of the cost), with bit 31
// set for pairs that need the large layout (they come first).  *nbig counts
// those pairs.
__global__ void wfa_key_kernel(Batch b, const int32_t *me_arr, int32_t me_scalar, int32_t me_cap,
                               WfaPen p, WfaGeom gm, uint32_t *keys, int32_t *idx,
                               int32_t *nbig) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= b.n_pairs) return;
    const int me = me_arr ? me_arr[i] : me_scalar;
    const int ql = b.q_len[i], rl = b.r_len[i];
    const bool small = me >= 0 && me <= me_cap && ql >= 0 && rl >= 0 &&
                       wfa_fits_small(ql, rl, me, p, gm);
    keys[i] = (small ? 0u : 0x80000000u) | wfa_work_key(me, p, gm);
    idx[i] = (int32_t)i;
    const uint64_t m = __ballot(!small);
    if (m && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1)
        atomicAdd(nbig, __popcll(m));
}

// ranges[0..1] = [0, nbig), ranges[2..3] = [nbig, n) of the sorted order
__global__ void wfa_split_ranges_kernel(const int32_t *nbig, int64_t n, int32_t *ranges) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    ranges[0] = 0;
    ranges[1] = *nbig;
    ranges[2] = *nbig;
    ranges[3] = (int32_t)n;
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

int launch_wfa_wave_ranges(const Batch &b, const int32_t *max_error, int32_t max_error_scalar,
                           int32_t me_cap, WfaPen p, int32_t *out, const int32_t *order,
                           const int32_t *range_big, const int32_t *range_small, int32_t *queues,
                           int32_t *slab, int64_t slab_words, int grid_big, int grid_small,
                           hipStream_t stream) {
    WfaGeom gm;
    wfa_geometry(p, &gm.g, &gm.nM, &gm.nI);
    if (int rc = set_kernel_lds((const void *)wfa_wave_kernel<kTeamThreadsBig>, kWfaLdsBytes))
        return rc;
    if (int rc = set_kernel_lds((const void *)wfa_wave_kernel<kTeamThreads>, kWfaSmallLdsBytes))
        return rc;
    if (range_big && grid_big > 0)
        hipLaunchKernelGGL(wfa_wave_kernel<kTeamThreadsBig>, dim3((unsigned)grid_big),
                           dim3(kTeamThreadsBig), (size_t)kWfaLdsBytes, stream, b, max_error,
                           max_error_scalar, me_cap, p, gm, out, order, range_big, queues, kWfaLdsBytes,
                           slab, slab_words);
    if (range_small && grid_small > 0)
        hipLaunchKernelGGL(wfa_wave_kernel<kTeamThreads>, dim3((unsigned)grid_small),
                           dim3(kTeamThreads), (size_t)kWfaSmallLdsBytes, stream, b, max_error,
                           max_error_scalar, me_cap, p, gm, out, order, range_small, queues + 1,
                           kWfaSmallLdsBytes, nullptr, (int64_t)0);
    CGM_HIP(hipGetLastError());
    return CGM_OK;
}

int launch_wfa_wave(const Batch &b, const int32_t *max_error, int32_t max_error_scalar, WfaPen p,
                    int32_t *out, const int32_t *order, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_wfa_pen(p)) return rc;
    if (b.n_pairs == 0) return CGM_OK;
    int32_t me_max = max_error_scalar;
    if (max_error) {
        if (int rc = device_max_i32(max_error, b.n_pairs, stream, &me_max)) return rc;
    }
    if (me_max < 0) me_max = 0;
    if (me_max > kMaxWfaError) me_max = kMaxWfaError;
    WfaGeom gm;
    wfa_geometry(p, &gm.g, &gm.nM, &gm.nI);
    const int slots = gm.nM + 2 * gm.nI;
    const int Kmax = me_max < reach_host(me_max, p) ? me_max : reach_host(me_max, p);
    const int64_t slab_words = (int64_t)slots * (2 * (int64_t)Kmax + 1);
    int cus = 0;
    if (int rc = device_cu_count(&cus)) return rc;
    int64_t grid = (int64_t)cus * kWfaBlocksPerCu;
    if (grid > b.n_pairs) grid = b.n_pairs;
    // keep the HBM slabs within 4 GiB
    while (grid > 1 && grid * slab_words * 4 > (int64_t(4) << 30)) grid /= 2;
    int64_t grid_small = (int64_t)cus * kWfaSmallBlocksPerCu;
    if (grid_small > b.n_pairs) grid_small = b.n_pairs;
    int32_t *slab = nullptr, *queue = nullptr, *own_order = nullptr;
    uint32_t *keys = nullptr;
    int32_t *idx = nullptr;
    CGM_HIP(hipMallocAsync((void **)&queue, 64, stream));
    CGM_HIP(hipMallocAsync((void **)&slab, (size_t)(grid * slab_words * 4), stream));
    // queue[0] / queue[1]: counters of the two launches; queue[2]: pairs of
    // the large layout; queue[4..7]: the two launches' ranges
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, stream, queue, 16);
    int rc = CGM_OK;
    if (!order) {
        // Longest first (a late long pair would otherwise leave the GPU
        // idle), the large-layout pairs ahead of the others: they run first,
        // on 512-thread workgroups, then the rest on small ones.
        CGM_HIP(hipMallocAsync((void **)&own_order, (size_t)b.n_pairs * 4, stream));
        CGM_HIP(hipMallocAsync((void **)&keys, (size_t)b.n_pairs * 4, stream));
        CGM_HIP(hipMallocAsync((void **)&idx, (size_t)b.n_pairs * 4, stream));
        hipLaunchKernelGGL(wfa_key_kernel, dim3((unsigned)((b.n_pairs + 255) / 256)), dim3(256), 0,
                           stream, b, max_error, max_error_scalar, me_max, p, gm, keys, idx,
                           queue + 2);
        rc = sort_desc_by_key(keys, idx, own_order, b.n_pairs, stream);
        hipLaunchKernelGGL(wfa_split_ranges_kernel, dim3(1), dim3(64), 0, stream, queue + 2,
                           b.n_pairs, queue + 4);
        if (rc == CGM_OK)
            rc = launch_wfa_wave_ranges(b, max_error, max_error_scalar, me_max, p, out, own_order,
                                        queue + 4, queue + 6, queue, slab, slab_words, (int)grid,
                                        (int)grid_small, stream);
    } else {
        // caller's order: every pair on the large layout, in that order
        // (queue[3] is 0, so queue[6..7] = [0, n))
        hipLaunchKernelGGL(wfa_split_ranges_kernel, dim3(1), dim3(64), 0, stream, queue + 3,
                           b.n_pairs, queue + 4);
        rc = launch_wfa_wave_ranges(b, max_error, max_error_scalar, me_max, p, out, order,
                                    queue + 6, nullptr, queue, slab, slab_words, (int)grid, 0,
                                    stream);
    }
    hipError_t le = hipGetLastError();
    CGM_HIP(hipFreeAsync(slab, stream));
    CGM_HIP(hipFreeAsync(queue, stream));
    if (own_order) {
        CGM_HIP(hipFreeAsync(own_order, stream));
        CGM_HIP(hipFreeAsync(keys, stream));
        CGM_HIP(hipFreeAsync(idx, stream));
    }
    if (rc) return rc;
    CGM_HIP(le);
    return CGM_OK;
}

}  // namespace cgm
