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

#include "cgm_internal.h"

namespace cgm {

namespace {

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
                                                                  int32_t *scratch, int64_t slab) {
    __shared__ int found;
    int32_t *M = scratch + (int64_t)blockIdx.x * slab;
    const int tid = threadIdx.x;
    const int m_hist = max(p.x, p.o + p.e) + 1;
    const int ie_hist = p.e + 1;
    for (int64_t pair = blockIdx.x; pair < b.n_pairs; pair += gridDim.x) {
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

}  // namespace

int launch_wfa_plain(const Batch &b, const int32_t *max_error, int32_t max_error_scalar, WfaPen p,
                     int32_t *out, hipStream_t stream) {
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
                       max_error, max_error_scalar, me_max, p, out, scratch, slab);
    hipError_t le = hipGetLastError();
    CGM_HIP(hipFreeAsync(scratch, stream));
    CGM_HIP(le);
    return CGM_OK;
}

}  // namespace cgm
