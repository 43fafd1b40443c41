// cgm_cigar.hip -- gap-affine WFA with traceback (CIGAR) for pair batches.
//
// Semantics of wfa_affine_cigar_cpu (genomics/cpu/align_cpu.mojo:380-623):
// the optimal global cost (as wfa_affine_cpu) and the CIGAR its backtrace
// picks, as run-length operations run = (count << 2) | op with op 0 '=',
// 1 'X', 2 'I' (query base), 3 'D' (reference base), first operation first.
//
// Pipeline:
//  1. costs: the automatic WFA layout (persistent kernel, cached plan);
//  2. per pair, the size of its full wavefront history: one row of
//     2K + 1 diagonals (K = the reach of its cost) per score step, for M
//     (after extension), I and D, int16 when every row offset fits;
//  3. pairs ordered by history size, largest first (the longest pairs
//     start first, so the launch does not end on one of them), in chunks
//     whose histories fit the workspace budget (half the free device
//     memory, at least 8 GiB and at most 96 GiB); per chunk,
//     one 256-thread workgroup per pair recomputes the forward pass (the
//     same exact reductions as the cost kernels: only multiples of
//     gcd(x, o, e), only diagonals in the score's range
//     [max(-rlen, -me, -R(s)), min(qlen, me, R(s))], R(s) = (s - o) / e),
//     keeping the score rows it still reads in an LDS ring (as the cost
//     kernels do) and writing each row's range to its HBM history, then
//     thread 0 walks the reference's backtrace.  Pairs whose ring does not
//     fit LDS read the earlier rows back from the history instead.
// A diagonal outside a row's range holds no reachable state, so the
// backtrace reads it as invalid (the reference's NEG_INF) without touching
// memory; rows are written only inside their range.  M before extension
// (the reference's M_pre) is recomputed from the stored rows with the
// forward formula.

#include <algorithm>
#include <mutex>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cgm_wfa_dev.h"

namespace cgm {

namespace {

constexpr int kCigarThreads = 512;       // workgroups of the pairs beyond kCigarSmallLds
constexpr int kCigarSmallThreads = 256;  // workgroups of the pairs within it
constexpr int kCigarUnroll = 4;  // diagonal groups per wavefront and pass of a score

// Per-device history workspace kept between calls (see cgm_wfa_cigar_launch).
constexpr int kCigarMaxDevices = 64;
struct CigarWorkspace {
    std::mutex mu;
    uint8_t *ptr = nullptr;
    size_t bytes = 0;
};
CigarWorkspace g_cigar_ws[kCigarMaxDevices];
constexpr int kCigarLds = 52 * 1024;  // ring + sequences per workgroup (3 per CU)
// pairs whose branch-free forward layout fits this run on 256-thread
// workgroups with it (8 per CU)
constexpr int kCigarSmallLds = 20 * 1024;
constexpr int64_t kCigarBudgetMin = int64_t(8) << 30;
constexpr int64_t kCigarBudgetMax = int64_t(96) << 30;

struct CigarGeom {
    int64_t K, W, rows, esize;  // K = 0 and rows = 0: nothing to store
};

__host__ __device__ __forceinline__ int64_t reach64(int64_t s, const WfaPen &p) {
    return s >= p.o + p.e ? (s - p.o) / p.e : 0;
}

__host__ __device__ __forceinline__ CigarGeom cigar_geom(int qlen, int rlen, int me, int cost,
                                                         const WfaPen &p, int g) {
    CigarGeom c{0, 0, 0, 0};
    if (cost < 0 || (qlen == 0 && rlen == 0)) return c;
    c.K = min(min((int64_t)me, (int64_t)max(qlen, rlen)), reach64(cost, p));
    c.W = 2 * c.K + 1;
    c.rows = cost / g + 1;
    // rows grow past qlen only by +1 steps (mismatch x, insertion e)
    c.esize = ((int64_t)qlen + cost / min(p.x, p.e) + 2 <= 32767) ? 2 : 4;
    return c;
}

// History bytes of each pair, and small[i] = 1 when its branch-free forward
// layout fits small_lds (the pair then runs on a small workgroup).
__global__ void cigar_size_kernel(Batch b, const int32_t *me_arr, int32_t me_scalar, WfaPen p,
                                  wfadev::WfaGeom gm, int32_t small_lds, const int32_t *cost,
                                  int64_t *bytes, int32_t *small) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= b.n_pairs) return;
    const uint64_t *q, *r;
    int qlen = 0, rlen = 0;
    int64_t v = 0;
    int s = 0;
    if (load_pair(b, i, q, r, qlen, rlen)) {
        const CigarGeom c =
            cigar_geom(qlen, rlen, me_arr ? me_arr[i] : me_scalar, cost[i], p, gm.g);
        v = 3 * c.rows * c.W * c.esize;
        v = (v + 255) & ~(int64_t)255;
        s = c.esize == 2 && c.W > 0 &&
            wfadev::team_fast_bytes(qlen, rlen, (int)c.W, gm) <= small_lds;
    }
    bytes[i] = v;
    small[i] = s;
}

template <typename T>
__device__ __forceinline__ int hist_get(const T *row, int64_t K, int64_t k) {
    if (k < -K || k > K) return -1;
    return (int)row[k + K];
}

// Diagonal range of score step si (score si * g): the only entries of the
// row that the forward pass writes.
struct RowRange {
    int qlen, rlen, me, g;
    WfaPen p;
    __device__ __forceinline__ bool has(int si, int64_t k) const {
        const int64_t R = reach64((int64_t)si * g, p);
        const int64_t lo = max(max((int64_t)-rlen, (int64_t)-me), -R);
        const int64_t hi = min(min((int64_t)qlen, (int64_t)me), R);
        return k >= lo && k <= hi;
    }
};

// Backtrace of align_cpu.mojo:510-600 over the stored rows; writes the runs
// (reversed, then flipped) and returns their number, or -2 past `cap`.
template <typename T>
__device__ int cigar_backtrace(const T *M, const T *I, const T *D, int64_t K, int64_t W,
                               int qlen, int rlen, int me, int cost, const WfaPen &p, int g,
                               uint32_t *out, int64_t cap) {
    const int dx = p.x / g, dO = (p.o + p.e) / g, de = p.e / g;
    const RowRange rr{qlen, rlen, me, g, p};
    // entry k of step si of A, or -1 outside the step's range
    struct Row {
        const T *base;
        int si;
    };
    auto row_of = [&](const T *A, int si) { return Row{A + (int64_t)si * W, si}; };
    auto hist_get = [&](Row row, int64_t KK, int64_t k) -> int {
        if (!rr.has(row.si, k)) return -1;
        return cgm::hist_get(row.base, KK, k);
    };
    // M before extension at (si, k): the forward formula over stored rows
    auto m_pre = [&](int si, int64_t k) -> int {
        if (si == 0) return k == 0 ? 0 : -1;
        int v = -1;
        if (si >= dx) {
            const int a = hist_get(row_of(M, si - dx), K, k);
            if (a >= 0) v = max(v, a + 1);
        }
        const int iv = hist_get(row_of(I, si), K, k);
        if (iv >= 0) v = max(v, iv);
        const int dv = hist_get(row_of(D, si), K, k);
        if (dv >= 0) v = max(v, dv);
        return v;
    };
    int64_t nr = 0;
    int run_op = -1;
    int64_t run_cnt = 0;
    auto emit = [&](int op, int64_t cnt) {
        if (cnt <= 0) return;
        if (op == run_op) {
            run_cnt += cnt;
            return;
        }
        if (run_op >= 0) {
            if (nr < cap) out[nr] = (uint32_t)((run_cnt << 2) | run_op);
            ++nr;
        }
        run_op = op;
        run_cnt = cnt;
    };
    int si = cost / g;
    int64_t cur_d = qlen - rlen;
    int cur_matrix = 0;
    int64_t cur_row = hist_get(row_of(M, si), K, cur_d);
    for (;;) {
        if (cur_matrix == 0) {
            const int pre_r = m_pre(si, cur_d);
            emit(0, cur_row - pre_r);
            cur_row = pre_r;
            if (si == 0 && cur_d == 0) break;
            if (si >= dx) {
                const int prev_m = hist_get(row_of(M, si - dx), K, cur_d);
                if (prev_m >= 0 && prev_m + 1 == cur_row) {
                    emit(1, 1);
                    si -= dx;
                    cur_row = prev_m;
                    continue;
                }
            }
            const int i_here = hist_get(row_of(I, si), K, cur_d);
            if (i_here >= 0 && i_here == cur_row) {
                cur_matrix = 1;
                continue;
            }
            const int d_here = hist_get(row_of(D, si), K, cur_d);
            if (d_here >= 0 && d_here == cur_row) {
                cur_matrix = 2;
                continue;
            }
            break;
        } else if (cur_matrix == 1) {
            emit(2, 1);
            const int64_t prev_d = cur_d - 1;
            if (si >= dO) {
                const int mv = hist_get(row_of(M, si - dO), K, prev_d);
                if (mv >= 0 && mv + 1 == cur_row) {
                    si -= dO;
                    cur_d = prev_d;
                    cur_row = mv;
                    cur_matrix = 0;
                    continue;
                }
            }
            if (si >= de) {
                const int iv = hist_get(row_of(I, si - de), K, prev_d);
                if (iv >= 0 && iv + 1 == cur_row) {
                    si -= de;
                    cur_d = prev_d;
                    cur_row = iv;
                    continue;
                }
            }
            break;
        } else {
            emit(3, 1);
            const int64_t prev_d = cur_d + 1;
            if (si >= dO) {
                const int mv = hist_get(row_of(M, si - dO), K, prev_d);
                if (mv >= 0 && mv == cur_row) {
                    si -= dO;
                    cur_d = prev_d;
                    cur_row = mv;
                    cur_matrix = 0;
                    continue;
                }
            }
            if (si >= de) {
                const int dv = hist_get(row_of(D, si - de), K, prev_d);
                if (dv >= 0 && dv == cur_row) {
                    si -= de;
                    cur_d = prev_d;
                    cur_row = dv;
                    continue;
                }
            }
            break;
        }
    }
    if (run_op >= 0) {
        if (nr < cap) out[nr] = (uint32_t)((run_cnt << 2) | run_op);
        ++nr;
    }
    if (nr > cap) return -2;
    for (int64_t a = 0, z = nr - 1; a < z; ++a, --z) {
        const uint32_t t = out[a];
        out[a] = out[z];
        out[z] = t;
    }
    return (int)nr;
}

// Forward pass with the score rows it still reads in an LDS ring of
// nM M-rows and nI I- and D-rows (slot of step si: si mod n), as the cost
// kernels keep them; every row's range is also written to the HBM history
// (row si of M, I and D).  Ranges grow with the score, so a ring slot's
// entries outside the current range were never written and still hold the
// initial invalid value.
template <typename T>
__device__ void cigar_forward_ring(const uint64_t *q, int qlen, const uint64_t *r, int rlen, int me,
                                   int cost, const WfaPen &p, const wfadev::WfaGeom &gm, T *M,
                                   T *I, T *D, int64_t K, int64_t W, T *ring) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const T neg = wfadev::RowNeg<T>::v;
    const int w = (int)W, kk = (int)K;
    T *RM = ring, *RI = ring + gm.nM * w, *RD = RI + gm.nI * w;
    for (int i = tid; i < (gm.nM + 2 * gm.nI) * w; i += nt) ring[i] = neg;
    __syncthreads();
    if (tid == 0) {
        const T v = (T)wfadev::extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
        RM[kk] = v;
        M[kk] = v;
        I[kk] = neg;
        D[kk] = neg;
    }
    __syncthreads();
    wfadev::WfaSteps st;
    st.init(p, gm);
    const int nrows = cost / gm.g + 1;
    for (int si = 1; si < nrows; ++si) {
        st.next(p, gm);
        const int lo = max(max(-rlen, -me), -st.R), hi = min(min(qlen, me), st.R);
        const int ms = st.ms * w, is = st.is * w;
        const int mx = st.mx(gm) >= 0 ? st.mx(gm) * w : -1;
        const int mo = st.mo(gm) >= 0 ? st.mo(gm) * w : -1;
        const int ie = st.ie(gm) >= 0 ? st.ie(gm) * w : -1;
        T *Mr = M + (int64_t)si * W, *Ir = I + (int64_t)si * W, *Dr = D + (int64_t)si * W;
        for (int k = lo + tid; k <= hi; k += nt) {
            const int di = k + kk;
            int a = wfadev::kNegRow, c = wfadev::kNegRow, bb = wfadev::kNegRow,
                d = wfadev::kNegRow, m = wfadev::kNegRow;
            if (mo >= 0) {
                if (di > 0) a = (int)RM[mo + di - 1];
                if (di < w - 1) bb = (int)RM[mo + di + 1];
            }
            if (ie >= 0) {
                if (di > 0) c = (int)RI[ie + di - 1];
                if (di < w - 1) d = (int)RD[ie + di + 1];
            }
            if (mx >= 0) m = (int)RM[mx + di];
            const int iv = max(a, c) + 1;
            const int dv = max(bb, d);
            int mv = max(max(m + 1, iv), dv);
            const T ti = iv < 0 ? neg : (T)iv, td = dv < 0 ? neg : (T)dv;
            RI[is + di] = ti;
            RD[is + di] = td;
            Ir[di] = ti;
            Dr[di] = td;
            if (mv >= 0) mv = wfadev::extend_diag(q, qlen, qw, r, rlen, rw, mv, k);
            const T tm = mv < 0 ? neg : (T)mv;
            RM[ms + di] = tm;
            Mr[di] = tm;
        }
        __syncthreads();
    }
}

// Forward pass in the branch-free LDS layout of the cost kernels
// (wfadev::wfa_pair_team_fast: LSB-first sequence words, guard-padded int16
// ring rows, I and D side by side), writing every step's range to the int16
// HBM history rows as cigar_forward_ring does.  `lds` holds at least
// wfadev::team_fast_bytes(qlen, rlen, W, gm).
template <int U, int NT>
__device__ void cigar_forward_fast(const uint64_t *qg, int qlen, const uint64_t *rg, int rlen,
                                   int me, int cost, const WfaPen &p, const wfadev::WfaGeom &gm,
                                   int16_t *M, int16_t *I, int16_t *D, int K, int W, uint8_t *lds) {
    using namespace wfadev;
    const int tid = threadIdx.x, lane = tid & 63;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const int nq = 2 * qw + 4, nr = 2 * rw + 4;
    uint32_t *qs = (uint32_t *)lds;
    uint32_t *rs = qs + nq;
    for (int i = tid; i < nq + nr; i += NT) {
        const bool isq = i < nq;
        const int j = isq ? i : i - nq;
        uint32_t v = 0;
        if ((j >> 1) < (isq ? qw : rw)) {
            const uint64_t w = (isq ? qg : rg)[j >> 1];
            v = __builtin_bitreverse32((j & 1) ? (uint32_t)w : (uint32_t)(w >> 32));
        }
        qs[i] = v;
    }
    const int WP = W + 2;
    const int m_ents = (gm.nM * WP + 1) & ~1;
    int16_t *Mb = (int16_t *)(lds + 4 * (nq + nr));
    int16_t *IDb = Mb + m_ents;
    const int16_t neg = RowNeg<int16_t>::v;
    {
        const int words = (m_ents + 2 * gm.nI * WP + kFastSlack / 2) >> 1;
        uint32_t *w32 = (uint32_t *)Mb;
        for (int i = tid; i < words; i += NT) w32[i] = 0x80008000u;
    }
    __syncthreads();
    if (tid == 0) {
        const int16_t v = (int16_t)extend_lsb(qs, qlen, rs, rlen, 0, 0);
        Mb[K + 1] = v;
        M[K] = v;
        I[K] = neg;
        D[K] = neg;
    }
    __syncthreads();
    const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);
    WfaSteps st;
    st.init(p, gm);
    const int nrows = cost / gm.g + 1;
    for (int si = 1; si < nrows; ++si) {
        st.next(p, gm);
        const int lo = max(max(-rlen, -me), -st.R), hi = min(min(qlen, me), st.R);
        const int16_t *Mo = Mb + WfaSteps::back(st.ms, st.dO, gm.nM) * WP;
        const int16_t *Mx = Mb + WfaSteps::back(st.ms, st.dx, gm.nM) * WP;
        const int16_t *IDe = IDb + WfaSteps::back(st.is, st.de, gm.nI) * 2 * WP;
        int16_t *Ms = Mb + st.ms * WP;
        int16_t *IDs = IDb + st.is * 2 * WP;
        int16_t *Mr = M + (int64_t)si * W, *Ir = I + (int64_t)si * W, *Dr = D + (int64_t)si * W;
        for (int kb = lo + wave0; kb <= hi; kb += U * NT) {
            const int di0 = kb + lane + K;  // row entry di + 1 holds diagonal di - K
            int mv[U], mr[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kb + u * NT > hi) break;
                const int di = di0 + u * NT;
                const int m = Mx[di + 1];
                __builtin_amdgcn_sched_barrier(0);  // keep the M read first
                const int a = Mo[di], b = Mo[di + 2];
                const int c = IDe[2 * di], d = IDe[2 * di + 5];
                const int iv = max(a, c) + 1;
                const int dv = max(b, d);  // negative only when both are `neg`
                mv[u] = max(max(m + 1, iv), dv);
                if (di - K <= hi) {
                    const int16_t ti = iv < 0 ? neg : (int16_t)iv;
                    *(uint32_t *)(IDs + 2 * di + 2) = (uint32_t)(uint16_t)ti | ((uint32_t)dv << 16);
                    Ir[di] = ti;
                    Dr[di] = (int16_t)dv;
                }
            }
            uint64_t more = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kb + u * NT > hi) break;
                const int k = di0 + u * NT - K;
                const int row = mv[u], col = row - k;
                uint32_t a0, a1, b0, b1;
                window_lsb(qs, med3(row, 0, qlen), a0, a1);
                window_lsb(rs, med3(col, 0, rlen), b0, b1);
                const int lim = min(qlen - row, rlen - col);
                const int n = med3(match32(a0, a1, b0, b1), lim, 0);  // match >= 0
                mr[u] = row + n;
                if (__ballot(row >= 0 && n == 32 && lim > 32)) more |= 1ull << u;
            }
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!(more >> u & 1)) continue;
                    const int k = di0 + u * NT - K;
                    const int row = mv[u], col = row - k;
                    if (row >= 0 && mr[u] == row + 32 && min(qlen - row, rlen - col) > 32 && k <= hi)
                        mr[u] = extend_lsb(qs, qlen, rs, rlen, row + 32, col + 32);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kb + u * NT > hi) break;
                const int di = di0 + u * NT;
                if (di - K <= hi) {
                    const int16_t tm = mv[u] < 0 ? neg : (int16_t)mr[u];
                    Ms[di + 1] = tm;
                    Mr[di] = tm;
                }
            }
        }
        __syncthreads();
    }
}

// Forward pass with full history, one workgroup per pair; rows of score
// step si hold diagonals -K..K (entries outside the step's range invalid).
template <typename T>
__device__ void cigar_forward(const uint64_t *q, int qlen, const uint64_t *r, int rlen, int me,
                              int cost, const WfaPen &p, int g, T *M, T *I, T *D, int64_t K,
                              int64_t W) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const T neg = wfadev::RowNeg<T>::v;
    const int dx = p.x / g, dO = (p.o + p.e) / g, de = p.e / g;
    for (int64_t i = tid; i < W; i += nt) {
        M[i] = neg;
        I[i] = neg;
        D[i] = neg;
    }
    __syncthreads();
    if (tid == 0) M[K] = (T)wfadev::extend_diag(q, qlen, qw, r, rlen, rw, 0, 0);
    __syncthreads();
    const int nrows = cost / g + 1;
    for (int si = 1; si < nrows; ++si) {
        const int64_t s = (int64_t)si * g;
        const int64_t Rs = reach64(s, p);
        const int64_t lo = max(max((int64_t)-rlen, (int64_t)-me), -Rs);
        const int64_t hi = min(min((int64_t)qlen, (int64_t)me), Rs);
        T *Mr = M + si * W, *Ir = I + si * W, *Dr = D + si * W;
        const T *Mx = si >= dx ? M + (si - dx) * W : nullptr;
        const T *Mo = si >= dO ? M + (si - dO) * W : nullptr;
        const T *Ie = si >= de ? I + (si - de) * W : nullptr;
        const T *De = si >= de ? D + (si - de) * W : nullptr;
        for (int64_t k = -K + tid; k <= K; k += nt) {
            const int64_t di = k + K;
            if (k < lo || k > hi) {
                Mr[di] = neg;
                Ir[di] = neg;
                Dr[di] = neg;
                continue;
            }
            int a = -1, c = -1, bb = -1, d = -1, m = -1;
            if (Mo) {
                if (di > 0) a = (int)Mo[di - 1];
                if (di < W - 1) bb = (int)Mo[di + 1];
            }
            if (Ie) {
                if (di > 0) c = (int)Ie[di - 1];
                if (di < W - 1) d = (int)De[di + 1];
            }
            if (Mx) m = (int)Mx[di];
            const int iv = max(a < 0 ? -1 : a + 1, c < 0 ? -1 : c + 1);
            const int dv = max(bb, d);
            int mv = max(max(m < 0 ? -1 : m + 1, iv), dv);
            Ir[di] = iv < 0 ? neg : (T)iv;
            Dr[di] = dv < 0 ? neg : (T)dv;
            if (mv >= 0) mv = wfadev::extend_diag(q, qlen, qw, r, rlen, rw, mv, (int)k);
            Mr[di] = mv < 0 ? neg : (T)mv;
        }
        __syncthreads();
    }
}

template <int NT, int LDS>
__global__ __launch_bounds__(NT) void wfa_cigar_kernel(
    Batch b, const int32_t *me_arr, int32_t me_scalar, WfaPen p, wfadev::WfaGeom gm,
    const int32_t *cost_arr, const int64_t *hist_off, uint8_t *ws, const int32_t *order,
    int64_t p0, uint32_t *cigar, const int64_t *cigar_off, int32_t *cigar_len) {
    extern __shared__ uint64_t lds[];  // LDS bytes: ring, then sequences
    const int g = gm.g;
    const int64_t pair = order[p0 + blockIdx.x];
    const int cost = cost_arr[pair];
    const uint64_t *qg, *rg;
    int qlen = 0, rlen = 0;
    if (cost < 0 || !load_pair(b, pair, qg, rg, qlen, rlen) || (qlen == 0 && rlen == 0)) {
        if (threadIdx.x == 0) cigar_len[pair] = 0;
        return;
    }
    const int me = me_arr ? me_arr[pair] : me_scalar;
    const CigarGeom cg = cigar_geom(qlen, rlen, me, cost, p, g);
    uint8_t *h = ws + hist_off[pair];
    const int64_t plane = cg.rows * cg.W;
    uint32_t *out = cigar + cigar_off[pair];
    const int64_t cap = cigar_off[pair + 1] - cigar_off[pair];
    if (cg.esize == 2 && wfadev::team_fast_bytes(qlen, rlen, (int)cg.W, gm) <= LDS) {
        int16_t *M = (int16_t *)h;
        cigar_forward_fast<kCigarUnroll, NT>(qg, qlen, rg, rlen, me, cost, p, gm, M,
                                                        M + plane, M + 2 * plane, (int)cg.K,
                                                        (int)cg.W, (uint8_t *)lds);
        if (threadIdx.x == 0)
            cigar_len[pair] = cigar_backtrace<int16_t>(M, M + plane, M + 2 * plane, cg.K, cg.W,
                                                       qlen, rlen, me, cost, p, g, out, cap);
        return;
    }
    const int qw = (qlen + 31) >> 5, rw = (rlen + 31) >> 5;
    const uint64_t *q = qg, *r = rg;
    // LDS: the ring when it fits, then the sequences when they fit after it
    const int64_t ring_bytes =
        (((int64_t)gm.nM + 2 * gm.nI) * cg.W * cg.esize + 15) & ~(int64_t)15;
    const bool ring = ring_bytes <= LDS;
    const int64_t seq_at = ring ? ring_bytes : 0;
    if (seq_at + 8 * (qw + rw) <= LDS) {
        uint64_t *seq = lds + seq_at / 8;
        for (int i = threadIdx.x; i < qw + rw; i += blockDim.x) seq[i] = i < qw ? qg[i] : rg[i - qw];
        q = seq;
        r = seq + qw;
        __syncthreads();
    }
    if (cg.esize == 2) {
        int16_t *M = (int16_t *)h;
        if (ring)
            cigar_forward_ring<int16_t>(q, qlen, r, rlen, me, cost, p, gm, M, M + plane,
                                        M + 2 * plane, cg.K, cg.W, (int16_t *)lds);
        else
            cigar_forward<int16_t>(q, qlen, r, rlen, me, cost, p, g, M, M + plane, M + 2 * plane,
                                   cg.K, cg.W);
        if (threadIdx.x == 0)
            cigar_len[pair] = cigar_backtrace<int16_t>(M, M + plane, M + 2 * plane, cg.K, cg.W,
                                                       qlen, rlen, me, cost, p, g, out, cap);
    } else {
        int32_t *M = (int32_t *)h;
        if (ring)
            cigar_forward_ring<int32_t>(q, qlen, r, rlen, me, cost, p, gm, M, M + plane,
                                        M + 2 * plane, cg.K, cg.W, (int32_t *)lds);
        else
            cigar_forward<int32_t>(q, qlen, r, rlen, me, cost, p, g, M, M + plane, M + 2 * plane,
                                   cg.K, cg.W);
        if (threadIdx.x == 0)
            cigar_len[pair] = cigar_backtrace<int32_t>(M, M + plane, M + 2 * plane, cg.K, cg.W,
                                                       qlen, rlen, me, cost, p, g, out, cap);
    }
}

}  // namespace

}  // namespace cgm

using namespace cgm;

extern "C" {

int cgm_wfa_cigar_launch(const uint64_t *q_packed, const uint64_t *r_packed,
                         const int32_t *q_len, const int32_t *r_len, const int32_t *q_off,
                         const int32_t *r_off, int64_t n_pairs, int64_t q_words, int64_t r_words,
                         const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                         int32_t e, int32_t *out_cost, uint32_t *cigar, const int64_t *cigar_off,
                         int32_t *cigar_len, void *stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    const Batch b = {q_packed, r_packed, q_len, r_len, q_off, r_off, n_pairs, q_words, r_words};
    const WfaPen p = {x, o, e};
    if (int rc = check_batch(b)) return rc;
    if (int rc = check_wfa_pen(p)) return rc;
    if (n_pairs == 0) return CGM_OK;
    if (!out_cost || !cigar || !cigar_off || !cigar_len)
        return set_error(CGM_ERR_INVALID, "cigar: null output buffer");
    if (!max_error && (max_error_scalar < 0 || max_error_scalar > kMaxWfaError))
        return set_error(CGM_ERR_INVALID, "cigar: max_error out of range");
    int g = 0, nM = 0, nI = 0;
    if (int rc = wfa_geometry(p, &g, &nM, &nI)) return rc;
    // 1. costs (the automatic WFA launcher)
    if (int rc = launch_wfa_wave(b, max_error, max_error_scalar, p, out_cost, true, stream))
        return rc;
    // 2. history sizes -> host, chunks within the budget
    int64_t *d_bytes = nullptr, *d_off = nullptr;
    int32_t *d_small = nullptr;
    CGM_HIP(hipMallocAsync((void **)&d_bytes, (size_t)n_pairs * 8, stream));
    CGM_HIP(hipMallocAsync((void **)&d_small, (size_t)n_pairs * 4, stream));
    hipLaunchKernelGGL(cigar_size_kernel, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0,
                       stream, b, max_error, max_error_scalar, p, wfadev::WfaGeom{g, nM, nI},
                       (int32_t)kCigarSmallLds, out_cost, d_bytes, d_small);
    std::vector<int64_t> bytes((size_t)n_pairs), off((size_t)n_pairs);
    std::vector<int32_t> small((size_t)n_pairs);
    hipError_t err = hipMemcpyAsync(bytes.data(), d_bytes, (size_t)n_pairs * 8,
                                    hipMemcpyDeviceToHost, stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(small.data(), d_small, (size_t)n_pairs * 4, hipMemcpyDeviceToHost,
                             stream);
    if (err == hipSuccess) err = hipStreamSynchronize(stream);
    (void)hipFreeAsync(d_bytes, stream);
    (void)hipFreeAsync(d_small, stream);
    if (err != hipSuccess) return hip_error(err, "cigar: history sizes");
    // pairs by history size, largest first; chunks [p0, p1) of that order,
    // a pair larger than the budget forming a chunk of its own
    std::vector<int32_t> order((size_t)n_pairs);
    for (int64_t i = 0; i < n_pairs; ++i) order[(size_t)i] = (int32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t c) { return bytes[(size_t)a] > bytes[(size_t)c]; });
    // The history workspace stays allocated per device between calls: it
    // is tens of GB for thousands of kb-long pairs, and a fresh allocation
    // of that size on every call costs more than the alignment once the
    // driver has to hand out recycled memory.  It grows when a call needs
    // more; calls on one device take turns on it.
    int dev = 0;
    if ((err = hipGetDevice(&dev)) != hipSuccess) return hip_error(err, "cigar: device");
    if (dev < 0 || dev >= kCigarMaxDevices) return set_error(CGM_ERR_INVALID, "cigar: device index");
    CigarWorkspace &cw = g_cigar_ws[dev];
    std::lock_guard<std::mutex> ws_lock(cw.mu);
    size_t free_b = 0, total_b = 0;
    int64_t budget = kCigarBudgetMin;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        budget = std::min(kCigarBudgetMax,
                          std::max(kCigarBudgetMin, (int64_t)((free_b + cw.bytes) / 2)));
    std::vector<std::pair<int64_t, int64_t>> chunks;
    int64_t ws_need = 0;
    for (int64_t i = 0; i < n_pairs;) {
        int64_t j = i, acc = 0;
        while (j < n_pairs && (j == i || acc + bytes[(size_t)order[(size_t)j]] <= budget)) {
            off[(size_t)order[(size_t)j]] = acc;
            acc += bytes[(size_t)order[(size_t)j]];
            ++j;
        }
        // within the chunk: the pairs for the large workgroups first, then
        // those for the small ones, each still largest first
        std::stable_partition(order.begin() + i, order.begin() + j,
                              [&](int32_t a) { return small[(size_t)a] == 0; });
        chunks.push_back({i, j});
        ws_need = std::max(ws_need, acc);
        i = j;
    }
    // workspaces first, uploads after: every exit below waits for the stream,
    // so the host vectors outlive their asynchronous copies
    int32_t *d_order = nullptr;
    int rc = CGM_OK;
    if ((size_t)ws_need > cw.bytes) {
        // the previous user of the workspace synchronised before returning
        if (cw.ptr) (void)hipFree(cw.ptr);
        cw.ptr = nullptr;
        cw.bytes = 0;
        if ((err = hipMalloc((void **)&cw.ptr, (size_t)std::max<int64_t>(ws_need, 256))) !=
            hipSuccess) {
            cw.ptr = nullptr;
            return hip_error(err, "cigar: workspace");
        }
        cw.bytes = (size_t)std::max<int64_t>(ws_need, 256);
    }
    uint8_t *ws = cw.ptr;
    if ((err = hipMallocAsync((void **)&d_order, (size_t)n_pairs * 4, stream)) != hipSuccess ||
        (err = hipMallocAsync((void **)&d_off, (size_t)n_pairs * 8, stream)) != hipSuccess)
        rc = hip_error(err, "cigar: order buffers");
    if (!rc && (err = hipMemcpyAsync(d_order, order.data(), (size_t)n_pairs * 4,
                                     hipMemcpyHostToDevice, stream)) != hipSuccess)
        rc = hip_error(err, "cigar: order");
    if (!rc && (err = hipMemcpyAsync(d_off, off.data(), (size_t)n_pairs * 8,
                                     hipMemcpyHostToDevice, stream)) != hipSuccess)
        rc = hip_error(err, "cigar: offsets");
    // 3. forward with history + backtrace, chunk by chunk (stream order
    // keeps a chunk's workspace intact until its kernel is done)
    for (size_t c = 0; !rc && c < chunks.size(); ++c) {
        const int64_t p0 = chunks[c].first, p1 = chunks[c].second;
        int64_t nbig = 0;
        while (p0 + nbig < p1 && small[(size_t)order[(size_t)(p0 + nbig)]] == 0) ++nbig;
        if (nbig > 0)
            hipLaunchKernelGGL((wfa_cigar_kernel<kCigarThreads, kCigarLds>), dim3((unsigned)nbig),
                               dim3(kCigarThreads), (size_t)kCigarLds, stream, b, max_error,
                               max_error_scalar, p, wfadev::WfaGeom{g, nM, nI}, out_cost, d_off,
                               ws, d_order, p0, cigar, cigar_off, cigar_len);
        if (p1 - p0 - nbig > 0)
            hipLaunchKernelGGL((wfa_cigar_kernel<kCigarSmallThreads, kCigarSmallLds>),
                               dim3((unsigned)(p1 - p0 - nbig)), dim3(kCigarSmallThreads),
                               (size_t)kCigarSmallLds, stream, b, max_error, max_error_scalar,
                               p, wfadev::WfaGeom{g, nM, nI}, out_cost, d_off, ws, d_order,
                               p0 + nbig, cigar, cigar_off, cigar_len);
        err = hipGetLastError();
        if (err != hipSuccess) rc = hip_error(err, "cigar: kernel launch");
    }
    // the host copies of `off` and `order` must outlive the asynchronous uploads
    err = hipStreamSynchronize(stream);
    if (!rc && err != hipSuccess) rc = hip_error(err, "cigar: synchronize");
    if (d_off) (void)hipFreeAsync(d_off, stream);
    if (d_order) (void)hipFreeAsync(d_order, stream);
    return rc;
}

int cgm_wfa_cigar_host(const uint64_t *q_packed, const uint64_t *r_packed, const int32_t *q_len,
                       const int32_t *r_len, const int32_t *q_off, const int32_t *r_off,
                       int64_t n_pairs, int64_t q_words, int64_t r_words,
                       const int32_t *max_error, int32_t max_error_scalar, int32_t x, int32_t o,
                       int32_t e, int32_t *out_cost, uint32_t *cigar, const int64_t *cigar_off,
                       int32_t *cigar_len) {
    if (n_pairs < 0) return set_error(CGM_ERR_INVALID, "n_pairs < 0");
    if (n_pairs == 0) return CGM_OK;
    if (!cigar_off) return set_error(CGM_ERR_INVALID, "cigar: null cigar_off");
    const int64_t runs = cigar_off[n_pairs];
    if (runs < 0) return set_error(CGM_ERR_INVALID, "cigar: negative capacity");
    hipStream_t s = nullptr;
    CGM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct Guard {
        hipStream_t s;
        std::vector<void *> bufs;
        ~Guard() {
            (void)hipStreamSynchronize(s);
            for (void *p : bufs) (void)hipFree(p);
            (void)hipStreamDestroy(s);
        }
    } guard{s, {}};
    auto up = [&](const void *src, size_t bytes, void **d) -> int {
        CGM_HIP(hipMalloc(d, bytes > 0 ? bytes : 8));
        guard.bufs.push_back(*d);
        if (src && bytes) CGM_HIP(hipMemcpyAsync(*d, src, bytes, hipMemcpyHostToDevice, s));
        return CGM_OK;
    };
    void *qp, *rp, *ql, *rl, *qo, *ro, *me = nullptr, *oc, *cg, *co, *cl;
    int rc = CGM_OK;
    const size_t n4 = (size_t)n_pairs * 4;
    if ((rc = up(q_packed, (size_t)q_words * 8, &qp)) || (rc = up(r_packed, (size_t)r_words * 8, &rp)) ||
        (rc = up(q_len, n4, &ql)) || (rc = up(r_len, n4, &rl)) || (rc = up(q_off, n4, &qo)) ||
        (rc = up(r_off, n4, &ro)) || (max_error && (rc = up(max_error, n4, &me))) ||
        (rc = up(nullptr, n4, &oc)) || (rc = up(nullptr, (size_t)runs * 4, &cg)) ||
        (rc = up(cigar_off, (size_t)(n_pairs + 1) * 8, &co)) || (rc = up(nullptr, n4, &cl)))
        return rc;
    if ((rc = cgm_wfa_cigar_launch((const uint64_t *)qp, (const uint64_t *)rp, (const int32_t *)ql,
                                   (const int32_t *)rl, (const int32_t *)qo, (const int32_t *)ro,
                                   n_pairs, q_words, r_words, (const int32_t *)me,
                                   max_error_scalar, x, o, e, (int32_t *)oc, (uint32_t *)cg,
                                   (const int64_t *)co, (int32_t *)cl, s)))
        return rc;
    CGM_HIP(hipMemcpyAsync(out_cost, oc, n4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipMemcpyAsync(cigar_len, cl, n4, hipMemcpyDeviceToHost, s));
    if (runs > 0) CGM_HIP(hipMemcpyAsync(cigar, cg, (size_t)runs * 4, hipMemcpyDeviceToHost, s));
    CGM_HIP(hipStreamSynchronize(s));
    return CGM_OK;
}

}  // extern "C"
