// Block one-sided Jacobi SVD for large problems (many work-groups per step).
//
// The nv rows of X are split into blocks of kB rows.  One sweep pairs every
// two blocks once (round-robin over the blocks); in a step the pairs are
// disjoint, so each runs on its own work-group:
//   S = Y Y^T for the 2kB rows Y of the pair (f64 MFMA, operands from L2),
//   S = W diag W^T by cyclic two-sided Jacobi in LDS (kInner sweeps at most),
//   Y <- W^T Y and the same rows of rt <- W^T rt (f64 MFMA).
// A pair whose scaled off-diagonal max |S_ik| / sqrt(S_ii S_kk) is below tol
// is left alone; the sweep loop stops after a sweep without rotations.  This
// is the block form of Hestenes' method (Hari & Veselic; Becka, Oksa &
// Vajtersic).  Rows end orthogonal with x_in = rt^T x_out; norms and the sort
// by decreasing norm finish the decomposition.
#include "common.hpp"
#include "prof.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace syn {

constexpr int kChunk = 128;   // columns of the pair staged in LDS at a time

__device__ __forceinline__ void rr_pair(int64_t step, int64_t w, int64_t np, int64_t &a, int64_t &b) {
    if (w == 0) {
        a = step;
        b = np - 1;
    } else {
        a = (step + w) % (np - 1);
        b = (step - w + np - 1) % (np - 1);
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// Stage rows[0..P) x columns [c0, c0+kChunk) of M into Ys.  All of a
// thread's loads are issued before the first LDS store, so their latencies
// overlap instead of adding up.
template <int P, int NT>
__device__ __forceinline__ void stage_chunk(double (*Ys)[kChunk + 4], const int *rows, const double *M, int64_t ld,
                                            int64_t cols, int64_t c0, int tid) {
    constexpr int PER = P * kChunk / NT;
    static_assert(PER * NT == P * kChunk, "chunk / threads mismatch");
    double v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = tid + q * NT, i = e / kChunk, t = e % kChunk;
        const int r = rows[i];
        v[q] = (r >= 0 && c0 + t < cols) ? M[(int64_t)r * ld + c0 + t] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = tid + q * NT;
        Ys[e / kChunk][e % kChunk] = v[q];
    }
}

// MFMA f64 16x16x4 layout: A[i = l%16][k = l/16], B[k = l/16][j = l%16],
// D reg r of lane l = D[l/16 + 4r][l%16].
//
// KB rows per block, NT threads per work-group (NT/64 waves).  KB = 8 with a
// single wave keeps the inner 16 x 16 problem free of cross-wave barriers.
template <int KB, int NT>
__device__ __forceinline__ void bjacobi_pair_step(double *X, int64_t nv, int64_t len, int64_t ldx, double *R,
                                                  int64_t ldr, int64_t nblk, int64_t step, int64_t pair,
                                                  double tol, int *counter, int inner_sweeps) {
    constexpr int P = 2 * KB;        // rows of a pair
    constexpr int RT = P / 16;       // 16-row MFMA tiles
    constexpr int NWV = NT / 64;     // waves
    __shared__ double S[P][P + 1];
    __shared__ double W[P][P + 1];
    __shared__ double Ys[P][kChunk + 4];
    __shared__ double cs[P / 2][2];
    __shared__ int rows[P];
    __shared__ unsigned char pq[P - 1][P / 2][2];  // inner round-robin pairs
    __shared__ int do_rot;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int64_t np = nblk + (nblk & 1);
    int64_t ba, bb;
    rr_pair(step, pair, np, ba, bb);
    if (ba >= nblk && bb >= nblk) return;  // a block paired with the padding slot still
    if (tid < P) {                           // gets its own rows orthogonalised
        const int64_t blk = tid < KB ? ba : bb;
        const int64_t r = blk * KB + (tid % KB);
        rows[tid] = (blk < nblk && r < nv) ? (int)r : -1;
    }
    for (int e = tid; e < P * P; e += NT) W[e / P][e % P] = (e / P == e % P) ? 1.0 : 0.0;
    for (int e = tid; e < (P - 1) * (P / 2); e += NT) {
        const int r = e / (P / 2), pw = e % (P / 2);
        int64_t a, b;
        rr_pair(r, pw, P, a, b);
        pq[r][pw][0] = (unsigned char)(a < b ? a : b);
        pq[r][pw][1] = (unsigned char)(a < b ? b : a);
    }
    if (tid == 0) do_rot = 0;
    __syncthreads();

    // S = Y Y^T on the matrix cores, rows staged through LDS in chunks
    v4d sacc[(RT * RT + NWV - 1) / NWV];
#pragma unroll
    for (int q = 0; q < (RT * RT + NWV - 1) / NWV; ++q) sacc[q] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int64_t c0 = 0; c0 < len; c0 += kChunk) {
        stage_chunk<P, NT>(Ys, rows, X, ldx, len, c0, tid);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < (RT * RT + NWV - 1) / NWV; ++q) {
            const int tile = w + q * NWV;
            if (tile < RT * RT) {
                const int ta = tile / RT, tb = tile % RT;
#pragma unroll 8
                for (int k = 0; k < kChunk; k += 4)
                    sacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ys[ta * 16 + li][k + lk], Ys[tb * 16 + li][k + lk],
                                                                   sacc[q], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < (RT * RT + NWV - 1) / NWV; ++q) {
        const int tile = w + q * NWV;
        if (tile < RT * RT) {
            const int ta = tile / RT, tb = tile % RT;
#pragma unroll
            for (int r = 0; r < 4; ++r) S[ta * 16 + lk + 4 * r][tb * 16 + li] = sacc[q][r];
        }
    }
    __syncthreads();
    for (int e = tid; e < P * P; e += NT) {
        const int i = e / P, k = e % P;
        if (i < k && fabs(S[i][k]) > tol * sqrt(S[i][i]) * sqrt(S[k][k])) do_rot = 1;
    }
    __syncthreads();
    if (!do_rot) return;

    // cyclic two-sided Jacobi on S, accumulating W (S_final = W^T S W).
    // One round rotates P/2 disjoint index pairs; thread (pa, pb) owns the
    // 2x2 block of S at the rows of pair pa and the columns of pair pb and
    // applies both rotations to it at once, so a round needs two barriers.
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
        int any = 0;
        for (int r = 0; r < P - 1; ++r) {
            if (tid < P / 2) {
                const int p = pq[r][tid][0], q = pq[r][tid][1];
                const double apq = S[p][q], app = S[p][p], aqq = S[q][q];
                double c = 1.0, s = 0.0;
                if (fabs(apq) > tol * sqrt(fabs(app)) * sqrt(fabs(aqq)) && apq != 0.0) {
                    const double zeta = (aqq - app) / (2.0 * apq);
                    const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
                    c = 1.0 / sqrt(fma(t, t, 1.0));
                    s = c * t;
                }
                cs[tid][0] = c;
                cs[tid][1] = s;
            }
            __syncthreads();
            for (int e = tid; e < (P / 2) * (P / 2); e += NT) {
                const int pa = e / (P / 2), pb = e % (P / 2);
                const double ca = cs[pa][0], sa = cs[pa][1], cb = cs[pb][0], sb = cs[pb][1];
                if (sa != 0.0 || sb != 0.0) {
                    const int p = pq[r][pa][0], q = pq[r][pa][1], u = pq[r][pb][0], v = pq[r][pb][1];
                    const double spu = S[p][u], spv = S[p][v], squ = S[q][u], sqv = S[q][v];
                    // rows: [p; q] <- J_a^T [p; q]
                    const double rpu = ca * spu - sa * squ, rpv = ca * spv - sa * sqv;
                    const double rqu = sa * spu + ca * squ, rqv = sa * spv + ca * sqv;
                    // columns: [u, v] <- [u, v] J_b
                    S[p][u] = cb * rpu - sb * rpv;
                    S[p][v] = sb * rpu + cb * rpv;
                    S[q][u] = cb * rqu - sb * rqv;
                    S[q][v] = sb * rqu + cb * rqv;
                }
            }
            for (int e = tid; e < (P / 2) * P; e += NT) {  // columns p, q of W
                const int pr = e / P, i = e % P;
                const double c = cs[pr][0], s = cs[pr][1];
                if (s != 0.0) {
                    const int p = pq[r][pr][0], q = pq[r][pr][1];
                    const double u = W[i][p], v = W[i][q];
                    W[i][p] = c * u - s * v;
                    W[i][q] = s * u + c * v;
                }
            }
            for (int pr = 0; pr < P / 2; ++pr) any |= (cs[pr][1] != 0.0);
            __syncthreads();
        }
        if (!any) break;
    }

    // Y <- W^T Y on the matrix cores, chunk by chunk through LDS; a wave forms
    // all 16-row output tiles of its 16-column blocks before writing them.
    for (int pass = 0; pass < 2; ++pass) {
        double *M = pass ? R : X;
        const int64_t ld = pass ? ldr : ldx, cols = pass ? nv : len;
        for (int64_t c0 = 0; c0 < cols; c0 += kChunk) {
            stage_chunk<P, NT>(Ys, rows, M, ld, cols, c0, tid);
            __syncthreads();
            for (int cb = w; cb < kChunk / 16; cb += NWV) {
                const int tc = cb * 16 + li;
                v4d acc[RT];
#pragma unroll
                for (int ti = 0; ti < RT; ++ti) acc[ti] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < P / 4; ++kk) {
                    const int kr = kk * 4 + lk;
                    const double b = Ys[kr][tc];
#pragma unroll
                    for (int ti = 0; ti < RT; ++ti)
                        acc[ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(W[kr][ti * 16 + li], b, acc[ti], 0, 0, 0);
                }
                const int64_t col = c0 + tc;
                if (col < cols) {
#pragma unroll
                    for (int ti = 0; ti < RT; ++ti)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int rr = rows[ti * 16 + lk + 4 * r];
                            if (rr >= 0) M[(int64_t)rr * ld + col] = acc[ti][r];
                        }
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) atomicAdd(counter, 1);
}

template <int KB, int NT>
__global__ __launch_bounds__(NT) void bjacobi_step_kernel(double *X, int64_t nv, int64_t len, int64_t ldx, double *R,
                                                          int64_t ldr, int64_t nblk, int64_t step, double tol,
                                                          int *counter, int inner_sweeps) {
    bjacobi_pair_step<KB, NT>(X, nv, len, ldx, R, ldr, nblk, step, blockIdx.x, tol, counter, inner_sweeps);
}

// One step for many problems: work-group -> (problem, block pair).  A problem
// with fewer blocks than the largest one repeats its round-robin schedule.
__global__ __launch_bounds__(256) void bjacobi_many_step_kernel(const syn_bjac_item *__restrict__ items, int n,
                                                                int64_t step, double tol, int inner_sweeps) {
    const int64_t bid = blockIdx.x;
    const syn_bjac_item &P = items[find_item(items, n, bid)];
    const int64_t nblk = (P.nv + 15) / 16;
    const int64_t np = nblk + (nblk & 1);
    bjacobi_pair_step<16, 256>(P.x, P.nv, P.len, P.ldx, P.rt, P.ldr, nblk, step % (np - 1), bid - P.block_start,
                               tol, P.counter, inner_sweeps);
}

__global__ __launch_bounds__(256) void bjac_many_init_kernel(const syn_bjac_item *__restrict__ items) {
    const syn_bjac_item &P = items[blockIdx.x];
    const int64_t nel = P.nv * P.nv;
    for (int64_t e = threadIdx.x; e < nel; e += 256) {
        const int64_t i = e / P.nv, j = e - i * P.nv;
        P.rt[i * P.ldr + j] = (i == j) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void bjac_many_norms_kernel(const syn_bjac_item *__restrict__ items) {
    const syn_bjac_item &P = items[blockIdx.x];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t i = w; i < P.nv; i += 4) {
        double a = 0.0;
        for (int64_t t = lane; t < P.len; t += 64) a = fma(P.x[i * P.ldx + t], P.x[i * P.ldx + t], a);
        a = wave_sum(a);
        if (lane == 0) P.s[i] = P.gram ? sqrt(sqrt(a)) : sqrt(a);
    }
}

// Sort rows of problem b by perm (device, offsets in poff[b]) through the
// workspace (xoff/roff[b]).  pass 0 gathers into the workspace, pass 1 copies back.
__global__ __launch_bounds__(256) void bjac_many_permute_kernel(const syn_bjac_item *__restrict__ items,
                                                                const int64_t *__restrict__ perm,
                                                                const int64_t *__restrict__ poff, double *work,
                                                                const int64_t *__restrict__ woff, int pass) {
    const syn_bjac_item &P = items[blockIdx.x];
    const int64_t *pm = perm + poff[blockIdx.x];
    double *wx = work + woff[blockIdx.x];
    double *wr = wx + P.nv * P.len;
    double *ws = wr + P.nv * P.nv;
    const int64_t nx = P.nv * P.len, nr = P.nv * P.nv;
    for (int64_t e = threadIdx.x; e < nx + nr + P.nv; e += 256) {
        if (e < nx) {
            const int64_t i = e / P.len, t = e - i * P.len;
            if (pass == 0)
                wx[e] = P.x[pm[i] * P.ldx + t];
            else
                P.x[i * P.ldx + t] = wx[e];
        } else if (e < nx + nr) {
            const int64_t f = e - nx, i = f / P.nv, t = f - i * P.nv;
            if (pass == 0)
                wr[f] = P.rt[pm[i] * P.ldr + t];
            else
                P.rt[i * P.ldr + t] = wr[f];
        } else {
            const int64_t i = e - nx - nr;
            if (pass == 0)
                ws[i] = P.s[pm[i]];
            else
                P.s[i] = ws[i];
        }
    }
}

__global__ __launch_bounds__(256) void zero_ints_kernel(int32_t *c, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) c[i] = 0;
}

__global__ __launch_bounds__(256) void identity_kernel(double *R, int64_t n, int64_t ld) {
    const int64_t nel = n * n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nel; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / n, j = e - i * n;
        R[i * ld + j] = (i == j) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(256) void zero_int_kernel(int *c) { *c = 0; }

__global__ __launch_bounds__(256) void rownorm_kernel(const double *X, int64_t nv, int64_t len, int64_t ldx,
                                                      double *s, int gram) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = blockIdx.x * 4 + w;
    if (i >= nv) return;
    double a = 0.0;
    for (int64_t t = lane; t < len; t += 64) a = fma(X[i * ldx + t], X[i * ldx + t], a);
    a = wave_sum(a);
    if (lane == 0) s[i] = gram ? sqrt(sqrt(a)) : sqrt(a);
}

// dst[i, :] = src[perm[i], :]
__global__ __launch_bounds__(256) void gather_rows_kernel(const double *src, int64_t lds, double *dst, int64_t ldd,
                                                          int64_t rows, int64_t cols, const int64_t *perm) {
    const int64_t nel = rows * cols;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nel; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / cols, j = e - i * cols;
        dst[i * ldd + j] = src[perm[i] * lds + j];
    }
}

static unsigned grid_for(int64_t nel) {
    int64_t b = (nel + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

}  // namespace syn

extern "C" int64_t syn_jacobi_blocked_workspace(int64_t nv, int64_t len) {
    // sorted copies of X and rt, the norms, the permutation and a counter
    return nv * len + nv * nv + 2 * nv + 8;
}

// Rows per block: 16 (four waves per pair) from 384 rows up, else 8 (one wave).
// One inner Jacobi sweep on the 2KB x 2KB pair Gram (the outer sweep repeats).
static constexpr int kInnerSweeps = 1;

extern "C" int syn_jacobi_blocked_f64(double *X, int64_t nv, int64_t len, int64_t ldx, double *rt, int64_t ldr,
                                      double *s, double *work, double tol, int max_sweeps, int gram, int *sweeps,
                                      void *stream) {
    const int block_rows = nv >= 384 ? 16 : 8;
    using namespace syn;
    if (nv <= 0 || len <= 0 || ldx < len || ldr < nv || !X || !rt || !s || !work || max_sweeps < 0 ||
        !(tol >= 0.0)) {
        set_error("syn_jacobi_blocked_f64: bad arguments (nv=%lld len=%lld)", (long long)nv, (long long)len);
        return SYN_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    double *Xw = work;
    double *Rw = Xw + nv * len;
    double *sw = Rw + nv * nv;
    int64_t *perm = (int64_t *)(sw + nv);
    int *counter = (int *)(perm + nv);

    hipLaunchKernelGGL(identity_kernel, dim3(grid_for(nv * nv)), dim3(256), 0, st, rt, nv, ldr);
    SYN_LAUNCH_CHECK("identity");
    const int kb = (block_rows == 16) ? 16 : 8;
    const int64_t nblk = (nv + kb - 1) / kb;
    const int64_t np = nblk + (nblk & 1);
    int used = 0;
    for (; used < max_sweeps && nblk > 0; ++used) {
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, st, counter);
        for (int64_t step = 0; step < np - 1; ++step) {
            // flops of a step: the pair Grams and the row updates of X and rt
            ProfScope ps(kTagBJacobiStep,
                         (double)(np / 2) * 2.0 * (2.0 * kb) * (2.0 * kb) * (double)(2 * len + nv), st);
            if (kb == 16)
                hipLaunchKernelGGL((bjacobi_step_kernel<16, 256>), dim3((unsigned)(np / 2)), dim3(256), 0, st, X, nv,
                                   len, ldx, rt, ldr, nblk, step, tol, counter, kInnerSweeps);
            else
                hipLaunchKernelGGL((bjacobi_step_kernel<8, 64>), dim3((unsigned)(np / 2)), dim3(64), 0, st, X, nv,
                                   len, ldx, rt, ldr, nblk, step, tol, counter, kInnerSweeps);
        }
        SYN_LAUNCH_CHECK("bjacobi_step");
        int host_count = 0;
        if (int rc = hip_check(hipMemcpyAsync(&host_count, counter, sizeof(int), hipMemcpyDeviceToHost, st),
                               "bjacobi counter copy"))
            return rc;
        if (int rc = hip_check(hipStreamSynchronize(st), "bjacobi sync")) return rc;
        if (host_count == 0) {
            ++used;
            break;
        }
    }
    if (sweeps) *sweeps = used;

    // norms, host argsort (descending, stable), gather into the workspace and back
    hipLaunchKernelGGL(rownorm_kernel, dim3((unsigned)((nv + 3) / 4)), dim3(256), 0, st, X, nv, len, ldx, sw, gram);
    SYN_LAUNCH_CHECK("rownorm");
    std::vector<double> hs(nv);
    std::vector<int64_t> hp(nv);
    if (int rc = hip_check(hipMemcpyAsync(hs.data(), sw, nv * sizeof(double), hipMemcpyDeviceToHost, st), "norms"))
        return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "norms sync")) return rc;
    std::iota(hp.begin(), hp.end(), 0);
    std::stable_sort(hp.begin(), hp.end(), [&](int64_t a, int64_t b) { return hs[a] > hs[b]; });
    if (int rc = hip_check(hipMemcpyAsync(perm, hp.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice, st), "perm"))
        return rc;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(nv * len)), dim3(256), 0, st, X, ldx, Xw, len, nv, len, perm);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(nv * nv)), dim3(256), 0, st, rt, ldr, Rw, nv, nv, nv, perm);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(nv)), dim3(256), 0, st, sw, 1, s, 1, nv, 1, perm);
    SYN_LAUNCH_CHECK("gather");
    // copy the sorted rows back
    if (int rc = hip_check(hipMemcpy2DAsync(X, ldx * sizeof(double), Xw, len * sizeof(double), len * sizeof(double),
                                            nv, hipMemcpyDeviceToDevice, st),
                           "copy X"))
        return rc;
    if (int rc = hip_check(hipMemcpy2DAsync(rt, ldr * sizeof(double), Rw, nv * sizeof(double), nv * sizeof(double),
                                            nv, hipMemcpyDeviceToDevice, st),
                           "copy rt"))
        return rc;
    // the host vectors are read by the copies above: wait before they go out of scope
    return hip_check(hipStreamSynchronize(st), "bjacobi final sync");
}

extern "C" int syn_jacobi_blocked_many_f64(syn_bjac_item *items, const syn_bjac_item *items_host, int n_items,
                                           double *work, int32_t *counters, double tol, int max_sweeps, int *sweeps,
                                           void *stream) {
    using namespace syn;
    if (n_items < 0 || !items || (n_items && (!items_host || !work || !counters)) || max_sweeps < 0 ||
        !(tol >= 0.0)) {
        set_error("syn_jacobi_blocked_many_f64: bad arguments");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    hipStream_t st = (hipStream_t)stream;
    std::vector<syn_bjac_item> all(items_host, items_host + n_items);
    for (int i = 0; i < n_items; ++i) {
        if (all[i].nv <= 0 || all[i].len <= 0 || all[i].ldx < all[i].len || all[i].ldr < all[i].nv) {
            set_error("syn_jacobi_blocked_many_f64: bad problem %d (nv=%lld len=%lld)", i, (long long)all[i].nv,
                      (long long)all[i].len);
            return SYN_EINVAL;
        }
        all[i].counter = counters + i;
        all[i].active = all[i].nv > 1;
        all[i].block_start = 0;
    }
    // full table on the device for init / norms / permute; the copy goes
    // through the caller's stream and is waited for, so it is ordered with
    // that stream's earlier kernels and the host vector may change afterwards
    auto upload = [&](const std::vector<syn_bjac_item> &v) -> int {
        if (int rc = hip_check(hipMemcpyAsync(items, v.data(), v.size() * sizeof(syn_bjac_item),
                                              hipMemcpyHostToDevice, st),
                               "bjac table upload"))
            return rc;
        return hip_check(hipStreamSynchronize(st), "bjac table upload");
    };
    if (int rc = upload(all)) return rc;
    hipLaunchKernelGGL(bjac_many_init_kernel, dim3(n_items), dim3(256), 0, st, items);
    SYN_LAUNCH_CHECK("bjac init");

    std::vector<int32_t> hc(n_items);
    std::vector<int> idx;  // problems still rotating
    for (int i = 0; i < n_items; ++i)
        if (all[i].active) idx.push_back(i);
    int used = 0;
    for (; used < max_sweeps && !idx.empty(); ++used) {
        std::vector<syn_bjac_item> act;
        int64_t start = 0, max_np = 2;
        for (int i : idx) {
            syn_bjac_item it = all[i];
            const int64_t nblk = (it.nv + 15) / 16, np = nblk + (nblk & 1);
            it.block_start = start;
            start += np / 2;
            if (np > max_np) max_np = np;
            act.push_back(it);
        }
        if (int rc = upload(act)) return rc;
        hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(256), 0, st, counters, n_items);
        for (int64_t step = 0; step < max_np - 1; ++step) {
            ProfScope ps(kTagBJacobiStep, 0.0, st);
            hipLaunchKernelGGL(bjacobi_many_step_kernel, dim3((unsigned)start), dim3(256), 0, st, items,
                               (int)act.size(), step, tol, kInnerSweeps);
        }
        SYN_LAUNCH_CHECK("bjac many step");
        if (int rc = hip_check(hipMemcpyAsync(hc.data(), counters, n_items * sizeof(int32_t), hipMemcpyDeviceToHost,
                                              st),
                               "bjac counters"))
            return rc;
        if (int rc = hip_check(hipStreamSynchronize(st), "bjac sync")) return rc;
        std::vector<int> next;
        for (int i : idx)
            if (hc[i] != 0) next.push_back(i);
        idx.swap(next);
        if (idx.empty()) {
            ++used;
            break;
        }
    }
    if (sweeps) *sweeps = used;

    // norms, per-problem argsort on the host, permute through the workspace
    if (int rc = upload(all)) return rc;
    hipLaunchKernelGGL(bjac_many_norms_kernel, dim3(n_items), dim3(256), 0, st, items);
    SYN_LAUNCH_CHECK("bjac norms");
    int64_t tot_nv = 0, tot_w = 0;
    std::vector<int64_t> poff(n_items), woff(n_items);
    for (int i = 0; i < n_items; ++i) {
        poff[i] = tot_nv;
        tot_nv += all[i].nv;
        woff[i] = tot_w;
        tot_w += all[i].nv * all[i].len + all[i].nv * all[i].nv + all[i].nv;
    }
    std::vector<double> hs(tot_nv);
    for (int i = 0; i < n_items; ++i)
        if (int rc = hip_check(hipMemcpyAsync(hs.data() + poff[i], all[i].s, all[i].nv * sizeof(double),
                                              hipMemcpyDeviceToHost, st),
                               "bjac norms copy"))
            return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "bjac norms sync")) return rc;
    std::vector<int64_t> hp(tot_nv);
    for (int i = 0; i < n_items; ++i) {
        int64_t *p = hp.data() + poff[i];
        const double *sv = hs.data() + poff[i];
        std::iota(p, p + all[i].nv, 0);
        std::stable_sort(p, p + all[i].nv, [&](int64_t a, int64_t b) { return sv[a] > sv[b]; });
    }
    // device scratch for perm and offsets lives after the row workspace
    int64_t *dperm = (int64_t *)(work + tot_w);
    int64_t *dpoff = dperm + tot_nv;
    int64_t *dwoff = dpoff + n_items;
    if (int rc = hip_check(hipMemcpyAsync(dperm, hp.data(), tot_nv * sizeof(int64_t), hipMemcpyHostToDevice, st),
                           "perm"))
        return rc;
    if (int rc = hip_check(hipMemcpyAsync(dpoff, poff.data(), n_items * sizeof(int64_t), hipMemcpyHostToDevice, st),
                           "poff"))
        return rc;
    if (int rc = hip_check(hipMemcpyAsync(dwoff, woff.data(), n_items * sizeof(int64_t), hipMemcpyHostToDevice, st),
                           "woff"))
        return rc;
    hipLaunchKernelGGL(bjac_many_permute_kernel, dim3(n_items), dim3(256), 0, st, items, dperm, dpoff, work, dwoff, 0);
    hipLaunchKernelGGL(bjac_many_permute_kernel, dim3(n_items), dim3(256), 0, st, items, dperm, dpoff, work, dwoff, 1);
    SYN_LAUNCH_CHECK("bjac permute");
    return hip_check(hipStreamSynchronize(st), "bjac final sync");
}

extern "C" int64_t syn_jacobi_blocked_many_workspace(const int64_t *nv, const int64_t *len, int n_items) {
    int64_t tot = 0;
    for (int i = 0; i < n_items; ++i) tot += nv[i] * len[i] + nv[i] * nv[i] + nv[i] + 2 * nv[i] + 2;
    return tot + 2 * (int64_t)n_items + 16;
}
