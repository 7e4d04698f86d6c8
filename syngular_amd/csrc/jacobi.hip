// Batched one-sided Jacobi SVD (Hestenes), one work-group per problem.
//
// The problem is a set of nv vectors of length len stored as rows.  Pairs of
// rows are rotated until all rows are mutually orthogonal (relative
// tolerance `tol` on |x_p . x_q| / (||x_p|| ||x_q||)); the same rotations are
// accumulated in rt (nv x nv).  Rotation angles follow Rutishauser's formula
// as used by Demmel & Veselic (SIAM J. Matrix Anal. Appl. 13, 1992):
//   zeta = (b - a) / (2 g),  t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   c = 1 / sqrt(1 + t^2),   s = c t.
// Each sweep visits every pair once in the round-robin (circle) order, so the
// nv/2 rotations of one round touch disjoint rows and run on different waves
// without synchronisation.  Data stay in LDS when they fit (the small-bond
// case) and in the caller's buffers (L2) otherwise.  Finally rows are sorted
// by decreasing norm with in-place swaps.
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kJacThreads = 256;
constexpr int kJacWaves = kJacThreads / 64;
constexpr int64_t kJacLdsDoubles = (128 * 1024) / 8;

template <bool IN_LDS>
__device__ void jacobi_one(const syn_jacobi_item &P, double *lds, double tol, int max_sweeps) {
    __shared__ int rotated;
    __shared__ double sred[kJacWaves];
    __shared__ int sidx[kJacWaves];
    const int64_t nv = P.nv, len = P.len;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double *X, *R;
    int64_t ldx, ldr;
    if (IN_LDS) {
        X = lds;
        ldx = len;
        R = lds + nv * len;
        ldr = nv;
        for (int64_t e = tid; e < nv * len; e += kJacThreads) {
            const int64_t i = e / len, t = e - i * len;
            X[e] = P.x[i * P.ldx + t];
        }
    } else {
        X = P.x;
        ldx = P.ldx;
        R = P.rt;
        ldr = P.ldr;
    }
    for (int64_t e = tid; e < nv * nv; e += kJacThreads) {
        const int64_t i = e / nv, t = e - i * nv;
        R[i * ldr + t] = (i == t) ? 1.0 : 0.0;
    }
    __syncthreads();

    const int64_t np = nv + (nv & 1);
    int sweep = 0;
    for (; sweep < max_sweeps && nv > 1; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int64_t r = 0; r < np - 1; ++r) {
            for (int64_t k = w; k < np / 2; k += kJacWaves) {
                int64_t a, b;
                if (k == 0) {
                    a = r;
                    b = np - 1;
                } else {
                    a = (r + k) % (np - 1);
                    b = (r - k + np - 1) % (np - 1);
                }
                if (a >= nv || b >= nv) continue;
                const int64_t p = a < b ? a : b, q = a < b ? b : a;
                double *xp = X + p * ldx, *xq = X + q * ldx;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int64_t t = lane; t < len; t += 64) {
                    const double u = xp[t], v = xq[t];
                    al = fma(u, u, al);
                    be = fma(v, v, be);
                    ga = fma(u, v, ga);
                }
                al = wave_sum(al);
                be = wave_sum(be);
                ga = wave_sum(ga);
                if (!(fabs(ga) > tol * sqrt(al) * sqrt(be))) continue;
                const double zeta = (be - al) / (2.0 * ga);
                const double tt = copysign(1.0, zeta) / (fabs(zeta) + hypot(1.0, zeta));
                const double c = 1.0 / sqrt(fma(tt, tt, 1.0)), s = c * tt;
                for (int64_t t = lane; t < len; t += 64) {
                    const double u = xp[t], v = xq[t];
                    xp[t] = c * u - s * v;
                    xq[t] = s * u + c * v;
                }
                double *rp = R + p * ldr, *rq = R + q * ldr;
                for (int64_t t = lane; t < nv; t += 64) {
                    const double u = rp[t], v = rq[t];
                    rp[t] = c * u - s * v;
                    rq[t] = s * u + c * v;
                }
                if (lane == 0) rotated = 1;
            }
            __syncthreads();
        }
        const int any = rotated;
        __syncthreads();
        if (!any) break;
    }

    // norms, then selection sort by decreasing norm (in-place row swaps)
    for (int64_t i = w; i < nv; i += kJacWaves) {
        double a = 0.0;
        for (int64_t t = lane; t < len; t += 64) a = fma(X[i * ldx + t], X[i * ldx + t], a);
        a = wave_sum(a);
        if (lane == 0) P.s[i] = P.gram ? sqrt(sqrt(a)) : sqrt(a);
    }
    __syncthreads();
    for (int64_t i = 0; i + 1 < nv; ++i) {
        double best = -1.0;
        int bi = (int)i;
        for (int64_t j = i + tid; j < nv; j += kJacThreads) {
            const double v = P.s[j];
            if (v > best) {
                best = v;
                bi = (int)j;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) {
            sred[w] = best;
            sidx[w] = bi;
        }
        __syncthreads();
        best = sred[0];
        bi = sidx[0];
        for (int q = 1; q < kJacWaves; ++q)
            if (sred[q] > best || (sred[q] == best && sidx[q] < bi)) {
                best = sred[q];
                bi = sidx[q];
            }
        __syncthreads();
        if (bi != i) {
            for (int64_t t = tid; t < len; t += kJacThreads) {
                const double u = X[i * ldx + t];
                X[i * ldx + t] = X[bi * ldx + t];
                X[bi * ldx + t] = u;
            }
            for (int64_t t = tid; t < nv; t += kJacThreads) {
                const double u = R[i * ldr + t];
                R[i * ldr + t] = R[bi * ldr + t];
                R[bi * ldr + t] = u;
            }
            if (tid == 0) {
                const double u = P.s[i];
                P.s[i] = P.s[bi];
                P.s[bi] = u;
            }
        }
        __syncthreads();
    }

    if (IN_LDS) {
        for (int64_t e = tid; e < nv * len; e += kJacThreads) {
            const int64_t i = e / len, t = e - i * len;
            P.x[i * P.ldx + t] = X[e];
        }
        for (int64_t e = tid; e < nv * nv; e += kJacThreads) {
            const int64_t i = e / nv, t = e - i * nv;
            P.rt[i * P.ldr + t] = R[e];
        }
    }
    if (tid == 0 && P.sweeps) *P.sweeps = sweep;
}

__global__ __launch_bounds__(kJacThreads) void jacobi_kernel(const syn_jacobi_item *__restrict__ items, double tol,
                                                             int max_sweeps) {
    extern __shared__ double lds[];
    const syn_jacobi_item &P = items[blockIdx.x];
    if (P.nv * P.len + P.nv * P.nv <= kJacLdsDoubles)
        jacobi_one<true>(P, lds, tol, max_sweeps);
    else
        jacobi_one<false>(P, lds, tol, max_sweeps);
}

// ---------------------------------------------------------------------------
// Two-sided (symmetric) Jacobi for Gram matrices in LDS.
//
// G (g x g, symmetric) and the accumulated rotations R = V^T stay in LDS.  A
// round of the round-robin order rotates g/2 disjoint planes (p, q): every
// angle comes from the 2 x 2 diagonal block [G_pp G_pq; G_pq G_qq] alone
// (Rutishauser's formula, as in the one-sided kernel), so no dot products or
// reductions are needed.  The update G <- J^T G J is applied as independent
// 2 x 2 block updates (rows of plane a, columns of plane b: both rotations at
// once), and R <- J^T R row pair by row pair; a round costs two barriers.
// Convergence: no plane with |G_pq| > tol * sqrt(|G_pp G_qq|) in a sweep.
// On exit the eigenvalues lambda_i = G_ii are sorted decreasingly; s_i =
// sqrt(max(lambda_i, 0)) are the singular values of C when G = C^T C, and the
// rows of rt are the matching eigenvectors.
// ---------------------------------------------------------------------------
// threads per problem by the largest order of the launch: one wave up to
// 16 (at most 8 planes a round), two up to 32, four up to 64, so small
// problems do not wait at four-wave barriers and more of them share a CU
constexpr int kGramThreads = 256;
constexpr int kGramMax = 64;

template <int NT>
__global__ __launch_bounds__(NT) void jacobi_gram_kernel(const syn_jacobi_item *__restrict__ items,
                                                                   int nv_max, double tol, int max_sweeps) {
    extern __shared__ double lds[];
    __shared__ double cs[kGramMax / 2][2];
    __shared__ int pp[kGramMax / 2], qq[kGramMax / 2];
    __shared__ int rank[kGramMax];
    __shared__ int rotated;
    const syn_jacobi_item &P = items[blockIdx.x];
    const int tid = threadIdx.x;
    if (P.nv < 1 || P.nv > nv_max || P.len != P.nv) {
        // not a problem this kernel takes: flag it in s and leave the rest
        for (int64_t i = tid; i < P.nv; i += NT) P.s[i] = __builtin_nan("");
        return;
    }
    const int g = (int)P.nv, ld = g + 1;
    double *A = lds;
    double *R = lds + g * ld;
    for (int e = tid; e < g * g; e += NT) {
        const int i = e / g, j = e - i * g;
        A[i * ld + j] = P.x[i * P.ldx + j];
        R[i * ld + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    // absolute rotation floor eps * max_i G_ii: entries below it are at the
    // rounding level of forming G (a rank-deficient G has a null space of
    // such entries, which a purely relative test would keep rotating)
    __shared__ double floor_abs;
    if (tid < 64) {
        double m = 0.0;
        for (int i = tid; i < g; i += 64) m = fmax(m, fabs(A[i * ld + i]));
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        if (tid == 0) floor_abs = __DBL_EPSILON__ * m;
    }
    const int np = g + (g & 1), half = np / 2;
    int sweep = 0;
    for (; sweep < max_sweeps && g > 1; ++sweep) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int r = 0; r < np - 1; ++r) {
            if (tid < half) {
                int a, b;
                if (tid == 0) {
                    a = r;
                    b = np - 1;
                } else {
                    a = (r + tid) % (np - 1);
                    b = (r - tid + np - 1) % (np - 1);
                }
                int p = a < b ? a : b, q = a < b ? b : a;
                double c = 1.0, sn = 0.0;
                if (q < g) {
                    const double apq = A[p * ld + q], app = A[p * ld + p], aqq = A[q * ld + q];
                    if (fabs(apq) > fmax(tol * sqrt(fabs(app) * fabs(aqq)), floor_abs)) {
                        const double zeta = (aqq - app) / (2.0 * apq);
                        const double t = copysign(1.0, zeta) / (fabs(zeta) + hypot(1.0, zeta));
                        c = 1.0 / sqrt(fma(t, t, 1.0));
                        sn = c * t;
                        rotated = 1;
                    }
                } else {
                    q = -1;  // p is paired with the padding slot
                }
                cs[tid][0] = c;
                cs[tid][1] = sn;
                pp[tid] = p;
                qq[tid] = q;
            }
            __syncthreads();
            for (int e = tid; e < half * half; e += NT) {
                const int ka = e / half, kb = e - ka * half;
                const double ca = cs[ka][0], sa = cs[ka][1], cb = cs[kb][0], sb = cs[kb][1];
                if (sa == 0.0 && sb == 0.0) continue;
                const int p = pp[ka], q = qq[ka], u = pp[kb], v = qq[kb];
                const bool r2 = q >= 0, c2 = v >= 0;
                const double apu = A[p * ld + u], apv = c2 ? A[p * ld + v] : 0.0;
                const double aqu = r2 ? A[q * ld + u] : 0.0, aqv = (r2 && c2) ? A[q * ld + v] : 0.0;
                const double rpu = ca * apu - sa * aqu, rpv = ca * apv - sa * aqv;
                const double rqu = sa * apu + ca * aqu, rqv = sa * apv + ca * aqv;
                A[p * ld + u] = cb * rpu - sb * rpv;
                if (c2) A[p * ld + v] = sb * rpu + cb * rpv;
                if (r2) {
                    A[q * ld + u] = cb * rqu - sb * rqv;
                    if (c2) A[q * ld + v] = sb * rqu + cb * rqv;
                }
                if (ka == kb && sa != 0.0) {  // the annihilated pair
                    A[p * ld + q] = 0.0;
                    A[q * ld + p] = 0.0;
                }
            }
            for (int e = tid; e < half * g; e += NT) {
                const int k = e / g, i = e - k * g;
                const double sn = cs[k][1];
                if (sn != 0.0) {
                    const double c = cs[k][0];
                    const int p = pp[k], q = qq[k];
                    const double u = R[p * ld + i], v = R[q * ld + i];
                    R[p * ld + i] = c * u - sn * v;
                    R[q * ld + i] = sn * u + c * v;
                }
            }
            __syncthreads();
        }
        const int any = rotated;
        __syncthreads();
        if (!any) break;
    }
    __syncthreads();
    // decreasing order by eigenvalue (ties by index): rank of each row
    for (int i = tid; i < g; i += NT) {
        const double li = A[i * ld + i];
        int rk = 0;
        for (int j = 0; j < g; ++j) {
            const double lj = A[j * ld + j];
            rk += (lj > li) || (lj == li && j < i);
        }
        rank[i] = rk;
        P.s[rk] = sqrt(fmax(li, 0.0));
    }
    __syncthreads();
    for (int e = tid; e < g * g; e += NT) {
        const int i = e / g, j = e - i * g;
        P.rt[(int64_t)rank[i] * P.ldr + j] = R[i * ld + j];
    }
    if (tid == 0 && P.sweeps) *P.sweeps = sweep;
}

}  // namespace syn

extern "C" int syn_jacobi_gram_f64(const syn_jacobi_item *items, int n_items, int64_t nv_max, double tol,
                                   int max_sweeps, void *stream) {
    using namespace syn;
    if (n_items < 0 || nv_max < 1 || nv_max > kGramMax || max_sweeps < 0 || !(tol >= 0.0)) {
        set_error("syn_jacobi_gram_f64: bad arguments (nv_max=%lld, limit %d)", (long long)nv_max, kGramMax);
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    static bool attr_set = false;
    if (!attr_set) {   // orders above 32 take more than 64 KiB of LDS
        if (int rc = hip_check(hipFuncSetAttribute((const void *)jacobi_gram_kernel<kGramThreads>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)(2 * kGramMax * (kGramMax + 1) * sizeof(double))),
                               "syn_jacobi_gram_f64: hipFuncSetAttribute"))
            return rc;
        attr_set = true;
    }
    const size_t shm = 2 * (size_t)nv_max * (size_t)(nv_max + 1) * sizeof(double);
    ProfScope ps(kTagJacobi, 0.0, (hipStream_t)stream);
    auto kern = nv_max <= 16 ? jacobi_gram_kernel<64>
                             : (nv_max <= 32 ? jacobi_gram_kernel<128> : jacobi_gram_kernel<kGramThreads>);
    const int nt = nv_max <= 16 ? 64 : (nv_max <= 32 ? 128 : kGramThreads);
    hipLaunchKernelGGL(kern, dim3(n_items), dim3(nt), shm, (hipStream_t)stream, items,
                       (int)nv_max, tol, max_sweeps);
    SYN_LAUNCH_CHECK("syn_jacobi_gram_f64 launch");
    return SYN_OK;
}

extern "C" int syn_jacobi_f64(const syn_jacobi_item *items, int n_items, double tol, int max_sweeps,
                              void *stream) {
    using namespace syn;
    if (n_items < 0 || max_sweeps < 0 || !(tol >= 0.0)) {
        set_error("syn_jacobi_f64: bad arguments");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)jacobi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(kJacLdsDoubles * sizeof(double)));
        attr_set = true;
    }
    ProfScope ps(kTagJacobi, 0.0, (hipStream_t)stream);
    hipLaunchKernelGGL(jacobi_kernel, dim3(n_items), dim3(kJacThreads), kJacLdsDoubles * sizeof(double),
                       (hipStream_t)stream, items, tol, max_sweeps);
    SYN_LAUNCH_CHECK("syn_jacobi_f64 launch");
    return SYN_OK;
}
