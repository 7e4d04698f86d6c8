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

}  // namespace syn

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
