// Batched Householder QR, one work-group per matrix.
//
// Used wherever the reference calls numpy.linalg.qr on a site unfolding: the
// QR-truncation ">>" sweep, decompose, and left/right orthonormalisation.
// The leading nb = min(n, keep) columns are factored (the later columns only
// enter R = Q^T A, which the caller forms with the GEMM), then the first kq
// columns of the complete orthogonal factor are accumulated backwards,
// Q = H_0 H_1 ... H_{p-1} [I; 0]  (Golub & Van Loan, Alg. 5.1.2 / 5.2.1).
//
// Parallel layout: the 64 lanes of a wave own 64 consecutive columns and the
// 4 waves own 4 contiguous row segments, so every global access of the
// row-major W and Q is coalesced along a row; the per-column partial dot
// products of the 4 segments are combined through LDS.  The working copy of
// the leading block lives in LDS when it fits and in the caller's workspace
// (L2-resident at these sizes) otherwise.
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kQrThreads = 256;
constexpr int kQrWaves = kQrThreads / 64;
constexpr int64_t kQrLdsDoubles = (96 * 1024) / 8;

__device__ __forceinline__ void seg_bounds(int64_t lo, int64_t hi, int w, int64_t &a, int64_t &b) {
    const int64_t len = hi > lo ? hi - lo : 0;
    const int64_t per = (len + kQrWaves - 1) / kQrWaves;
    a = lo + per * w;
    b = a + per;
    if (a > hi) a = hi;
    if (b > hi) b = hi;
}

// Apply H = I - tau v v^T (v = [1; M[j+1:, j]] taken from V) to columns
// [c_lo, c_hi) of M (rows j..m-1).  V and M may alias (factorisation step).
__device__ __forceinline__ void apply_reflector(const double *V, int64_t ldv, int64_t j, double tau, double *M,
                                                int64_t ldm, int64_t m, int64_t c_lo, int64_t c_hi,
                                                double (*red)[64]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t r0, r1;
    seg_bounds(j + 1, m, w, r0, r1);
    for (int64_t c0 = c_lo; c0 < c_hi; c0 += 64) {
        const int64_t c = c0 + lane;
        const bool act = c < c_hi;
        double part = 0.0;
        if (act)
            for (int64_t i = r0; i < r1; ++i) part = fma(V[i * ldv + j], M[i * ldm + c], part);
        red[w][lane] = part;
        __syncthreads();
        double dot = 0.0;
        if (act) {
            dot = M[j * ldm + c];
#pragma unroll
            for (int q = 0; q < kQrWaves; ++q) dot += red[q][lane];
        }
        const double f = tau * dot;
        __syncthreads();  // every wave has read red[] and M[j, c] before row j changes
        if (act) {
            if (w == 0) M[j * ldm + c] -= f;
            for (int64_t i = r0; i < r1; ++i) M[i * ldm + c] = fma(-f, V[i * ldv + j], M[i * ldm + c]);
        }
    }
}

template <bool IN_LDS>
__device__ void qr_one(const syn_qr_item &P, double *lds_w) {
    __shared__ double red[kQrWaves][64];
    __shared__ double shv[4];
    const int64_t m = P.m, n = P.n, keep = P.keep;
    const int64_t nb = n < keep ? n : keep;
    const int64_t kq = keep < m ? keep : m;
    const int64_t p = m < nb ? m : nb;
    double *W = IN_LDS ? lds_w : P.work;
    const int tid = threadIdx.x;

    for (int64_t e = tid; e < m * nb; e += kQrThreads) {
        const int64_t i = e / nb, j = e - i * nb;
        W[e] = P.a[i * P.rs + j * P.cs];
    }
    __syncthreads();

    for (int64_t j = 0; j < p; ++j) {
        double part = 0.0;
        for (int64_t i = j + 1 + tid; i < m; i += kQrThreads) {
            const double v = W[i * nb + j];
            part = fma(v, v, part);
        }
        part = wave_sum(part);
        if ((tid & 63) == 0) red[tid >> 6][0] = part;
        __syncthreads();
        if (tid == 0) {
            double xn2 = 0.0;
            for (int q = 0; q < kQrWaves; ++q) xn2 += red[q][0];
            const double alpha = W[j * nb + j];
            double tau = 0.0, scal = 0.0;
            if (xn2 > 0.0) {
                const double beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
                tau = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
                W[j * nb + j] = beta;
            }
            shv[0] = tau;
            shv[1] = scal;
            P.tau[j] = tau;
        }
        __syncthreads();
        const double tau = shv[0], scal = shv[1];
        if (tau != 0.0) {
            for (int64_t i = j + 1 + tid; i < m; i += kQrThreads) W[i * nb + j] *= scal;
            __syncthreads();
            apply_reflector(W, nb, j, tau, W, nb, m, j + 1, nb, red);
        }
        __syncthreads();
    }

    // Q is accumulated row-major in P.q (or, for q_trans, in the workspace
    // tail and transposed on the way out).
    double *Q = P.q_trans ? P.work + m * nb : P.q;
    const int64_t ldq = P.q_trans ? kq : P.ldq;
    for (int64_t e = tid; e < m * kq; e += kQrThreads) {
        const int64_t i = e / kq, c = e - i * kq;
        Q[i * ldq + c] = (i == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int64_t j = p - 1; j >= 0; --j) {
        const double tau = P.tau[j];
        if (tau != 0.0) apply_reflector(W, nb, j, tau, Q, ldq, m, j, kq, red);
        __syncthreads();
    }
    if (P.q_trans) {
        for (int64_t e = tid; e < m * kq; e += kQrThreads) {
            const int64_t c = e / m, i = e - c * m;
            P.q[c * P.ldq + i] = Q[i * ldq + c];
        }
    }
}

// lds_cap: doubles of dynamic LDS the launch holds; a matrix whose leading
// block fits works in LDS, a larger one in its workspace.
__global__ __launch_bounds__(kQrThreads) void qr_kernel(const syn_qr_item *__restrict__ items, int n_items,
                                                         int64_t lds_cap) {
    extern __shared__ double lds_w[];
    const syn_qr_item &P = items[blockIdx.x];
    const int64_t nb = P.n < P.keep ? P.n : P.keep;
    if (P.m * nb <= lds_cap)
        qr_one<true>(P, lds_w);
    else
        qr_one<false>(P, lds_w);
}

}  // namespace syn

extern "C" int syn_qr_f64(const syn_qr_item *items, int n_items, void *stream) {
    using namespace syn;
    if (n_items < 0) {
        set_error("syn_qr_f64: negative item count");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    static bool attr_set = false;
    if (!attr_set) {
        if (int rc = hip_check(hipFuncSetAttribute((const void *)qr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)(kQrLdsDoubles * sizeof(double))),
                               "syn_qr_f64: hipFuncSetAttribute"))
            return rc;
        attr_set = true;
    }
    ProfScope ps(kTagQr, 0.0, (hipStream_t)stream);
    hipLaunchKernelGGL(qr_kernel, dim3(n_items), dim3(kQrThreads), kQrLdsDoubles * sizeof(double),
                       (hipStream_t)stream, items, n_items, kQrLdsDoubles);
    SYN_LAUNCH_CHECK("syn_qr_f64 launch");
    return SYN_OK;
}

extern "C" int syn_qr_sized_f64(const syn_qr_item *items, int n_items, int64_t lds_doubles, void *stream) {
    using namespace syn;
    if (n_items < 0 || lds_doubles < 0 || lds_doubles > kQrLdsDoubles) {
        set_error("syn_qr_sized_f64: bad item count %d or LDS size %lld (at most %lld doubles)", n_items,
                  (long long)lds_doubles, (long long)kQrLdsDoubles);
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    static bool attr_set = false;
    if (!attr_set) {
        if (int rc = hip_check(hipFuncSetAttribute((const void *)qr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)(kQrLdsDoubles * sizeof(double))),
                               "syn_qr_sized_f64: hipFuncSetAttribute"))
            return rc;
        attr_set = true;
    }
    ProfScope ps(kTagQr, 0.0, (hipStream_t)stream);
    hipLaunchKernelGGL(qr_kernel, dim3(n_items), dim3(kQrThreads), (size_t)lds_doubles * sizeof(double),
                       (hipStream_t)stream, items, n_items, lds_doubles);
    SYN_LAUNCH_CHECK("syn_qr_sized_f64 launch");
    return SYN_OK;
}
