// Panel factorisation of the blocked Householder QR.
//
// The blocked QR (LAPACK dgeqrf / dorgqr structure) produces the same
// reflectors as the one-work-group kernel of qr.hip, but only a panel of at
// most 32 columns is factored column by column; the rest of the matrix is
// updated with the compact WY form H_j0 ... H_{j0+b-1} = I - V T V^T on the
// MFMA GEMM, and Q is accumulated panel by panel the same way.  This kernel
// factors one panel in place (one work-group of 1024 threads, rows spread
// over the threads, the panel in L2) and writes the explicit V (unit
// diagonal, zeros above) and T (LAPACK dlarft, forward, columnwise:
// T[0:j, j] = -tau_j T[0:j, 0:j] V[:, 0:j]^T v_j, T[j, j] = tau_j).
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kPanelThreads = 1024;
constexpr int kPanelWaves = kPanelThreads / 64;
constexpr int kPanelMax = 32;
// the dot products and updates use a (row group, column) layout: lane c of
// a 32-thread group owns panel column c, the 32 groups interleave the rows,
// so every pass over the panel reads whole row segments (coalesced) and the
// per-column sums need one LDS reduction over the groups
constexpr int kRowGroups = kPanelThreads / kPanelMax;

// sum over rows i = i0, i0 + kRowGroups, ... < m of a[i * lda] * b[i * ldb],
// four rows in flight (two partial sums), so the loads of a thread overlap
__device__ __forceinline__ double strided_dot(const double *__restrict__ a, int64_t lda,
                                              const double *__restrict__ b, int64_t ldb, int64_t i0, int64_t m) {
    double s0 = 0.0, s1 = 0.0;
    constexpr int64_t G = kRowGroups;
    const double *pa = a + i0 * lda, *pb = b + i0 * ldb;
    int64_t i = i0;
    for (; i + 3 * G < m; i += 4 * G, pa += 4 * G * lda, pb += 4 * G * ldb) {
        const double x0 = pa[0], x1 = pa[G * lda], x2 = pa[2 * G * lda], x3 = pa[3 * G * lda];
        const double y0 = pb[0], y1 = pb[G * ldb], y2 = pb[2 * G * ldb], y3 = pb[3 * G * ldb];
        s0 = fma(x0, y0, s0);
        s1 = fma(x1, y1, s1);
        s0 = fma(x2, y2, s0);
        s1 = fma(x3, y3, s1);
    }
    for (; i < m; i += G, pa += G * lda, pb += G * ldb) s0 = fma(pa[0], pb[0], s0);
    return s0 + s1;
}

// w[i * ldw] -= f * v[i * ldv] for rows i = i0, i0 + kRowGroups, ... < m,
// four rows in flight
__device__ __forceinline__ void strided_axpy(double *__restrict__ w, int64_t ldw, const double *__restrict__ v,
                                             int64_t ldv, double f, int64_t i0, int64_t m) {
    constexpr int64_t G = kRowGroups;
    double *pw = w + i0 * ldw;
    const double *pv = v + i0 * ldv;
    int64_t i = i0;
    for (; i + 3 * G < m; i += 4 * G, pw += 4 * G * ldw, pv += 4 * G * ldv) {
        const double x0 = pv[0], x1 = pv[G * ldv], x2 = pv[2 * G * ldv], x3 = pv[3 * G * ldv];
        const double y0 = pw[0], y1 = pw[G * ldw], y2 = pw[2 * G * ldw], y3 = pw[3 * G * ldw];
        pw[0] = fma(-f, x0, y0);
        pw[G * ldw] = fma(-f, x1, y1);
        pw[2 * G * ldw] = fma(-f, x2, y2);
        pw[3 * G * ldw] = fma(-f, x3, y3);
    }
    for (; i < m; i += G, pw += G * ldw, pv += G * ldv) pw[0] = fma(-f, pv[0], pw[0]);
}

__global__ __launch_bounds__(kPanelThreads) void hh_panel_kernel(double *__restrict__ W, int64_t ldw,
                                                                 double *__restrict__ V, int64_t ldv,
                                                                 double *__restrict__ tau, double *__restrict__ T,
                                                                 int64_t ldt, int64_t m, int64_t j0, int b) {
    __shared__ double red[kPanelWaves];
    __shared__ double part[kRowGroups][kPanelMax + 1];
    __shared__ double fvec[kPanelMax];
    __shared__ double stau[kPanelMax];
    __shared__ double Ts[kPanelMax][kPanelMax + 1];
    __shared__ double shv[2];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = tid & (kPanelMax - 1), rg = tid / kPanelMax;

    for (int jj = 0; jj < b; ++jj) {
        const int64_t j = j0 + jj;
        // the column below the diagonal: its squared norm
        double sq = 0.0;
        for (int64_t i = j + 1 + tid; i < m; i += kPanelThreads) {
            const double v = W[i * ldw + j];
            sq = fma(v, v, sq);
        }
        sq = wave_sum(sq);
        if (lane == 0) red[w] = sq;
        __syncthreads();
        if (tid == 0) {
            double xn2 = 0.0;
            for (int q = 0; q < kPanelWaves; ++q) xn2 += red[q];
            const double alpha = W[j * ldw + j];
            double t = 0.0, scal = 0.0;
            if (xn2 > 0.0) {
                const double beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
                t = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
                W[j * ldw + j] = beta;
            }
            shv[0] = t;
            shv[1] = scal;
            stau[jj] = t;
            tau[j] = t;
        }
        __syncthreads();
        const double t = shv[0], scal = shv[1];
        // v = [0; 1; x * scal]: scale the column in place and write V
        for (int64_t i = j0 + tid; i < m; i += kPanelThreads) {
            double vi;
            if (i < j) {
                vi = 0.0;
            } else if (i == j) {
                vi = 1.0;
            } else {
                vi = (t != 0.0) ? W[i * ldw + j] * scal : 0.0;
                if (t != 0.0) W[i * ldw + j] = vi;
            }
            V[i * ldv + j] = vi;
        }
        __syncthreads();
        const int nc = b - jj - 1;
        if (t != 0.0 && nc > 0) {
            // dots of v with the panel columns right of j
            double acc = 0.0;
            if (c < nc) acc = strided_dot(V + j, ldv, W + j + 1 + c, ldw, j + 1 + rg, m);
            part[rg][c] = acc;
            __syncthreads();
            if (tid < nc) {
                double dot = W[j * ldw + j + 1 + tid];
                for (int g = 0; g < kRowGroups; ++g) dot += part[g][tid];
                fvec[tid] = t * dot;
            }
            __syncthreads();
            if (tid < nc) W[j * ldw + j + 1 + tid] -= fvec[tid];
            if (c < nc) strided_axpy(W + j + 1 + c, ldw, V + j, ldv, fvec[c], j + 1 + rg, m);
        }
        __syncthreads();
    }

    // T, column by column: T[0:jj, jj] = -tau_jj T[0:jj, 0:jj] (V[:, j0:j0+jj]^T v_jj)
    for (int jj = 0; jj < b; ++jj) {
        const int64_t j = j0 + jj;
        double acc = 0.0;
        if (c < jj) acc = strided_dot(V + j0 + c, ldv, V + j, ldv, j + rg, m);
        part[rg][c] = acc;
        __syncthreads();
        if (tid < jj) {
            double s = 0.0;
            for (int g = 0; g < kRowGroups; ++g) s += part[g][tid];
            fvec[tid] = s;
        }
        __syncthreads();
        if (tid < jj) {
            double s = 0.0;
            for (int l = tid; l < jj; ++l) s = fma(Ts[tid][l], fvec[l], s);
            Ts[tid][jj] = -stau[jj] * s;
        }
        if (tid == 0) Ts[jj][jj] = stau[jj];
        __syncthreads();
    }
    for (int e = tid; e < b * b; e += kPanelThreads) {
        const int r = e / b, cc = e - r * b;
        T[r * ldt + cc] = (r <= cc) ? Ts[r][cc] : 0.0;
    }
}

}  // namespace syn

extern "C" int syn_hh_panel_f64(double *w, int64_t ldw, double *v, int64_t ldv, double *tau, double *t, int64_t ldt,
                                int64_t m, int64_t j0, int64_t b, void *stream) {
    using namespace syn;
    if (!w || !v || !tau || !t || m <= 0 || j0 < 0 || j0 >= m || b <= 0 || b > kPanelMax || ldw < j0 + b ||
        ldv < j0 + b || ldt < b) {
        set_error("syn_hh_panel_f64: bad arguments (m=%lld j0=%lld b=%lld)", (long long)m, (long long)j0,
                  (long long)b);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    ProfScope ps(kTagQr, 0.0, s);
    hipLaunchKernelGGL(hh_panel_kernel, dim3(1), dim3(kPanelThreads), 0, s, w, ldw, v, ldv, tau, t, ldt, m, j0,
                       (int)b);
    SYN_LAUNCH_CHECK("syn_hh_panel_f64 launch");
    return SYN_OK;
}
