// Blocked Cholesky and CholeskyQR for the left-orthonormalisation sweep.
//
// syn_cholqr_f64 performs one CholeskyQR pass on a tall m x n matrix A:
//   G = A^T A (MFMA GEMM), G = L L^T (blocked right-looking Cholesky),
//   Q = A L^{-T} (blocked right-looking triangular solve).
// Two passes (CholeskyQR2, Yamamoto et al., ETNA 44, 2015) give Q orthonormal
// to working precision whenever cond(A) < ~1e8; the sweeps of the workloads
// here stay below 1e5.
//
// Semidefinite Grams: a pivot d_j <= tolrel * max_i G_ii is dropped.  Its
// column of L is set to zero and, in exact arithmetic, the Schur complement
// row/column of a positive semidefinite matrix with a zero pivot vanishes, so
// the factorisation continues.  The matching column of Q comes out as zero
// and the matching row of R = L^T as zero; Q R = A still holds up to the
// dropped weight, which is below tolrel * max_i ||A e_i||^2.
//
// Every step is a kernel launched from the host with its arguments by value;
// the 64 x 64 diagonal blocks are factored and inverted in LDS by one
// work-group, the panel solves and trailing updates run on the GEMM.
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kNb = 64;

__global__ __launch_bounds__(256) void fill_kernel(double *x, int64_t rows, int64_t cols, int64_t ld, double v) {
    const int64_t nel = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nel; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e - i * cols;
        x[i * ld + j] = v;
    }
}

__global__ __launch_bounds__(256) void maxdiag_kernel(const double *g, int64_t n, int64_t ld, double *out,
                                                      int *dropped) {
    __shared__ double red[4];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, g[i * ld + i]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        *out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        *dropped = 0;
    }
}

// Factor the diagonal block G[j0:j0+nb, j0:j0+nb] = L11 L11^T in LDS, write
// L11 (lower, zeros above) into L and Z = inv(L~11) with dropped columns
// zeroed into zb (nb x nb, ld nb).  L~11 equals L11 with every dropped row
// replaced by the unit row, so that a panel solve X L11^T = B over the kept
// columns is X = B Z^T with zero columns at the dropped positions.
__global__ __launch_bounds__(256) void chol_diag_kernel(const double *G, int64_t ldg, double *L, int64_t ldl,
                                                        int64_t n, int64_t j0, double *zb, const double *maxdiag,
                                                        double tolrel, int *dropped) {
    __shared__ double a[kNb][kNb + 1];
    __shared__ double z[kNb][kNb + 1];
    __shared__ int keep[kNb];
    const int nb = (int)((n - j0) < kNb ? (n - j0) : kNb);
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const double thr = tolrel * (*maxdiag);
    {
        double v[kNb * kNb / 256];
#pragma unroll
        for (int q = 0; q < kNb * kNb / 256; ++q) {
            const int e = tid + 256 * q, i = e >> 6, k = e & 63;
            v[q] = (i < nb && k <= i) ? G[(j0 + i) * ldg + (j0 + k)] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kNb * kNb / 256; ++q) {
            const int e = tid + 256 * q, i = e >> 6, k = e & 63;
            a[i][k] = v[q];
            z[i][k] = 0.0;
        }
    }
    __syncthreads();
    int ndrop = 0;
    for (int j = 0; j < nb; ++j) {
        const double d = a[j][j];  // every thread decides the same way
        const bool kp = d > thr;
        const double inv = kp ? 1.0 / sqrt(d) : 0.0;
        ndrop += !kp;
        // column j below the diagonal (rows j+1..nb-1) and the pivot itself
        if (tid < nb && tid > j) a[tid][j] *= inv;
        __syncthreads();
        if (tid == 0) {
            keep[j] = kp;
            a[j][j] = kp ? sqrt(d) : 0.0;
        }
        // trailing lower triangle: a[i][k] -= a[i][j] a[k][j], j < k <= i,
        // thread (ty, tx) owns rows j+1+ty+16u and columns j+1+tx+16v
        for (int i = j + 1 + ty; i < nb; i += 16) {
            const double aij = a[i][j];
            for (int k = j + 1 + tx; k <= i; k += 16) a[i][k] = fma(-aij, a[k][j], a[i][k]);
        }
        __syncthreads();
    }
    if (tid == 0 && ndrop) atomicAdd(dropped, ndrop);
    // L~ = L with every dropped row replaced by the unit row (in place)
    for (int e = tid; e < nb * nb; e += 256) {
        const int i = e / nb, k = e % nb;
        if (!keep[i] && k <= i) {
            z[i][k] = a[i][k];  // keep L's row for the write-out below
            a[i][k] = (i == k) ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    // Z = inv(L~) by rows: z_i,c = -(sum_{c<=k<i} L~_ik z_kc) / L~_ii
    __shared__ double zz[kNb][kNb + 1];
    for (int e = tid; e < kNb * kNb; e += 256) zz[e >> 6][e & 63] = 0.0;
    __syncthreads();
    if (tid < nb) zz[tid][tid] = 1.0 / a[tid][tid];
    __syncthreads();
    for (int i = 1; i < nb; ++i) {
        // columns c < i in parallel; four threads per column split the k range
        const int c = tid >> 2, part = tid & 3;
        double acc = 0.0;
        if (c < i)
            for (int k = c + part; k < i; k += 4) acc = fma(a[i][k], zz[k][c], acc);
        acc += __shfl_xor(acc, 1, 64);
        acc += __shfl_xor(acc, 2, 64);
        if (c < i && part == 0) zz[i][c] = -acc / a[i][i];
        __syncthreads();
    }
    for (int e = tid; e < nb * nb; e += 256) {
        const int i = e / nb, k = e % nb;
        const double lv = keep[i] ? a[i][k] : z[i][k];
        L[(j0 + i) * ldl + (j0 + k)] = (k <= i) ? lv : 0.0;
        zb[i * kNb + k] = keep[k] ? zz[i][k] : 0.0;
    }
}

static int fill(double *x, int64_t rows, int64_t cols, int64_t ld, double v, hipStream_t s) {
    const int64_t nel = rows * cols;
    if (nel <= 0) return SYN_OK;
    int64_t blocks = (nel + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, rows, cols, ld, v);
    SYN_LAUNCH_CHECK("fill");
    return SYN_OK;
}

// Blocked Cholesky of G (destroyed) into L; Z holds the inverted diagonal
// blocks (ceil(n/64) blocks of 64 x 64).  scal[0] = max diag, *dropped counts
// dropped pivots.
static int potrf(double *G, int64_t ldg, double *L, int64_t ldl, int64_t n, double *Z, double *scal, int *dropped,
                 double tolrel, hipStream_t s) {
    int rc = fill(L, n, n, ldl, 0.0, s);
    if (rc) return rc;
    hipLaunchKernelGGL(maxdiag_kernel, dim3(1), dim3(256), 0, s, G, n, ldg, scal, dropped);
    SYN_LAUNCH_CHECK("maxdiag");
    for (int64_t j0 = 0, b = 0; j0 < n; j0 += kNb, ++b) {
        const int64_t nb = (n - j0) < kNb ? (n - j0) : kNb;
        const int64_t j1 = j0 + nb;
        double *zb = Z + b * kNb * kNb;
        ProfScope ps(kTagCholDiag, (double)nb * nb * nb * (1.0 / 3.0 + 1.0 / 3.0), s);
        hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, s, G, ldg, L, ldl, n, j0, zb, scal, tolrel,
                           dropped);
        SYN_LAUNCH_CHECK("chol_diag");
        if (j1 >= n) break;
        // L21 = G21 Z^T
        rc = gemm_f64_launch(make_gemm(G + j1 * ldg + j0, ldg, 0, zb, kNb, 1, L + j1 * ldl + j0, ldl, n - j1, nb, nb,
                                       1.0, 0.0),
                             s);
        if (rc) return rc;
        // G22 -= L21 L21^T
        rc = gemm_f64_launch(make_gemm(L + j1 * ldl + j0, ldl, 0, L + j1 * ldl + j0, ldl, 1, G + j1 * ldg + j1, ldg,
                                       n - j1, n - j1, nb, -1.0, 1.0),
                             s);
        if (rc) return rc;
    }
    return SYN_OK;
}

// Q = A L^{-T} by block columns (A destroyed): Q_J = A_J Z_J^T,
// A_{J+1:} -= Q_J L_{J+1:,J}^T.
static int trsm_right_lt(double *A, int64_t lda, const double *L, int64_t ldl, const double *Z, double *Q,
                         int64_t ldq, int64_t m, int64_t n, hipStream_t s) {
    for (int64_t j0 = 0, b = 0; j0 < n; j0 += kNb, ++b) {
        const int64_t nb = (n - j0) < kNb ? (n - j0) : kNb;
        const int64_t j1 = j0 + nb;
        int rc = gemm_f64_launch(make_gemm(A + j0, lda, 0, Z + b * kNb * kNb, kNb, 1, Q + j0, ldq, m, nb, nb, 1.0, 0.0), s);
        if (rc) return rc;
        if (j1 >= n) break;
        rc = gemm_f64_launch(make_gemm(Q + j0, ldq, 0, L + j1 * ldl + j0, ldl, 1, A + j1, lda, m, n - j1, nb, -1.0, 1.0), s);
        if (rc) return rc;
    }
    return SYN_OK;
}

}  // namespace syn

extern "C" int64_t syn_cholqr_workspace(int64_t n) {
    const int64_t nblk = (n + syn::kNb - 1) / syn::kNb;
    return n * n + nblk * syn::kNb * syn::kNb + 8;
}

extern "C" int syn_potrf_f64(double *G, int64_t n, int64_t ldg, double *L, int64_t ldl, double *work, double tolrel,
                             int *dropped, void *stream) {
    using namespace syn;
    if (n <= 0 || ldg < n || ldl < n || !G || !L || !work || !dropped || !(tolrel >= 0.0)) {
        set_error("syn_potrf_f64: bad arguments (n=%lld)", (long long)n);
        return SYN_EINVAL;
    }
    const int64_t nblk = (n + kNb - 1) / kNb;
    double *Z = work;
    double *scal = work + nblk * kNb * kNb;
    return potrf(G, ldg, L, ldl, n, Z, scal, dropped, tolrel, (hipStream_t)stream);
}

extern "C" int syn_cholqr_f64(double *A, int64_t m, int64_t n, int64_t lda, double *Q, int64_t ldq, double *L,
                              int64_t ldl, double *work, double tolrel, int *dropped, void *stream) {
    using namespace syn;
    if (m <= 0 || n <= 0 || lda < n || ldq < n || ldl < n || !A || !Q || !L || !work || !dropped ||
        !(tolrel >= 0.0)) {
        set_error("syn_cholqr_f64: bad arguments (m=%lld n=%lld)", (long long)m, (long long)n);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblk = (n + kNb - 1) / kNb;
    double *G = work;
    double *Z = work + n * n;
    double *scal = Z + nblk * kNb * kNb;
    int rc = gemm_f64_launch(make_gemm(A, lda, 1, A, lda, 0, G, n, n, n, m, 1.0, 0.0), s);
    if (rc) return rc;
    rc = potrf(G, n, L, ldl, n, Z, scal, dropped, tolrel, s);
    if (rc) return rc;
    return trsm_right_lt(A, lda, L, ldl, Z, Q, ldq, m, n, s);
}
