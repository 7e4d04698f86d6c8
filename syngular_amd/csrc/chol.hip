// Blocked Cholesky and CholeskyQR for the left-orthonormalisation sweep.
//
// syn_cholqr_f64 performs one CholeskyQR pass on a tall m x n matrix A:
//   G = A^T A (MFMA GEMM), G = L L^T together with X = L^{-1} (recursive,
//   GEMM-rich), Q = A X^T (one MFMA GEMM).
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
// work-group, everything above them runs on the GEMM.
#include "common.hpp"
#include "prof.hpp"

#include <algorithm>

namespace syn {

constexpr int kNb = 64;

// 1 / x for normal positive x from the hardware estimate (v_rcp_f64) and two
// Newton-Raphson steps r <- r + r (1 - x r), each squaring the relative
// error, to within a unit or two in the last place: shorter than the IEEE
// division sequence on a panel step's critical path
__device__ __forceinline__ double recip_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Work-group barrier that orders LDS only: global stores issued before it
// stay in flight.  For phases whose global writes no thread of the
// work-group reads back (a plain __syncthreads waits for them to land).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
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
// columns is X = B Z^T with zero columns at the dropped positions.  a and lo
// are two kNb x (kNb + 1) LDS tiles of the caller (free again on return):
// the Schur complements, then L and L~ (dropped rows -> unit rows).
template <int TG>
__device__ __forceinline__ void chol_diag_body(const double *G, int64_t ldg, double *L, int64_t ldl, int64_t n,
                                               int64_t j0, double *zb, int64_t ldz, const double *maxdiag,
                                               double tolrel, int *dropped, double (*a)[kNb + 1],
                                               double (*lo)[kNb + 1]) {
    __shared__ int keep[kNb];
    const int nb = (int)((n - j0) < kNb ? (n - j0) : kNb);
    constexpr int NT = TG * TG;
    const int tid = threadIdx.x;
    const double thr = tolrel * (*maxdiag);
    {
        double v[kNb * kNb / NT];
#pragma unroll
        for (int q = 0; q < kNb * kNb / NT; ++q) {
            const int e = tid + NT * q, i = e >> 6, k = e & 63;
            v[q] = (i < nb && k <= i) ? G[(j0 + i) * ldg + (j0 + k)] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kNb * kNb / NT; ++q) {
            const int e = tid + NT * q, i = e >> 6, k = e & 63;
            a[i][k] = v[q];
            lo[i][k] = 0.0;
        }
    }
    lds_barrier();
    // Blocked right-looking factorisation in four panels of 16 columns.  Wave
    // 0 factors a panel in registers (lane r holds row r of the panel, the
    // pivot and the column entries reach every lane as LDS broadcasts: no
    // work-group barrier inside a panel); then the four waves subtract
    // its rank-16 product from the trailing lower tiles of a on the matrix
    // cores.  lo receives the unscaled columns (pivot on the diagonal) and
    // rdk the pivots' reciprocals (0 for dropped and padding columns), as the
    // column-by-column form leaves them.
    __shared__ double rdk[kNb], colb[kNb];
    int ndrop = 0;
    {
        typedef double v4d __attribute__((ext_vector_type(4)));
        const int w = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
#pragma unroll 1
        for (int k = 0; 16 * k < nb; ++k) {
            const int c0 = 16 * k;
            if (w == 0) {
                double pv[16];
#pragma unroll
                for (int m = 0; m < 16; ++m) pv[m] = a[l][c0 + m];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int jc = c0 + j;
                    // column j reaches every lane through LDS (one wave: its
                    // LDS operations run in order, no barrier needed)
                    colb[l] = pv[j];
                    const double d = colb[jc];
                    double cm[16];
#pragma unroll
                    for (int m = j + 1; m < 16; ++m) cm[m] = colb[c0 + m];
                    const bool valid = jc < nb, kp = valid && d > thr;
                    ndrop += valid && !kp;
                    const double rd = kp ? recip_nr(d) : 0.0;
                    if (valid && l >= jc && l < nb) lo[l][jc] = pv[j];
                    if (l == 0) {
                        rdk[jc] = rd;
                        if (valid) keep[jc] = kp;
                    }
                    const double sr = pv[j] * rd;
#pragma unroll
                    for (int m = j + 1; m < 16; ++m) pv[m] = fma(-sr, cm[m], pv[m]);
                }
            }
            lds_barrier();
            // trailing tiles (I, J), k < J <= I < 4, numbered row by row
            const int nt = (3 - k) * (4 - k) / 2;
            for (int t = w; t < nt; t += 4) {
                int I = k + 1, J = k + 1, q = t;
                while (q > I - k - 1) {
                    q -= I - k;
                    ++I;
                }
                J = k + 1 + q;
                v4d acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = a[16 * I + lk + 4 * r][16 * J + li];
#pragma unroll
                for (int k0 = 0; k0 < 16; k0 += 4) {
                    const int kc = c0 + k0 + lk;
                    // A[i][kk] = -s_(16I+i, kc), B[kk][m] = c_(16J+m, kc)
                    const double av = -(lo[16 * I + li][kc] * rdk[kc]);
                    const double bv = lo[16 * J + li][kc];
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) a[16 * I + lk + 4 * r][16 * J + li] = acc[r];
            }
            lds_barrier();
        }
    }
    if (tid == 0 && ndrop) atomicAdd(dropped, ndrop);
    // L_jj = sqrt(d_j), L_ij = col_ij / L_jj for kept pivots, zero columns for
    // dropped ones; 1 / L~_jj for the inversion below (1 on dropped rows)
    __shared__ double sdiag[kNb], rdiag[kNb];
    if (tid < nb) {
        const double sd = keep[tid] ? sqrt(lo[tid][tid]) : 0.0;
        sdiag[tid] = sd;
        rdiag[tid] = keep[tid] ? 1.0 / sd : 1.0;
    } else if (tid < kNb) {
        rdiag[tid] = 1.0;   // rows past nb: zero rows of L~, unit rows of its inverse
    }
    lds_barrier();
    // L (kept columns scaled by the pivot's reciprocal square root, dropped
    // ones zero) to global memory, and lo turned into L~ (dropped rows ->
    // unit rows): each thread rewrites the entries it has just read, so no
    // barrier separates the two
#pragma unroll
    for (int q = 0; q < kNb * kNb / NT; ++q) {
        const int e = tid + NT * q, i = e >> 6, k = e & 63;
        if (i < nb && k < nb) {
            double v = 0.0;
            if (k <= i) v = (i == k) ? sdiag[k] : (keep[k] ? lo[i][k] * rdiag[k] : 0.0);
            L[(j0 + i) * ldl + (j0 + k)] = v;
            if (k <= i) lo[i][k] = keep[i] ? v : (i == k ? 1.0 : 0.0);
        }
    }
    lds_barrier();
    // Z = inv(L~) by blocks.  The four waves first invert the four 16 x 16
    // diagonal blocks of L~ (wave w: rows and columns 16w .. 16w + 15) by
    // forward substitution: lane (g, c) holds column c of rows g, g + 4, g + 8,
    // g + 12, and a finished row reaches the lanes of its columns by a
    // cross-lane read, so these 16 steps need no work-group barrier.  Two
    // doubling levels follow on the matrix cores, Z21 = -Z22 (L21 Z11):
    // 16 -> 32 for both 32-blocks, then 32 -> 64.  Their products T = L21 Z11
    // are kept in the upper triangle of zs, which Z (lower) does not use, so
    // the Z operands are masked to their lower triangle where they are read.
    static_assert(kNb == 64 && NT == 256, "the blocked inverse assumes four waves and 64 x 64 leaves");
    double(*zs)[kNb + 1] = a;   // the Schur complements are no longer needed
    const int w = tid >> 6, l = tid & 63;
    {
        const int c = l & 15, g = l >> 4, b0 = 16 * w;
        double zr[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) zr[u] = (g + 4 * u == c) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int gk = k & 3, uk = k >> 2;
            if (g == gk) zr[uk] *= rdiag[b0 + k];   // row k is final
            const double zk = __shfl(zr[uk], gk * 16 + c, 64);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (g + 4 * u > k) zr[u] = fma(-lo[b0 + g + 4 * u][b0 + k], zk, zr[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) zs[b0 + g + 4 * u][b0 + c] = zr[u];
    }
    lds_barrier();
    // MFMA f64 16x16x4 operands: A[i = l%16][k = l/16], B[k = l/16][j = l%16];
    // D register r of lane l is D[l/16 + 4r][l%16]
    const int li = l & 15, lk = l >> 4;
    typedef double v4d __attribute__((ext_vector_type(4)));
    // 16 -> 32: waves 0 and 1 take the 32-blocks at o = 0 and o = 32
    if (w < 2) {
        const int o = 32 * w;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {   // T = L21 Z11
            const int k = k0 + lk;
            const double za = lo[o + 16 + li][o + k];
            const double zb_ = (k >= li) ? zs[o + k][o + li] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za, zb_, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) zs[o + lk + 4 * r][o + 16 + li] = acc[r];
    }
    lds_barrier();
    if (w < 2) {
        const int o = 32 * w;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 16; k0 += 4) {   // Z21 = -Z22 T
            const int k = k0 + lk;
            const double za = (k <= li) ? zs[o + 16 + li][o + 16 + k] : 0.0;
            const double zb_ = zs[o + k][o + 16 + li];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za, zb_, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) zs[o + 16 + lk + 4 * r][o + li] = -acc[r];
    }
    lds_barrier();
    // 32 -> 64: wave w takes output tile (ti, tj) of the 32 x 32 products
    {
        const int ti = w >> 1, tj = w & 1;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 4) {   // T = L21 Z11
            const int k = k0 + lk;
            const double za = lo[32 + 16 * ti + li][k];
            const double zb_ = (k >= 16 * tj + li) ? zs[k][16 * tj + li] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za, zb_, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) zs[16 * ti + lk + 4 * r][32 + 16 * tj + li] = acc[r];
    }
    lds_barrier();
    {
        const int ti = w >> 1, tj = w & 1;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 4) {   // Z21 = -Z22 T
            const int k = k0 + lk;
            const double za = (k <= 16 * ti + li) ? zs[32 + 16 * ti + li][32 + k] : 0.0;
            const double zb_ = zs[k][32 + 16 * tj + li];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(za, zb_, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) zs[32 + 16 * ti + lk + 4 * r][16 * tj + li] = -acc[r];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < kNb * kNb / NT; ++q) {
        const int e = tid + NT * q, k = e >> 6, c = e & 63;
        if (k < nb && c < nb) zb[k * ldz + c] = (keep[c] && c <= k) ? zs[k][c] : 0.0;
    }
}

// 16 x 16 threads (four waves): wave 0 factors each 16-column panel, the
// four waves share the trailing tile updates, the blocked inverse and the
// loads and stores.
constexpr int kCholTG = 16;
constexpr int kCholThreads = kCholTG * kCholTG;

__global__ __launch_bounds__(kCholThreads) void chol_diag_kernel(const double *G, int64_t ldg, double *L,
                                                                 int64_t ldl, int64_t n, int64_t j0, double *zb,
                                                                 int64_t ldz, const double *maxdiag, double tolrel,
                                                                 int *dropped) {
    __shared__ double a[kNb][kNb + 1], lo[kNb][kNb + 1];
    chol_diag_body<kCholTG>(G, ldg, L, ldl, n, j0, zb, ldz, maxdiag, tolrel, dropped, a, lo);
}

// One leaf (n <= 64) per work-group, for many problems at once.
__global__ __launch_bounds__(kCholThreads) void chol_leaf_many_kernel(const syn_chol_item *__restrict__ items) {
    __shared__ double a[kNb][kNb + 1], lo[kNb][kNb + 1];
    const syn_chol_item &P = items[blockIdx.x];
    chol_diag_body<kCholTG>(P.g, P.ldg, P.l, P.ldl, P.n, 0, P.x, P.ldx, P.maxdiag, P.tolrel, P.dropped, a, lo);
}

// C (m x nn, ldc) = alpha op(A) op(B) + beta C for 1 <= m, nn, kk <= 64
// inside one work-group of kCholThreads: A (m x kk, lda) and op(B) (kk x nn;
// B^T when bt, B then stored nn x kk) are staged whole into the LDS tiles As
// and Bs, zero-padded to 64, and wave w computes rows 16w .. 16w + 15 with
// f64 16x16x4 MFMA.  Every global load of a thread is issued before the first
// is used (addresses clamped into the operand, padding selected to zero
// afterwards), so their latencies overlap.  lower: only entries j <= i are
// written.  The caller's writes of A, B and C must be visible (a work-group
// barrier); ends with one.
__device__ __forceinline__ void node_gemm(const double *A, int64_t lda, const double *B, int64_t ldb, bool bt,
                                          double *C, int64_t ldc, int m, int nn, int kk, double alpha, double beta,
                                          bool lower, double (*As)[kNb + 1], double (*Bs)[kNb + 1]) {
    static_assert(kNb == 64 && kCholThreads == 256, "the node products assume 64 x 64 tiles and four waves");
    constexpr int Q = kNb * kNb / kCholThreads;
    const int tid = threadIdx.x;
    // B's stored shape: rows x cols = nn x kk (bt) or kk x nn
    const int br = bt ? nn : kk, bc = bt ? kk : nn;
    double va[Q], vb[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = tid + kCholThreads * q, i = e >> 6, k = e & 63;   // consecutive threads: consecutive columns
        va[q] = A[(int64_t)min(i, m - 1) * lda + min(k, kk - 1)];
        vb[q] = B[(int64_t)min(i, br - 1) * ldb + min(k, bc - 1)];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = tid + kCholThreads * q, i = e >> 6, k = e & 63;
        As[i][k] = (i < m && k < kk) ? va[q] : 0.0;
        const double b = (i < br && k < bc) ? vb[q] : 0.0;
        if (bt)
            Bs[k][i] = b;
        else
            Bs[i][k] = b;
    }
    // MFMA f64 16x16x4 operands: A[i = l%16][k = l/16], B[k = l/16][j = l%16];
    // D register r of lane l is D[l/16 + 4r][l%16]
    typedef double v4d __attribute__((ext_vector_type(4)));
    const int w = tid >> 6, l = tid & 63, li = l & 15, lk = l >> 4;
    // C's old values (beta != 0) are loaded now, under the products
    double cv[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * w + lk + 4 * r, j = 16 * t + li;
            cv[t][r] = (beta != 0.0) ? C[(int64_t)min(i, m - 1) * ldc + min(j, nn - 1)] : 0.0;
        }
    __syncthreads();
    v4d acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    // the padding is zero, so the full 64 of K adds nothing past kk
#pragma unroll
    for (int k0 = 0; k0 < kNb; k0 += 4) {
        const double av = As[16 * w + li][k0 + lk];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Bs[k0 + lk][16 * t + li], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * w + lk + 4 * r, j = 16 * t + li;
            if (i < m && j < nn && (!lower || j <= i)) {
                const double v = alpha * acc[t][r];
                C[(int64_t)i * ldc + j] = (beta != 0.0) ? fma(beta, cv[t][r], v) : v;
            }
        }
    __syncthreads();
}

// One node of the recursive Cholesky-inverse with 64 < n <= 128 (split
// 64 + (n - 64)) in one work-group: the leaf of G11, L21 = G21 X11^T,
// G22 -= L21 L21^T (lower), T = L21 X11 (T: (n - 64) x 64, ld 64), the leaf of
// G22 and X21 = -X22 T, the five launches of the recursion at that size.
// Each stage reads what the one before wrote to global memory after a
// work-group barrier; the leaves and the products share the two LDS tiles.
__global__ __launch_bounds__(kCholThreads) void chol_node_kernel(double *G, int64_t ldg, double *L, int64_t ldl,
                                                                 double *X, int64_t ldx, int64_t n, double *T,
                                                                 const double *maxdiag, double tolrel,
                                                                 int *dropped) {
    __shared__ double a[kNb][kNb + 1], lo[kNb][kNb + 1];
    const int n2 = (int)(n - kNb);
    double *G22 = G + kNb * ldg + kNb, *L21 = L + kNb * ldl, *X21 = X + kNb * ldx;
    chol_diag_body<kCholTG>(G, ldg, L, ldl, kNb, 0, X, ldx, maxdiag, tolrel, dropped, a, lo);
    __syncthreads();
    node_gemm(G + kNb * ldg, ldg, X, ldx, true, L21, ldl, n2, kNb, kNb, 1.0, 0.0, false, a, lo);
    node_gemm(L21, ldl, L21, ldl, true, G22, ldg, n2, n2, kNb, -1.0, 1.0, true, a, lo);
    node_gemm(L21, ldl, X, ldx, false, T, kNb, n2, kNb, kNb, 1.0, 0.0, false, a, lo);
    chol_diag_body<kCholTG>(G22, ldg, L21 + kNb, ldl, n2, 0, X21 + kNb, ldx, maxdiag, tolrel, dropped, a, lo);
    __syncthreads();
    node_gemm(X21 + kNb, ldx, T, kNb, false, X21, ldx, n2, kNb, n2, -1.0, 0.0, false, a, lo);
}

__global__ __launch_bounds__(256) void diagmax_many_kernel(const syn_diagmax_item *__restrict__ items) {
    const syn_diagmax_item &P = items[blockIdx.x];
    __shared__ double red[4];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < P.n; i += 256) m = fmax(m, P.g[i * P.ld + i]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        *P.out = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        *P.dropped = 0;
    }
}

// Zero an m x n block (leading dimension ld) on the stream.  These routines
// are not meant for graph capture, so a memset is fine here.
// Zero fill of a rows x cols block (leading dimension ld): 256 columns per
// work-group along x, rows strided over y.  A kernel rather than a memset
// node, so that a stream capture of the CholeskyQR records no memset.
__global__ __launch_bounds__(256) void zero2d_kernel(double *x, int64_t rows, int64_t cols, int64_t ld) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= cols) return;
    for (int64_t i = blockIdx.y; i < rows; i += gridDim.y) x[i * ld + j] = 0.0;
}

static int zero(double *x, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return SYN_OK;
    const int64_t gx = (cols + 255) / 256, gy = rows < 1024 ? rows : 1024;
    if (gx > 0x7fffffffLL) {
        set_error("zero fill: %lld columns", (long long)cols);
        return SYN_EINVAL;
    }
    hipLaunchKernelGGL(zero2d_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, x, rows, cols, ld);
    SYN_LAUNCH_CHECK("zero fill");
    return SYN_OK;
}

// Blocked Cholesky of G (destroyed) into L; Z holds the inverted diagonal
// blocks (ceil(n/64) blocks of 64 x 64).  scal[0] = max diag, *dropped counts
// dropped pivots.
static int potrf(double *G, int64_t ldg, double *L, int64_t ldl, int64_t n, double *Z, double *scal, int *dropped,
                 double tolrel, hipStream_t s) {
    int rc = zero(L, n, n, ldl, s);
    if (rc) return rc;
    hipLaunchKernelGGL(maxdiag_kernel, dim3(1), dim3(256), 0, s, G, n, ldg, scal, dropped);
    SYN_LAUNCH_CHECK("maxdiag");
    for (int64_t j0 = 0, b = 0; j0 < n; j0 += kNb, ++b) {
        const int64_t nb = (n - j0) < kNb ? (n - j0) : kNb;
        const int64_t j1 = j0 + nb;
        double *zb = Z + b * kNb * kNb;
        ProfScope ps(kTagCholDiag, (double)nb * nb * nb * (1.0 / 3.0 + 1.0 / 3.0), s);
        hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(kCholThreads), 0, s, G, ldg, L, ldl, n, j0, zb, (int64_t)kNb,
                           scal, tolrel, dropped);
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

// Recursive Cholesky with inverse: G = L L^T and X = inv(L~) with the
// dropped columns zeroed (the convention of chol_diag_kernel), by
//   (L11, X11) = chol_inv(G11);  L21 = G21 X11^T;  G22 -= L21 L21^T;
//   (L22, X22) = chol_inv(G22);  X21 = -X22 (L21 X11).
// Every step above the 64 x 64 leaves is a GEMM; G22 -= L21 L21^T and
// T = L21 X11 need only L21 and X11 and share one launch.  L and X must be
// zero above the diagonal on entry; T is scratch of chol_inv_scratch(n)
// doubles (this level's n2 x n1 product, the second recursion's past it).
// ws (ws_cap doubles), when given, lets the products split along K; the
// CholeskyQR callers pass none: on the 8192 x 4096 chain the split saved
// under 1% of a step there.
// Nodes with 64 < n <= 128 run in one launch (chol_node_kernel).

static int64_t chol_inv_scratch(int64_t n) {
    if (n <= kNb) return 0;
    const int64_t n1 = ((n / kNb + 1) / 2) * kNb, n2 = n - n1;
    return std::max(chol_inv_scratch(n1), n2 * n1 + chol_inv_scratch(n2));
}

static int chol_inv(double *G, int64_t ldg, double *L, int64_t ldl, double *X, int64_t ldx, int64_t n, double *T,
                    const double *scal, int *dropped, double tolrel, hipStream_t s, double *ws = nullptr,
                    int64_t ws_cap = 0) {
    if (n <= kNb) {
        ProfScope ps(kTagCholDiag, (double)n * n * n * (2.0 / 3.0), s);
        hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(kCholThreads), 0, s, G, ldg, L, ldl, n, (int64_t)0, X, ldx, scal,
                           tolrel, dropped);
        SYN_LAUNCH_CHECK("chol_diag");
        return SYN_OK;
    }
    if (n <= 2 * kNb) {
        ProfScope ps(kTagCholNode, (double)n * n * n * (2.0 / 3.0), s);
        hipLaunchKernelGGL(chol_node_kernel, dim3(1), dim3(kCholThreads), 0, s, G, ldg, L, ldl, X, ldx, n, T, scal,
                           tolrel, dropped);
        SYN_LAUNCH_CHECK("chol_node");
        return SYN_OK;
    }
    const int64_t n1 = ((n / kNb + 1) / 2) * kNb, n2 = n - n1;
    int rc = chol_inv(G, ldg, L, ldl, X, ldx, n1, T, scal, dropped, tolrel, s, ws, ws_cap);
    if (rc) return rc;
    double *L21 = L + n1 * ldl;
    // L21 = G21 X11^T (op(B) = X11^T upper)
    rc = gemm_f64_split_launch(make_gemm(G + n1 * ldg, ldg, 0, X, ldx, 1, L21, ldl, n2, n1, n1, 1.0, 0.0, 8), s, ws,
                               ws_cap);
    if (rc) return rc;
    // G22 -= L21 L21^T (lower triangle only) and T = L21 X11 (op(B) = X11
    // lower), side by side; with a split-K workspace they go separately
    const syn_gemm_item upd = make_gemm(L21, ldl, 0, L21, ldl, 1, G + n1 * ldg + n1, ldg, n2, n2, n1, -1.0, 1.0, 1);
    const syn_gemm_item tl = make_gemm(L21, ldl, 0, X, ldx, 0, T, n1, n2, n1, n1, 1.0, 0.0, 16);
    if (ws) {
        if ((rc = gemm_f64_split_launch(upd, s, ws, ws_cap))) return rc;
        if ((rc = gemm_f64_split_launch(tl, s, ws, ws_cap))) return rc;
    } else if ((rc = gemm_f64_pair_launch(upd, tl, s))) {
        return rc;
    }
    rc = chol_inv(G + n1 * ldg + n1, ldg, L + n1 * ldl + n1, ldl, X + n1 * ldx + n1, ldx, n2, T + n2 * n1, scal,
                  dropped, tolrel, s, ws, ws_cap);
    if (rc) return rc;
    // X21 = -X22 T (op(A) = X22 lower)
    return gemm_f64_split_launch(
        make_gemm(X + n1 * ldx + n1, ldx, 0, T, n1, 0, X + n1 * ldx, ldx, n2, n1, n2, -1.0, 0.0, 2), s, ws, ws_cap);
}

// max over the lower triangle of |G_ij - delta_ij|, grid-wide: every block
// reduces its share and combines with an atomic max on the bit pattern (for
// non-negative doubles the IEEE order is the unsigned integer order).  *out
// must be zero on entry.
__global__ __launch_bounds__(256) void offdiag_identity_kernel(const double *g, int64_t n, int64_t ld,
                                                               unsigned long long *out) {
    __shared__ double red[4];
    double m = 0.0;
    const int64_t nel = n * n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nel; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / n, j = e - i * n;
        if (j <= i) m = fmax(m, fabs(g[i * ld + j] - (i == j ? 1.0 : 0.0)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

// First-order Cholesky of G = I + E: F = strict_lower(E) + diag(E)/2,
// L = I + F, X = I - F (lower; zeros above the diagonal).
__global__ __launch_bounds__(256) void near_identity_chol_kernel(const double *g, int64_t n, int64_t ldg, double *L,
                                                                 int64_t ldl, double *X, int64_t ldx) {
    const int64_t nel = n * n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nel; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / n, j = e - i * n;
        double f = 0.0, d = 0.0;
        if (j < i) f = g[i * ldg + j];
        if (j == i) {
            f = 0.5 * (g[i * ldg + i] - 1.0);
            d = 1.0;
        }
        L[i * ldl + j] = d + f;
        X[i * ldx + j] = d - f;
    }
}

}  // namespace syn

extern "C" int64_t syn_cholqr_workspace(int64_t n) {
    // G (n x n), X = inv(L) (n x n), the recursion scratch, scalars, and
    // three n x n partial products of the split-K Gram (up to four K slices)
    const int64_t nblk = (n + syn::kNb - 1) / syn::kNb;
    return 5 * n * n + syn::chol_inv_scratch(n) + nblk * syn::kNb * syn::kNb + 16;
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
                              int64_t ldl, double *work, double tolrel, int *dropped, int mixed, void *stream) {
    using namespace syn;
    if (m <= 0 || n <= 0 || lda < n || ldq < n || ldl < n || !A || !Q || !L || !work || !dropped ||
        !(tolrel >= 0.0)) {
        set_error("syn_cholqr_f64: bad arguments (m=%lld n=%lld)", (long long)m, (long long)n);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblk = (n + kNb - 1) / kNb;
    double *G = work;
    double *X = G + n * n;
    double *T = X + n * n;
    double *scal = T + chol_inv_scratch(n) + nblk * kNb * kNb;
    // G = A^T A, lower triangle only (the Cholesky reads nothing else)
    double *splitws = scal + 16;
    int rc = gemm_f64_split_launch(make_gemm(A, lda, 1, A, lda, 0, G, n, n, n, m, 1.0, 0.0, 1), s, splitws,
                                   3 * n * n);
    if (rc) return rc;
    if ((rc = zero(L, n, n, ldl, s))) return rc;
    if ((rc = zero(X, n, n, n, s))) return rc;
    hipLaunchKernelGGL(maxdiag_kernel, dim3(1), dim3(256), 0, s, G, n, n, scal, dropped);
    SYN_LAUNCH_CHECK("maxdiag");
    if ((rc = chol_inv(G, n, L, ldl, X, n, n, T, scal, dropped, tolrel, s))) return rc;
    // Q = A X^T  (X = L^{-1} with dropped columns zeroed; op(B) = X^T upper)
    return gemm_f64_launch(make_gemm(A, lda, 0, X, n, 1, Q, ldq, m, n, n, 1.0, 0.0, 8), s, mixed != 0);
}

extern "C" int syn_chol_leaf_f64(const syn_chol_item *items, int n_items, void *stream) {
    using namespace syn;
    if (n_items < 0 || (n_items && !items)) {
        set_error("syn_chol_leaf_f64: bad arguments");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    hipStream_t s = (hipStream_t)stream;
    ProfScope ps(kTagCholDiag, 0.0, s);
    hipLaunchKernelGGL(chol_leaf_many_kernel, dim3(n_items), dim3(kCholThreads), 0, s, items);
    SYN_LAUNCH_CHECK("syn_chol_leaf_f64 launch");
    return SYN_OK;
}

extern "C" int syn_diagmax_f64(const syn_diagmax_item *items, int n_items, void *stream) {
    using namespace syn;
    if (n_items < 0 || (n_items && !items)) {
        set_error("syn_diagmax_f64: bad arguments");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    hipLaunchKernelGGL(diagmax_many_kernel, dim3(n_items), dim3(256), 0, (hipStream_t)stream, items);
    SYN_LAUNCH_CHECK("syn_diagmax_f64 launch");
    return SYN_OK;
}

extern "C" int syn_cholqr_refine_f64(const double *Q1, int64_t m, int64_t n, int64_t ldq1, double *Q, int64_t ldq,
                                     double *L, int64_t ldl, double *work, double thresh, double tolrel,
                                     int *dropped, int *full, int mixed, void *stream) {
    using namespace syn;
    if (m <= 0 || n <= 0 || ldq1 < n || ldq < n || ldl < n || !Q1 || !Q || !L || !work || !dropped ||
        !(thresh >= 0.0) || !(tolrel >= 0.0)) {
        set_error("syn_cholqr_refine_f64: bad arguments (m=%lld n=%lld)", (long long)m, (long long)n);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblk = (n + kNb - 1) / kNb;
    double *G = work;
    double *X = G + n * n;
    double *T = X + n * n;
    double *scal = T + chol_inv_scratch(n) + nblk * kNb * kNb;
    // G = Q1^T Q1 (lower triangle) and max |G - I| on it
    double *splitws = scal + 16;
    int rc = gemm_f64_split_launch(make_gemm(Q1, ldq1, 1, Q1, ldq1, 0, G, n, n, n, m, 1.0, 0.0, 1), s, splitws,
                                   3 * n * n);
    if (rc) return rc;
    if ((rc = hip_check(hipMemsetAsync(scal + 1, 0, sizeof(double), s), "refine emax reset"))) return rc;
    {
        int64_t blocks = (n * n + 4 * 256 - 1) / (4 * 256);
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(offdiag_identity_kernel, dim3((unsigned)blocks), dim3(256), 0, s, G, n, n,
                           (unsigned long long *)(scal + 1));
        SYN_LAUNCH_CHECK("cholqr refine check");
    }
    double emax = 0.0;
    if ((rc = hip_check(hipMemcpyAsync(&emax, scal + 1, sizeof(double), hipMemcpyDeviceToHost, s), "refine emax")))
        return rc;
    if ((rc = hip_check(hipStreamSynchronize(s), "refine sync"))) return rc;
    if (emax <= thresh) {
        // L = I + F, X = I - F: exact to O(n emax^2)
        int64_t blocks = (n * n + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(near_identity_chol_kernel, dim3((unsigned)blocks), dim3(256), 0, s, G, n, n, L, ldl, X,
                           n);
        SYN_LAUNCH_CHECK("cholqr refine factor");
        hipLaunchKernelGGL(maxdiag_kernel, dim3(1), dim3(256), 0, s, G, n, n, scal, dropped);   // zeroes *dropped
        SYN_LAUNCH_CHECK("cholqr refine dropped");
        if (full) *full = 0;
    } else {
        if ((rc = zero(L, n, n, ldl, s))) return rc;
        if ((rc = zero(X, n, n, n, s))) return rc;
        hipLaunchKernelGGL(maxdiag_kernel, dim3(1), dim3(256), 0, s, G, n, n, scal, dropped);
        SYN_LAUNCH_CHECK("maxdiag");
        if ((rc = chol_inv(G, n, L, ldl, X, n, n, T, scal, dropped, tolrel, s))) return rc;
        if (full) *full = 1;
    }
    // Q = Q1 X^T  (op(B) = X^T upper)
    return gemm_f64_launch(make_gemm(Q1, ldq1, 0, X, n, 1, Q, ldq, m, n, n, 1.0, 0.0, 8), s, mixed != 0);
}
