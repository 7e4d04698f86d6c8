// Householder signs of a CholeskyQR factor, and column scaling.
//
// The QR of a tall well-conditioned matrix is cheapest as CholeskyQR2 on the
// MFMA GEMM, which gives A = Q R with R's diagonal positive.  The reference
// takes numpy.linalg.qr (LAPACK Householder, beta = -sign(alpha) ||x||), whose
// factors are (Q D, D R) for a sign matrix D.  With the compact-WY form of the
// Householder factor, Q D - [I; 0] = -Y T Y1^T is an LU factorisation (Y unit
// lower trapezoidal, T Y1^T upper), so D falls out of the LU without pivoting
// of Q[:n, :n] - D where D_jj is chosen when column j is reached:
// D_jj = -sign(c_j), c_j the Schur-complement diagonal entry, which makes the
// pivot c_j - D_jj = -tau_j D_jj with |pivot| in [1, 2] (Ballard, Demmel,
// Grigori, Jacquelin, Knight, Nguyen, "Reconstructing Householder vectors from
// tall-skinny QR", JPDC 85, 2015).
//
// lu_sign_leaf_kernel factors one diagonal block of at most 64 columns in LDS
// (one work-group per item) and returns D and the inverses of its L and U;
// the host drives the blocked right-looking LU with the GEMM (L21 = M21 U^-1,
// U12 = L^-1 M12, M22 -= L21 U12).
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kLuLeaf = 64;
constexpr int kLuLd = kLuLeaf + 1;  // padded rows: column walks hit distinct banks
constexpr int kLuThreads = 256;
constexpr size_t kLuLdsBytes = 2 * kLuLeaf * kLuLd * sizeof(double);

__global__ __launch_bounds__(kLuThreads) void lu_sign_leaf_kernel(const syn_lusign_item *__restrict__ items,
                                                                  double amb) {
    extern __shared__ double lds_lu[];
    double *A = lds_lu;                    // the block, becomes L\U
    double *X = lds_lu + kLuLeaf * kLuLd;  // inverse columns
    __shared__ double piv_s;
    const syn_lusign_item &P = items[blockIdx.x];
    const int bs = (int)P.bs, tid = threadIdx.x;
    for (int e = tid; e < bs * bs; e += kLuThreads) {
        const int i = e / bs, j = e - i * bs;
        A[i * kLuLd + j] = P.a[i * P.lda + j];
    }
    __syncthreads();
    int ambiguous = 0;
    for (int j = 0; j < bs; ++j) {
        if (tid == 0) {
            const double c = A[j * kLuLd + j];
            double d;
            if (j == P.free_col) {
                d = (c >= 0.0) ? 1.0 : -1.0;  // no reflector: R_jj keeps the sign of alpha
            } else {
                d = (c >= 0.0) ? -1.0 : 1.0;
                if (fabs(c) > amb) ambiguous = 1;
            }
            P.d[j] = d;
            A[j * kLuLd + j] = c - d;
            piv_s = c - d;
        }
        __syncthreads();
        const double p = piv_s;
        const int r = bs - j - 1;
        for (int i = tid; i < r; i += kLuThreads) A[(j + 1 + i) * kLuLd + j] /= p;
        __syncthreads();
        for (int e = tid; e < r * r; e += kLuThreads) {
            const int i = j + 1 + e / r, k = j + 1 + e % r;
            A[i * kLuLd + k] = fma(-A[i * kLuLd + j], A[j * kLuLd + k], A[i * kLuLd + k]);
        }
        __syncthreads();
    }
    if (P.linv != nullptr) {
        // L^-1 (unit lower): thread c solves L x = e_c by forward substitution
        if (tid < bs) {
            const int c = tid;
            for (int i = 0; i < c; ++i) X[i * kLuLd + c] = 0.0;
            X[c * kLuLd + c] = 1.0;
            for (int i = c + 1; i < bs; ++i) {
                double s = 0.0;
                for (int k = c; k < i; ++k) s = fma(A[i * kLuLd + k], X[k * kLuLd + c], s);
                X[i * kLuLd + c] = -s;
            }
        }
        __syncthreads();
        for (int e = tid; e < bs * bs; e += kLuThreads) {
            const int i = e / bs, j = e - i * bs;
            P.linv[i * P.ldl + j] = X[i * kLuLd + j];
        }
        __syncthreads();
        // U^-1 (upper): thread c solves U x = e_c by back substitution
        if (tid < bs) {
            const int c = tid;
            for (int i = c + 1; i < bs; ++i) X[i * kLuLd + c] = 0.0;
            X[c * kLuLd + c] = 1.0 / A[c * kLuLd + c];
            for (int i = c - 1; i >= 0; --i) {
                double s = 0.0;
                for (int k = i + 1; k <= c; ++k) s = fma(A[i * kLuLd + k], X[k * kLuLd + c], s);
                X[i * kLuLd + c] = -s / A[i * kLuLd + i];
            }
        }
        __syncthreads();
        for (int e = tid; e < bs * bs; e += kLuThreads) {
            const int i = e / bs, j = e - i * bs;
            P.uinv[i * P.ldu + j] = X[i * kLuLd + j];
        }
    }
    if (tid == 0 && ambiguous) *P.flag = 1;
}

constexpr int kCsTile = 64;
constexpr int kCsThreads = 256;

__global__ __launch_bounds__(kCsThreads) void colscale_kernel(const syn_colscale_item *__restrict__ items, int n,
                                                              int64_t total) {
    __shared__ double T[kCsTile][kCsTile + 1];
    const int64_t bid = blockIdx.x;
    const syn_colscale_item &P = items[find_item(items, n, bid)];
    const int64_t t = bid - P.block_start;
    const int64_t tc = (P.cols + kCsTile - 1) / kCsTile;
    const int64_t r0 = (t / tc) * kCsTile, c0 = (t % tc) * kCsTile;
    const int tx = threadIdx.x & (kCsTile - 1), ty = threadIdx.x / kCsTile;
    constexpr int kRowsPerPass = kCsThreads / kCsTile;
    const int64_t c = c0 + tx;
    const double sc = c < P.cols ? (P.s != nullptr ? P.s[c] : 1.0) : 0.0;
    if (!P.trans) {
        for (int i = ty; i < kCsTile; i += kRowsPerPass) {
            const int64_t r = r0 + i;
            if (r < P.rows && c < P.cols) P.dst[r * P.ldd + c] = P.src[r * P.lds + c] * sc;
        }
        return;
    }
    for (int i = ty; i < kCsTile; i += kRowsPerPass) {
        const int64_t r = r0 + i;
        if (r < P.rows && c < P.cols) T[i][tx] = P.src[r * P.lds + c] * sc;
    }
    __syncthreads();
    for (int i = ty; i < kCsTile; i += kRowsPerPass) {
        const int64_t cc = c0 + i, r = r0 + tx;
        if (cc < P.cols && r < P.rows) P.dst[cc * P.ldd + r] = T[tx][i];
    }
}

}  // namespace syn

extern "C" int syn_lu_sign_leaf_f64(const syn_lusign_item *items, int n_items, double amb, void *stream) {
    using namespace syn;
    if (n_items < 0) {
        set_error("syn_lu_sign_leaf_f64: negative item count");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    static bool attr_set = false;
    if (!attr_set) {
        // 66,560 B of dynamic LDS: above the 64 KiB a launch gets without the attribute
        if (int rc = hip_check(hipFuncSetAttribute((const void *)lu_sign_leaf_kernel,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLuLdsBytes),
                               "syn_lu_sign_leaf_f64: hipFuncSetAttribute(MaxDynamicSharedMemorySize)"))
            return rc;
        attr_set = true;
    }
    hipLaunchKernelGGL(lu_sign_leaf_kernel, dim3(n_items), dim3(kLuThreads), kLuLdsBytes, (hipStream_t)stream, items,
                       amb);
    SYN_LAUNCH_CHECK("syn_lu_sign_leaf_f64 launch");
    return SYN_OK;
}

extern "C" int syn_colscale_f64(const syn_colscale_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (n_items < 0 || total_blocks < 0 || total_blocks > 0x7fffffffLL) {
        set_error("syn_colscale_f64: bad item count %d or block count %lld", n_items, (long long)total_blocks);
        return SYN_EINVAL;
    }
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(colscale_kernel, dim3((unsigned)total_blocks), dim3(kCsThreads), 0, (hipStream_t)stream, items,
                       n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_colscale_f64 launch");
    return SYN_OK;
}
