// Device-side truncation decision for the SVD rounding sweep.
//
// For sorted singular values s_0 >= s_1 >= ... the kept count is the
// smallest c whose discarded tail sum_{i>=c} s_i^2 is <= cutoff * sum_i s_i^2
// (when cutoff > 0), capped by chi_max (when > 0) and at least 1.  All sums
// run in FP64 from the smallest value upward, one wave per item.
#include "common.hpp"

namespace syn {

__device__ __forceinline__ double wave_suffix_incl(double v, int lane) {
    // inclusive suffix sum across the 64 lanes: out[l] = sum_{j>=l} v[j]
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_down(v, off, 64);
        if (lane + off < 64) v += o;
    }
    return v;
}

__global__ __launch_bounds__(64) void trunc_kernel(const syn_trunc_item *__restrict__ items) {
    const syn_trunc_item &P = items[blockIdx.x];
    const int lane = threadIdx.x;
    const int64_t n = P.n;
    const double *s = P.s;
    double tot = 0.0;
    for (int64_t i = lane; i < n; i += 64) tot = fma(s[i], s[i], tot);
    tot = wave_sum(tot);

    int64_t c = n;
    if (P.cutoff > 0.0 && n > 0) {
        const double thr = P.cutoff * tot;
        double run = 0.0;  // sum of w over indices beyond the current chunk
        int64_t best = n;
        for (int64_t start = ((n - 1) / 64) * 64; start >= 0; start -= 64) {
            const int64_t i = start + lane;
            const double wv = (i < n) ? s[i] * s[i] : 0.0;
            const double tail = run + wave_suffix_incl(wv, lane);
            const bool ok = (i < n) && tail <= thr;
            const unsigned long long mask = __ballot(ok);
            if (mask) best = start + __ffsll((long long)mask) - 1;
            run += wave_sum(wv);
            if (!(mask & 1ull)) break;  // the chunk start fails -> no earlier index can pass
        }
        c = best;
    }
    if (P.chi_max > 0 && c > P.chi_max) c = P.chi_max;
    if (c < 1) c = 1;
    double disc = 0.0;
    for (int64_t i = c + lane; i < n; i += 64) disc = fma(s[i], s[i], disc);
    disc = wave_sum(disc);
    if (lane == 0) {
        *P.keep = (int32_t)c;
        *P.discarded = disc;
    }
}

}  // namespace syn

extern "C" int syn_truncate_f64(const syn_trunc_item *items, int n_items, void *stream) {
    using namespace syn;
    if (n_items < 0) {
        set_error("syn_truncate_f64: negative item count");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    hipLaunchKernelGGL(trunc_kernel, dim3(n_items), dim3(64), 0, (hipStream_t)stream, items);
    SYN_LAUNCH_CHECK("syn_truncate_f64 launch");
    return SYN_OK;
}

// ---------------------------------------------------------------------------
// Euclidean norms of the columns (axis 0) or rows (axis 1) of row-major
// matrices.  In the rounding sweep the singular values are read off the
// products that carry them: the columns of C V = U S (tall C) or the rows of
// U^T C = S V^T (wide C).  A norm taken there has a relative error of about
// eps * s_max / s_i (the rounding of C itself), where the square root of an
// eigenvalue of the Gram matrix has eps * (s_max / s_i)^2.
// axis 0: one work-group per 64 columns, four row phases per column summed
// in a fixed order; axis 1: one wave per row.  Both orders are fixed, so the
// result does not depend on the launch.
namespace syn {

constexpr int kNormThreads = 256;

__global__ __launch_bounds__(kNormThreads) void vecnorm_kernel(const syn_vecnorm_item *__restrict__ items, int n,
                                                               int64_t total) {
    __shared__ double part[kNormThreads];
    const int64_t bid = blockIdx.x;
    const syn_vecnorm_item &P = items[find_item(items, n, bid)];
    const int64_t t = bid - P.block_start;
    const int tid = threadIdx.x, lane = tid & 63, ph = tid >> 6;
    const int64_t ld = P.ld;
    if (P.axis == 0) {
        const int64_t j = t * 64 + lane;
        double a0 = 0.0, a1 = 0.0;
        if (j < P.cols) {
            const double *col = P.x + j;
            int64_t i = ph;
            for (; i + 4 < P.rows; i += 8) {
                const double u = col[i * ld], v = col[(i + 4) * ld];
                a0 = fma(u, u, a0);
                a1 = fma(v, v, a1);
            }
            if (i < P.rows) {
                const double u = col[i * ld];
                a0 = fma(u, u, a0);
            }
        }
        part[tid] = a0 + a1;
        __syncthreads();
        if (ph == 0 && j < P.cols) P.out[j] = sqrt(((part[lane] + part[64 + lane]) + part[128 + lane]) + part[192 + lane]);
        return;
    }
    const int64_t i = t * 4 + ph;
    if (i >= P.rows) return;
    const double *row = P.x + i * ld;
    double a = 0.0;
    for (int64_t j = lane; j < P.cols; j += 64) a = fma(row[j], row[j], a);
    a = wave_sum(a);
    if (lane == 0) P.out[i] = sqrt(a);
}

}  // namespace syn

extern "C" int syn_vecnorms_f64(const syn_vecnorm_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (n_items < 0 || total_blocks < 0 || total_blocks > 0x7fffffffLL) {
        set_error("syn_vecnorms_f64: bad item count %d or block count %lld", n_items, (long long)total_blocks);
        return SYN_EINVAL;
    }
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(vecnorm_kernel, dim3((unsigned)total_blocks), dim3(kNormThreads), 0, (hipStream_t)stream,
                       items, n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_vecnorms_f64 launch");
    return SYN_OK;
}
