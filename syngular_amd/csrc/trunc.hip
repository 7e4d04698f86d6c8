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
