// Shared helpers for the gfx950 kernels of syngular_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/syngular_amd.h"

namespace syn {

// Thread-local message of the last failure (returned by syn_last_error).
void set_error(const char *fmt, ...);

inline int hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return SYN_EHIP;
    }
    return SYN_OK;
}

#define SYN_LAUNCH_CHECK(what)                                   \
    do {                                                         \
        hipError_t _e = hipGetLastError();                       \
        if (_e != hipSuccess) return ::syn::hip_check(_e, what); \
    } while (0)

// Number of XCDs on MI355X.  Consecutive work-group ids are dealt round-robin
// to the XCDs; remapping so that each XCD receives a contiguous range of
// logical blocks keeps neighbouring tiles (which share operand panels) in the
// same L2.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    if (nblocks < 2 * kXcds) return bid;
    int64_t q = nblocks / kXcds, r = nblocks % kXcds;
    int64_t x = bid % kXcds, j = bid / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// Items are sorted by block_start; find the item that owns logical block b.
template <typename Item>
__device__ __forceinline__ int find_item(const Item *items, int n, int64_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].block_start <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Host-side helpers shared by the multi-step routines (gemm.hip).
int gemm_f64_launch(syn_gemm_item P, hipStream_t s, bool mixed = false);
int gemm_f64_split_launch(syn_gemm_item P, hipStream_t s, double *ws, int64_t ws_cap);
// two independent FP64 problems in one launch (one tile size for both)
int gemm_f64_pair_launch(syn_gemm_item P0, syn_gemm_item P1, hipStream_t s);
syn_gemm_item make_gemm(const double *A, int64_t lda, int ta, const double *B, int64_t ldb, int tb, double *C,
                        int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta, int tri = 0);

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace syn
