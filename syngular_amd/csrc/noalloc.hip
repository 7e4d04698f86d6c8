// Allocation-free CholeskyQR (syn_cholqr_noalloc_f64): the pass of
// syn_cholqr_f64 with its workspace in this file's device array.
#include "common.hpp"

#include <mutex>

// ---------------------------------------------------------------------------
// Allocation-free CholeskyQR.  The same pass as syn_cholqr_f64, with the
// workspace taken from `scratch`, a device array of this file that every GPU
// the library runs on holds until the process ends (48 MiB: the workspace of
// up to 1078 columns; m is not limited, A, Q and L are the caller's).  Its
// device address is looked up once per device, with that device current, on
// the entry point's first call there; host code only ever passes that
// address.  All calls on one GPU share the array, so calls that overlap on
// one GPU (different streams, different host threads) race; everything the
// pass reads from it is written earlier in the same call (the Gram matrix
// and the partial products are stored, not accumulated; L^-1 is zero-filled
// by a kernel), so nothing carries over from an earlier call.
namespace syn {

constexpr int64_t kScratchDoubles = 6LL << 20;   // 48 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kScratchDevices = 64;
static double *g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

static int scratch_address(double **out) {
    int dev = 0;
    if (int rc = hip_check(hipGetDevice(&dev), "syn_cholqr_noalloc_f64: hipGetDevice")) return rc;
    if (dev < 0 || dev >= kScratchDevices) {
        set_error("syn_cholqr_noalloc_f64: device %d beyond the %d this library tracks", dev, kScratchDevices);
        return SYN_ENOTSUP;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                               "syn_cholqr_noalloc_f64: hipGetSymbolAddress(scratch)"))
            return rc;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return SYN_OK;
}

}  // namespace syn

extern "C" int64_t syn_cholqr_noalloc_max_n(void) {
    int64_t n = 1;
    while (syn_cholqr_workspace(n + 1) <= syn::kScratchDoubles) ++n;
    return n;
}

extern "C" int syn_cholqr_noalloc_f64(double *A, int64_t m, int64_t n, int64_t lda, double *Q, int64_t ldq,
                                      double *L, int64_t ldl, double tolrel, int *dropped, int mixed, void *stream) {
    using namespace syn;
    if (n <= 0 || syn_cholqr_workspace(n) > kScratchDoubles) {
        set_error("syn_cholqr_noalloc_f64: n=%lld columns; this entry point takes 1 <= n <= %lld (its "
                  "%lld-double device scratch); use syn_cholqr_f64 with a workspace",
                  (long long)n, (long long)syn_cholqr_noalloc_max_n(), (long long)kScratchDoubles);
        return SYN_EINVAL;
    }
    double *work = nullptr;
    if (int rc = scratch_address(&work)) return rc;
    return syn_cholqr_f64(A, m, n, lda, Q, ldq, L, ldl, work, tolrel, dropped, mixed, stream);
}
