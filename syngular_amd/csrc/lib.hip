// Library-level entry points of the C ABI: error reporting and device query.
#include <stdarg.h>

#include "common.hpp"

namespace syn {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace syn

extern "C" const char *syn_last_error(void) { return syn::g_err; }

extern "C" int syn_version(void) { return 100; }

extern "C" int syn_device_count(int *count) {
    if (!count) {
        syn::set_error("syn_device_count: null pointer");
        return SYN_EINVAL;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return syn::hip_check(e, "hipGetDeviceCount");
    if (n <= 0) {
        syn::set_error("no HIP device visible");
        return SYN_EHIP;
    }
    *count = n;
    return SYN_OK;
}

extern "C" int syn_struct_sizes(int64_t *out, int n) {
    const int64_t sizes[] = {(int64_t)sizeof(syn_gemm_item),    (int64_t)sizeof(syn_apply_item),
                             (int64_t)sizeof(syn_mpo_mpo_item), (int64_t)sizeof(syn_dsum_item),
                             (int64_t)sizeof(syn_qr_item),      (int64_t)sizeof(syn_jacobi_item),
                             (int64_t)sizeof(syn_trunc_item),   (int64_t)sizeof(syn_rowscale_item),
                             (int64_t)sizeof(syn_bjac_item),    (int64_t)sizeof(syn_eye_item),
                             (int64_t)sizeof(syn_copy_item),    (int64_t)sizeof(syn_diagmax_item),
                             (int64_t)sizeof(syn_chol_item),    (int64_t)sizeof(syn_gbj_item),
                             (int64_t)sizeof(syn_lusign_item),  (int64_t)sizeof(syn_colscale_item),
                             (int64_t)sizeof(syn_vecnorm_item)};
    const int count = (int)(sizeof(sizes) / sizeof(sizes[0]));
    if (!out) return count;
    for (int i = 0; i < n && i < count; ++i) out[i] = sizes[i];
    return count;
}
