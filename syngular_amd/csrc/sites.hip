// Site-level kernels: MPO x MPS and MPO x MPO core contractions, direct-sum
// packing.  All three are streaming kernels bounded by HBM bandwidth (one
// output element costs d multiply-adds against 8 bytes written), so each
// work-group covers SYN_APPLY_BLOCK consecutive output elements of one item,
// writes them coalesced, and reads the small operands through L2.
#include "common.hpp"
#include "prof.hpp"

namespace syn {

constexpr int kSiteThreads = 256;
constexpr int kPerThread = SYN_APPLY_BLOCK / kSiteThreads;

// P[(a*dl+c), o, (b*dr+e)] = sum_s X[a,s,b] * O[c,s,o,e]
template <typename Idx>
__device__ __forceinline__ void apply_block(const syn_apply_item &P, int64_t local_block) {
    const Idx dl = (Idx)P.dl, din = (Idx)P.din, dout = (Idx)P.dout, dr = (Idx)P.dr, chir = (Idx)P.chir;
    const Idx ncol = chir * dr;
    const Idx total = (Idx)P.chil * dl * dout * ncol;
    const double *__restrict__ x = P.x;
    const double *__restrict__ o = P.o;
    double *__restrict__ p = P.p;
    const Idx xs = chir;                 // stride of s in X
    const Idx os = dout * dr;            // stride of s in O
    const Idx base = (Idx)local_block * SYN_APPLY_BLOCK + threadIdx.x;
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
        const Idx idx = base + q * kSiteThreads;
        if (idx >= total) break;
        const Idx row = idx / ncol, col = idx - row * ncol;
        const Idx b = col / dr, e = col - b * dr;
        const Idx rc = row / dout, oo = row - rc * dout;
        const Idx a = rc / dl, c = rc - a * dl;
        const double *xp = x + (a * din) * xs + b;
        const double *op = o + (c * din) * os + oo * dr + e;
        double acc = 0.0;
        for (Idx s = 0; s < din; ++s) acc = fma(xp[s * xs], op[s * os], acc);
        p[idx] = acc;
    }
}

__global__ __launch_bounds__(kSiteThreads) void apply_kernel(const syn_apply_item *__restrict__ items, int n,
                                                             int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_apply_item &P = items[find_item(items, n, bid)];
    const int64_t nel = P.chil * P.dl * P.dout * P.chir * P.dr;
    if (nel < 0x7fffffffLL)
        apply_block<uint32_t>(P, bid - P.block_start);
    else
        apply_block<uint64_t>(P, bid - P.block_start);
}

// P[(a*bl+c), i, p, (b*br+e)] = sum_j A[a,i,j,b] * B[c,j,p,e]
__global__ __launch_bounds__(kSiteThreads) void mpo_mpo_kernel(const syn_mpo_mpo_item *__restrict__ items, int n,
                                                               int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_mpo_mpo_item &P = items[find_item(items, n, bid)];
    const int64_t al = P.al, din = P.din, dm = P.dmid, ar = P.ar, bl = P.bl, dout = P.dout, br = P.br;
    const int64_t ncol = ar * br;
    const int64_t nel = al * bl * din * dout * ncol;
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx >= nel) break;
        int64_t row = idx / ncol, col = idx - row * ncol;
        const int64_t b = col / br, e = col - b * br;
        const int64_t pp = row % dout;
        row /= dout;
        const int64_t i = row % din;
        row /= din;
        const int64_t c = row % bl, a = row / bl;
        const double *ap = P.a + ((a * din + i) * dm) * ar + b;       // A[a,i,j,b], stride of j = ar
        const double *bp = P.b + ((c * dm) * dout + pp) * br + e;     // B[c,j,p,e], stride of j = dout*br
        double acc = 0.0;
        for (int64_t j = 0; j < dm; ++j) acc = fma(ap[j * ar], bp[j * dout * br], acc);
        P.p[idx] = acc;
    }
}

// Kronecker product of the bond legs for every physical index pair:
// P[(a*bl+c), i, p, (b*br+e)] = A[a,i,p,b] * B[c,i,p,e]  (dmid unused).
__global__ __launch_bounds__(kSiteThreads) void kron_kernel(const syn_mpo_mpo_item *__restrict__ items, int n,
                                                            int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_mpo_mpo_item &P = items[find_item(items, n, bid)];
    const int64_t al = P.al, din = P.din, ar = P.ar, bl = P.bl, dout = P.dout, br = P.br;
    const int64_t ncol = ar * br;
    const int64_t nel = al * bl * din * dout * ncol;
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx >= nel) break;
        int64_t row = idx / ncol, col = idx - row * ncol;
        const int64_t b = col / br, e = col - b * br;
        const int64_t pp = row % dout;
        row /= dout;
        const int64_t i = row % din;
        row /= din;
        const int64_t c = row % bl, a = row / bl;
        P.p[idx] = P.a[((a * din + i) * dout + pp) * ar + b] * P.b[((c * din + i) * dout + pp) * br + e];
    }
}

// Direct sum of two cores along their bond legs.
__global__ __launch_bounds__(kSiteThreads) void dsum_kernel(const syn_dsum_item *__restrict__ items, int n,
                                                            int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_dsum_item &P = items[find_item(items, n, bid)];
    const int64_t lo = P.lo, ro = P.ro, mid = P.mid;
    const int64_t nel = lo * mid * ro;
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx >= nel) break;
        const int64_t r = idx % ro, t = idx / ro;
        const int64_t md = t % mid, l = t / mid;
        double v = 0.0;
        if (P.edge == 1) {  // [A | B]: rows shared
            if (r < P.ra)
                v = P.a[(l * mid + md) * P.ra + r];
            else
                v = P.b[(l * mid + md) * P.rb + (r - P.ra)];
        } else if (P.edge == 2) {  // [A ; B]: columns shared
            if (l < P.la)
                v = P.a[(l * mid + md) * P.ra + r];
            else
                v = P.b[((l - P.la) * mid + md) * P.rb + r];
        } else {
            if (l < P.la && r < P.ra)
                v = P.a[(l * mid + md) * P.ra + r];
            else if (l >= P.la && r >= P.ra)
                v = P.b[((l - P.la) * mid + md) * P.rb + (r - P.ra)];
        }
        P.out[idx] = v;
    }
}

__global__ __launch_bounds__(kSiteThreads) void rowscale_kernel(const syn_rowscale_item *__restrict__ items,
                                                                int n, int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_rowscale_item &P = items[find_item(items, n, bid)];
    const int64_t nel = P.rows * P.cols;
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx >= nel) break;
        const int64_t i = idx / P.cols, j = idx - i * P.cols;
        double &v = P.x[i * P.ld + j];
        const double s = P.s[i];
        if (P.mode == 0)
            v *= s;
        else
            v = (s != 0.0) ? v / s : 0.0;
    }
}

__global__ __launch_bounds__(kSiteThreads) void eye_kernel(const syn_eye_item *__restrict__ items, int n,
                                                           int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_eye_item &P = items[find_item(items, n, bid)];
    const int64_t nel = P.n * P.n;
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx >= nel) break;
        const int64_t i = idx / P.n, j = idx - i * P.n;
        P.x[i * P.ld + j] = (i == j) ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(kSiteThreads) void copy_kernel(const syn_copy_item *__restrict__ items, int n,
                                                            int64_t total) {
    const int64_t bid = blockIdx.x;
    const syn_copy_item &P = items[find_item(items, n, bid)];
    const int64_t base = (bid - P.block_start) * SYN_APPLY_BLOCK + threadIdx.x;
    for (int q = 0; q < kPerThread; ++q) {
        const int64_t idx = base + q * kSiteThreads;
        if (idx < P.n) P.dst[idx] = P.src[idx];
    }
}

static int grid_check(int n_items, int64_t total_blocks, const char *who) {
    if (n_items < 0 || total_blocks < 0 || total_blocks > 0x7fffffffLL) {
        set_error("%s: bad item count %d or block count %lld", who, n_items, (long long)total_blocks);
        return SYN_EINVAL;
    }
    return SYN_OK;
}

}  // namespace syn

extern "C" int syn_mpo_apply_f64(const syn_apply_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_mpo_apply_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    ProfScope ps(kTagApply, -1.0, (hipStream_t)stream);
    hipLaunchKernelGGL(apply_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream,
                       items, n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_mpo_apply_f64 launch");
    return SYN_OK;
}

extern "C" int syn_mpo_mpo_f64(const syn_mpo_mpo_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_mpo_mpo_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(mpo_mpo_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream,
                       items, n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_mpo_mpo_f64 launch");
    return SYN_OK;
}

extern "C" int syn_direct_sum_f64(const syn_dsum_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_direct_sum_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(dsum_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream, items,
                       n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_direct_sum_f64 launch");
    return SYN_OK;
}

extern "C" int syn_rowscale_f64(const syn_rowscale_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_rowscale_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(rowscale_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream,
                       items, n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_rowscale_f64 launch");
    return SYN_OK;
}

extern "C" int syn_kron_bonds_f64(const syn_mpo_mpo_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_kron_bonds_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(kron_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream, items,
                       n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_kron_bonds_f64 launch");
    return SYN_OK;
}

extern "C" int syn_eye_f64(const syn_eye_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_eye_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(eye_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream, items,
                       n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_eye_f64 launch");
    return SYN_OK;
}

extern "C" int syn_copy_f64(const syn_copy_item *items, int n_items, int64_t total_blocks, void *stream) {
    using namespace syn;
    if (int rc = grid_check(n_items, total_blocks, "syn_copy_f64")) return rc;
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)total_blocks), dim3(kSiteThreads), 0, (hipStream_t)stream, items,
                       n_items, total_blocks);
    SYN_LAUNCH_CHECK("syn_copy_f64 launch");
    return SYN_OK;
}
