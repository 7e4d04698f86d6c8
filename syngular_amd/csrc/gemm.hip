// Grouped FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// C = alpha * op(A) * op(B) + beta * C for a table of independent problems of
// any shape.  A work-group of 4 waves owns one BM x BN output tile; each wave
// owns a WM x WN sub-tile held as (WM/16) x (WN/16) MFMA accumulators.  Operand
// tiles are staged through LDS stored k-major ([k][i] and [k][j]) so that the
// 16 lanes that feed one MFMA row read 16 consecutive doubles; rows are padded
// by one double to spread the transposed stores over the banks.  The K loop is
// double-buffered: the next global tile is loaded into registers while the
// current LDS tile feeds the MFMAs.  Each operand orientation is a separate
// instance (no per-element branches), which keeps a 128 x 128 tile within the
// register budget of two work-groups per compute unit, so one work-group's
// barrier and load waits are covered by the other's MFMAs.
//
// MFMA f64 16x16x4 operand layout (lane l of 64):
//   A: A[i = l%16][k = l/16]     B: B[k = l/16][j = l%16]
//   D: reg r of lane l holds D[i = l/16 + 4*r][j = l%16]  (checked against
//      NumPy on the device: rows 4g+r and g+4r differ unless g == r)
#include "common.hpp"

#include <algorithm>
#include "prof.hpp"

namespace syn {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Compute types: double (v_mfma_f64_16x16x4f64) or float, the mixed form
// (operands rounded to FP32 when staged into LDS, v_mfma_f32_16x16x4f32 with
// FP32 accumulation; the result is stored as FP64).  The two MFMAs take the
// same operand layout; the output rows of lane l, register r are
// l/16 + 4r (f64) and 4(l/16) + r (f32).
template <typename CT>
struct Mfma;
template <>
struct Mfma<double> {
    typedef v4d acc_t;
    static __device__ __forceinline__ acc_t zero() { return v4d{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t zero() { return v4f{0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// BF16 compute type (the precision form's contractions): operands rounded
// to bfloat16 (round to nearest even) when staged into LDS, stored as their
// 16-bit patterns, v_mfma_f32_16x16x16_bf16 with FP32 accumulation.  One
// MFMA covers K = 16: lane l holds A[i = l%16][k = 4(l/16) + v] and
// B[k = 4(l/16) + v][j = l%16] for v = 0..3; the output rows are those of
// the f32 MFMA, 4(l/16) + r.
struct bf16s {
    unsigned short bits;
};
typedef short v4s __attribute__((ext_vector_type(4)));
template <>
struct Mfma<bf16s> {
    typedef v4f acc_t;
    static __device__ __forceinline__ acc_t zero() { return v4f{0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc_t mac4(v4s a, v4s b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// FP64 -> staged operand: FP64, FP32 (round to nearest) or BF16 (FP32 first,
// then round to nearest even on the upper 16 bits; the inputs are finite).
template <typename CT>
__device__ __forceinline__ CT stage(double x) {
    return (CT)x;
}
template <>
__device__ __forceinline__ bf16s stage<bf16s>(double x) {
    unsigned u = __builtin_bit_cast(unsigned, (float)x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16s{(unsigned short)(u >> 16)};
}

template <int BM, int BN, int BK, typename CT>
struct GemmSmem {
    CT As[2][BK][BM + 1];
    CT Bs[2][BK][BN + 1];
};

template <int BM, int BN, int BK, int WM, int WN, bool TA, bool TB, typename CT>
__device__ __forceinline__ void gemm_f64_tile_t(const syn_gemm_item &P, int64_t t, GemmSmem<BM, BN, BK, CT> &sm,
                                                int64_t klo, int64_t khi) {
    constexpr int NW = (BM / WM) * (BN / WN);
    constexpr int NT = NW * 64;
    constexpr int TI = WM / 16, TJ = WN / 16;
    constexpr int EA = BM * BK / NT, EB = BK * BN / NT;
    static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "tile/threads mismatch");
    auto &As = sm.As;
    auto &Bs = sm.Bs;

    const int64_t row0 = (t / P.tiles_n) * BM, col0 = (t % P.tiles_n) * BN;
    const int tri = P.tri;
    if ((tri & 1) && col0 >= row0 + BM) return;  // tile strictly above the diagonal of C
    const int64_t m = P.m, n = P.n, lda = P.lda, ldb = P.ldb;
    // K band of the structurally non-zero products for this tile
    int64_t kmin = 0, k = P.k;
    if (tri & 2) k = k < row0 + BM ? k : row0 + BM;
    if (tri & 8) k = k < col0 + BN ? k : col0 + BN;
    if (tri & 4) kmin = row0;
    if (tri & 16) kmin = kmin > col0 ? kmin : col0;
    if (kmin < klo) kmin = klo;   // this work-group's K slice (split-K); klo is a multiple of BK
    if (k > khi) k = khi;
    kmin = (kmin / BK) * BK;
    constexpr bool ta = TA, tb = TB;
    const double *__restrict__ A = (const double *)P.A;
    const double *__restrict__ B = (const double *)P.B;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm0 = (w / (BN / WN)) * WM, wn0 = (w % (BN / WN)) * WN;

    double ra[EA], rb[EB];
    // Interior tiles (all BM rows, BN columns and BK steps of the K tile in
    // range) load through buffer instructions: element s of a thread sits at
    // base(k0) + voff + s * stride, with voff one VGPR computed once, s *
    // stride a scalar offset and the K tile's base in the scalar descriptor,
    // so the steady state spends no vector instructions on addresses or
    // bounds (the guarded form below costs about two per MFMA).  Offsets are
    // 32-bit: the path is taken for leading dimensions below 2^20 doubles,
    // which keeps every offset of a tile under 2^31.
    static_assert(NT % BK == 0 && NT % BM == 0 && NT % BN == 0, "thread/tile shapes");
    constexpr int kStepA = ta ? NT / BM : NT / BK;   // rows of the stored A per element step
    constexpr int kStepB = tb ? NT / BK : NT / BN;   // rows of the stored B per element step
    const bool ld_small = lda < (1 << 20) && ldb < (1 << 20);
    const int voff_a = ta ? ((tid / BM) * (int)lda + tid % BM) * 8 : ((tid / BK) * (int)lda + tid % BK) * 8;
    const int voff_b = tb ? ((tid / BK) * (int)ldb + tid % BK) * 8 : ((tid / BN) * (int)ldb + tid % BN) * 8;
    const int sstr_a = kStepA * (int)lda * 8, sstr_b = kStepB * (int)ldb * 8;
    auto load_fast = [&](int64_t k0) {
        const double *pa = ta ? A + k0 * lda + row0 : A + row0 * lda + k0;
        const double *pb = tb ? B + col0 * ldb + k0 : B + k0 * ldb + col0;
        const __amdgpu_buffer_rsrc_t sa = __builtin_amdgcn_make_buffer_rsrc((void *)pa, (short)0, 0x7fffffff, 0x00020000);
        const __amdgpu_buffer_rsrc_t sb = __builtin_amdgcn_make_buffer_rsrc((void *)pb, (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int s = 0; s < EA; ++s)
            ra[s] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(sa, voff_a, s * sstr_a, 0));
#pragma unroll
        for (int s = 0; s < EB; ++s)
            rb[s] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(sb, voff_b, s * sstr_b, 0));
    };
    const bool tile_in = row0 + BM <= m && col0 + BN <= n && ld_small;
    auto load = [&](int64_t k0) {
        if (tile_in && k0 + BK <= k) {
            load_fast(k0);
            return;
        }
#pragma unroll
        for (int s = 0; s < EA; ++s) {
            int e = tid + NT * s;
            int r, c;
            if constexpr (!ta) { r = e / BK; c = e % BK; } else { c = e / BM; r = e % BM; }
            int64_t gi = row0 + r, gk = k0 + c;
            if constexpr (ta)
                ra[s] = (gi < m && gk < k) ? A[gk * lda + gi] : 0.0;
            else
                ra[s] = (gi < m && gk < k) ? A[gi * lda + gk] : 0.0;
        }
#pragma unroll
        for (int s = 0; s < EB; ++s) {
            int e = tid + NT * s;
            int j, c;
            if constexpr (!tb) { c = e / BN; j = e % BN; } else { j = e / BK; c = e % BK; }
            int64_t gj = col0 + j, gk = k0 + c;
            if constexpr (tb)
                rb[s] = (gj < n && gk < k) ? B[gj * ldb + gk] : 0.0;
            else
                rb[s] = (gj < n && gk < k) ? B[gk * ldb + gj] : 0.0;
        }
    };
    // Element s of a thread lands at a fixed LDS base plus a compile-time
    // multiple of s (the same (row, k) map as the loads: e = tid + NT s), so
    // the stores take immediate offsets instead of per-element arithmetic.
    constexpr int kLdsStepA = ta ? (NT / BM) * (BM + 1) : NT / BK;
    constexpr int kLdsStepB = tb ? NT / BK : (NT / BN) * (BN + 1);
    const int lds_a = ta ? (tid / BM) * (BM + 1) + tid % BM : (tid % BK) * (BM + 1) + tid / BK;
    const int lds_b = tb ? (tid % BK) * (BN + 1) + tid / BK : (tid / BN) * (BN + 1) + tid % BN;
    auto store = [&](int buf) {
        CT *sa = &As[buf][0][0] + lds_a;
        CT *sb = &Bs[buf][0][0] + lds_b;
#pragma unroll
        for (int s = 0; s < EA; ++s) sa[s * kLdsStepA] = stage<CT>(ra[s]);
#pragma unroll
        for (int s = 0; s < EB; ++s) sb[s * kLdsStepB] = stage<CT>(rb[s]);
    };

    typedef Mfma<CT> M;
    typename M::acc_t acc[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = M::zero();

    const int64_t nk = k > kmin ? (k - kmin + BK - 1) / BK : 0;
    load(kmin);
    store(0);
    __syncthreads();
    const int li = lane & 15, lk = lane >> 4;
    for (int64_t kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < nk) load(kmin + (kt + 1) * BK);
        if constexpr (__is_same(CT, bf16s)) {
            static_assert(BK % 16 == 0, "the bf16 MFMA takes K = 16");
#pragma unroll
            for (int kk = 0; kk < BK / 16; ++kk) {
                v4s a[TI], b[TJ];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
#pragma unroll
                    for (int i = 0; i < TI; ++i) a[i][v] = (short)As[cur][kk * 16 + 4 * lk + v][wm0 + i * 16 + li].bits;
#pragma unroll
                    for (int j = 0; j < TJ; ++j) b[j][v] = (short)Bs[cur][kk * 16 + 4 * lk + v][wn0 + j * 16 + li].bits;
                }
#pragma unroll
                for (int i = 0; i < TI; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j) acc[i][j] = M::mac4(a[i], b[j], acc[i][j]);
            }
            if (kt + 1 < nk) store(cur ^ 1);
            __syncthreads();
            continue;
        } else {
        // the operands of step kk + 1 are read from LDS while the MFMAs of
        // step kk run
        CT a[2][TI], b[2][TJ];
        auto read_ops = [&](int kk, int h) {
#pragma unroll
            for (int i = 0; i < TI; ++i) a[h][i] = As[cur][kk * 4 + lk][wm0 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < TJ; ++j) b[h][j] = Bs[cur][kk * 4 + lk][wn0 + j * 16 + li];
        };
        read_ops(0, 0);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            if (kk + 1 < BK / 4) read_ops(kk + 1, (kk + 1) & 1);
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) acc[i][j] = M::mac(a[kk & 1][i], b[kk & 1][j], acc[i][j]);
        }
        if (kt + 1 < nk) store(cur ^ 1);
        __syncthreads();
        }
    }

    double *__restrict__ C = (double *)P.C;
    const int64_t ldc = P.ldc;
    const double alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int64_t gi = row0 + wm0 + i * 16 + M::row(lk, r);
                int64_t gj = col0 + wn0 + j * 16 + li;
                if (gi < m && gj < n) {
                    double v = alpha * (double)acc[i][j][r];
                    if (beta != 0.0) v += beta * C[gi * ldc + gj];
                    C[gi * ldc + gj] = v;
                }
            }
}

// The operand orientation is fixed at compile time in each instance, so the
// loads carry no per-element branches and the register budget fits two
// work-groups per compute unit.
template <int BM, int BN, int BK, int WM, int WN, typename CT>
__device__ __forceinline__ void gemm_f64_tile(const syn_gemm_item &P, int64_t t, int64_t klo = 0,
                                              int64_t khi = INT64_MAX) {
    __shared__ GemmSmem<BM, BN, BK, CT> sm;
    if (!P.transA) {
        if (!P.transB)
            gemm_f64_tile_t<BM, BN, BK, WM, WN, false, false, CT>(P, t, sm, klo, khi);
        else
            gemm_f64_tile_t<BM, BN, BK, WM, WN, false, true, CT>(P, t, sm, klo, khi);
    } else {
        if (!P.transB)
            gemm_f64_tile_t<BM, BN, BK, WM, WN, true, false, CT>(P, t, sm, klo, khi);
        else
            gemm_f64_tile_t<BM, BN, BK, WM, WN, true, true, CT>(P, t, sm, klo, khi);
    }
}

// Tile order of a single problem.  Work-groups are dealt to the XCDs round
// robin (id % 8), and each XCD has its own L2.  XCD x gets whole tile rows
// (or columns), zig-zag paired within every 16: rows x and 15 - x, then 16 + x
// and 31 - x, ...  A whole row reuses its A panel in one L2, and when the
// work of a tile grows or shrinks along the rows (a triangular hint) each XCD
// gets the same sum of long and short rows.  Columns are dealt instead when
// the hint makes the work depend on the column (op(B) triangular).  Shapes
// that do not split into such groups keep contiguous ranges per XCD.
__device__ __forceinline__ int64_t tile_of_block(int64_t bid, int64_t total, int64_t tiles_n, int tri) {
    const int64_t tiles_m = total / tiles_n;
    const bool by_col = (tri & (8 | 16)) && !(tri & (1 | 2 | 4));
    const int64_t lines = by_col ? tiles_n : tiles_m, len = by_col ? tiles_m : tiles_n;
    if (lines % (2 * kXcds) != 0 || total < 2 * kXcds) return xcd_remap(bid, total);
    const int64_t x = bid % kXcds, j = bid / kXcds, u = j / len, g = u / 2;
    const int64_t line = 2 * kXcds * g + ((u & 1) ? 2 * kXcds - 1 - x : x), pos = j % len;
    return by_col ? pos * tiles_n + line : line * tiles_n + pos;
}

template <int BM, int BN, int BK, int WM, int WN, typename CT = double>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void gemm_f64_grouped(
    const syn_gemm_item *__restrict__ items, int n_items, int64_t total) {
    if (n_items == 1) {
        gemm_f64_tile<BM, BN, BK, WM, WN, CT>(items[0],
                                              tile_of_block(blockIdx.x, total, items[0].tiles_n, items[0].tri));
        return;
    }
    const int64_t bid = xcd_remap(blockIdx.x, total);
    const syn_gemm_item &P = items[find_item(items, n_items, bid)];
    gemm_f64_tile<BM, BN, BK, WM, WN, CT>(P, bid - P.block_start);
}

// Persistent form of the grouped launch: a fixed grid of work-groups (a
// multiple of the 8 XCDs) walks the logical tiles b = blockIdx.x + i * grid,
// so every tile of a work-group stays on its XCD and the XCD keeps the same
// contiguous tile range as the one-tile-per-work-group launch.  For launches
// of many small tiles (thousands of chains' products) it saves the per-tile
// work-group start and drain; a tile computes exactly what it computes there.
template <int BM, int BN, int BK, int WM, int WN, typename CT = double>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void gemm_f64_grouped_persistent(
    const syn_gemm_item *__restrict__ items, int n_items, int64_t total) {
    for (int64_t b = blockIdx.x; b < total; b += gridDim.x) {
        if (b != (int64_t)blockIdx.x) __syncthreads();   // LDS of the previous tile is free
        const int64_t bid = xcd_remap(b, total);
        const syn_gemm_item &P = items[find_item(items, n_items, bid)];
        gemm_f64_tile<BM, BN, BK, WM, WN, CT>(P, bid - P.block_start);
    }
}

template <int BM, int BN, int BK, int WM, int WN, typename CT = double>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void gemm_f64_single(syn_gemm_item P) {
    const int64_t total = (int64_t)gridDim.x;
    gemm_f64_tile<BM, BN, BK, WM, WN, CT>(P, tile_of_block(blockIdx.x, total, P.tiles_n, P.tri));
}

// Two independent problems in one launch: work-groups [0, tiles0) take the
// first, the rest the second (each in its own tile order).  For short
// dependent chains of small products (the recursive Cholesky-inverse) this
// saves a launch and lets the two run side by side.
template <int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void gemm_f64_pair(syn_gemm_item P0, int64_t tiles0,
                                                                             syn_gemm_item P1, int64_t tiles1) {
    const int64_t bid = blockIdx.x;
    if (bid < tiles0)
        gemm_f64_tile<BM, BN, BK, WM, WN, double>(P0, tile_of_block(bid, tiles0, P0.tiles_n, P0.tri));
    else
        gemm_f64_tile<BM, BN, BK, WM, WN, double>(P1, tile_of_block(bid - tiles0, tiles1, P1.tiles_n, P1.tri));
}

// Split-K form of a single problem: work-group b computes K slice
// b / tiles of tile b % tiles; slice 0 writes C (with beta), slice s > 0
// writes the partial product into ws + (s - 1) m ws_ld, and
// gemm_splitk_reduce adds the partials into C.  Used where the tiles alone
// would leave most of the chip idle for a last partial round (a Gram
// matrix's lower triangle: 528 tiles of 128 on 512 work-group slots).
template <int BM, int BN, int BK, int WM, int WN>
__global__ __launch_bounds__((BM / WM) * (BN / WN) * 64, 2) void gemm_f64_splitk(syn_gemm_item P, int64_t tiles,
                                                                               int64_t kc, double *ws,
                                                                               int64_t ws_ld) {
    const int64_t bid = blockIdx.x;
    const int64_t slice = bid / tiles, tloc = bid - slice * tiles;
    const int64_t t = tile_of_block(tloc, tiles, P.tiles_n, P.tri);
    const int64_t klo = slice * kc;
    const int64_t khi = klo + kc < P.k ? klo + kc : P.k;
    if (slice > 0) {
        P.C = ws + (slice - 1) * P.m * ws_ld;
        P.ldc = ws_ld;
        P.beta = 0.0;
    }
    gemm_f64_tile<BM, BN, BK, WM, WN, double>(P, t, klo, khi);
}

__global__ __launch_bounds__(256) void gemm_splitk_reduce(double *C, int64_t ldc, const double *ws, int64_t ws_ld,
                                                          int64_t m, int64_t n, int extra, int lower) {
    const int64_t total = m * n;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
        const int64_t i = e / n, j = e - i * n;
        if (lower && j > i) continue;
        double v = C[i * ldc + j];
        for (int q = 0; q < extra; ++q) v += ws[q * m * ws_ld + i * ws_ld + j];
        C[i * ldc + j] = v;
    }
}

// One GEMM launched with its descriptor passed by value (no device table):
// used by the library's own multi-step routines (Cholesky, CholeskyQR, ...).
int gemm_f64_pair_launch(syn_gemm_item P0, syn_gemm_item P1, hipStream_t s) {
    const bool live0 = P0.m > 0 && P0.n > 0 && !(P0.k <= 0 && P0.beta == 1.0);
    const bool live1 = P1.m > 0 && P1.n > 0 && !(P1.k <= 0 && P1.beta == 1.0);
    if (!live0 || !live1) {
        if (live0) return gemm_f64_launch(P0, s);
        if (live1) return gemm_f64_launch(P1, s);
        return SYN_OK;
    }
    // one tile size for both, chosen by the larger output
    const int64_t area = std::max(P0.m * P0.n, P1.m * P1.n);
    const int tile = (area >= (int64_t)256 * 128 * 128) ? 1 : 0;
    const int64_t bm = tile ? 128 : 64;
    P0.tiles_n = (P0.n + bm - 1) / bm;
    P1.tiles_n = (P1.n + bm - 1) / bm;
    P0.block_start = P1.block_start = 0;
    const int64_t t0 = ((P0.m + bm - 1) / bm) * P0.tiles_n, t1 = ((P1.m + bm - 1) / bm) * P1.tiles_n;
    double fl = 0.0;
    for (const syn_gemm_item *P : {&P0, &P1}) {
        double f = 2.0 * (double)P->m * (double)P->n * (double)P->k;
        const int nbits = __builtin_popcount((unsigned)P->tri);
        if (nbits == 1) f *= 0.5;
        if (nbits >= 2) f /= 3.0;
        fl += f;
    }
    ProfScope ps(tile ? kTagGemmPair128 : kTagGemmPair64, fl, s);
    if (tile)
        hipLaunchKernelGGL((gemm_f64_pair<128, 128, 16, 64, 64>), dim3((unsigned)(t0 + t1)), dim3(256), 0, s, P0, t0,
                           P1, t1);
    else
        hipLaunchKernelGGL((gemm_f64_pair<64, 64, 16, 32, 32>), dim3((unsigned)(t0 + t1)), dim3(256), 0, s, P0, t0,
                           P1, t1);
    SYN_LAUNCH_CHECK("gemm_f64_pair_launch");
    return SYN_OK;
}

int gemm_f64_launch(syn_gemm_item P, hipStream_t s, bool mixed) {
    if (P.m <= 0 || P.n <= 0) return SYN_OK;
    if (P.k <= 0 && P.beta == 1.0) return SYN_OK;
    const int tile = (P.m * P.n >= (int64_t)256 * 128 * 128) ? 1 : 0;
    const int64_t bm = tile ? 128 : 64;
    P.tiles_n = (P.n + bm - 1) / bm;
    const int64_t tiles = ((P.m + bm - 1) / bm) * P.tiles_n;
    P.block_start = 0;
    // algorithmic flops: a structural-zero hint halves the product (a third
    // for triangle x triangle)
    double fl = 2.0 * (double)P.m * (double)P.n * (double)P.k;
    const int nbits = __builtin_popcount((unsigned)P.tri);
    if (nbits == 1) fl *= 0.5;
    if (nbits >= 2) fl /= 3.0;
    if (mixed) {
        ProfScope ps(tile ? kTagGemmMixed128 : kTagGemmMixed64, fl, s);
        if (tile)
            hipLaunchKernelGGL((gemm_f64_single<128, 128, 16, 64, 64, float>), dim3((unsigned)tiles), dim3(256), 0,
                               s, P);
        else
            hipLaunchKernelGGL((gemm_f64_single<64, 64, 16, 32, 32, float>), dim3((unsigned)tiles), dim3(256), 0, s,
                               P);
        SYN_LAUNCH_CHECK("gemm_f64_launch (mixed)");
        return SYN_OK;
    }
    ProfScope ps(tile ? kTagGemmSingle128 : kTagGemmSingle64, fl, s);
    if (tile)
        hipLaunchKernelGGL((gemm_f64_single<128, 128, 16, 64, 64>), dim3((unsigned)tiles), dim3(256), 0, s, P);
    else
        hipLaunchKernelGGL((gemm_f64_single<64, 64, 16, 32, 32>), dim3((unsigned)tiles), dim3(256), 0, s, P);
    SYN_LAUNCH_CHECK("gemm_f64_launch");
    return SYN_OK;
}

// One problem split along K when its tiles alone would leave work-group
// slots idle in the last round (a Gram matrix's lower triangle: 528 tiles of
// 128 on 512 slots).  ws holds ws_cap doubles for the partial products (the K slices
// after the first).  With 2 work-groups per CU the chip holds 2 * CUs tiles
// at a time; for `live` computed tiles (the lower triangle for tri bit 1),
// s slices take ceil(live s / slots) rounds of 1/s of the work; s <= 4 and
// every slice keeps at least 256 of K.
int gemm_f64_split_launch(syn_gemm_item P, hipStream_t s, double *ws, int64_t ws_cap) {
    if (P.m <= 0 || P.n <= 0) return SYN_OK;
    if (P.k <= 0 && P.beta == 1.0) return SYN_OK;
    const int tile = (P.m * P.n >= (int64_t)256 * 128 * 128) ? 1 : 0;
    if (ws == nullptr || P.k < 512) return gemm_f64_launch(P, s);
    static int slots = 0;
    if (slots == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        slots = 2 * cus;
    }
    const int64_t bm = tile ? 128 : 64;
    const int64_t tm = (P.m + bm - 1) / bm, tn = (P.n + bm - 1) / bm;
    int64_t live = 0;
    if (P.tri & 1)
        for (int64_t r = 0; r < tm; ++r) live += (r + 1 < tn ? r + 1 : tn);
    else
        live = tm * tn;
    int best = 1;
    double best_cost = (double)((live + slots - 1) / slots);
    for (int sl = 2; sl <= 4; ++sl) {
        const int64_t kc = ((P.k + sl - 1) / sl + 15) / 16 * 16;
        if (kc < 256 || (int64_t)(sl - 1) * P.m * P.n > ws_cap) break;
        const double cost = (double)((live * sl + slots - 1) / slots) / sl;
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = sl;
        }
    }
    if (best == 1) return gemm_f64_launch(P, s);
    const int64_t kc = ((P.k + best - 1) / best + 15) / 16 * 16;
    P.tiles_n = tn;
    P.block_start = 0;
    const int64_t tiles = tm * tn;
    double fl = 2.0 * (double)P.m * (double)P.n * (double)P.k;
    const int nbits = __builtin_popcount((unsigned)P.tri);
    if (nbits == 1) fl *= 0.5;
    if (nbits >= 2) fl /= 3.0;
    ProfScope ps(tile ? kTagGemmSplitK128 : kTagGemmSplitK64, fl, s);
    if (tile)
        hipLaunchKernelGGL((gemm_f64_splitk<128, 128, 16, 64, 64>), dim3((unsigned)(tiles * best)), dim3(256), 0, s,
                           P, tiles, kc, ws, P.n);
    else
        hipLaunchKernelGGL((gemm_f64_splitk<64, 64, 16, 32, 32>), dim3((unsigned)(tiles * best)), dim3(256), 0, s,
                           P, tiles, kc, ws, P.n);
    SYN_LAUNCH_CHECK("gemm_f64_split_launch");
    int64_t blocks = (P.m * P.n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(gemm_splitk_reduce, dim3((unsigned)blocks), dim3(256), 0, s, (double *)P.C, P.ldc,
                       (const double *)ws, P.n, P.m, P.n, best - 1, (P.tri & 1) ? 1 : 0);
    SYN_LAUNCH_CHECK("gemm_f64_split_launch reduce");
    return SYN_OK;
}

syn_gemm_item make_gemm(const double *A, int64_t lda, int ta, const double *B, int64_t ldb, int tb, double *C,
                        int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta, int tri) {
    syn_gemm_item P;
    memset(&P, 0, sizeof(P));
    P.A = A;
    P.B = B;
    P.C = C;
    P.m = m;
    P.n = n;
    P.k = k;
    P.lda = lda;
    P.ldb = ldb;
    P.ldc = ldc;
    P.alpha = alpha;
    P.beta = beta;
    P.transA = ta;
    P.transB = tb;
    P.tri = tri;
    return P;
}

}  // namespace syn

namespace syn {

// Grouped launches of more tiles than this many work-group slots run
// persistent (kPersistentGrid work-groups: 512 CUs' worth of two each would
// be 512; four per CU for the 64-tile kernels).
constexpr int64_t kPersistentMin = 4096;
constexpr int64_t kPersistentGrid = 1024;

template <typename CT>
static void launch_grouped(const syn_gemm_item *items, int n_items, int64_t total, int tile, hipStream_t s) {
    const bool pers = total >= kPersistentMin && n_items > 1;
    const unsigned grid = (unsigned)(pers ? kPersistentGrid : total);
    if (tile == 0) {
        if (pers)
            hipLaunchKernelGGL((gemm_f64_grouped_persistent<64, 64, 16, 32, 32, CT>), dim3(grid), dim3(256), 0, s,
                               items, n_items, total);
        else
            hipLaunchKernelGGL((gemm_f64_grouped<64, 64, 16, 32, 32, CT>), dim3(grid), dim3(256), 0, s, items, n_items,
                               total);
    } else {
        if (pers)
            hipLaunchKernelGGL((gemm_f64_grouped_persistent<128, 128, 16, 64, 64, CT>), dim3(grid), dim3(256), 0, s,
                               items, n_items, total);
        else
            hipLaunchKernelGGL((gemm_f64_grouped<128, 128, 16, 64, 64, CT>), dim3(grid), dim3(256), 0, s, items,
                               n_items, total);
    }
}

}  // namespace syn

extern "C" int syn_gemm_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile,
                            void *stream) {
    using namespace syn;
    if (n_items < 0 || total_blocks < 0) {
        set_error("syn_gemm_f64: negative item or block count");
        return SYN_EINVAL;
    }
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    if (total_blocks > 0x7fffffffLL) {
        set_error("syn_gemm_f64: %lld tiles exceed the grid limit", (long long)total_blocks);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (tile != 0 && tile != 1) {
        set_error("syn_gemm_f64: unknown tile configuration %d", tile);
        return SYN_EINVAL;
    }
    ProfScope ps(tile == 1 ? kTagGemmGrouped128 : kTagGemmGrouped64, -1.0, s);
    launch_grouped<double>(items, n_items, total_blocks, tile, s);
    SYN_LAUNCH_CHECK("syn_gemm_f64 launch");
    return SYN_OK;
}


extern "C" int syn_gemm_bf16_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile,
                                 void *stream) {
    using namespace syn;
    if (n_items < 0 || total_blocks < 0) {
        set_error("syn_gemm_bf16_f64: negative item or block count");
        return SYN_EINVAL;
    }
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    if (total_blocks > 0x7fffffffLL) {
        set_error("syn_gemm_bf16_f64: %lld tiles exceed the grid limit", (long long)total_blocks);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (tile != 0 && tile != 1) {
        set_error("syn_gemm_bf16_f64: unknown tile configuration %d", tile);
        return SYN_EINVAL;
    }
    ProfScope ps(tile == 1 ? kTagGemmBf16128 : kTagGemmBf1664, -1.0, s);
    launch_grouped<bf16s>(items, n_items, total_blocks, tile, s);
    SYN_LAUNCH_CHECK("syn_gemm_bf16_f64 launch");
    return SYN_OK;
}

extern "C" int syn_gemm_mixed_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile,
                                  void *stream) {
    using namespace syn;
    if (n_items < 0 || total_blocks < 0) {
        set_error("syn_gemm_mixed_f64: negative item or block count");
        return SYN_EINVAL;
    }
    if (n_items == 0 || total_blocks == 0) return SYN_OK;
    if (total_blocks > 0x7fffffffLL) {
        set_error("syn_gemm_mixed_f64: %lld tiles exceed the grid limit", (long long)total_blocks);
        return SYN_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    if (tile != 0 && tile != 1) {
        set_error("syn_gemm_mixed_f64: unknown tile configuration %d", tile);
        return SYN_EINVAL;
    }
    ProfScope ps(tile == 1 ? kTagGemmMixed128 : kTagGemmMixed64, -1.0, s);
    launch_grouped<float>(items, n_items, total_blocks, tile, s);
    SYN_LAUNCH_CHECK("syn_gemm_mixed_f64 launch");
    return SYN_OK;
}
