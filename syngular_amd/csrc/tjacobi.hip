// Two-sided block Jacobi eigen-decomposition of symmetric Gram matrices.
//
// For the right sweep of the SVD rounding the matrix to decompose is a Gram
// matrix G = C^T C (or C C^T) of size g = 65 .. a few thousand.  The rows and
// columns of G are split into blocks of 16; a sweep pairs every two blocks
// once in the round-robin order, and in one step the pairs are disjoint:
//
//   solve   one work-group per pair: the 32 x 32 diagonal tile of the pair
//           (rows and columns of both blocks) is built in LDS and rotated by
//           two-sided Jacobi, accumulating the orthogonal W of the pair: a
//           full round-robin sweep of the tile at the first step of a pass
//           (which meets every block), only the planes between the pair's
//           two blocks at the others, so a pass treats every plane of G once;
//   update  G_ab <- W_a^T G_ab W_b for every pair of pairs a <= b (written
//           with its transpose) and rt_a <- W_a^T rt_a for every 64-column
//           slab of the accumulated eigenvector rows, on the f64 MFMA with
//           the operands in registers.
//
// A step is one launch, the solves of step s beside the updates of step
// s - 1, and several problems share every launch.  A pair whose tile has no
// |G_ik| > max(tol sqrt(|G_ii G_kk|), eps max_j G_jj) among the planes it
// treats is not rotated; a problem stops after a pass that rotated nothing
// or when no off-diagonal entry of G passes that test at the end of a pass,
// so on exit every off-diagonal entry of G meets it, G is diagonal to it and
// G_in = rt^T diag(G_out) rt.  This is the two-sided block Jacobi method
// (Hansen; Van Loan; Becka, Oksa & Vajtersic) in its parallel round-robin
// form, with the planes in the cyclic-by-blocks order.
#include "common.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace syn {

// Blocks of TB = 16 rows (the helpers keep TB as a parameter); a pair tile
// has TP = 2 TB rows.  The pair solve uses one thread per 2 x 2 block of a
// round, (TP/2)^2 threads; an update work-group four waves.
constexpr int kTUpdThreads = 256;    // four waves per update work-group
constexpr int kTCol = 64;            // columns of an rt slab

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Blocks (a, b) of pair w at a step of the round-robin order over np blocks.
// 32-bit arithmetic (a 64-bit remainder is a long instruction sequence):
// step < np - 1 and w < np / 2 at every call, and np is far below 2^31.
__device__ __forceinline__ void tj_pair(int64_t step, int64_t w, int64_t np, int64_t &a, int64_t &b) {
    const int s = (int)step, ww = (int)w, m = (int)np - 1;
    if (ww == 0) {
        a = s;
        b = m;
    } else {
        a = (s + ww) % m;
        b = (s - ww + m) % m;
    }
}

// Matrix row of entry i (0 <= i < 2 TB) of pair w at a step, -1 if padding.
template <int TB>
__device__ __forceinline__ int tj_row(int64_t step, int64_t w, int64_t np, int64_t nblk, int64_t nv, int i) {
    int64_t a, b;
    tj_pair(step, w, np, a, b);
    const int64_t blk = i < TB ? a : b;
    const int64_t r = blk * TB + (i % TB);
    return (blk < nblk && r < nv) ? (int)r : -1;
}

// Plane k of the round-robin order over the TP indices of a pair tile:
// in round r it joins a = r, b = m (k = 0) or a = (r + k) mod m,
// b = (r - k) mod m (k > 0), m = TP - 1.  From one round to the next both
// move up by one modulo m (b stays at m for k = 0), so a thread steps its
// planes along instead of dividing by m every round.
template <int TP>
struct RrPlane {
    int a, b;
    __device__ __forceinline__ void start(int k) {
        constexpr int m = TP - 1;
        a = k;
        b = k == 0 ? m : m - k;
    }
    __device__ __forceinline__ void next() {
        constexpr int m = TP - 1;
        a = a + 1 == m ? 0 : a + 1;
        if (b != m) b = b + 1 == m ? 0 : b + 1;
    }
    __device__ __forceinline__ int p() const { return a < b ? a : b; }
    __device__ __forceinline__ int q() const { return a < b ? b : a; }
    static constexpr int kRounds = TP - 1;
};

// Rotation test of plane (p, q): |G_pq| > max(tol sqrt(|G_pp G_qq|), floor),
// in squares (no square root on the critical path of a round).
__device__ __forceinline__ bool rotates(double apq, double app, double aqq, double tol, double floor_abs) {
    const double a2 = apq * apq;
    return a2 > tol * tol * fabs(app * aqq) && fabs(apq) > floor_abs;
}

// Jacobi rotation annihilating G_pq (Rutishauser's angle, Golub & Van Loan
// 8.5.2): with d = G_qq - G_pp and e = 2 G_pq, tau = d / e and
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)) = sign(d e) |e| / (|d| + sqrt(d^2 + e^2)),
// c = 1 / sqrt(1 + t^2), s = c t.  The scaled form needs one division and one
// square root for t; operands far from 1 (|d|, |e| outside 1e-150 .. 1e150)
// take the tau form, which cannot overflow.
// 1 / x and 1 / sqrt(x) for normal positive x from the hardware estimates
// (v_rcp_f64, v_rsq_f64) and two Newton-Raphson steps each: r <- r + r (1 - x r)
// and y <- y + y (1 - x y^2) / 2, each step squaring the relative error, to
// within a unit or two in the last place.  The rotation needs c^2 + s^2 = 1 to
// that level, not correctly rounded c and s; the steps are shorter than the
// IEEE division and square root sequences on a round's critical path.
__device__ __forceinline__ double recip_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

__device__ __forceinline__ void rotation(double apq, double app, double aqq, double &c, double &s) {
    const double d = aqq - app, e = 2.0 * apq;
    const double ad = fabs(d), ae = fabs(e), m = fmax(ad, ae);
    double t;
    if (m < 1e150 && m > 1e-150) {
        const double h2 = fma(d, d, e * e);
        t = copysign(ae, d * e) * recip_nr(ad + h2 * rsqrt_nr(h2));
    } else {
        const double tau = d / e, at = fabs(tau);
        t = copysign(1.0, tau) / (at < 1e150 ? at + sqrt(fma(at, at, 1.0)) : 2.0 * at);
    }
    c = rsqrt_nr(fma(t, t, 1.0));
    s = c * t;
}

struct TjDims {
    int64_t nblk, np, half, ncs;
};

// rt slab width: the fused 16-row steps give each wave two groups of 16
// columns (fewer, fuller work-groups, so a launch of a few large problems
// fits one round of resident work-groups); the 32-row updates one
template <int TB>
__host__ __device__ constexpr int tj_slab_cols() {
    return TB == 16 ? 2 * kTCol : kTCol;
}

template <int TB>
__host__ __device__ __forceinline__ TjDims tj_dims(int64_t nv) {
    TjDims d;
    d.nblk = (nv + TB - 1) / TB;
    d.np = d.nblk + (d.nblk & 1);
    d.half = d.np / 2;
    d.ncs = (nv + tj_slab_cols<TB>() - 1) / tj_slab_cols<TB>();
    return d;
}

// Update work of a problem at a step: the upper-triangle G tiles (a, b),
// a <= b, two per work-group (each by two waves, one per 16-column half),
// then one work-group per rt slab of tj_slab_cols columns.
__host__ __device__ __forceinline__ int64_t tj_gtiles(int64_t half) { return half * (half + 1) / 2; }
// update work-groups of a problem at a step: 16-row blocks put two 32 x 32
// G tiles in a work-group, 32-row blocks one 64 x 64 tile; one per rt slab
template <int TB>
__host__ __device__ __forceinline__ int64_t tj_ugroups(const TjDims &D) {
    return (TB == 16 ? (tj_gtiles(D.half) + 1) / 2 : tj_gtiles(D.half)) + D.half * D.ncs;
}

// Work-group owner table of a sweep's fused steps: entry b < ss names the
// problem of solve b, entry ss + t the problem of tile t (one work-group per
// problem fills its runs), so a work-group finds its problem with one load.
template <int TB>
__global__ __launch_bounds__(256) void tj_owner_kernel(const syn_gbj_item *__restrict__ items, int64_t ss,
                                                        int32_t *__restrict__ owner) {
    const int i = blockIdx.x;
    const syn_gbj_item &P = items[i];
    const TjDims D = tj_dims<TB>(P.nv);
    for (int64_t b = threadIdx.x; b < D.half; b += 256) owner[P.solve_start + b] = i;
    const int64_t nt = tj_ugroups<TB>(D);
    for (int64_t t = threadIdx.x; t < nt; t += 256) owner[ss + P.tile_start + t] = i;
}

// owner entries (int32) of a problem in the fused steps' table
template <int TB>
static int64_t tj_owner_count(int64_t nv) {
    const TjDims D = tj_dims<TB>(nv);
    return D.half + tj_ugroups<TB>(D);
}

// Per-problem scratch layout (doubles): two sets of W tiles (half * TP^2
// each) and of rotation flags (half ints each), selected by step parity so a
// step's pair solves can write theirs while the previous step's update reads
// its own; the rotation floor; the sort scratch (rt copy, s); a second
// nv x nv buffer for the Gram matrix (the fused steps read one buffer and
// write the other).
template <int TB>
__device__ __forceinline__ double *tj_wset(const syn_gbj_item &P, int64_t half, int par) {
    return P.w + par * half * (2 * TB) * (2 * TB);
}

template <int TB>
__device__ __forceinline__ int *tj_flags(const syn_gbj_item &P, int64_t half, int par = 0) {
    return (int *)(P.w + 2 * half * (2 * TB) * (2 * TB)) + par * half;
}

// Absolute rotation floor eps * max_i G_ii (set by the init kernel): Gram
// entries below it are at the rounding level of forming G, so rotating them
// changes nothing that the Gram route can resolve, and in a rank-deficient G
// (a null space at noise level) a purely relative test would never settle.
template <int TB>
__host__ __device__ __forceinline__ int64_t tj_floor_off(int64_t half) {
    return 2 * half * (2 * TB) * (2 * TB) + half;
}

// the second Gram buffer (ld nv), after the sort scratch
template <int TB>
__host__ __device__ __forceinline__ int64_t tj_g2_off(int64_t half, int64_t nv) {
    return tj_floor_off<TB>(half) + 1 + nv * nv + nv + 4;
}

// The pair solve on a tile S already in LDS: W <- I, the rotation test of
// the tile, then inner_sweeps round-robin sweeps of two-sided rotations;
// writes W to Wout and the pair's flag (1: rotated) and counts a rotating
// pair in P.counter.  Every thread of the work-group calls it.
template <int TB>
__device__ __forceinline__ void tj_pair_solve(const syn_gbj_item &P, const TjDims &D, int64_t pw,
                                              double (*S)[2 * TB + 1], double (*W)[2 * TB + 1], v2d *cs,
                                              int &do_rot, int &any_rot, double tol, int inner_sweeps, double *Wout,
                                              int *flags, bool cross = false) {
    constexpr int TP = 2 * TB, NT = TB * TB;
    const int tid = threadIdx.x;
    for (int e = tid; e < TP * TP; e += NT) W[e / TP][e % TP] = (e / TP == e % TP) ? 1.0 : 0.0;
    if (tid == 0) do_rot = 0;
    __syncthreads();
    const double floor_abs = P.w[tj_floor_off<TB>(D.half)];
    for (int e = tid; e < TP * TP; e += NT) {
        const int i = e / TP, k = e % TP;
        if (i < k && (!cross || (i < TB && k >= TB)) && rotates(S[i][k], S[i][i], S[k][k], tol, floor_abs))
            do_rot = 1;
    }
    __syncthreads();
    if (!do_rot) {
        if (tid == 0) flags[pw] = 0;
        return;
    }
    // thread (ka, kb) = (tid / (TP/2), tid % (TP/2)) owns the 2 x 2 block of
    // planes ka (rows) and kb (columns) of each round, and row wi of the
    // columns of planes wk and wk + TP/4 of W
    const int ka = tid / (TP / 2), kb = tid % (TP / 2);
    const int wk = tid / TP, wi = tid % TP;
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
        if (tid == 0) any_rot = 0;
        __syncthreads();
        // one pass of the round-robin plane order Pl over every plane of the tile
        auto rounds = [&](auto tag) __attribute__((always_inline)) {
            using Pl = decltype(tag);
            Pl pa, pb, pw2[2];
            pa.start(ka);
            pb.start(kb);
            pw2[0].start(wk);
            pw2[1].start(wk + TP / 4);
#pragma unroll 1
            for (int r = 0; r < Pl::kRounds; ++r) {
                if (r) {
                    pa.next();
                    pb.next();
                    pw2[0].next();
                    pw2[1].next();
                }
                const int p = pa.p(), q = pa.q(), u = pb.p(), v = pb.q();
                // the rotation's operands are read first (their wave waits for
                // them alone); this round's update operands were final at the last
                // barrier and phase 1 only reads S, so they are loaded next and
                // their latency overlaps the rotation
                const bool rot_thread = tid < TP / 2;   // here ka == 0 and kb == tid: plane tid is (u, v)
                double apq = 0.0, app = 0.0, aqq = 0.0;
                if (rot_thread) {
                    apq = S[u][v];
                    app = S[u][u];
                    aqq = S[v][v];
                }
                const double spu = S[p][u], spv = S[p][v], squ = S[q][u], sqv = S[q][v];
                int wp[2], wq[2];
                double wx[2], wy[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    wp[h] = pw2[h].p();
                    wq[h] = pw2[h].q();
                    wx[h] = W[wi][wp[h]];
                    wy[h] = W[wi][wq[h]];
                }
                if (rot_thread) {
                    double c = 1.0, sn = 0.0;
                    if (rotates(apq, app, aqq, tol, floor_abs)) {
                        rotation(apq, app, aqq, c, sn);
                        any_rot = 1;
                    }
                    cs[tid] = v2d{c, sn};
                }
                __syncthreads();
                // every rotation this thread applies, read in one batch of 16-byte
                // loads (one LDS round trip instead of one per use)
                const v2d csa = cs[ka], csb = cs[kb], csw[2] = {cs[wk], cs[wk + TP / 4]};
                {
                    const double ca = csa.x, sa = csa.y, cb = csb.x, sb = csb.y;
                    if (sa != 0.0 || sb != 0.0) {
                        const double rpu = ca * spu - sa * squ, rpv = ca * spv - sa * sqv;
                        const double rqu = sa * spu + ca * squ, rqv = sa * spv + ca * sqv;
                        S[p][u] = cb * rpu - sb * rpv;
                        S[p][v] = sb * rpu + cb * rpv;
                        S[q][u] = cb * rqu - sb * rqv;
                        S[q][v] = sb * rqu + cb * rqv;
                        if (ka == kb) {  // the annihilated entry of this plane
                            S[p][q] = 0.0;
                            S[q][p] = 0.0;
                        }
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) {  // W <- W J (columns of plane wk + h TP/4)
                    const double c = csw[h].x, sn = csw[h].y;
                    if (sn != 0.0) {
                        W[wi][wp[h]] = c * wx[h] - sn * wy[h];
                        W[wi][wq[h]] = sn * wx[h] + c * wy[h];
                    }
                }
                __syncthreads();
            }
        };
        // The cross-block pass (TB rounds): plane k joins p = k of the first
        // block with q = TB + (k + r) mod TB of the second in round r, so the
        // rounds meet every (first block, second block) plane once.  Here p =
        // ka, u = kb and the W columns wk, wk + TB/2 stay fixed: S[ka][kb] and
        // W[wi][wk], W[wi][wk + TB/2] are touched by this thread alone and live
        // in registers for the pass (the diagonal owners also store theirs to
        // LDS each round, where the rotation threads read it); written back at
        // the end.
        auto cross_rounds = [&]() __attribute__((always_inline)) {
            constexpr int M = TB - 1;
            const bool rot_thread = tid < TB;   // ka == 0, kb == tid: plane tid
            double spu = S[ka][kb];
            double wx0 = W[wi][wk], wx1 = W[wi][wk + TB / 2];
#pragma unroll 1
            for (int r = 0; r < TB; ++r) {
                const int q = TB + ((ka + r) & M), v = TB + ((kb + r) & M);
                const int wq0 = TB + ((wk + r) & M), wq1 = TB + ((wk + TB / 2 + r) & M);
                double apq = 0.0, app = 0.0, aqq = 0.0;
                if (rot_thread) {
                    apq = S[kb][v];
                    app = S[kb][kb];
                    aqq = S[v][v];
                }
                const double spv = S[ka][v], squ = S[q][kb], sqv = S[q][v];
                const double wy0 = W[wi][wq0], wy1 = W[wi][wq1];
                if (rot_thread) {
                    double c = 1.0, sn = 0.0;
                    if (rotates(apq, app, aqq, tol, floor_abs)) {
                        rotation(apq, app, aqq, c, sn);
                        any_rot = 1;
                    }
                    cs[tid] = v2d{c, sn};
                }
                __syncthreads();
                const v2d csa = cs[ka], csb = cs[kb], csw0 = cs[wk], csw1 = cs[wk + TB / 2];
                {
                    const double ca = csa.x, sa = csa.y, cb = csb.x, sb = csb.y;
                    if (sa != 0.0 || sb != 0.0) {
                        const double rpu = ca * spu - sa * squ, rpv = ca * spv - sa * sqv;
                        const double rqu = sa * spu + ca * squ, rqv = sa * spv + ca * sqv;
                        spu = cb * rpu - sb * rpv;
                        S[ka][v] = sb * rpu + cb * rpv;
                        S[q][kb] = cb * rqu - sb * rqv;
                        S[q][v] = sb * rqu + cb * rqv;
                        if (ka == kb) {  // the diagonal entry, and the annihilated one
                            S[ka][kb] = spu;
                            S[ka][q] = 0.0;
                            S[q][ka] = 0.0;
                        }
                    }
                }
                if (csw0.y != 0.0) {  // W <- W J (columns of planes wk, wk + TB/2)
                    const double x = wx0;
                    wx0 = csw0.x * x - csw0.y * wy0;
                    W[wi][wq0] = csw0.y * x + csw0.x * wy0;
                }
                if (csw1.y != 0.0) {
                    const double x = wx1;
                    wx1 = csw1.x * x - csw1.y * wy1;
                    W[wi][wq1] = csw1.y * x + csw1.x * wy1;
                }
                __syncthreads();
            }
            S[ka][kb] = spu;
            W[wi][wk] = wx0;
            W[wi][wk + TB / 2] = wx1;
        };
        if (cross)
            cross_rounds();
        else
            rounds(RrPlane<TP>{});
        // any_rot was last written before the final barrier of the sweep
        const int any = any_rot;
        __syncthreads();
        if (!any) break;
    }
    for (int e = tid; e < TP * TP; e += NT) Wout[e] = W[e / TP][e % TP];
    if (tid == 0) {
        flags[pw] = 1;
        atomicAdd(P.counter, 1);
    }
}


// Pair of step `step` - 1 holding block x, and the half (0: first 16 rows,
// 1: second) it sits in: the inverse of tj_pair's round-robin map.
// 32-bit arithmetic as in tj_pair (x, sp < np, np far below 2^31).
__device__ __forceinline__ void tj_owner(int64_t x, int64_t sp, int64_t np, int64_t half, int64_t &w, int &h) {
    const int xi = (int)x, si = (int)sp, np1 = (int)np - 1;
    if (xi == np1) {
        w = 0;
        h = 1;
    } else if (xi == si) {
        w = 0;
        h = 0;
    } else {
        const int w1 = ((xi - si) % np1 + np1) % np1;
        if (w1 >= 1 && w1 <= (int)half - 1) {
            w = w1;
            h = 0;
        } else {
            w = ((si - xi) % np1 + np1) % np1;
            h = 1;
        }
    }
}

// One fused step: the tile updates of step - 1 (has_update) and the pair
// solves of step in one launch, so the updates' work-groups run beside the
// latency-bound solves.  The Gram matrix is read from buffer `cur` (0: P.g,
// 1: the workspace copy) and the updates write the other buffer, every tile
// (unrotated ones are copied), so no buffer is read and written in the same
// launch.  A solve builds its tile of G_step = W^T G_(step-1) W from buffer
// `cur` and the previous step's rotations: for its blocks a and b, with
// V = diag(V_A, V_B), V_A = W_A(:, half of a) over the rows of a's pair A at
// step - 1 (and V_B likewise), S = V^T G_old V.  Solves come first in the
// grid (they are the long ones).  W tiles and flags alternate between two
// sets by step parity.
//
// Work-groups are dealt to the eight XCDs round-robin by launch index; the
// two sections of the grid (solves, then updates) are each padded to a
// multiple of 8 and read in XCD-major order, so XCD x takes a contiguous run
// of each section: the solves and tiles of neighbouring problems, whose W
// tiles and Gram rows then meet in that XCD's L2 instead of being fetched by
// all eight.
__device__ __forceinline__ int64_t tj_xcd_order(int64_t b, int64_t padded) {
    return (b & 7) * (padded >> 3) + (b >> 3);
}

// The register products below use the f64 MFMA 16x16x4 layout: lane
// l = 16 lk + li supplies A[li][k] and B[k][li] of k-step j (k = 4 j + lk)
// and holds D[lk + 4r][li] in register r, so a product's result is the B
// operand of a next product that contracts over its rows.  Gram tiles enter
// as A operands through their transposes: A[li][k] = G(row of k, row of li),
// sixteen consecutive doubles of one Gram row per load.
template <int TB>
__global__ __launch_bounds__(TB * TB, 5) void tjf_step_kernel(const syn_gbj_item *__restrict__ items,
                                                          const int32_t *__restrict__ owner_s,
                                                          const int32_t *__restrict__ owner_t, int64_t step,
                                                          int has_update, int64_t solve_pad, int64_t solve_count,
                                                          int64_t tile_pad, int64_t tile_count, int cur, double tol,
                                                          int passes, unsigned long long *bytes) {
    constexpr int TP = 2 * TB, NT = TB * TB;
    static_assert(TB == 16 && NT == kTUpdThreads, "fused steps: 16-row blocks, one block size for both roles");
    // LDS: the solve role's tile S and rotation W (17 KiB); the update role
    // works in registers
    __shared__ double S[TP][TP + 1];
    __shared__ double W[TP][TP + 1];
    __shared__ v2d cs[TP / 2];
    __shared__ int rows[TP];
    __shared__ int do_rot, any_rot;
    const int64_t bid = blockIdx.x;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    auto gbuf = [](const syn_gbj_item &P, int which, int64_t half, int64_t &ld) -> double * {
        if (which == 0) {
            ld = P.ldg;
            return P.g;
        }
        ld = P.nv;
        return P.w + tj_g2_off<TB>(half, P.nv);
    };
    if (bid < solve_pad) {
        // ---- pair solve of `step` ------------------------------------------
        const int64_t sb = tj_xcd_order(bid, solve_pad);
        if (sb >= solve_count) return;
        const syn_gbj_item &P = items[owner_s[sb]];
        const TjDims D = tj_dims<TB>(P.nv);
        const int64_t pw = sb - P.solve_start;
        const int64_t st = step % (D.np - 1);
        int64_t ldc;
        const double *Gc = gbuf(P, cur, D.half, ldc);
        if (!has_update) {
            if (tid < TP) rows[tid] = tj_row<TB>(st, pw, D.np, D.nblk, P.nv, tid);
            __syncthreads();
            for (int e = tid; e < TP * TP; e += NT) {
                const int i = e / TP, k = e % TP, ri = rows[i], rk = rows[k];
                S[i][k] = (ri >= 0 && rk >= 0) ? Gc[(int64_t)ri * ldc + rk] : 0.0;
            }
        } else {
            // wave (si, sj) builds tile (si, sj) of S: with the old pair wr
            // of new block si and wc of new block sj,
            //   Z = G_old(rows of wr, rows of wc) V_sj   (32 x 16)
            //   S(si, sj) = V_si^T Z                     (16 x 16)
            // so the four waves read the four quadrants of G_old once.
            const int64_t sp = (step - 1) % (D.np - 1);
            const int si = wv >> 1, sj = wv & 1;
            int64_t ba, bb, wA, wB;
            int hA, hB;
            tj_pair(st, pw, D.np, ba, bb);
            tj_owner(ba, sp, D.np, D.half, wA, hA);
            tj_owner(bb, sp, D.np, D.half, wB, hB);
            const int64_t wr = si ? wB : wA, wc = sj ? wB : wA;
            const int hr = si ? hB : hA, hc = sj ? hB : hA;
            const int pp = (int)((step - 1) & 1);
            const int *fl = tj_flags<TB>(P, D.half, pp);
            const int fr = fl[wr], fc = fl[wc];
            const double *Wr = tj_wset<TB>(P, D.half, pp) + wr * TP * TP;
            const double *Wc = tj_wset<TB>(P, D.half, pp) + wc * TP * TP;
            // G_old(rows of wr, columns of wc), read through its transpose
            // (G is kept symmetric) so that lanes li take 16 consecutive
            // doubles of one row; padding reads as zero
            int rr[2];
#pragma unroll
            for (int rho = 0; rho < 2; ++rho) rr[rho] = tj_row<TB>(sp, wr, D.np, D.nblk, P.nv, 16 * rho + li);
            double g[2][8], v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int cr = tj_row<TB>(sp, wc, D.np, D.nblk, P.nv, 4 * j + lk);
#pragma unroll
                for (int rho = 0; rho < 2; ++rho) {
                    const bool ok = rr[rho] >= 0 && cr >= 0;
                    const double val = Gc[ok ? (int64_t)cr * ldc + rr[rho] : 0];
                    g[rho][j] = ok ? val : 0.0;
                }
                const int e = 4 * j + lk;
                v[j] = fc ? Wc[e * TP + hc * TB + li] : (e == hc * TB + li ? 1.0 : 0.0);
            }
            v4d z[2];
#pragma unroll
            for (int rho = 0; rho < 2; ++rho) {
                z[rho] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int j = 0; j < 8; ++j) z[rho] = __builtin_amdgcn_mfma_f64_16x16x4f64(g[rho][j], v[j], z[rho], 0, 0, 0);
            }
            v4d sacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int rho = 0; rho < 2; ++rho)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = 16 * rho + 4 * r + lk;
                    const double a = fr ? Wr[e * TP + hr * TB + li] : (e == hr * TB + li ? 1.0 : 0.0);
                    sacc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, z[rho][r], sacc, 0, 0, 0);
                }
#pragma unroll
            for (int r = 0; r < 4; ++r) S[TB * si + lk + 4 * r][TB * sj + li] = sacc[r];
        }
        const int ps = (int)(step & 1);
        // the first step of a pass meets every block and treats the planes
        // within blocks; the others only the planes between their two blocks,
        // so a pass treats every plane of G once
        tj_pair_solve<TB>(P, D, pw, S, W, cs, do_rot, any_rot, tol, passes,
                          tj_wset<TB>(P, D.half, ps) + pw * TP * TP, tj_flags<TB>(P, D.half, ps), st != 0);
        return;
    }
    if (!has_update) return;
    // ---- updates of step - 1: buffer cur -> buffer cur ^ 1, in registers ----
    // (no barrier from here on: waves return independently)
    constexpr unsigned long long kTileBytes = (unsigned long long)TP * TP * sizeof(double);
    constexpr int kCols = tj_slab_cols<TB>();
    constexpr unsigned long long kSlabBytes = 2ull * TP * kCols * sizeof(double);
    const int64_t ub = tj_xcd_order(bid - solve_pad, tile_pad);
    if (ub >= tile_count) return;
    const syn_gbj_item &P = items[owner_t[ub]];
    const TjDims D = tj_dims<TB>(P.nv);
    const int64_t su = step - 1, st = su % (D.np - 1);
    const int pu = (int)(su & 1);
    const int64_t t = ub - P.tile_start;
    const int64_t ng = tj_gtiles(D.half), gw = (ng + 1) / 2;
    const int *flags = tj_flags<TB>(P, D.half, pu);
    const double *wset = tj_wset<TB>(P, D.half, pu);
    if (t < gw) {
        // G tile (a, b), a <= b, and its transpose (b, a): waves 2h, 2h + 1
        // take tile 2t + h, wave parity the 16 output columns tj
        const int64_t gt = 2 * t + (wv >> 1);
        if (gt >= ng) return;
        int64_t a = 0, rem = gt;
        while (rem >= D.half - a) {
            rem -= D.half - a;
            ++a;
        }
        const int64_t b = a + rem;
        const int tj = wv & 1;
        const int fa = flags[a], fb = flags[b];
        int64_t ldc, ldn;
        const double *Gc = gbuf(P, cur, D.half, ldc);
        double *Gn = gbuf(P, cur ^ 1, D.half, ldn);
        // x[ti][j] = X(16 ti + li, 4 j + lk), X = G(rows of pair a, rows of
        // pair b), read through its transpose (G is kept symmetric) so that
        // lanes li take 16 consecutive doubles of one row
        int ra2[2], cr[8];
        double x[2][8];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) ra2[ti] = tj_row<TB>(st, a, D.np, D.nblk, P.nv, TB * ti + li);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            cr[j] = tj_row<TB>(st, b, D.np, D.nblk, P.nv, 4 * j + lk);
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) {
                const bool ok = ra2[ti] >= 0 && cr[j] >= 0;
                const double val = Gc[ok ? (int64_t)cr[j] * ldc + ra2[ti] : 0];
                x[ti][j] = ok ? val : 0.0;
            }
        }
        // bytes moved for this tile: one tile read, one or two written
        if (bytes && tj == 0 && lane == 0) atomicAdd(bytes + owner_t[ub], (a == b ? 2ull : 3ull) * kTileBytes);
        if (!fa && !fb) {
            // nothing rotated: wave tj copies rows 16 tj .. 16 tj + 15 of X
            // (and their transpose)
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) {   // static register indices only
                const int gr = ra2[ti];
                if (ti != tj || gr < 0) continue;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (cr[j] >= 0) {
                        Gn[(int64_t)gr * ldn + cr[j]] = x[ti][j];
                        if (a != b) Gn[(int64_t)cr[j] * ldn + gr] = x[ti][j];
                    }
                }
            }
            return;
        }
        const double *wa = wset + a * TP * TP, *wb = wset + b * TP * TP;
        // U = X Wb(:, tj): two 16 x 16 tiles (rows 0-15, 16-31)
        double vb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = 4 * j + lk, c = TB * tj + li;
            vb[j] = fb ? wb[e * TP + c] : (e == c ? 1.0 : 0.0);
        }
        v4d u[2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
            u[ti] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 8; ++j) u[ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[ti][j], vb[j], u[ti], 0, 0, 0);
        }
        // X_new(ti, tj) = Wa(:, 16 ti ..)^T U, contraction rows 4 j + lk
        const int gc = tj_row<TB>(st, b, D.np, D.nblk, P.nv, TB * tj + li);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int e = 4 * j + lk, c = TB * ti + li;
                const double av = fa ? wa[e * TP + c] : (e == c ? 1.0 : 0.0);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, u[j >> 2][j & 3], acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = tj_row<TB>(st, a, D.np, D.nblk, P.nv, TB * ti + lk + 4 * r);
                if (gr >= 0 && gc >= 0) {
                    Gn[(int64_t)gr * ldn + gc] = acc[r];
                    if (a != b) Gn[(int64_t)gc * ldn + gr] = acc[r];
                }
            }
        }
        return;
    }
    // rt slab (a, columns [c0, c0 + kCols)), in place: wave wv takes columns
    // c0 + 64 h + 16 wv .. c0 + 64 h + 16 wv + 15 (h = 0, 1) of every row of
    // pair a; the W operand of a k-step serves both column groups
    const int64_t u = t - gw;
    const int64_t a = u / D.ncs, c0 = (u % D.ncs) * kCols;
    if (!flags[a]) return;
    if (bytes && tid == 0) atomicAdd(bytes + owner_t[ub], kSlabBytes);
    const double *wa = wset + a * TP * TP;
    int64_t col[2];
    bool cok[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        col[h] = c0 + kTCol * h + TB * wv + li;
        cok[h] = col[h] < P.nv;
    }
    double xs[2][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int r = tj_row<TB>(st, a, D.np, D.nblk, P.nv, 4 * j + lk);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool ok = r >= 0 && cok[h];
            const double val = P.rt[ok ? (int64_t)r * P.ldr + col[h] : 0];
            xs[h][j] = ok ? val : 0.0;
        }
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        v4d acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double w = wa[(4 * j + lk) * TP + TB * ti + li];
#pragma unroll
            for (int h = 0; h < 2; ++h) acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(w, xs[h][j], acc[h], 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gr = tj_row<TB>(st, a, D.np, D.nblk, P.nv, TB * ti + lk + 4 * r);
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (gr >= 0 && cok[h]) P.rt[(int64_t)gr * P.ldr + col[h]] = acc[h][r];
        }
    }
}

// Wide blocks (32 rows, 64 x 64 pair tiles), for launches of many problems:
// a sweep takes half the steps of 16-row blocks, so the Gram matrix and rt
// stream through HBM half as often for the same MFMA work.  Solves and
// updates are launches of their own, one after the other: the solve reads
// its tile straight from G (as the previous step's updates left it) and
// rotates it in LDS with 1024 threads; the update applies the step's
// rotations to G and rt in place, in registers.  W tiles and flags use the
// first parity set only.
template <int TB>
__global__ __launch_bounds__(TB * TB) void tj_solve_kernel(const syn_gbj_item *__restrict__ items,
                                                           const int32_t *__restrict__ owner_s, int64_t step,
                                                           int64_t solve_pad, int64_t solve_count, double tol) {
    constexpr int TP = 2 * TB, NT = TB * TB;
    __shared__ double S[TP][TP + 1];
    __shared__ double W[TP][TP + 1];
    __shared__ v2d cs[TP / 2];
    __shared__ int rows[TP];
    __shared__ int do_rot, any_rot;
    const int64_t sb = tj_xcd_order(blockIdx.x, solve_pad);
    if (sb >= solve_count) return;
    const syn_gbj_item &P = items[owner_s[sb]];
    const TjDims D = tj_dims<TB>(P.nv);
    const int64_t pw = sb - P.solve_start;
    const int64_t st = step % (D.np - 1);
    const int tid = threadIdx.x;
    if (tid < TP) rows[tid] = tj_row<TB>(st, pw, D.np, D.nblk, P.nv, tid);
    __syncthreads();
    for (int e = tid; e < TP * TP; e += NT) {
        const int i = e / TP, k = e % TP, ri = rows[i], rk = rows[k];
        S[i][k] = (ri >= 0 && rk >= 0) ? P.g[(int64_t)ri * P.ldg + rk] : 0.0;
    }
    // the first step of a pass sweeps the whole tile, the others the planes
    // between its two blocks
    tj_pair_solve<TB>(P, D, pw, S, W, cs, do_rot, any_rot, tol, 1, tj_wset<TB>(P, D.half, 0) + pw * TP * TP,
                      tj_flags<TB>(P, D.half, 0), st != 0);
}

// Updates of a step with 32-row blocks, in place: a work-group takes one
// 64 x 64 G tile (a, b), a <= b (wave w: output columns 16 w .. 16 w + 15;
// every wave reads the whole tile, so the writes wait for a barrier), or one
// rt slab of 64 rows x 64 columns (wave w: columns 16 w .. 16 w + 15, which
// it alone reads and writes).  Tiles of unrotated pairs stay as they are.
__global__ __launch_bounds__(256) void tjw_update_kernel(const syn_gbj_item *__restrict__ items,
                                                         const int32_t *__restrict__ owner_t, int64_t step,
                                                         int64_t tile_pad, int64_t tile_count,
                                                         unsigned long long *bytes) {
    constexpr int TB = 32, TP = 64;
    constexpr unsigned long long kTileBytes = (unsigned long long)TP * TP * sizeof(double);
    constexpr unsigned long long kSlabBytes = 2ull * TP * kTCol * sizeof(double);
    const int64_t ub = tj_xcd_order(blockIdx.x, tile_pad);
    if (ub >= tile_count) return;
    const syn_gbj_item &P = items[owner_t[ub]];
    const TjDims D = tj_dims<TB>(P.nv);
    const int64_t st = step % (D.np - 1);
    const int64_t t = ub - P.tile_start, ng = tj_gtiles(D.half);
    const int *flags = tj_flags<TB>(P, D.half, 0);
    const double *wset = tj_wset<TB>(P, D.half, 0);
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    if (t < ng) {
        int64_t a = 0, rem = t;
        while (rem >= D.half - a) {
            rem -= D.half - a;
            ++a;
        }
        const int64_t b = a + rem;
        const int fa = flags[a], fb = flags[b];
        if (!fa && !fb) return;
        if (bytes && tid == 0) atomicAdd(bytes + owner_t[ub], (a == b ? 2ull : 3ull) * kTileBytes);
        double *G = P.g;
        const int64_t ld = P.ldg;
        const double *wa = wset + a * TP * TP, *wb = wset + b * TP * TP;
        const int tj = wv;
        int ra[4];
#pragma unroll
        for (int rho = 0; rho < 4; ++rho) ra[rho] = tj_row<TB>(st, a, D.np, D.nblk, P.nv, 16 * rho + li);
        // U = X Wb(:, 16 tj ..): X(i, k) read as G(row of k, row of i)
        v4d u[4];
#pragma unroll
        for (int rho = 0; rho < 4; ++rho) u[rho] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int j = 0; j < 16; ++j) {
            const int e = 4 * j + lk, c = 16 * tj + li;
            const int cr = tj_row<TB>(st, b, D.np, D.nblk, P.nv, e);
            const double vb = fb ? wb[e * TP + c] : (e == c ? 1.0 : 0.0);
#pragma unroll
            for (int rho = 0; rho < 4; ++rho) {
                const bool ok = ra[rho] >= 0 && cr >= 0;
                const double xv = G[ok ? (int64_t)cr * ld + ra[rho] : 0];
                u[rho] = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? xv : 0.0, vb, u[rho], 0, 0, 0);
            }
        }
        // X_new(16 ti .., 16 tj ..) = Wa(:, 16 ti ..)^T U
        v4d acc[4];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            acc[ti] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int e = 4 * j + lk, c = 16 * ti + li;
                const double av = fa ? wa[e * TP + c] : (e == c ? 1.0 : 0.0);
                acc[ti] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, u[j >> 2][j & 3], acc[ti], 0, 0, 0);
            }
        }
        __syncthreads();   // every wave has read the tile
        const int gc = tj_row<TB>(st, b, D.np, D.nblk, P.nv, 16 * tj + li);
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = tj_row<TB>(st, a, D.np, D.nblk, P.nv, 16 * ti + lk + 4 * r);
                if (gr >= 0 && gc >= 0) {
                    G[(int64_t)gr * ld + gc] = acc[ti][r];
                    if (a != b) G[(int64_t)gc * ld + gr] = acc[ti][r];
                }
            }
        return;
    }
    const int64_t u2 = t - ng;
    const int64_t a = u2 / D.ncs, c0 = (u2 % D.ncs) * kTCol;
    if (!flags[a]) return;
    if (bytes && tid == 0) atomicAdd(bytes + owner_t[ub], kSlabBytes);
    const double *wa = wset + a * TP * TP;
    const int64_t col = c0 + 16 * wv + li;
    const bool cok = col < P.nv;
    double xs[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int r = tj_row<TB>(st, a, D.np, D.nblk, P.nv, 4 * j + lk);
        const bool ok = r >= 0 && cok;
        const double val = P.rt[ok ? (int64_t)r * P.ldr + col : 0];
        xs[j] = ok ? val : 0.0;
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 16; ++j)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[(4 * j + lk) * TP + 16 * ti + li], xs[j], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gr = tj_row<TB>(st, a, D.np, D.nblk, P.nv, 16 * ti + lk + 4 * r);
            if (gr >= 0 && cok) P.rt[(int64_t)gr * P.ldr + col] = acc[r];
        }
    }
}

// Copy the Gram matrix back from the workspace buffer (ld nv) into P.g for
// the problems whose last fused step left it there (which[i] != 0).
template <int TB>
__global__ __launch_bounds__(256) void tjf_copyback_kernel(const syn_gbj_item *__restrict__ items,
                                                           const int32_t *__restrict__ which) {
    const syn_gbj_item &P = items[blockIdx.x];
    if (!which[blockIdx.x]) return;
    const double *g2 = P.w + tj_g2_off<TB>(tj_dims<TB>(P.nv).half, P.nv);
    const int64_t nel = P.nv * P.nv;
    for (int64_t e = (int64_t)blockIdx.y * 256 + threadIdx.x; e < nel; e += (int64_t)gridDim.y * 256) {
        const int64_t i = e / P.nv, j = e - i * P.nv;
        P.g[i * P.ldg + j] = g2[e];
    }
}

template <int TB>
__global__ __launch_bounds__(256) void tj_init_kernel(const syn_gbj_item *__restrict__ items) {
    __shared__ double red[4];
    const syn_gbj_item &P = items[blockIdx.x];
    const int64_t nel = P.nv * P.nv;
    for (int64_t e = threadIdx.x; e < nel; e += 256) {
        const int64_t i = e / P.nv, j = e - i * P.nv;
        P.rt[i * P.ldr + j] = (i == j) ? 1.0 : 0.0;
    }
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < P.nv; i += 256) m = fmax(m, fabs(P.g[i * P.ldg + i]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        P.w[tj_floor_off<TB>(tj_dims<TB>(P.nv).half)] =
            __DBL_EPSILON__ * fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// End-of-sweep test: chk[i] = 1 when some off-diagonal entry of problem i's
// Gram matrix (buffer `cur`: 0 P.g, 1 the workspace copy) passes the
// rotation test, i.e. when the next sweep would rotate.  A sweep's first
// step reads its tiles straight from G, and later tiles stay exact copies of
// G while nothing has rotated, so "no entry passes" means the next sweep
// rotates nothing.  Both triangles are tested: a pair solve tests G_pq or
// G_qp depending on its tile order, and G is symmetric only to rounding.
template <int TB>
__global__ __launch_bounds__(256) void tj_check_kernel(const syn_gbj_item *__restrict__ items, int cur, double tol,
                                                       const int32_t *counters, int32_t *chk) {
    const syn_gbj_item &P = items[blockIdx.x];
    const int64_t half = tj_dims<TB>(P.nv).half;
    int64_t ld = P.ldg;
    const double *G = P.g;
    if (cur) {
        ld = P.nv;
        G = P.w + tj_g2_off<TB>(half, P.nv);
    }
    const double floor_abs = P.w[tj_floor_off<TB>(half)];
    bool hit = false;
    const int64_t nel = P.nv * P.nv;
    for (int64_t e = (int64_t)blockIdx.y * 256 + threadIdx.x; e < nel && !hit; e += (int64_t)gridDim.y * 256) {
        const int64_t i = e / P.nv, k = e - i * P.nv;
        if (i != k && rotates(G[i * ld + k], G[i * ld + i], G[k * ld + k], tol, floor_abs)) hit = true;
    }
    if (hit) chk[P.counter - counters] = 1;
}

__global__ __launch_bounds__(256) void tj_diag_kernel(const syn_gbj_item *__restrict__ items) {
    const syn_gbj_item &P = items[blockIdx.x];
    for (int64_t i = threadIdx.x; i < P.nv; i += 256) P.s[i] = sqrt(fmax(P.g[i * P.ldg + i], 0.0));
}

// rows of rt and entries of s in the order perm (problem blockIdx.x, slice
// blockIdx.y), through tmp; pass 0 gathers into tmp, pass 1 copies back
__global__ __launch_bounds__(256) void tj_permute_kernel(const syn_gbj_item *__restrict__ items,
                                                         const int64_t *__restrict__ perm,
                                                         const int64_t *__restrict__ poff, double *tmp,
                                                         const int64_t *__restrict__ toff, int pass) {
    const syn_gbj_item &P = items[blockIdx.x];
    const int64_t *pm = perm + poff[blockIdx.x];
    double *tr = tmp + toff[blockIdx.x];
    double *ts = tr + P.nv * P.nv;
    const int64_t nr = P.nv * P.nv;
    for (int64_t e = blockIdx.y * 256 + threadIdx.x; e < nr + P.nv; e += (int64_t)gridDim.y * 256) {
        if (e < nr) {
            const int64_t i = e / P.nv, j = e - i * P.nv;
            if (pass == 0)
                tr[e] = P.rt[pm[i] * P.ldr + j];
            else
                P.rt[i * P.ldr + j] = tr[e];
        } else {
            const int64_t i = e - nr;
            if (pass == 0)
                ts[i] = P.s[pm[i]];
            else
                P.s[i] = ts[i];
        }
    }
}

// Decreasing order of each problem's s (ties by index, as a stable sort):
// rank[i] = #{j : s_j > s_i or (s_j == s_i and j < i)}, perm[rank[i]] = i.
// s is staged through LDS in chunks of kRankChunk values.
constexpr int kRankChunk = 2048;
__global__ __launch_bounds__(256) void tj_rank_kernel(const syn_gbj_item *__restrict__ items,
                                                      const int64_t *__restrict__ poff, int64_t *__restrict__ perm) {
    __shared__ double sv[kRankChunk];
    const syn_gbj_item &P = items[blockIdx.x];
    const int64_t nv = P.nv;
    int64_t *pm = perm + poff[blockIdx.x];
    for (int64_t i0 = 0; i0 < nv; i0 += 256) {
        const int64_t i = i0 + threadIdx.x;
        // a NaN orders last (as -inf), so the ranks always form a permutation
        const double s0 = i < nv ? P.s[i] : 0.0, si = s0 == s0 ? s0 : -__builtin_huge_val();
        int64_t rk = 0;
        for (int64_t c0 = 0; c0 < nv; c0 += kRankChunk) {
            const int64_t cn = nv - c0 < kRankChunk ? nv - c0 : kRankChunk;
            __syncthreads();
            for (int64_t e = threadIdx.x; e < cn; e += 256) {
                const double x = P.s[c0 + e];
                sv[e] = x == x ? x : -__builtin_huge_val();
            }
            __syncthreads();
            if (i < nv)
                for (int64_t e = 0; e < cn; ++e) {
                    const double sj = sv[e];
                    rk += (sj > si) || (sj == si && c0 + e < i);
                }
        }
        if (i < nv) pm[rk] = i;
    }
}

__global__ __launch_bounds__(256) void tj_zero_kernel(int32_t *c, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) c[i] = 0;
}

// per-problem workspace: W tiles and flags (two sets), floor, the sort
// scratch (rt copy, s) and the second Gram buffer
template <int TB>
static int64_t tj_work(int64_t nv) {
    const int64_t half = tj_dims<TB>(nv).half;
    return tj_g2_off<TB>(half, nv) + nv * nv;
}

// Gram entries from which a call counts as many problems: its launches are
// bound by the tile updates' bytes rather than by the pair solves' latency.
constexpr int64_t kManyEntries = int64_t(1) << 23;

// the wide (32-row) blocks take the problems of order kWideMinOrder and
// above when together they hold at least kWideMinEntries Gram entries (below
// that a call is latency-bound and the fused 16-row steps finish first)
constexpr int64_t kWideMinOrder = 384;
constexpr int64_t kWideMinEntries = kManyEntries;

// Pinned host staging of the host loop's small transfers.  Copies from and
// to pageable memory go through the runtime's own staging and wait for the
// stream, so each readback of a sweep's flags would cost a round trip of its
// own; from pinned memory they are plain DMA on the stream, and a sweep
// waits once.  Buffers come from a process-wide pool (any host thread may
// call the library) and go back to it when the lease ends, after its stream
// has finished every copy that reads or writes them; they are never freed.
class PinnedLease {
  public:
    PinnedLease(size_t bytes, hipStream_t st) : st_(st) {
        {
            std::lock_guard<std::mutex> g(mu());
            auto &pool = free_list();
            size_t best = pool.size();
            for (size_t i = 0; i < pool.size(); ++i)
                if (pool[i].second >= bytes && (best == pool.size() || pool[i].second < pool[best].second)) best = i;
            if (best != pool.size()) {
                buf_ = pool[best].first;
                cap_ = pool[best].second;
                pool.erase(pool.begin() + (ptrdiff_t)best);
                return;
            }
        }
        cap_ = std::max(bytes, (size_t)65536);
        void *p = nullptr;
        rc_ = hip_check(hipHostMalloc(&p, cap_, hipHostMallocDefault), "pinned staging");
        buf_ = rc_ ? nullptr : (char *)p;
    }
    ~PinnedLease() {
        if (!buf_) return;
        (void)hipStreamSynchronize(st_);
        std::lock_guard<std::mutex> g(mu());
        free_list().emplace_back(buf_, cap_);
    }
    PinnedLease(const PinnedLease &) = delete;
    PinnedLease &operator=(const PinnedLease &) = delete;
    char *data() const { return buf_; }
    int rc() const { return rc_; }

  private:
    static std::mutex &mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<std::pair<char *, size_t>> &free_list() {
        static std::vector<std::pair<char *, size_t>> v;
        return v;
    }
    hipStream_t st_;
    char *buf_ = nullptr;
    size_t cap_ = 0;
    int rc_ = 0;
};

template <int TB>
static int gram_blocked(syn_gbj_item *items, std::vector<syn_gbj_item> &all, int n_items, double *work,
                        int32_t *counters, double tol, int max_sweeps, int *sweeps, hipStream_t st) {
    // passes of the pair solves' rounds (while they rotate): a launch of many
    // problems is bound by its tile updates, beside which a second pass costs
    // little and saves sweeps (batch of 4096 chains: 7272 -> 5996 launches a
    // step, 3.5% faster); a launch of a few problems waits on its solves,
    // where it costs more than it saves (8 order-512 problems: 8.6 -> 10.9 ms)
    int64_t entries = 0;
    for (int i = 0; i < n_items; ++i) entries += all[i].nv * all[i].nv;
    const int passes = entries >= kManyEntries ? 2 : 1;
    std::vector<int64_t> woff(n_items);
    int64_t tot_w = 0;
    for (int i = 0; i < n_items; ++i) {
        woff[i] = tot_w;
        all[i].w = work + tot_w;
        tot_w += tj_work<TB>(all[i].nv);
        all[i].counter = counters + i;
        all[i].active = all[i].nv > 1;
        all[i].solve_start = all[i].tile_start = 0;
    }
    // pinned regions, one per kind of transfer: the item table (first and
    // last upload), the active table, the offsets, the Gram parities, the
    // per-sweep readback (counters, then the sweep test flags) and the byte
    // counters.  An upload is copied into its region and sent on the stream
    // without a wait (stream order puts it after the kernels that read the
    // old table); a region is written again only after a wait on the stream
    // that covers its previous copy (the active table's after the sweep's
    // readback wait), and the lease waits for the stream before it ends.
    const size_t n = (size_t)n_items, tbl = (n * sizeof(syn_gbj_item) + 63) / 64 * 64;
    const size_t r_all0 = 0, r_all1 = tbl, r_act = 2 * tbl, r_off = 3 * tbl, r_par = r_off + (16 * n + 63) / 64 * 64,
                 r_back = r_par + (4 * n + 63) / 64 * 64, r_bytes = r_back + (8 * n + 63) / 64 * 64,
                 r_end = r_bytes + 8 * n;
    PinnedLease pin(r_end, st);
    if (pin.rc()) return pin.rc();
    auto upload = [&](size_t region, const void *src, void *dst, size_t bytes, const char *what) -> int {
        memcpy(pin.data() + region, src, bytes);
        return hip_check(hipMemcpyAsync(dst, pin.data() + region, bytes, hipMemcpyHostToDevice, st), what);
    };
    if (int rc = upload(r_all0, all.data(), items, all.size() * sizeof(syn_gbj_item), "gram table")) return rc;
    hipLaunchKernelGGL(tj_init_kernel<TB>, dim3(n_items), dim3(256), 0, st, items);
    SYN_LAUNCH_CHECK("gram init");
    // byte counter of the update launches, after the sort scratch
    int64_t tot_nv = 0;
    for (int i = 0; i < n_items; ++i) tot_nv += all[i].nv;
    unsigned long long *dbytes = (unsigned long long *)(work + tot_w + tot_nv + 2 * (int64_t)n_items);
    hipLaunchKernelGGL(tj_zero_kernel, dim3((n_items + 127) / 128), dim3(256), 0, st, (int32_t *)dbytes,
                       2 * n_items);
    SYN_LAUNCH_CHECK("gram byte counter");
    // the update work-groups count the Gram and rt bytes they move, one
    // counter per problem of the sweep, only when a profile asks for them
    unsigned long long *kbytes = prof_wants(kTagGramBytes) ? dbytes : nullptr;

    // end-of-sweep test flags, in the workspace's spare tail after the byte
    // counter (at least tot_nv + 15 doubles are left there)
    int32_t *chk = (int32_t *)(dbytes + n_items);
    // owner table of the fused steps, after n_items doubles for the flags
    int32_t *owner = (int32_t *)(dbytes + 2 * (int64_t)n_items);
    // order and permutation scratch of the end: each problem's offset in the
    // permutation array and its rt-copy scratch
    int64_t *dperm = (int64_t *)(work + tot_w);
    int64_t *dpoff = dperm + tot_nv;
    int64_t *dtoff = dpoff + n_items;
    {
        std::vector<int64_t> pt(2 * (size_t)n_items);
        int64_t acc_nv = 0;
        for (int i = 0; i < n_items; ++i) {
            pt[i] = acc_nv;
            acc_nv += all[i].nv;
            pt[n_items + i] = woff[i] + tj_floor_off<TB>(tj_dims<TB>(all[i].nv).half) + 1;
        }
        if (int rc = upload(r_off, pt.data(), dpoff, pt.size() * sizeof(int64_t), "gram offsets")) return rc;
    }
    // the sweep's readback: rotation counters, then the sweep test flags
    const int32_t *hc = (const int32_t *)(pin.data() + r_back), *hk = hc + n_items;
    // buffer holding each problem's Gram matrix (0: P.g, 1: the workspace
    // copy); all problems of a sweep share it
    std::vector<int32_t> gpar(n_items, 0);
    int cur = 0;
    std::vector<int> idx;
    for (int i = 0; i < n_items; ++i)
        if (all[i].active) idx.push_back(i);
    int used = 0;
    // the table of a sweep's active problems and its owner table are built
    // again only when the active set changed (a sweep that ends with every
    // problem still going leaves both as they are)
    std::vector<int> built;
    int64_t ss = 0, ts = 0, max_np = 2;
    int n_act = 0;
    for (; used < max_sweeps && !idx.empty(); ++used) {
        if (idx != built) {
            std::vector<syn_gbj_item> act;
            ss = ts = 0;
            max_np = 2;
            for (int i : idx) {
                syn_gbj_item it = all[i];
                const TjDims D = tj_dims<TB>(it.nv);
                it.solve_start = ss;
                it.tile_start = ts;
                ss += D.half;
                ts += tj_ugroups<TB>(D);
                if (D.np > max_np) max_np = D.np;
                act.push_back(it);
            }
            if (int rc = upload(r_act, act.data(), items, act.size() * sizeof(syn_gbj_item), "gram active table"))
                return rc;
            n_act = (int)act.size();
            hipLaunchKernelGGL(tj_owner_kernel<TB>, dim3((unsigned)n_act), dim3(256), 0, st, items, ss, owner);
            SYN_LAUNCH_CHECK("gram owner table");
            built = idx;
        }
        hipLaunchKernelGGL(tj_zero_kernel, dim3(1), dim3(256), 0, st, counters, n_items);
        if constexpr (TB == 16) {
            for (int64_t step = 0; step < max_np; ++step) {
                // solves of step (all but the last launch) beside the
                // updates of step - 1 (all but the first)
                const bool hs = step < max_np - 1, hu = step > 0;
                const int64_t sp = hs ? (ss + 7) / 8 * 8 : 0, tp = hu ? (ts + 7) / 8 * 8 : 0;
                ProfScope ps(kTagGramStep, 0.0, st);
                hipLaunchKernelGGL(tjf_step_kernel<TB>, dim3((unsigned)(sp + tp)), dim3(256), 0, st, items, owner,
                                   owner + ss, step, hu ? 1 : 0, sp, hs ? ss : (int64_t)0, tp, ts, cur, tol, passes,
                                   kbytes);
                if (hu) cur ^= 1;
            }
        } else {
            // wide blocks: the solves of a step, then its updates, in place
            const int64_t sp = (ss + 7) / 8 * 8, tp = (ts + 7) / 8 * 8;
            for (int64_t step = 0; step < max_np - 1; ++step) {
                {
                    ProfScope ps(kTagGramStep, 0.0, st);
                    hipLaunchKernelGGL(tj_solve_kernel<TB>, dim3((unsigned)sp), dim3(TB * TB), 0, st, items, owner,
                                       step, sp, ss, tol);
                }
                ProfScope ps(kTagGramStep, 0.0, st);
                hipLaunchKernelGGL(tjw_update_kernel, dim3((unsigned)tp), dim3(256), 0, st, items, owner + ss, step,
                                   tp, ts, kbytes);
            }
        }
        for (int i : idx) gpar[i] = cur;
        SYN_LAUNCH_CHECK("gram step");
        // a problem goes on when its sweep rotated and the end-of-sweep test
        // finds an entry that would still rotate: this saves the last sweep,
        // which only confirmed convergence, with the same result
        hipLaunchKernelGGL(tj_zero_kernel, dim3(1), dim3(256), 0, st, chk, n_items);
        hipLaunchKernelGGL(tj_check_kernel<TB>, dim3((unsigned)n_act, 64), dim3(256), 0, st, items,
                           cur, tol, counters, chk);
        SYN_LAUNCH_CHECK("gram sweep test");
        if (int rc = hip_check(hipMemcpyAsync((void *)hc, counters, n * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                               "gram counters"))
            return rc;
        if (int rc = hip_check(hipMemcpyAsync((void *)hk, chk, n * sizeof(int32_t), hipMemcpyDeviceToHost, st),
                               "gram sweep test flags"))
            return rc;
        if (int rc = hip_check(hipStreamSynchronize(st), "gram sync")) return rc;
        std::vector<int> next;
        for (int i : idx)
            if (hc[i] != 0 && hk[i] != 0) next.push_back(i);
        idx.swap(next);
        if (idx.empty()) {
            ++used;
            break;
        }
    }
    if (sweeps) *sweeps = used;

    // eigenvalues, their decreasing order (ties by index) on the device, the
    // permutation of rt rows and s
    if (int rc = upload(r_all1, all.data(), items, all.size() * sizeof(syn_gbj_item), "gram table")) return rc;
    // the Gram matrix is diagonalised in place: bring back the ones the last
    // fused step left in the workspace copy
    if (int rc = upload(r_par, gpar.data(), counters, n * sizeof(int32_t), "gram parity")) return rc;
    hipLaunchKernelGGL(tjf_copyback_kernel<TB>, dim3(n_items, 64), dim3(256), 0, st, items, counters);
    SYN_LAUNCH_CHECK("gram copy back");
    hipLaunchKernelGGL(tj_diag_kernel, dim3(n_items), dim3(256), 0, st, items);
    SYN_LAUNCH_CHECK("gram diag");
    hipLaunchKernelGGL(tj_rank_kernel, dim3(n_items), dim3(256), 0, st, items, dpoff, dperm);
    SYN_LAUNCH_CHECK("gram order");
    // grid.y spreads each problem's rows over 64 work-groups
    hipLaunchKernelGGL(tj_permute_kernel, dim3(n_items, 64), dim3(256), 0, st, items, dperm, dpoff, work, dtoff, 0);
    hipLaunchKernelGGL(tj_permute_kernel, dim3(n_items, 64), dim3(256), 0, st, items, dperm, dpoff, work, dtoff, 1);
    SYN_LAUNCH_CHECK("gram permute");
    if (kbytes) {
        const unsigned long long *hb = (const unsigned long long *)(pin.data() + r_bytes);
        if (int rc = hip_check(hipMemcpyAsync((void *)hb, dbytes, n * sizeof(unsigned long long),
                                              hipMemcpyDeviceToHost, st),
                               "gram bytes"))
            return rc;
        if (int rc = hip_check(hipStreamSynchronize(st), "gram bytes sync")) return rc;
        double tot_b = 0.0;
        for (int i = 0; i < n_items; ++i) tot_b += (double)hb[i];
        prof_add_work(kTagGramBytes, tot_b);
    }
    return hip_check(hipStreamSynchronize(st), "gram final sync");
}

}  // namespace syn

extern "C" int64_t syn_jacobi_gram_blocked_workspace(const int64_t *nv, int n_items) {
    int64_t tot = 0;
    for (int i = 0; i < n_items; ++i)
        tot += std::max(syn::tj_work<16>(nv[i]), syn::tj_work<32>(nv[i])) + 2 * nv[i] +
               (std::max(syn::tj_owner_count<16>(nv[i]), syn::tj_owner_count<32>(nv[i])) + 1) / 2;
    return tot + 4 * (int64_t)n_items + 16;
}

extern "C" int syn_jacobi_gram_blocked_f64(syn_gbj_item *items, const syn_gbj_item *items_host, int n_items,
                                           double *work, int32_t *counters, double tol, int max_sweeps, int *sweeps,
                                           void *stream) {
    using namespace syn;
    if (n_items < 0 || !items || (n_items && (!items_host || !work || !counters)) || max_sweeps < 0 ||
        !(tol >= 0.0)) {
        set_error("syn_jacobi_gram_blocked_f64: bad arguments");
        return SYN_EINVAL;
    }
    if (n_items == 0) return SYN_OK;
    std::vector<syn_gbj_item> all(items_host, items_host + n_items);
    for (int i = 0; i < n_items; ++i) {
        if (all[i].nv <= 0 || all[i].ldg < all[i].nv || all[i].ldr < all[i].nv || !all[i].g || !all[i].rt ||
            !all[i].s) {
            set_error("syn_jacobi_gram_blocked_f64: bad problem %d (nv=%lld)", i, (long long)all[i].nv);
            return SYN_EINVAL;
        }
    }
    // many large problems: throughput, where the wide blocks halve the bytes
    // a sweep moves; small or few problems: latency, where the fused 16-row
    // steps overlap each step's solves with the previous step's updates.  The
    // two groups run one after the other in the same workspace (each pass
    // waits for its stream before it returns).
    std::vector<syn_gbj_item> wide, narrow;
    int64_t wide_entries = 0;
    for (const syn_gbj_item &it : all) {
        const bool w = it.nv >= kWideMinOrder;
        (w ? wide : narrow).push_back(it);
        if (w) wide_entries += it.nv * it.nv;
    }
    if (wide_entries < kWideMinEntries) {
        narrow.swap(all);
        wide.clear();
    }
    int sw_w = 0, sw_n = 0;
    if (!wide.empty())
        if (int rc = gram_blocked<32>(items, wide, (int)wide.size(), work, counters, tol, max_sweeps, &sw_w,
                                      (hipStream_t)stream))
            return rc;
    if (!narrow.empty())
        if (int rc = gram_blocked<16>(items, narrow, (int)narrow.size(), work, counters, tol, max_sweeps, &sw_n,
                                      (hipStream_t)stream))
            return rc;
    if (sweeps) *sweeps = std::max(sw_w, sw_n);
    return SYN_OK;
}
