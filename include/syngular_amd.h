/*
 * syngular_amd.h -- C ABI of the MI355X (gfx950) matrix-product core.
 *
 * The reference (antoine311200/Syngular) is pure Python on NumPy/opt_einsum;
 * every function below replaces one NumPy/LAPACK/opt_einsum call site on the
 * path of syngular.tensor.MatrixProductState / MatrixProductOperator.  The
 * reference file:line each entry point replaces is given beside it.
 *
 * Conventions
 *   - All matrices are dense, row-major, addressed by a base pointer and a row
 *     stride ("ld", in elements).  A core (l, d, r) is the row-major array
 *     l*d*r; its left unfolding is the (l*d) x r matrix with ld = r, its right
 *     unfolding the l x (d*r) matrix with ld = d*r.
 *   - Every batched entry point takes a DEVICE pointer to an array of item
 *     descriptors (the caller owns that memory), the number of items and the
 *     total number of work-groups the items need (the caller fills each item's
 *     `block_start` with the running sum, in item order).  One launch serves one
 *     problem or thousands of independent problems of uneven shape.
 *   - Pointers inside items are device pointers.  Nothing here allocates device
 *     memory except where an entry point says so; callers pass workspaces.
 *   - Every function returns 0 on success or a negative SYN_E* code; the
 *     message of the last failure on the calling thread is syn_last_error().
 *   - `stream` is a hipStream_t passed as void* (0 = the null stream).
 */
#ifndef SYNGULAR_AMD_H
#define SYNGULAR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SYN_OK 0
#define SYN_EINVAL (-1)   /* bad argument / shape (the reference raises ValueError/Exception) */
#define SYN_EHIP (-2)     /* HIP runtime error */
#define SYN_ENOTSUP (-3)  /* unsupported dtype / configuration */

/* ------------------------------------------------------------------ */
/* library                                                              */
/* ------------------------------------------------------------------ */
const char *syn_last_error(void);
int syn_version(void);
/* number of gfx950 devices visible; fails with SYN_EHIP when none */
int syn_device_count(int *count);
/* ABI self-check: writes sizeof() of the item structs below, in declaration
 * order (gemm, apply, mpo_mpo, dsum, qr, jacobi, trunc, rowscale, bjac,
 * eye, copy, diagmax, chol, gbj, lusign, colscale, vecnorm), into out[0..n-1];
 * returns the number of structs. */
int syn_struct_sizes(int64_t *out, int n);

/* ------------------------------------------------------------------ */
/* GEMM  C = alpha * op(A) * op(B) + beta * C  (grouped, MFMA)           */
/* Replaces the absorb products  S @ R, B @ R, L @ U.T                   */
/*   tensor/matrix_product_state.py:383,415,448,561,575                  */
/*   tensor/matrix_product_operator.py:493,526,560,682,705               */
/* and the Gram products L.T @ L, R @ R.T (matrix_product_state.py:626,630;*/
/* matrix_product_operator.py:773,777).                                  */
/* op(A) is m x k: A[i*lda+l] (transA=0) or A[l*lda+i] (transA=1).       */
/* op(B) is k x n: B[l*ldb+j] (transB=0) or B[j*ldb+l] (transB=1).       */
/* block_start counts tiles of the tile configuration `tile` (0: 64x64,  */
/* 1: 128x128); tiles_n = ceil(n / tile width).                          */
/* ------------------------------------------------------------------ */
/* tri: structural-zero hints (bit mask), letting a tile skip work:       */
/*   1  only the lower triangle of C is needed (tiles above it are skipped)*/
/*   2  op(A) is lower triangular    4  op(A) is upper triangular          */
/*   8  op(B) is upper triangular   16  op(B) is lower triangular          */
typedef struct {
    const void *A;
    const void *B;
    void *C;
    int64_t m, n, k;
    int64_t lda, ldb, ldc;
    double alpha, beta;
    int32_t transA, transB;
    int64_t block_start;
    int64_t tiles_n;
    int32_t tri;
} syn_gemm_item;

/* f64 x f64 -> f64 on v_mfma_f64_16x16x4 */
int syn_gemm_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile, void *stream);
/* The mixed form of the same call: FP64 operands and result, products on   */
/* the FP32 matrix cores (operands rounded to FP32, FP32 accumulation).     */
int syn_gemm_mixed_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile, void *stream);
/* The BF16 form: FP64 operands and result, operands rounded to bfloat16   */
/* (round to nearest even) when staged, v_mfma_f32_16x16x16_bf16 with FP32   */
/* accumulation: the contractions of the precision form, whose error is    */
/* stated against the FP64 oracle (a relative ~1e-3 per product).           */
int syn_gemm_bf16_f64(const syn_gemm_item *items, int n_items, int64_t total_blocks, int tile, void *stream);

/* ------------------------------------------------------------------ */
/* MPO x MPS site contraction, materialised                              */
/*   P[(a*Dl+c), o, (b*Dr+e)] = sum_s X[a,s,b] * O[c,s,o,e]               */
/* Replaces tensor/matrix_product_operator.py:184-190 and                */
/* tensor/utils.py:35-41,49-55 (contract + reshape).                    */
/* One work-group covers SYN_APPLY_BLOCK output elements.                */
/* ------------------------------------------------------------------ */
#define SYN_APPLY_BLOCK 1024
typedef struct {
    const double *x;  /* (chil, din, chir) */
    const double *o;  /* (dl, din, dout, dr) */
    double *p;        /* (chil*dl, dout, chir*dr) */
    int64_t chil, din, chir;
    int64_t dl, dout, dr;
    int64_t block_start;
} syn_apply_item;
int syn_mpo_apply_f64(const syn_apply_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* MPO x MPO site contraction                                            */
/*   P[(a*Cl+c), i, p, (b*Cr+e)] = sum_j A[a,i,j,b] * B[c,j,p,e]          */
/* Replaces tensor/matrix_product_operator.py:280-287 and                */
/* tensor/utils.py:62-69.                                                */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *a;  /* (al, din, dmid, ar) */
    const double *b;  /* (bl, dmid, dout, br) */
    double *p;        /* (al*bl, din, dout, ar*br) */
    int64_t al, din, dmid, ar;
    int64_t bl, dout, br;
    int64_t block_start;
} syn_mpo_mpo_item;
int syn_mpo_mpo_f64(const syn_mpo_mpo_item *items, int n_items, int64_t total_blocks, void *stream);

/* Bond-leg Kronecker product for every physical index pair (MPO Hadamard  */
/* product): P[(a*bl+c), i, p, (b*br+e)] = A[a,i,p,b] * B[c,i,p,e], with   */
/* A (al, din, dout, ar), B (bl, din, dout, br); the item's dmid is unused.*/
/* Replaces the np.kron loop of tensor/matrix_product_operator.py:136-154.*/
int syn_kron_bonds_f64(const syn_mpo_mpo_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* Direct-sum core packing (one output core per item)                   */
/*   edge==1 (first site):  out[:, m, :] = [A | B]                       */
/*   edge==2 (last site):   out[:, m, :] = [A ; B]                       */
/*   edge==0 (interior):    out[:, m, :] = blockdiag(A, B)               */
/* `mid` is d for an MPS core and d_in*d_out for an MPO core.            */
/* Replaces tensor/matrix_product_state.py:82-96 and                    */
/* tensor/matrix_product_operator.py:90-106.                             */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *a;  /* (la, mid, ra) */
    const double *b;  /* (lb, mid, rb) */
    double *out;      /* (lo, mid, ro) */
    int64_t la, ra, lb, rb, mid, lo, ro;
    int32_t edge;
    int64_t block_start;
} syn_dsum_item;
int syn_direct_sum_f64(const syn_dsum_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* Householder QR, one work-group per item.                              */
/* Factors the leading nb = min(n, keep) columns of the m x n matrix A   */
/* (element (i,j) at A[i*rs + j*cs]) and writes the first kq = min(keep, m)*/
/* columns of the complete orthogonal factor Q (m x kq, row-major, ldq). */
/* The caller forms R = Q^T A with syn_gemm_f64.  Reflectors follow      */
/* Golub & Van Loan 5.1.2 with beta = -sign(alpha) * ||x||, the LAPACK    */
/* convention, so Q agrees with numpy.linalg.qr up to rounding.          */
/* work: m*nb doubles (m*(nb+kq) when q_trans); tau: nb doubles.        */
/* Replaces numpy.linalg.qr at tensor/matrix_product_state.py:307,332,  */
/* 443,560,574 and tensor/matrix_product_operator.py:436,555,680,703.     */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *a;
    int64_t rs, cs;
    int64_t m, n, keep;
    double *q;
    int64_t ldq;
    int32_t q_trans;  /* 1: store Q^T (kq x m, row-major, ldq) instead of Q */
    double *work;
    double *tau;
} syn_qr_item;
int syn_qr_f64(const syn_qr_item *items, int n_items, void *stream);
/* The same with the launch's dynamic LDS sized by the caller: lds_doubles  */
/* (at most 12288, the 96 KB of syn_qr_f64); a matrix whose leading block  */
/* m * min(n, keep) fits works in LDS, a larger one in its workspace.  A   */
/* batch of small matrices launched with a small size keeps more           */
/* work-groups on each compute unit.                                        */
int syn_qr_sized_f64(const syn_qr_item *items, int n_items, int64_t lds_doubles, void *stream);

/* Panel of the blocked Householder QR (one work-group of 1024 threads). */
/* Factors columns [j0, j0+b) of the m-row matrix w (row-major, ldw;     */
/* rows j0..m-1, the columns left of j0 already factored and the panel   */
/* already updated by them) in place with the reflectors of syn_qr_f64   */
/* (LAPACK convention): R on and above the diagonal, the scaled vectors  */
/* below it; tau[j0..j0+b) the factors; v (row-major, ldv) receives in    */
/* rows j0..m-1 of columns [j0, j0+b) the explicit V (unit diagonal,      */
/* zeros above); t (b x b, ldt) the upper-triangular T of the compact WY  */
/* form H_j0 ... H_{j0+b-1} = I - V T V^T.  b <= 32.  The caller updates  */
/* the columns right of the panel and accumulates Q with syn_gemm_f64.   */
/* Replaces numpy.linalg.qr on large matrices at the same call sites as  */
/* syn_qr_f64.                                                            */
int syn_hh_panel_f64(double *w, int64_t ldw, double *v, int64_t ldv, double *tau, double *t, int64_t ldt,
                     int64_t m, int64_t j0, int64_t b, void *stream);

/* ------------------------------------------------------------------ */
/* Householder signs of a CholeskyQR factor (the fast path of the QR   */
/* above for tall, well-conditioned matrices).  If A = Q R with R's     */
/* diagonal positive (CholeskyQR2 on the MFMA GEMM), the Householder QR */
/* of A in the LAPACK convention is (Q D, D R), D = diag(+-1): the      */
/* compact-WY form Q_h - [I; 0] = -Y T Y1^T is an LU factorisation, so */
/* D follows from the LU without pivoting of Q[:n, :n] - D in which     */
/* each D_jj = -sign(c_j), c_j the diagonal entry of the Schur          */
/* complement when column j is reached (the pivot c_j - D_jj = -tau_j  */
/* D_jj has magnitude in [1, 2]).  Ballard et al., "Reconstructing     */
/* Householder vectors from tall-skinny QR", JPDC 85 (2015).            */
/* LAPACK does not reflect a column whose part below the diagonal is   */
/* exactly zero (tau = 0, D_jj = sign(c_j)); that is always the case   */
/* for row m-1 of an m-row matrix (free_col), and otherwise shows as    */
/* |c_j| = 1, so |c_j| > amb sets *flag and the caller falls back to    */
/* syn_qr_f64.                                                          */
/* syn_lu_sign_leaf_f64 factors diagonal blocks of at most 64 (one      */
/* work-group per item): a (bs x bs, lda) is read, d[0..bs) receives D, */
/* linv / uinv (bs x bs, ldl / ldu; skipped when linv is NULL) the      */
/* inverses of the unit lower L and of U.  The caller updates the rest  */
/* of the matrix with syn_gemm_f64 (right-looking blocked LU).          */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *a;
    int64_t lda, bs;
    int64_t free_col; /* column (within the block) that LAPACK cannot reflect, or -1 */
    double *d;
    double *linv;
    int64_t ldl;
    double *uinv;
    int64_t ldu;
    int32_t *flag;
} syn_lusign_item;
int syn_lu_sign_leaf_f64(const syn_lusign_item *items, int n_items, double amb, void *stream);

/* Column scaling with optional transpose: per item dst = src diag(s)    */
/* (rows x cols, trans = 0) or its transpose (cols x rows, trans = 1);  */
/* s = NULL copies without scaling.  Row-major with leading dimensions  */
/* lds / ldd.  64 x 64 tiles through LDS; block_start is the prefix sum */
/* of the tiles of the items before.                                    */
typedef struct {
    const double *src;
    int64_t lds;
    double *dst;
    int64_t ldd;
    int64_t rows, cols;
    const double *s;
    int32_t trans;
    int64_t block_start;
} syn_colscale_item;
int syn_colscale_f64(const syn_colscale_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* One-sided Jacobi SVD (Hestenes), one work-group per item.             */
/* x holds nv vectors of length len as rows (x[i*ldx + t]).  The rows are */
/* rotated pairwise until mutually orthogonal; the same rotations are    */
/* applied to rt (nv x nv, initialised to I by the kernel).  On exit the */
/* rows are sorted by decreasing norm, s[i] = ||x_i||, and               */
/*     x_in = rt^T * x_out   with   x_out rows orthogonal.               */
/* With gram = 1 the input is a symmetric Gram matrix G = C^T C and      */
/* s[i] = sqrt(||x_i||) are the singular values of C (rt = V^T).         */
/* No reference call site (an addition: the SVD form of the rounding).  */
/* ------------------------------------------------------------------ */
typedef struct {
    double *x;
    int64_t ldx;
    int64_t nv, len;
    double *rt;
    int64_t ldr;
    double *s;
    int32_t *sweeps;  /* out: sweeps used (may be null) */
    int32_t gram;
} syn_jacobi_item;
int syn_jacobi_f64(const syn_jacobi_item *items, int n_items, double tol, int max_sweeps, void *stream);

/* Two-sided Jacobi eigen-decomposition of symmetric Gram matrices held  */
/* in LDS, one work-group per item (x: nv x nv with len == nv, read only; */
/* nv <= nv_max <= 64 for every item).  On exit rt rows are the           */
/* eigenvectors (rt = V^T, G = V diag(lambda) V^T) sorted by decreasing    */
/* eigenvalue and s[i] = sqrt(max(lambda_i, 0)), the singular values of C */
/* when G = C^T C.  Items outside the contract get s = NaN.  No reference */
/* call site (an addition: the SVD form of the rounding, small bonds).  */
int syn_jacobi_gram_f64(const syn_jacobi_item *items, int n_items, int64_t nv_max, double tol, int max_sweeps,
                        void *stream);

/* ------------------------------------------------------------------ */
/* Block one-sided Jacobi for large problems: same contract as            */
/* syn_jacobi_f64 for one problem (x: nv x len rows, rt: nv x nv, s: nv),  */
/* but every step spreads the disjoint 16-row block pairs over many        */
/* work-groups.  The host loop synchronises the stream once per sweep.    */
/* work: syn_jacobi_blocked_workspace(nv, len) doubles.  *sweeps (host    */
/* pointer, may be null) receives the sweeps used.                        */
/* ------------------------------------------------------------------ */
int64_t syn_jacobi_blocked_workspace(int64_t nv, int64_t len);
int syn_jacobi_blocked_f64(double *x, int64_t nv, int64_t len, int64_t ldx, double *rt, int64_t ldr, double *s,
                           double *work, double tol, int max_sweeps, int gram, int *sweeps, void *stream);

/* Batched block Jacobi: one call orthogonalises the rows of many        */
/* problems of uneven size; every step launch covers the block pairs of  */
/* all problems still rotating (per-problem counters retire the          */
/* converged ones after each sweep).  `items` is a DEVICE table and      */
/* `items_host` the same table in host memory (the host loop reads the   */
/* sizes from it); block_start and counter/active are filled by the     */
/* library.  Same per-problem contract as syn_jacobi_f64.  work:          */
/* syn_jacobi_blocked_many_workspace(...) doubles; counters: n_items ints.*/
typedef struct {
    double *x;
    int64_t ldx, nv, len;
    double *rt;
    int64_t ldr;
    double *s;
    int64_t block_start;   /* first work-group of this problem in a step */
    int32_t *counter;      /* rotations in the current sweep */
    int32_t active;        /* 0 once converged */
    int32_t gram;
} syn_bjac_item;
int syn_jacobi_blocked_many_f64(syn_bjac_item *items, const syn_bjac_item *items_host, int n_items, double *work,
                                int32_t *counters, double tol, int max_sweeps, int *sweeps, void *stream);
/* work size in doubles for problems (nv[i], len[i]) */
int64_t syn_jacobi_blocked_many_workspace(const int64_t *nv, const int64_t *len, int n_items);

/* Two-sided block Jacobi eigen-decomposition of many symmetric Gram     */
/* matrices of uneven size (nv > 16), for the right sweep of the SVD      */
/* rounding.  Per problem: g (nv x nv, ldg) is diagonalised IN PLACE;      */
/* on exit rt (nv x nv, ldr) holds the eigenvectors as rows sorted by     */
/* decreasing eigenvalue (G_in = rt^T diag(lambda) rt) and s[i] =          */
/* sqrt(max(lambda_i, 0)).  Blocks of 16 rows: every step is one launch   */
/* holding its pair solves (LDS, one work-group each) beside the previous */
/* step's MFMA tile updates, for all problems still rotating.  The        */
/* problems of order 384 and more take blocks of 32 rows (a solve launch  */
/* and an update launch per step, half the steps) when together they hold */
/* at least 2^23 Gram entries; they run first, the rest after them.  The  */
/* 16-row pair solves make two passes of their rounds (while they rotate) */
/* when the call's 16-row problems hold at least 2^23 entries, else one.  */
/* A problem stops after a sweep that rotated nothing or when no          */
/* off-diagonal entry passes the rotation test |G_pq| > tol sqrt(G_pp     */
/* G_qq) at the end of a sweep (the next sweep would rotate nothing).     */
/* `items` is a DEVICE table of n_items entries and `items_host` the same */
/* table on the host; w, solve_start, tile_start, counter and active are  */
/* filled by the library.  work: syn_jacobi_gram_blocked_workspace(nv, n) */
/* doubles (enough for either block size); counters: n_items ints;        */
/* *sweeps (host, may be null) gets the most sweeps any problem took.     */
/* No reference call site (an addition: the SVD form of the rounding).  */
typedef struct {
    double *g;
    int64_t ldg;
    double *rt;
    int64_t ldr;
    double *s;
    double *w;
    int64_t nv;
    int64_t solve_start, tile_start;
    int32_t *counter;
    int32_t active;
} syn_gbj_item;
int syn_jacobi_gram_blocked_f64(syn_gbj_item *items, const syn_gbj_item *items_host, int n_items, double *work,
                                int32_t *counters, double tol, int max_sweeps, int *sweeps, void *stream);
int64_t syn_jacobi_gram_blocked_workspace(const int64_t *nv, int n_items);

/* Batched copy: dst[0..n) = src[0..n) for every item (compaction). */
typedef struct {
    const double *src;
    double *dst;
    int64_t n;
    int64_t block_start;
} syn_copy_item;
int syn_copy_f64(const syn_copy_item *items, int n_items, int64_t total_blocks, void *stream);

/* Batched identity: x (n x n, ld) = I for every item. */
typedef struct {
    double *x;
    int64_t n, ld;
    int64_t block_start;
} syn_eye_item;
int syn_eye_f64(const syn_eye_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* Cholesky / CholeskyQR (left-orthonormalisation of large unfoldings)   */
/* syn_potrf_f64: G = L L^T (lower) by blocks of 64; G is destroyed, L is */
/* written in full (zeros above the diagonal).  A pivot below            */
/* tolrel * max_i G_ii is dropped (zero column of L); *dropped (device   */
/* int) counts them.  work: syn_cholqr_workspace(n) doubles suffice.     */
/* syn_cholqr_f64: one CholeskyQR pass on the m x n matrix A (kept):     */
/* G = A^T A, G = L L^T, Q = A L^{-T} (Q: m x n, ldq).  Two passes give   */
/* the reduced QR A = Q (L1 L2)^T with Q orthonormal to working          */
/* precision.  Replaces numpy.linalg.qr on large left unfoldings          */
/* (tensor/matrix_product_state.py:560, the left sweep of the rounding). */
/* ------------------------------------------------------------------ */
int64_t syn_cholqr_workspace(int64_t n);

/* syn_cholqr_refine_f64: the second pass of CholeskyQR2 on Q1 (m x n,   */
/* the first pass's Q, kept).  G = Q1^T Q1 = I + E; when max |E_ij| <=    */
/* thresh, L = I + F and its inverse I - F with F = strict_lower(E) +     */
/* diag(E)/2 (exact to O(n max|E|^2); no Cholesky), else the recursive    */
/* Cholesky with the pivot-drop rule (tolrel).  Q = Q1 L^{-T} (m x n, ldq) */
/* and L (n x n, lower) are written; *full (host, may be null) = 1 when   */
/* the Cholesky ran.  work: syn_cholqr_workspace(n) doubles.               */
/* mixed (both calls): the product Q = A L^{-T} on the FP32 matrix cores   */
/* (the Gram matrix and the Cholesky stay FP64).                          */
int syn_cholqr_refine_f64(const double *q1, int64_t m, int64_t n, int64_t ldq1, double *q, int64_t ldq, double *l,
                          int64_t ldl, double *work, double thresh, double tolrel, int *dropped, int *full,
                          int mixed, void *stream);

/* Batched pieces of the same factorisation, for many problems of uneven  */
/* size (the caller drives the recursion with syn_gemm_f64):              */
/* syn_diagmax_f64: *out = max_i g[i*ld+i] and *dropped = 0, per item.    */
/* syn_chol_leaf_f64: per item, factor the n x n (n <= 64) block g        */
/* (lower triangle read) as L L^T with the pivot-drop rule above           */
/* (threshold tolrel * *maxdiag), write L (zeros above the diagonal) and  */
/* X = inv(L~) with dropped columns zeroed (X is n x n, ld ldx), and add  */
/* the dropped pivots to *dropped.  One work-group per item.              */
typedef struct {
    const double *g;
    int64_t n, ld;
    double *out;
    int32_t *dropped;
} syn_diagmax_item;
int syn_diagmax_f64(const syn_diagmax_item *items, int n_items, void *stream);

typedef struct {
    const double *g;
    int64_t ldg;
    double *l;
    int64_t ldl;
    double *x;
    int64_t ldx;
    int64_t n;
    const double *maxdiag;
    double tolrel;
    int32_t *dropped;
} syn_chol_item;
int syn_chol_leaf_f64(const syn_chol_item *items, int n_items, void *stream);
int syn_potrf_f64(double *g, int64_t n, int64_t ldg, double *l, int64_t ldl, double *work, double tolrel,
                  int *dropped, void *stream);
int syn_cholqr_f64(double *a, int64_t m, int64_t n, int64_t lda, double *q, int64_t ldq, double *l, int64_t ldl,
                   double *work, double tolrel, int *dropped, int mixed, void *stream);
/* Allocation-free form of syn_cholqr_f64 (no counterpart in the reference:  */
/* it is for callers that cannot allocate device memory -- code under HIP     */
/* graph capture, tight serving loops, harnesses that forbid hipMalloc).      */
/* Same arguments without `work`: the workspace is a 48 MiB device array of   */
/* the library, one per GPU, held until the process ends, so n is limited to */
/* syn_cholqr_noalloc_max_n() (1078 columns; m is not limited) and larger n  */
/* fails with SYN_EINVAL.  Nothing is allocated and no memset is issued.     */
/* Every call on one GPU shares that array: calls that overlap on one GPU    */
/* (other streams or host threads) race and must be serialised by the caller.*/
int syn_cholqr_noalloc_f64(double *a, int64_t m, int64_t n, int64_t lda, double *q, int64_t ldq, double *l,
                           int64_t ldl, double tolrel, int *dropped, int mixed, void *stream);
int64_t syn_cholqr_noalloc_max_n(void);

/* ------------------------------------------------------------------ */
/* Device-side truncation: for sorted singular values s[0..n-1] keep the */
/* smallest c with sum_{i>=c} s_i^2 <= cutoff * sum_i s_i^2 (cutoff>0),  */
/* then c = min(c, chi_max) (chi_max>0), c >= 1.  Writes c and the       */
/* discarded weight sum_{i>=c} s_i^2 (accumulated in FP64).              */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *s;
    int64_t n;
    int64_t chi_max;
    double cutoff;
    int32_t *keep;
    double *discarded;
} syn_trunc_item;
int syn_truncate_f64(const syn_trunc_item *items, int n_items, void *stream);

/* ------------------------------------------------------------------ */
/* Vector norms: out[j] = ||x[:, j]|| (axis 0, j < cols) or              */
/* out[i] = ||x[i, :]|| (axis 1, i < rows) of row-major matrices, summed */
/* in FP64 in a fixed order.  block_start counts ceil(cols / 64) blocks   */
/* per item for axis 0 and ceil(rows / 4) for axis 1.  The rounding      */
/* sweep reads the retained singular values off C V = U S (tall C) or    */
/* U^T C = S V^T (wide C); no reference call site (the reference has no  */
/* SVD rounding; the values stand beside numpy.linalg.svd's).            */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *x;
    int64_t rows, cols, ld;
    int32_t axis;
    double *out;
    int64_t block_start;
} syn_vecnorm_item;
int syn_vecnorms_f64(const syn_vecnorm_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* Row scaling  X[i, :] *= s[i]  (mode 0)  or  X[i, :] /= s[i]  (mode 1; */
/* rows with s[i] == 0 are zeroed).  Used to turn the Jacobi factors into */
/* U*S and S^-1*U^T*C in the rounding sweep.                             */
/* ------------------------------------------------------------------ */
typedef struct {
    double *x;
    int64_t rows, cols, ld;
    const double *s;
    int32_t mode;
    int64_t block_start;
} syn_rowscale_item;
int syn_rowscale_f64(const syn_rowscale_item *items, int n_items, int64_t total_blocks, void *stream);

/* ------------------------------------------------------------------ */
/* Live per-kernel timing (HIP events on the launch stream).             */
/* Between syn_prof_begin and syn_prof_end every launch of the heavy     */
/* kernels is bracketed by events; syn_prof_end fills, per kernel tag,   */
/* launches, milliseconds and algorithmic flops (arrays of ntags) and    */
/* returns the number of tags.  Tags: 0/1 GEMM by value 64/128 tiles,     */
/* 2/3 grouped GEMM 64/128 tiles, 4 block-Jacobi step, 5 Cholesky        */
/* diagonal block, 6 Householder QR, 7 LDS Jacobi, 8 MPO x MPS apply,    */
/* 9/10 Gram block-Jacobi pair solve / tile update, 11/12 mixed GEMM       */
/* 64/128 tiles, 13/14 split-K GEMM of 128/64 tiles (CholeskyQR Gram       */
/* matrices and the recursive Cholesky's products, partial-sum pass        */
/* included), 15/16 two independent GEMMs in one launch of 64/128 tiles   */
/* (the recursive Cholesky's G22 update beside T = L21 X11), 17 one      */
/* recursive-Cholesky node of 64 < n <= 128 (two leaves, three products), */
/* 18/19 BF16 GEMM 64/128 tiles.                                          */
/* syn_prof_note_flops sets the count attached to the next grouped GEMM */
/* or apply launch (their tables live on the device).                    */
/* syn_prof_select(mask) limits the events to the tags whose bit is set  */
/* (all by default; kept until changed): two event packets per launch    */
/* slow a stream of short kernels down by microseconds each.             */
/* ------------------------------------------------------------------ */
#define SYN_PROF_NTAGS 20
int syn_prof_begin(void);
int syn_prof_select(uint32_t mask);
int syn_prof_note_flops(double flops);
int syn_prof_end(int64_t *launches, double *ms, double *flops, int ntags);

#ifdef __cplusplus
}
#endif
#endif /* SYNGULAR_AMD_H */
