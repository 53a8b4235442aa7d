/*
 * cacg_kernels.h — kernel-level C-ABI: the launchers the host driver calls, one per
 * hot-path kernel, exported for tests and benchmarks.
 *
 * Every function enqueues work on `st` and returns 0 or a hipError_t value.
 * Pointers named d_* and every matrix/block argument are device pointers.
 * Dense blocks are row-major with an explicit row stride (ld, in elements),
 * so a column block of a wider array (S_j inside S = [S_0 | ... | S_{s-1}])
 * is addressed as (base + j*m, ld = s*m).
 */
#ifndef CACG_KERNELS_H
#define CACG_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* device CSR, int32 row pointers and column ids (nnz < 2^31).
 * Long rows: with long_min > 0, the SpMMs leave every row with at least
 * long_min nonzeros to a row-split kernel (one workgroup per row), and
 * long_rows (device, n_long ids) must list exactly those rows
 * (cacg_csr_long_rows builds it).  long_min = 0 (zero-initialised): every row
 * vector-per-row. */
typedef struct {
    int n;
    const int *rowptr;
    const int *col;
    const double *val;
    const int *long_rows;
    int n_long;
    int long_min;
} cacg_dcsr;

/* the same pattern with FP32 values (the FP32-basis precision form) */
typedef struct {
    int n;
    const int *rowptr;
    const int *col;
    const float *val;
    const int *long_rows;
    int n_long;
    int long_min;
} cacg_scsr;

/* The row length from which a row of this matrix goes to the row-split
 * kernel: max(256, 16 x the mean row length), so rows far above the mean.
 * Host-side, from the host row pointers (n + 1 of them, int64). */
int cacg_csr_long_min(long n, const long *row_start);
/* The rows with at least long_min nonzeros, ascending, into out (host; may be
 * NULL to count).  Returns their number. */
long cacg_csr_long_rows(long n, const long *row_start, int long_min, int *out);

enum { CACG_MODE_REFERENCE = 0, CACG_MODE_FAST = 1 };

/* Bytes of the per-device scratch array of cacg_bcbcg_device_noalloc. */
#define CACG_NOALLOC_SCRATCH_BYTES (128L << 20)
/* The calling thread's current device's copy of that array: its device
 * address (looked up once per device) and size.  Returns 0 or a hipError_t. */
int cacg_noalloc_scratch(void **d_ptr, long *bytes);

/* Gerschgorin interval (src/common.c:53-106): d_out[0] = min_i (a_ii - sum|a_ij|),
 * d_out[1] = max_i (a_ii + sum|a_ij|), the max seeded with DBL_MIN.
 * d_work holds 2 * CACG_GERSH_BLOCKS doubles. */
#define CACG_GERSH_BLOCKS 1024
int cacg_k_gerschgorin(const cacg_dcsr *A, double *d_work, double *d_out, hipStream_t st);

/* y(:, 0:k) = epilogue(A x(:, 0:k)), every sum in CSR order with separately
 * rounded products and sums (src/BLAS3.c:78-112):
 *   CACG_EPI_NONE : y = A x
 *   CACG_EPI_RESID: y = p1 - A x                        (R = B - A X)
 *   CACG_EPI_CHEB1: y = a (A x) + b x, and copy0 = x    (S_1 and S_0 = R)
 *   CACG_EPI_CHEB : y = (a (A x) + b p1) - p2            (S_j, p1 = S_{j-1}, p2 = S_{j-2}) */
enum { CACG_EPI_NONE = 0, CACG_EPI_RESID = 1, CACG_EPI_CHEB1 = 2, CACG_EPI_CHEB = 3 };
int cacg_k_spmm(const cacg_dcsr *A, int k, const double *x, long ldx, double *y, long ldy, int epi,
                double a, double b, const double *p1, long ld1, const double *p2, long ld2,
                double *copy0, long ldc, hipStream_t st);

/* FP32 forms for the FP32 basis, same CSR order with FP32 products and sums
 * (no reference counterpart: the reference is FP64 only):
 *   cacg_k_spmm_f32 : epi CACG_EPI_CHEB (S_j) or CACG_EPI_NONE (y = A x), all FP32;
 *   cacg_k_cheb1_f32: s0 = fl32(r), s1 = a (A s0) + b s0, r the FP64 residual. */
int cacg_k_spmm_f32(const cacg_scsr *A, int k, const float *x, long ldx, float *y, long ldy,
                    int epi, double a, double b, const float *p1, long ld1, const float *p2,
                    long ld2, hipStream_t st);
int cacg_k_cheb1_f32(const cacg_scsr *A, int k, const double *r, long ldr, float *s1, long ld1,
                     float *s0, long ld0, double a, double b, hipStream_t st);

/* Launch schedule of the SpMMs above (no effect on any result: each output is
 * summed by one lane group in CSR order).  grid: workgroups (0: 8192 for
 * k <= 32, 4096 above); chunk_rows: 0 deals each XCD a contiguous band of row
 * tiles, > 0 deals chunks of that many rows round-robin over the XCDs (an XCD
 * then owns a slice of every grid plane of a stencil matrix); passes (with
 * chunk_rows > 0; 0 or 1: one): each XCD sweeps its chunks in that many passes,
 * pass h taking its chunks h, h + passes, ... (a narrow column of the grid,
 * plane by plane; taken for k > 32, ignored for k = 4..32, whose kernel
 * sweeps one pass).  The _x forms take one (NULL: the defaults); the forms
 * above use the defaults. */
typedef struct {
    int grid;
    int chunk_rows;
    int passes;
} cacg_spmm_sched;
int cacg_k_spmm_x(const cacg_dcsr *A, int k, const double *x, long ldx, double *y, long ldy,
                  int epi, double a, double b, const double *p1, long ld1, const double *p2,
                  long ld2, double *copy0, long ldc, const cacg_spmm_sched *sc, hipStream_t st);
int cacg_k_spmm_f32_x(const cacg_scsr *A, int k, const float *x, long ldx, float *y, long ldy,
                      int epi, double a, double b, const float *p1, long ld1, const float *p2,
                      long ld2, const cacg_spmm_sched *sc, hipStream_t st);
int cacg_k_cheb1_f32_x(const cacg_scsr *A, int k, const double *r, long ldr, float *s1, long ld1,
                       float *s0, long ld0, double a, double b, const cacg_spmm_sched *sc,
                       hipStream_t st);
/* dst[e] = fl32(src[e]), e < count (the FP32 copy of the matrix values) */
int cacg_k_cvt_f64_f32(long count, const double *src, float *dst, hipStream_t st);
/* dst[e] = src[e] widened to FP64 (Q = S of a fresh FP32 basis) */
int cacg_k_cvt_f32_f64(long count, const float *src, double *dst, hipStream_t st);

/* [G1 | G2] = L^T [R1 | R2] over rows 0..n-1: G1 is kl x kr1 and G2 is kl x kr2,
 * both row-major and contiguous (G2 may be NULL when kr2 == 0).
 * Reference order: each entry is one sequential chain over the rows, products
 * and sums rounded separately (src/BLAS3.c:709-717, 741-749).
 * Fast order: split-K over row chunks with FP64 MFMA, then a fixed-order sum of
 * the chunk partials; d_work holds cacg_gram_workspace(kl, kr1 + kr2) doubles. */
int cacg_k_gram_ref(int n, const double *L, long ldl, int kl, const double *R1, long ld1, int kr1,
                    const double *R2, long ld2, int kr2, double *G1, double *G2, hipStream_t st);
long cacg_gram_workspace(int kl, int kr);
/* the doubles of d_work a Gram of n rows actually uses (<= cacg_gram_workspace) */
long cacg_gram_workspace_n(int n, int kl, int kr);
int cacg_k_gram_fast(int n, const double *L, long ldl, int kl, const double *R1, long ld1, int kr1,
                     const double *R2, long ld2, int kr2, double *G1, double *G2, double *d_work,
                     hipStream_t st);
/* The fast order with R1 block-major (column c of row i at
 * R1[(c / bw) * bs + i * bw + c % bw]: the basis S stored as its blocks S_j,
 * n x bw each, bs elements apart) and no R2; kr1 a multiple of bw.  Only the
 * shapes of the LDS-staged kernels (16-byte aligned rows: bw, bs even, R1
 * 16-byte aligned); others return hipErrorInvalidValue.  Same bits as
 * cacg_k_gram_fast on the row-major copy. */
int cacg_k_gram_fast_blk(int n, const double *L, long ldl, int kl, const double *R1, int bw,
                         long bs, int kr1, double *G1, double *d_work, hipStream_t st);
/* 1 when cacg_k_gram_fast_blk takes this shape (ldl = kl, bs = n bw, 16-byte
 * aligned L and R1), else 0 */
int cacg_gram_blk_supported(int n, int kl, int kr1, int bw);
/* The fast order with FP64 L and FP32 R1 (Q^T A S = AQ^T S against an FP32
 * basis): R1 widened to FP64 on load, FP64 products and accumulation. */
int cacg_k_gram_fast_f64f32(int n, const double *L, long ldl, int kl, const float *R1, long ld1,
                            int kr1, const double *R2, long ld2, int kr2, double *G1, double *G2,
                            double *d_work, hipStream_t st);

/* x (N x k) = lhs^{-1} rhs with the reference's Crout factorisation (unit-diagonal
 * U, no pivoting) and substitutions (src/DLSsolvers.c:106-208).  lhs is not
 * modified.  mode REFERENCE keeps the reference's ascending backward-substitution
 * order (bitwise equal); mode FAST runs the backward substitution column-oriented.
 * d_work (N*N doubles) receives the factor; with reuse_factor != 0 the factor
 * already in d_work (from a previous call on the same lhs) is used instead of
 * refactoring lhs, which gives the same bits.  *d_flag is set to 1 on a zero or
 * non-finite pivot. */
int cacg_k_lu_solve(int N, const double *lhs, int k, const double *rhs, double *x, double *d_work,
                    int *d_flag, int mode, int reuse_factor, hipStream_t st);

/* Q = S - Q beta, in place on Q (n x sm, row-major, ld = sm), beta sm x sm
 * (src/BLAS3.c:794-825 then 579-612). */
int cacg_k_qupdate(int n, int sm, const double *S, double *Q, const double *beta, int mode,
                   hipStream_t st);
/* Fast order with the X update deferred by one iteration, so Q is read once
 * per iteration (no reference counterpart; results are the bits of the
 * undeferred fast order), for cacg_k_split_supported(sm, m): sm a multiple of
 * 16 up to 128 and m <= 32.
 *   cacg_k_qxupdate      : X += Q alpha with the old Q, then Q = S - Q beta;
 *   cacg_k_xrupdate_split: which 1: X += Q alpha only (the last iteration's);
 *                          which 2: R -= AQ alpha and ||R(:,c)||^2 only. */
int cacg_k_split_supported(int sm, int m);
int cacg_k_qxupdate(int n, int sm, int m, const double *S, double *Q, const double *beta,
                    const double *alpha, double *X, hipStream_t st);
/* cacg_k_qxupdate with S block-major (blocks of m columns, sbs >= n m elements
 * apart); same bits */
int cacg_k_qxupdate_blk(int n, int sm, int m, const double *S, long sbs, double *Q,
                        const double *beta, const double *alpha, double *X, hipStream_t st);
/* the same with the FP32 basis S (the precision form) */
int cacg_k_qxupdate_s32(int n, int sm, int m, const float *S, double *Q, const double *beta,
                        const double *alpha, double *X, hipStream_t st);
int cacg_k_xrupdate_split(int n, int sm, int m, const double *Q, const double *AQ,
                          const double *alpha, double *X, double *R, double *d_norm2,
                          double *d_work, int which, hipStream_t st);
/* Q = -Q beta in place (S = 0; fast order), the tall-skinny product of the QR */
int cacg_k_qupdate_zero(int n, int sm, double *Q, const double *beta, hipStream_t st);
/* FP32 basis S, FP64 Q (the precision form's update): Q = S - Q beta in FP64. */
int cacg_k_qupdate_s32(int n, int sm, const float *S, double *Q, const double *beta,
                       hipStream_t st);

/* X += Q alpha; R -= AQ alpha (Q, AQ n x sm; alpha sm x m; X, R n x m) and
 * d_norm2[c] = sum_i R[i][c]^2 of the updated R.  Reference mode keeps the
 * reference's orders (the norms as one sequential chain per column,
 * src/MatrixOperations.c:110-123); fast mode uses MFMA products and a
 * fixed-order two-level norm reduction.  d_work: cacg_xrupdate_workspace(m). */
long cacg_xrupdate_workspace(int m);
int cacg_k_xrupdate(int n, int sm, int m, const double *Q, const double *AQ, const double *alpha,
                    double *X, double *R, double *d_norm2, double *d_work, int mode,
                    hipStream_t st);

/* d_norm2[c] = sum_i R[i][c]^2, one sequential chain per column (reference order) */
int cacg_k_colnorm2_ref(int n, int m, const double *R, long ldr, double *d_norm2, hipStream_t st);
/* d_norm2[c] = sum_i R[i][c]^2 of a contiguous n x m block, fixed-order
 * two-level reduction (fast order); d_work: cacg_xrupdate_workspace(m). */
int cacg_k_colnorm2_fast(int n, int m, const double *R, double *d_norm2, double *d_work,
                         hipStream_t st);

/* d[0:count] = v, as a kernel (safe inside captured graphs, unlike a memset node) */
int cacg_k_fill_i32(int *d, long count, int v, hipStream_t st);
int cacg_k_fill_f64(double *d, long count, double v, hipStream_t st);
/* Lanczos coefficients for the Chebyshev interval: `steps` steps of the
 * symmetric Lanczos recurrence on A from a fixed start vector (no reference
 * counterpart; the reference uses Gerschgorin discs).  d_ab receives
 * alpha_0..steps-1 then beta_0..steps-1; d_work holds
 * cacg_lanczos_workspace(n) doubles. */
long cacg_lanczos_workspace(long n);
int cacg_k_lanczos(const cacg_dcsr *A, int steps, double *d_work, double *d_ab, hipStream_t st);

/* d[0:count] *= a in place (a = -1: exact negation) */
int cacg_k_scale_f64(double *d, long count, double a, hipStream_t st);

/* Tall-skinny QR (no reference counterpart: the reference's TSQR.c is empty):
 * A (n x k, row stride lda, k <= 128) = Q R with Q (n x k, contiguous) having
 * orthonormal columns and R (k x k, row-major) upper triangular with a
 * positive diagonal.  CholeskyQR2: two passes of G = Q^T Q (split-K FP64 MFMA,
 * deterministic), single-workgroup LDS Cholesky and triangular inverse, and
 * Q = Q R^{-1}; orthogonality O(u) for cond(A) up to ~1e8.  *d_flag is set if
 * a Cholesky pivot falls below 64 u of its diagonal entry (A numerically rank
 * deficient or beyond that conditioning).  d_work
 * holds cacg_tsqr_workspace(k) doubles. */
long cacg_tsqr_workspace(int k);
int cacg_k_tsqr(int n, int k, const double *A, long lda, double *Q, double *R, double *d_work,
                int *d_flag, hipStream_t st);

/* ---- batched kernels (many independent systems per launch) ----------------
 * One descriptor per system slot, in device memory.  Work items are int2
 * {slot, index}: a 64-row tile (CACG_BT), a 2048-row Gram chunk (CACG_GR) or
 * a 256-entry block of a Gram result, depending on the kernel.  Every
 * per-system sum runs in an order fixed by the system alone, so a system's
 * results do not depend on which other systems share the launch.
 * No reference counterpart: the batched / serving form is new. */
#define CACG_BT 64
#define CACG_GR 2048
#define CACG_BATCH_MAX_M 32
#define CACG_BATCH_MAX_SM 128
typedef struct {
    int n, m, sm, ntile, nchunk, pad_;
    const int *rowptr, *col;
    const double *val;
    const float *val32;   /* FP32 basis only */
    const double *B;
    double *X, *R;        /* n x m */
    void *S;              /* n x sm basis: float (FP32 basis) or double */
    double *Q, *AQ;       /* n x sm */
    double *QtAQ, *QtAS, *beta, *lu;  /* sm x sm */
    double *QtR, *alpha;  /* sm x m */
    double *gpart;        /* nchunk x sm x (sm + m) Gram partials */
    double *npart;        /* ntile x m norm partials */
    double *norm2;        /* m: ||R(:,c)||^2 of the last true residual */
    int *flag;            /* set on a zero / non-finite pivot */
    double sa, sb;        /* Chebyshev interval map */
} cacg_bsys;

/* basis step j of the systems' tiles: j == 1 writes S_0 = R (rounded to the
 * basis type) and S_1, j >= 2 writes S_j; f32 selects the FP32 basis (FP32
 * values and arithmetic).  s == 1 systems take j == 0: S_0 = R only. */
int cacg_kb_basis(const cacg_bsys *d_sys, const int2 *d_tiles, int count, int j, int f32,
                  hipStream_t st);
/* Q = S widened to FP64 (first iteration after admission) */
int cacg_kb_q_from_s(const cacg_bsys *d_sys, const int2 *d_tiles, int count, int f32,
                     hipStream_t st);
/* AQ = A Q (FP64) */
int cacg_kb_aq(const cacg_bsys *d_sys, const int2 *d_tiles, int count, hipStream_t st);
/* Gram partials per 2048-row chunk with FP64 MFMA for the systems of one tile
 * class, tm = ceil(sm / 16) in 1 .. 8 and tn = ceil(kr / 16) in tm .. tm + 2,
 * kr = sm (which 0: QtAS = AQ^T S) or sm + m (which 1: [QtAQ | QtR] =
 * Q^T [AQ | R]); then the chunk-ordered sums by cacg_kb_gram_reduce ({slot,
 * 256-entry block} items).  Every admissible system (m <= 32, sm <= 128) is
 * in some class. */
int cacg_kb_gram_class(const cacg_bsys *d_sys, const int2 *d_chunks, int nchunks, int which,
                       int f32, int tm, int tn, hipStream_t st);
int cacg_kb_gram_reduce(const cacg_bsys *d_sys, const int2 *d_blocks, int nblocks, int which,
                        hipStream_t st);
/* which 0: beta = QtAQ^{-1} QtAS (factor from the last which-1 call); which 1:
 * alpha = QtAQ^{-1} QtR (factors QtAQ into lu); max_sm bounds every sm */
int cacg_k_lu_batch(const cacg_bsys *d_sys, const int2 *d_items, int count, int which, int max_sm,
                    hipStream_t st);
/* Q = S - Q beta with FP64 MFMA over 2048-row chunks ({slot, chunk} items) of
 * systems with smp = 16 ceil(sm / 16) (16 .. 128); exact != 0 when every listed
 * system has sm == smp (16-byte loads), else zero-padded guarded loads */
int cacg_kb_qupdate_class(const cacg_bsys *d_sys, const int2 *d_chunks, int count, int smp,
                          int exact, int f32, hipStream_t st);
/* X += Q alpha with FP64 MFMA over 2048-row chunks of systems with
 * smp = 16 ceil(sm / 16) and exact as for cacg_kb_qupdate_class; m selects the
 * instantiation (m <= 16, or 16 < m <= 32) and every listed system must be in
 * that m class */
int cacg_kb_xupdate_class(const cacg_bsys *d_sys, const int2 *d_chunks, int count, int smp,
                          int exact, int m, hipStream_t st);
/* R = B - A X (FP64) with per-tile column norm partials, then norm2 of every
 * listed system (d_sysitems: {slot, 0}) */
int cacg_kb_resid(const cacg_bsys *d_sys, const int2 *d_tiles, int count,
                  const int2 *d_sysitems, int nsys, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
