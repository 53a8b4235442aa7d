/*
 * CPU oracle for the s-step Chebyshev-basis (block) CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's algorithm (sclc/cacg, src/bcbcgKernel.c, src/cbcgKernel.c and
 * the kernels they call), kept in the reference's loop and summation orders
 * so that, at one rank, its floating-point results are the reference's bit
 * for bit.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only to check results or as the timed CPU baseline.  The
 * product library (cacg_amd/libcacg.so) never calls into it.
 *
 * Every function cites the reference file:line it restates.  All dense
 * blocks are row-major (src/GetRHS.c:9, src/BLAS3.c:103-104).
 */
#ifndef CACG_ORACLE_H
#define CACG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* A block of CSR rows: row_start/col_idx/csrdata of csrType_local
 * (src/DataTypes.h:22-42).  row_offset = csrType_local.start, the global id
 * of the first row; col ids are global. */
typedef struct {
    long n;
    long row_offset;
    const long *rowptr;
    const long *col;
    const double *val;
} orc_csr;

enum { ORC_STOP_REF_COL0_ABS = 0, ORC_STOP_REL_PER_COLUMN = 1 };
enum { ORC_OK = 0, ORC_MAXITER = 1, ORC_BREAKDOWN = 2, ORC_INVAL = 3 };

typedef struct {
    int stop_mode;       /* ORC_STOP_* */
    double eps;          /* absolute threshold on ||R(:,0)||_2 (ORC_STOP_REF_COL0_ABS) */
    double tol;          /* relative threshold per column (ORC_STOP_REL_PER_COLUMN) */
    long maxiter;        /* 100 for bcbcg_v1, 1000 for cbcg_v1 (src/bcbcgKernel.c:3, src/cbcgKernel.c:2) */
    int nranks;          /* row-block ranks emulated by threads; 1 = the reference at one rank */
    /* optional outputs (NULL to skip) */
    double *res_hist;    /* [maxiter][m]: ||R(:,j)||_2 after each outer iteration */
    double *x_hist;      /* [maxiter][n][m]: X after each outer iteration */
    double *bounds;      /* [2]: Gerschgorin min, max */
} orc_opts;

/* src/common.c:53-106 */
void orc_gerschgorin(const orc_csr *A, double bounds[2]);
/* src/BLAS3.c:20-155 / 157-232 / 913-976: y = A x, x and y n x k */
void orc_spmm(const orc_csr *A, const double *x, long k, double *y);
/* src/common.c:182-256 (m > 1) and 109-179 (m = 1): S = [S_0 | ... | S_{s-1}] */
void orc_cheb_basis(const orc_csr *A, const double *R, long m, long s,
                    double s_alpha, double s_beta, double *S);
/* src/BLAS3.c:699-758: G (kl x kr) = L^T R over n rows */
void orc_gram(long n, const double *L, long kl, const double *R, long kr, double *G);
/* src/BLAS3.c:794-825: Y (n x j) = L (n x k) W (k x j) */
void orc_gemm_ts(long n, const double *L, long k, const double *W, long j, double *Y);
/* src/DLSsolvers.c:106-208: x (N x k) = lhs^{-1} rhs, Crout, unit U, no pivot */
void orc_lu_solve(long N, const double *lhs, long k, const double *rhs, double *x);
/* src/BLAS3.c:579-612: out = a x + b y */
void orc_lincomb2(long len, double a, const double *x, double b, const double *y, double *out);
/* src/MatrixOperations.c:110-142: sum_i R[i, col]^2 */
double orc_col_norm2sq(long n, long m, const double *R, long col);

/* src/bcbcgKernel.c:27-556 (m RHS) and src/cbcgKernel.c:23-250 (m = 1).
 * X is in/out.  Returns ORC_*; *iters = outer iterations performed. */
int orc_bcbcg(const orc_csr *A, long m, long s, const double *B, double *X,
              const orc_opts *o, long *iters);

/* src/bcgKernel.c:22-462, O'Leary block CG (m RHS):
 *   R = B - A X, P = R, RtR = R^T R; per iteration
 *   AP = A P, PtAP = P^T AP, alpha = PtAP^{-1} RtR, X += P alpha,
 *   stop test on ||R(:,0)|| of the R the iteration started from,
 *   R -= AP alpha, beta = RtR^{-1} (R^T R), RtR = R^T R, P = R + P beta.
 * res_hist[it] holds ||R(:,c)|| of that starting R (what the stop test reads);
 * x_hist[it] the X after the iteration's update.  Returns ORC_*; *iters =
 * iterations performed (the last one updates X and stops). */
int orc_bcg(const orc_csr *A, long m, const double *B, double *X, const orc_opts *o, long *iters);

/* Wall-clock seconds of a bounded sample: setup (bounds, R = B - AX, work
 * allocation) and each of `iters` outer iterations (t_iters[iters]), no stop
 * test.  For the CPU baseline only. */
int orc_bcbcg_time_sample(const orc_csr *A, long m, long s, const double *B, double *X,
                          int nranks, long iters, double *t_setup, double *t_iters);

#ifdef __cplusplus
}
#endif
#endif
