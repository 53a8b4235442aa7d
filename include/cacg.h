/*
 * cacg.h — C-ABI of the MI355X (gfx950) s-step Chebyshev-basis CG library.
 *
 * Drop-in boundary: the two solver entry points of sclc/cacg, with the
 * reference's own argument types and meanings, plus an extended API with
 * status codes, options and a per-iteration report.  All pointers are plain
 * host pointers unless a name says `d_` (device memory, HBM).  All dense
 * blocks are row-major: B, X and R are n x m, element (i, c) at i*m + c
 * (src/GetRHS.c:9 of the reference).
 */
#ifndef CACG_H
#define CACG_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference types, byte-identical to src/DataTypes.h:14-61 ----------
 * Guarded by the reference header's own include guard, so a translation
 * unit that includes both sees one definition. */
#ifndef DATATYPES_H
#define DATATYPES_H
typedef struct {
    long num_rows;
    long num_cols;
    long nnz;
} matInfo;

typedef struct {
    long num_rows, num_cols, nnz; /* local rows, global columns, local nonzeros */
    long start;                   /* global id of the first local row */
    long *row_start;              /* num_rows + 1 offsets into col_idx / csrdata */
    long *col_idx;                /* global column ids, in the caller's order */
    double *csrdata;
} csrType_local;

typedef struct {
    long *rowIdx;    /* COO triplets as the reference's Matrix Market reader fills them */
    long *colIdx;
    double *coodata;
} cooType;

typedef struct {
    long local_num_row;
    long local_num_col;
    long global_num_row;
    long global_num_col;
    long start_idx;
    double *data;
} denseType;
#endif /* DATATYPES_H */

/* ---- drop-in entry points ------------------------------------------------
 * Replace src/cbcgKernel.h:21 and src/bcbcgKernel.h:22.  Same semantics as
 * the reference at one rank: X (n x m, row-major) is read as the initial
 * guess and overwritten with the solution; B and mat are not modified; the
 * solve stops when ||R(:,0)||_2 < epsilon (absolute, column 0 of the
 * recursively updated residual, src/bcbcgKernel.c:445-447, 524-526) or after
 * 100 (bcbcg_v1) / 1000 (cbcg_v1) outer iterations.  The Chebyshev interval
 * comes from Gerschgorin discs (src/common.c:53-106).
 *
 * The whole matrix must be local (myid == 0, numprocs == 1, mat.start == 0,
 * mat.num_rows == B.global_num_row): the multi-rank row-block distribution of
 * the reference becomes one matrix resident in one GPU's HBM.  Other values
 * are rejected with CACG_EINVAL.  The functions return nothing, as in the
 * reference; the outcome of the last call on the calling thread is
 * available from cacg_last_status() / cacg_last_iterations().  They never
 * call exit().  Arithmetic is FP64 in the reference's summation order
 * (CACG_ORDER_REFERENCE) unless the environment variable CACG_ORDER=fast
 * selects the split-K MFMA order. */
void cbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
             int numprocs);
void bcbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
              int numprocs);
/* src/bcgKernel.h:21 (src/bcgKernel.c:22-462): O'Leary block CG, 1000-iteration
 * cap (src/bcgKernel.c:2); prints the reference's "id:0, Loop: ..." lines. */
void bcg_v1(csrType_local mat, denseType B, denseType X, double epsilon, int myid, int numprocs);

/* ---- status codes ------------------------------------------------------- */
enum {
    CACG_OK = 0,          /* stop criterion met */
    CACG_MAXITER = 1,     /* iteration cap reached (the reference returns silently, src/bcbcgKernel.c:172) */
    CACG_BREAKDOWN = 2,   /* non-finite residual or zero / non-finite pivot in the coefficient solve */
    CACG_EINVAL = 3,      /* bad argument */
    CACG_ENOMEM = 4,      /* device allocation failed */
    CACG_EDEVICE = 5,     /* HIP runtime error or no usable gfx950 device */
    CACG_EIO = 6          /* a file could not be opened, read or written (include/cacg_io.h) */
};

int cacg_last_status(void);
long cacg_last_iterations(void);
const char *cacg_status_string(int status);

/* ---- extended API ------------------------------------------------------- */
enum { CACG_STOP_REF_COL0_ABS = 0, CACG_STOP_REL_PER_COLUMN = 1 };
/* CACG_ORDER_REFERENCE: every sum in the reference's order, products and sums
 * rounded separately (no FMA); results equal the reference at one rank bit
 * for bit.  The Gram sums then run as one sequential chain per entry, so this
 * order is for verification sizes.  CACG_ORDER_FAST: split-K FP64 MFMA Gram
 * and block products with a fixed (deterministic) reduction order. */
enum { CACG_ORDER_REFERENCE = 0, CACG_ORDER_FAST = 1 };
/* CACG_PREC_FP64: the reference's arithmetic, everything FP64.
 * CACG_PREC_FP32_REFINE: matrix values, basis S, Q and A Q stored in FP32 and
 * the Chebyshev / A Q SpMMs run in FP32 (half the bytes); Gram products, the
 * coefficient solves, X, R and the stop test stay FP64.  The true residual
 * R = B - A X is recomputed in FP64 (FP64 matrix) at refinement points (see
 * cacg_opts.inner_tol) and the stop test is on it.  No reference counterpart
 * (the reference is FP64 only); requires CACG_ORDER_FAST. */
enum { CACG_PREC_FP64 = 0, CACG_PREC_FP32_REFINE = 1 };
/* The Chebyshev interval [lo, hi] of the basis recurrence.
 * CACG_BOUNDS_GERSCHGORIN: the reference's Gerschgorin discs (src/common.c:53-106).
 * CACG_BOUNDS_LANCZOS: extreme Ritz values of lanczos_steps Lanczos steps on the
 * device (no reference counterpart), widened as cacg_lanczos_interval states;
 * computed once per solver and reused by its later solves.
 * CACG_BOUNDS_GIVEN: [bound_lo, bound_hi] from the caller. */
enum { CACG_BOUNDS_GERSCHGORIN = 0, CACG_BOUNDS_LANCZOS = 1, CACG_BOUNDS_GIVEN = 2 };

typedef struct {
    int stop_mode;       /* CACG_STOP_* */
    double eps;          /* absolute threshold on ||R(:,0)||_2, CACG_STOP_REF_COL0_ABS */
    double tol;          /* every column ||R(:,j)||_2 < tol * ||B(:,j)||_2, CACG_STOP_REL_PER_COLUMN */
    long maxiter;        /* outer-iteration cap, > 0 */
    int order;           /* CACG_ORDER_* */
    int precision;       /* CACG_PREC_* */
    int print_period;    /* > 0: print the reference's "Loop: %d, R[1]_norm=%30.30f" line every
                            print_period iterations, on stop and on the last iteration */
    /* CACG_PREC_FP32_REFINE only.  The true residual B - A X is recomputed in
     * FP64 when every column's recursive residual is below inner_tol times its
     * true residual at the previous recompute (or below the final target), or
     * after inner_maxiter iterations since it.  refine_restart = 0 (residual
     * replacement): the recomputed residual replaces the recursive one and the
     * recurrence continues with its Q, A Q and Gram factor; 1 (classic
     * iterative refinement): the s-step recurrence restarts from it.
     * 0 selects the defaults (inner_tol 1e-2 when replacing, 1e-5 when
     * restarting; no cap besides maxiter). */
    double inner_tol;
    long inner_maxiter;
    int refine_restart;
    int bounds;          /* CACG_BOUNDS_* (default GERSCHGORIN, the reference's) */
    int lanczos_steps;   /* CACG_BOUNDS_LANCZOS: 1..1000 steps; 0 selects 20 */
    double bound_lo, bound_hi;  /* CACG_BOUNDS_GIVEN: the interval, lo < hi */
    /* CACG_PREC_FP32_REFINE only: > 0: the basis is built in FP64 with the FP64
     * matrix (the solver then also holds an FP64 basis block) until every
     * column's true relative residual ||B - A X||/||B|| at a recompute is below
     * fp32_basis_below, in FP32 from then on (the recurrence continues).
     * 0: the FP32 basis throughout.  cacg_default_opts: 1e-3 (an FP32 basis in
     * the first three decades cost the 128^3 Laplacian with smooth right-hand
     * sides 8 of its 45 iterations; after them it costs none). */
    double fp32_basis_below;
} cacg_opts;

typedef struct {
    int status;          /* CACG_* */
    long iters;          /* outer iterations performed */
    double bounds[2];    /* Gerschgorin lower, upper */
    double *res_hist;    /* optional, caller-owned [res_hist_cap][m]: ||R(:,j)||_2 after each iteration */
    long res_hist_cap;
    double seconds;      /* device time of the solve (setup + iterations) */
    double phase_ms[12]; /* device time per phase, see CACG_PHASE_* */
    long phase_calls[12];
    long refinements;    /* CACG_PREC_FP32_REFINE: true-residual recomputes after the first */
    double true_relres;  /* CACG_PREC_FP32_REFINE: max over columns of ||B - A X||/||B|| at exit
                            (stop mode REL_PER_COLUMN); -1 when not computed */
} cacg_report;

/* Each phase is one kernel type (plus its small reduction kernel), so its time
 * divided by its launch count is that kernel's average launch duration. */
enum {
    CACG_PHASE_SETUP = 0,    /* Gerschgorin + R = B - A X */
    CACG_PHASE_BASIS = 1,    /* Chebyshev basis steps j >= 2 (s - 2 launches per iteration) */
    CACG_PHASE_GRAM = 2,     /* [Q^T A Q | Q^T R] */
    CACG_PHASE_SOLVE = 3,    /* small LU solves */
    CACG_PHASE_QUPDATE = 4,  /* Q = S - Q beta */
    CACG_PHASE_AQ = 5,       /* A Q SpMM */
    CACG_PHASE_XRUPDATE = 6, /* X += Q alpha, R -= A Q alpha, column norms */
    CACG_PHASE_OTHER = 7,    /* host gaps between iterations */
    CACG_PHASE_BASIS1 = 8,   /* Chebyshev basis step 1 (S_0 = R, S_1) */
    CACG_PHASE_GRAM_AS = 9,  /* Q^T A S */
    CACG_PHASE_REFINE = 10   /* true residual R = B - A X and its norms (CACG_PREC_FP32_REFINE) */
};

void cacg_default_opts(cacg_opts *o);

/* The interval CACG_BOUNDS_LANCZOS uses, from the Lanczos coefficients
 * (alpha_0..k-1, beta_0..k-1) and the Gerschgorin interval g: the extreme
 * eigenvalues t_min <= t_max of the tridiagonal T (Sturm bisection; T is cut
 * at the first negligible beta), widened to lo = t_min - 0.05 (t_max - t_min),
 * raised to t_min / 2 when t_min > 0 and lo is below that, and
 * hi = t_max + 0.05 (t_max - t_min) clipped at g[1], the Gerschgorin upper
 * bound.  Host code; returns CACG_EINVAL for
 * k <= 0 or non-finite coefficients. */
int cacg_lanczos_interval(const double *alpha, const double *beta, int k, const double g[2],
                          double out[2]);

/* Host-buffer solve: A from a csrType_local-like triple (row_start/col_idx of
 * n + 1 / nnz longs, values), B and X n x m row-major; X in/out.  s >= 1. */
int cacg_bcbcg(long n, const long *row_start, const long *col_idx, const double *val,
               long m, long s, const double *B, double *X, const cacg_opts *opts,
               cacg_report *report);

/* Host-buffer block CG (the reference's bcg_v1, src/bcgKernel.c:22-462):
 * CACG_PREC_FP64 only.  The stop test of iteration k reads the residual that
 * iteration started from, after X has been updated (src/bcgKernel.c:278-365),
 * and report->res_hist[k] holds those norms. */
int cacg_bcg(long n, const long *row_start, const long *col_idx, const double *val, long m,
             const double *B, double *X, const cacg_opts *opts, cacg_report *report);

/* Device-resident solver: the matrix is uploaded once (int32 indices when
 * nnz < 2^31) and the work blocks are allocated once for (m, s).  solve_device
 * takes device pointers to B and X (n x m row-major, X in/out) and runs on the
 * solver's own HIP stream. */
typedef struct cacg_solver cacg_solver;
int cacg_solver_create(cacg_solver **out, long n, const long *row_start, const long *col_idx,
                       const double *val, long m, long s, const cacg_opts *opts);
int cacg_solver_solve_device(cacg_solver *sv, const double *d_B, double *d_X, cacg_report *report);
/* Block CG (bcg_v1) with the solver's matrix and stream, for a solver created
 * with any s (it uses the leading n x m parts of the work blocks); phases:
 * AQ = A P, GRAM = P^T A P, GRAM_AS = R^T R, SOLVE = both LU solves,
 * XRUPDATE = X and R updates with the norms, QUPDATE = P = R + P beta. */
int cacg_solver_solve_bcg_device(cacg_solver *sv, const double *d_B, double *d_X,
                                 cacg_report *report);
int cacg_solver_destroy(cacg_solver *sv);
/* bytes of HBM the solver holds (matrix + work blocks) */
long cacg_solver_footprint(const cacg_solver *sv);

/* ---- the entry point that allocates nothing ------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the solve of cacg_solver_solve_device (FP64, the order
 * of opts->order, Gerschgorin or given bounds) with the matrix already in HBM
 * as int32 CSR (d_rowptr n + 1, d_col / d_val rowptr[n] entries) and B, X
 * device blocks (n x m row-major, X in/out), on `stream` (a hipStream_t; NULL
 * for the null stream).  Its work blocks live in a per-device scratch array
 * of CACG_NOALLOC_SCRATCH_BYTES (128 MiB, include/cacg_kernels.h) that the
 * library holds on every GPU it runs on until the process ends; it calls no
 * hipMalloc, hipFree or other allocating routine.  Limits: m <= 64 and
 * cacg_noalloc_bytes(n, m, s) <= that size (about n (3 s m + m) 8 bytes plus
 * the Gram partials), else CACG_EINVAL with a message; no SpMM schedule
 * tuning and no phase times (report->phase_ms stays zero).  Every call on a
 * GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * threads, other streams) race and must be serialised by the caller. */
#define CACG_NOALLOC_MAX_M 64
long cacg_noalloc_bytes(long n, long m, long s);
int cacg_bcbcg_device_noalloc(int n, const int *d_rowptr, const int *d_col, const double *d_val,
                              long m, long s, const double *d_B, double *d_X,
                              const cacg_opts *opts, cacg_report *report, void *stream);

/* ---- device memory helpers (for callers without their own HIP plumbing) -- */
int cacg_device_count(void);
int cacg_dev_alloc(void **d_ptr, long bytes);
int cacg_dev_free(void *d_ptr);
int cacg_h2d(void *d_dst, const void *src, long bytes);
int cacg_d2h(void *dst, const void *d_src, long bytes);
int cacg_d2d(void *d_dst, const void *d_src, long bytes);
int cacg_dev_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* CACG_H */
