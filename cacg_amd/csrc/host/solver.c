/*
 * Host driver of the s-step Chebyshev-basis (block) CG solver on one gfx950
 * GPU.  C11 over the HIP runtime's C API and the kernel launchers of
 * include/cacg_kernels.h.
 *
 * Also the reference's plain block CG, bcg_v1 (src/bcgKernel.c:22-462), on the
 * same kernels (cacg_solver_solve_bcg_device).
 *
 * One outer iteration (the reference's src/bcbcgKernel.c:172-526):
 *   S = [S_0 | ... | S_{s-1}], Chebyshev basis of R     s - 1 SpMMs, fused recurrence
 *   k == 0: Q = S (buffer swap)                           src/bcbcgKernel.c:199-203
 *   k >= 1: QtAS = AQ^T S; beta = QtAQ^{-1} QtAS;
 *           Q = S - Q beta                                :214, :229, :249-262
 *   AQ = A Q                                              :284
 *   [QtAQ | QtR] = Q^T [AQ | R]                           :307, :330 (one pass over Q)
 *   alpha = QtAQ^{-1} QtR                                 :344
 *   X += Q alpha; R -= AQ alpha; ||R(:,c)||^2             :376-447
 * (fast order: X += Q alpha is deferred into the next iteration's Q update,
 * which reads the same Q rows anyway, and applied once after the last one.)
 * The m column norms come back to the host once per iteration (one small
 * device-to-host copy, the only synchronisation in the loop) for the stop test.
 *
 * Device layout (HBM): A as CSR with int32 row pointers / column ids and
 * FP64 values; R n x m; S, Q, AQ n x sm row-major (sm = s*m), column block j of
 * S at offset j*m of every row, exactly the reference's basis layout
 * (src/MatrixOperations.c:70-92 output).  B and X are the caller's n x m
 * row-major blocks.
 *
 * Precision form (CACG_PREC_FP32_REFINE): an FP32 copy of the values shares
 * the index arrays and drives the Chebyshev SpMMs; the basis S is FP32 (same
 * layout).  Q = S - Q beta, A Q (FP64 matrix), R, X, the Gram blocks and the
 * coefficients stay FP64: Q is formed with cancellation and stored in FP32 it
 * slows the recurrence's convergence (measured).  R = B - A X is recomputed
 * in FP64 at refinement points and replaces the recursive residual.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "cacg_kernels.h"
#include "cacg.h"

#define NPHASE 12
#define MAXEV 20

struct cacg_solver {
    long n, nnz, m, s, sm;
    cacg_opts opts;
    hipStream_t st;
    cacg_dcsr A;
    int *d_rowptr, *d_col;
    int *d_long_rows;   /* rows with at least A.long_min nonzeros (row-split SpMM) */
    double *d_val;
    double *d_R, *d_S, *d_Q, *d_AQ;
    /* precision form: FP32 values and basis */
    cacg_scsr A32;
    float *f_val, *f_S;
    double *d_QtAQ, *d_QtAS, *d_QtR, *d_alpha, *d_beta;
    double *d_norm2, *d_bounds, *d_gwork, *d_gram_work, *d_xr_work, *d_lu_work;
    int *d_flag;
    double *h_pin;   /* pinned: [0, m) column norms^2, [m, 2m) ||B(:,c)||^2, [2m, 2m+2) bounds, then the flag */
    /* SpMM schedules of the basis steps and of A Q, picked by timing on the
     * first solve (tuned = 1) */
    cacg_spmm_sched sch_basis, sch_aq;
    int tuned;
    unsigned long long pattern_key;   /* key of the process-wide schedule cache */
    hipEvent_t tev[2];
    /* 1: keep the fast order's basis row-major (CACG_BASIS_ROWS=1 at creation) */
    int basis_rows;
    /* CACG_BOUNDS_LANCZOS: work space, coefficients, the interval once computed */
    double *d_lz_work, *d_lz_ab, *h_lz;
    int lz_steps, have_lz;
    double lz_iv[2];
    hipEvent_t ev[MAXEV];
    int ev_phase[MAXEV];
    int nev;
    int noev;   /* 1: no phase events (the solver of cacg_bcbcg_device_noalloc) */
    long footprint;
};

static _Thread_local int g_last_status = CACG_OK;
static _Thread_local long g_last_iters = 0;

int cacg_last_status(void) { return g_last_status; }
long cacg_last_iterations(void) { return g_last_iters; }

const char *cacg_status_string(int st)
{
    switch (st) {
    case CACG_OK: return "converged";
    case CACG_MAXITER: return "iteration cap reached";
    case CACG_BREAKDOWN: return "breakdown (zero/non-finite pivot or non-finite residual)";
    case CACG_EINVAL: return "invalid argument";
    case CACG_ENOMEM: return "device allocation failed";
    case CACG_EDEVICE: return "HIP runtime error or no gfx950 device";
    case CACG_EIO: return "file could not be opened, read or written";
    default: return "unknown status";
    }
}

void cacg_default_opts(cacg_opts *o)
{
    memset(o, 0, sizeof(*o));
    o->stop_mode = CACG_STOP_REL_PER_COLUMN;
    o->eps = 0.0;
    o->tol = 1e-8;
    o->maxiter = 100;
    o->order = CACG_ORDER_FAST;
    o->precision = CACG_PREC_FP64;
    o->print_period = 0;
    o->inner_tol = 0.0;
    o->inner_maxiter = 0;
    o->refine_restart = 0;
    o->bounds = CACG_BOUNDS_GERSCHGORIN;
    o->lanczos_steps = 20;
    o->bound_lo = 0.0;
    o->bound_hi = 0.0;
    o->fp32_basis_below = 1e-3;
}

/* number of eigenvalues of the tridiagonal (a, b) below x (Sturm sequence of
 * the LDL^T pivots) */
static int sturm_below(const double *a, const double *b, int k, double x)
{
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < k; i++) {
        d = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
        if (d == 0.0)
            d = -1e-300;
        if (d < 0.0)
            c++;
    }
    return c;
}

/* the idx-th smallest eigenvalue of the tridiagonal by bisection in [lo, hi] */
static double tridiag_eig(const double *a, const double *b, int k, int idx, double lo, double hi)
{
    for (int it = 0; it < 200; it++) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi)
            break;
        if (sturm_below(a, b, k, mid) > idx)
            hi = mid;
        else
            lo = mid;
    }
    return 0.5 * (lo + hi);
}

int cacg_lanczos_interval(const double *alpha, const double *beta, int k, const double g[2],
                          double out[2])
{
    if (!alpha || !beta || !g || !out || k <= 0)
        return CACG_EINVAL;
    /* cut T at the first negligible beta: the Krylov space is invariant there */
    int kk = k;
    for (int j = 0; j < k; j++) {
        if (!isfinite(alpha[j]) || !isfinite(beta[j]))
            return CACG_EINVAL;
        if (j + 1 < k &&
            !(fabs(beta[j]) > 1e-13 * (fabs(alpha[j]) + (j > 0 ? fabs(beta[j - 1]) : 0.0)))) {
            kk = j + 1;
            break;
        }
    }
    double lo = alpha[0], hi = alpha[0];
    for (int i = 0; i < kk; i++) {
        const double r = (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i + 1 < kk ? fabs(beta[i]) : 0.0);
        lo = fmin(lo, alpha[i] - r);
        hi = fmax(hi, alpha[i] + r);
    }
    const double pad = 1e-12 * (fabs(lo) + fabs(hi)) + 1e-300;
    const double tmin = tridiag_eig(alpha, beta, kk, 0, lo - pad, hi + pad);
    const double tmax = tridiag_eig(alpha, beta, kk, kk - 1, lo - pad, hi + pad);
    const double w = tmax - tmin;
    double a = tmin - 0.05 * w, b = tmax + 0.05 * w;
    /* a positive spectrum keeps a positive lower end: not below tmin / 2 (for
     * tmin <= 0 that clip would lift the end above tmin, so it is not applied) */
    if (tmin > 0.0 && a < 0.5 * tmin)
        a = 0.5 * tmin;
    if (b > g[1] && g[1] > tmax)
        b = g[1];
    if (!(b > a)) { /* a one-point spectrum: any interval around it */
        a = tmin - 0.5 * fabs(tmin) - 1.0;
        b = tmax + 0.5 * fabs(tmax) + 1.0;
    }
    out[0] = a;
    out[1] = b;
    return CACG_OK;
}

#define HIPCHK(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "cacg: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_),       \
                    __FILE__, __LINE__);                                                         \
            return CACG_EDEVICE;                                                                 \
        }                                                                                        \
    } while (0)

#define KCHK(call)                                                                               \
    do {                                                                                         \
        int e_ = (call);                                                                         \
        if (e_ != 0) {                                                                           \
            fprintf(stderr, "cacg: kernel launch %s failed: %s (%s:%d)\n", #call,                \
                    hipGetErrorString((hipError_t)e_), __FILE__, __LINE__);                      \
            return CACG_EDEVICE;                                                                 \
        }                                                                                        \
    } while (0)

/* ---- device helpers ---------------------------------------------------------- */

static int check_device(void)
{
    static int checked = 0, ok = 0;
    if (checked)
        return ok ? CACG_OK : CACG_EDEVICE;
    int dev = 0, cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) {
        fprintf(stderr, "cacg: no HIP device visible\n");
        checked = 1;
        return CACG_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        fprintf(stderr, "cacg: cannot query the HIP device\n");
        checked = 1;
        return CACG_EDEVICE;
    }
    ok = strncmp(prop.gcnArchName, "gfx950", 6) == 0;
    if (!ok)
        fprintf(stderr, "cacg: device %d is %s; this library is built for gfx950 only\n", dev,
                prop.gcnArchName);
    checked = 1;
    return ok ? CACG_OK : CACG_EDEVICE;
}

int cacg_device_count(void)
{
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess)
        return 0;
    return cnt;
}

int cacg_dev_alloc(void **d_ptr, long bytes)
{
    if (!d_ptr || bytes < 0)
        return CACG_EINVAL;
    *d_ptr = NULL;
    if (bytes == 0)
        return CACG_OK;
    return hipMalloc(d_ptr, (size_t)bytes) == hipSuccess ? CACG_OK : CACG_ENOMEM;
}

int cacg_dev_free(void *d_ptr)
{
    if (!d_ptr)
        return CACG_OK;
    return hipFree(d_ptr) == hipSuccess ? CACG_OK : CACG_EDEVICE;
}

int cacg_h2d(void *d_dst, const void *src, long bytes)
{
    if (bytes <= 0)
        return CACG_OK;
    return hipMemcpy(d_dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? CACG_OK
                                                                                     : CACG_EDEVICE;
}

int cacg_d2h(void *dst, const void *d_src, long bytes)
{
    if (bytes <= 0)
        return CACG_OK;
    return hipMemcpy(dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? CACG_OK
                                                                                     : CACG_EDEVICE;
}

int cacg_d2d(void *d_dst, const void *d_src, long bytes)
{
    if (bytes <= 0)
        return CACG_OK;
    return hipMemcpy(d_dst, d_src, (size_t)bytes, hipMemcpyDeviceToDevice) == hipSuccess
               ? CACG_OK
               : CACG_EDEVICE;
}

int cacg_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? CACG_OK : CACG_EDEVICE; }

/* ---- solver lifetime ----------------------------------------------------------- */

static int dalloc(cacg_solver *sv, void **p, long bytes)
{
    if (bytes <= 0)
        bytes = 8;
    if (hipMalloc(p, (size_t)bytes) != hipSuccess) {
        *p = NULL;
        return CACG_ENOMEM;
    }
    sv->footprint += bytes;
    return CACG_OK;
}

int cacg_solver_destroy(cacg_solver *sv)
{
    if (!sv)
        return CACG_OK;
    void *ptrs[] = {sv->d_rowptr, sv->d_col,   sv->d_val,   sv->d_R,       sv->d_S,
                    sv->d_Q,      sv->d_AQ,    sv->d_QtAQ,  sv->d_QtAS,    sv->d_QtR,
                    sv->d_alpha,  sv->d_beta,  sv->d_norm2, sv->d_bounds,  sv->d_gwork,
                    sv->d_gram_work, sv->d_xr_work, sv->d_lu_work, sv->d_flag,
                    sv->f_val,    sv->f_S,     sv->d_lz_work, sv->d_lz_ab, sv->d_long_rows};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
        if (ptrs[i])
            hipFree(ptrs[i]);
    for (int i = 0; i < MAXEV; i++)
        if (sv->ev[i])
            hipEventDestroy(sv->ev[i]);
    for (int i = 0; i < 2; i++)
        if (sv->tev[i])
            hipEventDestroy(sv->tev[i]);
    if (sv->st)
        hipStreamDestroy(sv->st);
    if (sv->h_pin)
        hipHostFree(sv->h_pin);
    free(sv->h_lz);
    free(sv);
    return CACG_OK;
}

long cacg_solver_footprint(const cacg_solver *sv) { return sv ? sv->footprint : 0; }

static int validate_csr(long n, const long *rp, const long *ci)
{
    if (n <= 0 || n >= (1L << 31) || !rp || rp[0] != 0)
        return CACG_EINVAL;
    const long nnz = rp[n];
    if (nnz < 0 || nnz >= (1L << 31) || (nnz > 0 && !ci))
        return CACG_EINVAL;
    for (long i = 0; i < n; i++)
        if (rp[i + 1] < rp[i])
            return CACG_EINVAL;
    for (long p = 0; p < nnz; p++)
        if (ci[p] < 0 || ci[p] >= n)
            return CACG_EINVAL;
    return CACG_OK;
}

static unsigned long long pattern_key(long n, const long *row_start, const long *col_idx, long m,
                                      long sm, int prec32);

int cacg_solver_create(cacg_solver **out, long n, const long *row_start, const long *col_idx,
                       const double *val, long m, long s, const cacg_opts *opts)
{
    if (!out)
        return CACG_EINVAL;
    *out = NULL;
    if (m <= 0 || s <= 0 || (long)s * m > 4096 || !val)
        return CACG_EINVAL;
    if (validate_csr(n, row_start, col_idx) != CACG_OK)
        return CACG_EINVAL;
    if (opts) {
        const int p32 = opts->precision == CACG_PREC_FP32_REFINE;
        if (opts->maxiter <= 0 || (opts->precision != CACG_PREC_FP64 && !p32) ||
            (p32 && opts->order != CACG_ORDER_FAST) || opts->inner_tol < 0.0 ||
            opts->inner_maxiter < 0 || opts->tol < 0.0 ||
            (opts->refine_restart != 0 && opts->refine_restart != 1) || opts->bounds < 0 ||
            opts->bounds > CACG_BOUNDS_GIVEN || opts->lanczos_steps < 0 ||
            opts->lanczos_steps > 1000 || !(opts->fp32_basis_below >= 0.0) ||
            (opts->bounds == CACG_BOUNDS_GIVEN &&
             !(isfinite(opts->bound_lo) && isfinite(opts->bound_hi) &&
               opts->bound_lo < opts->bound_hi)))
            return CACG_EINVAL;
    }
    int rc = check_device();
    if (rc != CACG_OK)
        return rc;
    cacg_solver *sv = (cacg_solver *)calloc(1, sizeof(*sv));
    if (!sv)
        return CACG_ENOMEM;
    sv->n = n;
    sv->nnz = row_start[n];
    sv->m = m;
    sv->s = s;
    sv->sm = s * m;
    if (opts)
        sv->opts = *opts;
    else
        cacg_default_opts(&sv->opts);
    const int prec32 = sv->opts.precision == CACG_PREC_FP32_REFINE;
    sv->pattern_key = pattern_key(n, row_start, col_idx, m, sv->sm, prec32);
    if (hipHostMalloc((void **)&sv->h_pin, sizeof(double) * (size_t)(2 * m + 4), 0) != hipSuccess) {
        sv->h_pin = NULL;
        cacg_solver_destroy(sv);
        return CACG_ENOMEM;
    }
    if (hipStreamCreate(&sv->st) != hipSuccess) {
        cacg_solver_destroy(sv);
        return CACG_EDEVICE;
    }
    for (int i = 0; i < MAXEV; i++)
        if (hipEventCreate(&sv->ev[i]) != hipSuccess) {
            cacg_solver_destroy(sv);
            return CACG_EDEVICE;
        }
    for (int i = 0; i < 2; i++)
        if (hipEventCreate(&sv->tev[i]) != hipSuccess) {
            cacg_solver_destroy(sv);
            return CACG_EDEVICE;
        }
    {
        const char *br = getenv("CACG_BASIS_ROWS");
        sv->basis_rows = br && br[0] == '1';
    }
    const long nm = n * m, nsm = n * sv->sm, sm = sv->sm;
    long gram_ws = cacg_gram_workspace((int)sm, (int)(sm + m));
    rc = CACG_OK;
    rc |= dalloc(sv, (void **)&sv->d_rowptr, (n + 1) * (long)sizeof(int));
    rc |= dalloc(sv, (void **)&sv->d_col, sv->nnz * (long)sizeof(int));
    rc |= dalloc(sv, (void **)&sv->d_val, sv->nnz * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_R, nm * (long)sizeof(double));
    if (prec32) {
        rc |= dalloc(sv, (void **)&sv->f_val, sv->nnz * (long)sizeof(float));
        rc |= dalloc(sv, (void **)&sv->f_S, nsm * (long)sizeof(float));
    }
    if (!prec32 || sv->opts.fp32_basis_below > 0.0)
        rc |= dalloc(sv, (void **)&sv->d_S, nsm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_Q, nsm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_AQ, nsm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_QtAQ, sm * sm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_QtAS, sm * sm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_QtR, sm * m * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_alpha, sm * m * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_beta, sm * sm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_norm2, m * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_bounds, 2 * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_gwork, 2L * CACG_GERSH_BLOCKS * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_gram_work, gram_ws * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_xr_work, cacg_xrupdate_workspace((int)m) * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_lu_work, sm * sm * (long)sizeof(double));
    rc |= dalloc(sv, (void **)&sv->d_flag, (long)sizeof(int));
    if (sv->opts.bounds == CACG_BOUNDS_LANCZOS) {
        int k = sv->opts.lanczos_steps > 0 ? sv->opts.lanczos_steps : 20;
        if (k > n)
            k = (int)n;
        sv->lz_steps = k;
        rc |= dalloc(sv, (void **)&sv->d_lz_work, cacg_lanczos_workspace(n) * (long)sizeof(double));
        rc |= dalloc(sv, (void **)&sv->d_lz_ab, 2L * k * (long)sizeof(double));
        sv->h_lz = (double *)malloc(sizeof(double) * 2 * (size_t)k);
        if (!sv->h_lz)
            rc = CACG_ENOMEM;
    }
    if (rc != CACG_OK) {
        cacg_solver_destroy(sv);
        return CACG_ENOMEM;
    }
    /* int32 CSR upload */
    int *tmp = (int *)malloc(sizeof(int) * (size_t)((n + 1) > sv->nnz ? (n + 1) : sv->nnz));
    if (!tmp) {
        cacg_solver_destroy(sv);
        return CACG_ENOMEM;
    }
    for (long i = 0; i <= n; i++)
        tmp[i] = (int)row_start[i];
    hipError_t e = hipMemcpy(sv->d_rowptr, tmp, sizeof(int) * (size_t)(n + 1), hipMemcpyHostToDevice);
    for (long p = 0; p < sv->nnz && e == hipSuccess; p++)
        tmp[p] = (int)col_idx[p];
    if (e == hipSuccess && sv->nnz > 0)
        e = hipMemcpy(sv->d_col, tmp, sizeof(int) * (size_t)sv->nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && sv->nnz > 0)
        e = hipMemcpy(sv->d_val, val, sizeof(double) * (size_t)sv->nnz, hipMemcpyHostToDevice);
    free(tmp);
    if (e != hipSuccess) {
        fprintf(stderr, "cacg: matrix upload failed: %s\n", hipGetErrorString(e));
        cacg_solver_destroy(sv);
        return CACG_EDEVICE;
    }
    sv->A.n = (int)n;
    sv->A.rowptr = sv->d_rowptr;
    sv->A.col = sv->d_col;
    sv->A.val = sv->d_val;
    {
        /* rows far longer than the mean go to the row-split SpMM */
        const int lmin = cacg_csr_long_min(n, row_start);
        const long nl = cacg_csr_long_rows(n, row_start, lmin, NULL);
        if (nl > 0) {
            int *ids = (int *)malloc(sizeof(int) * (size_t)nl);
            if (!ids || dalloc(sv, (void **)&sv->d_long_rows, nl * (long)sizeof(int)) != CACG_OK) {
                free(ids);
                cacg_solver_destroy(sv);
                return CACG_ENOMEM;
            }
            cacg_csr_long_rows(n, row_start, lmin, ids);
            e = hipMemcpy(sv->d_long_rows, ids, sizeof(int) * (size_t)nl, hipMemcpyHostToDevice);
            free(ids);
            if (e != hipSuccess) {
                cacg_solver_destroy(sv);
                return CACG_EDEVICE;
            }
            sv->A.long_rows = sv->d_long_rows;
            sv->A.n_long = (int)nl;
            sv->A.long_min = lmin;
        }
    }
    if (prec32) {
        if (cacg_k_cvt_f64_f32(sv->nnz, sv->d_val, sv->f_val, sv->st) != 0 ||
            hipStreamSynchronize(sv->st) != hipSuccess) {
            cacg_solver_destroy(sv);
            return CACG_EDEVICE;
        }
        sv->A32.n = (int)n;
        sv->A32.rowptr = sv->d_rowptr;
        sv->A32.col = sv->d_col;
        sv->A32.val = sv->f_val;
        sv->A32.long_rows = sv->A.long_rows;
        sv->A32.n_long = sv->A.n_long;
        sv->A32.long_min = sv->A.long_min;
    }
    *out = sv;
    return CACG_OK;
}

/* ---- long rows ------------------------------------------------------------------ */

int cacg_csr_long_min(long n, const long *row_start)
{
    if (n <= 0)
        return 256;
    const long nnz = row_start[n] - row_start[0];
    const long mean = (nnz + n - 1) / n;
    const long t = 16 * mean > 256 ? 16 * mean : 256;
    return t < 0x7fffffffL ? (int)t : 0x7fffffff;
}

long cacg_csr_long_rows(long n, const long *row_start, int long_min, int *out)
{
    long c = 0;
    for (long i = 0; i < n; i++)
        if (long_min > 0 && row_start[i + 1] - row_start[i] >= long_min) {
            if (out)
                out[c] = (int)i;
            c++;
        }
    return c;
}

/* ---- phase timing: events recorded between phases on the solver's stream ---- */

static void mark(cacg_solver *sv, int phase_done)
{
    if (sv->noev || sv->nev >= MAXEV)
        return;
    hipEventRecord(sv->ev[sv->nev], sv->st);
    sv->ev_phase[sv->nev] = phase_done;
    sv->nev++;
}

static void harvest(cacg_solver *sv, cacg_report *rep)
{
    if (sv->noev)
        return;
    for (int i = 1; i < sv->nev; i++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, sv->ev[i - 1], sv->ev[i]) == hipSuccess && rep) {
            const int ph = sv->ev_phase[i];
            rep->phase_ms[ph] += ms;
            rep->phase_calls[ph] += 1;
            rep->seconds += ms * 1e-3;
        }
    }
    if (sv->nev > 0) {
        hipEvent_t last = sv->ev[sv->nev - 1];
        sv->ev[sv->nev - 1] = sv->ev[0];
        sv->ev[0] = last;
        sv->ev_phase[0] = sv->ev_phase[sv->nev - 1];
        sv->nev = 1;
    }
}

/* ---- the solve ------------------------------------------------------------------ */

static int gram(cacg_solver *sv, int mode, const double *L, const double *R1, long ld1, int kr1,
                const double *R2, long ld2, int kr2, double *G1, double *G2)
{
    const int n = (int)sv->n, sm = (int)sv->sm;
    if (mode == CACG_MODE_REFERENCE)
        return cacg_k_gram_ref(n, L, sm, sm, R1, ld1, kr1, R2, ld2, kr2, G1, G2, sv->st);
    return cacg_k_gram_fast(n, L, sm, sm, R1, ld1, kr1, R2, ld2, kr2, G1, G2, sv->d_gram_work,
                            sv->st);
}

/* ---- SpMM schedule autotuner ------------------------------------------------------- */

#define TUNE_MIN_ROWS (1L << 17) /* below this the launches are latency-bound: defaults */

/* device time of `reps` launches of one SpMM shape under schedule sc, after
 * one untimed launch; a negative value when a launch failed */
static float time_spmm(cacg_solver *sv, int f32, int k, long ld, const void *x, void *y,
                       const void *p1, const void *p2, int epi, const cacg_spmm_sched *sc, int reps)
{
    float ms = -1.f;
    for (int r = -1; r < reps; r++) {
        if (r == 0 && hipEventRecord(sv->tev[0], sv->st) != hipSuccess)
            return -1.f;
        int rc = f32 ? cacg_k_spmm_f32_x(&sv->A32, k, (const float *)x, ld, (float *)y, ld, epi,
                                         0.5, 0.25, (const float *)p1, ld, (const float *)p2, ld,
                                         sc, sv->st)
                     : cacg_k_spmm_x(&sv->A, k, (const double *)x, ld, (double *)y, ld, epi, 0.5,
                                     0.25, (const double *)p1, ld, (const double *)p2, ld, NULL, 0,
                                     sc, sv->st);
        if (rc)
            return -1.f;
    }
    if (hipEventRecord(sv->tev[1], sv->st) != hipSuccess ||
        hipEventSynchronize(sv->tev[1]) != hipSuccess ||
        hipEventElapsedTime(&ms, sv->tev[0], sv->tev[1]) != hipSuccess)
        return -1.f;
    return ms;
}

/* Schedules tuned in this process, keyed by device, shape and a sample of the
 * sparsity pattern: a solver created for a matrix already tuned (the drop-in
 * entry points create one per call) takes its schedules instead of timing the
 * candidates again.  A schedule changes no result, so a key shared by two
 * different patterns can cost speed, never bits. */
#define SCHED_CACHE 8
static struct {
    unsigned long long key;
    int used;
    cacg_spmm_sched basis, aq;
} g_sched[SCHED_CACHE];
static int g_sched_next;
static pthread_mutex_t g_sched_mu = PTHREAD_MUTEX_INITIALIZER;

static unsigned long long key_mix(unsigned long long h, unsigned long long v)
{
    for (int b = 0; b < 8; b++) {   /* FNV-1a over the 8 bytes of v */
        h ^= (v >> (8 * b)) & 0xffu;
        h *= 1099511628211ULL;
    }
    return h;
}

/* device, n, nnz, m, s m, precision and 257 evenly spaced entries of the row
 * starts and of the column indices */
static unsigned long long pattern_key(long n, const long *row_start, const long *col_idx, long m,
                                      long sm, int prec32)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        dev = -1;
    const long nnz = row_start[n];
    unsigned long long h = 1469598103934665603ULL;
    h = key_mix(h, (unsigned long long)dev);
    h = key_mix(h, (unsigned long long)n);
    h = key_mix(h, (unsigned long long)nnz);
    h = key_mix(h, (unsigned long long)m);
    h = key_mix(h, (unsigned long long)sm);
    h = key_mix(h, (unsigned long long)prec32);
    for (int i = 0; i <= 256; i++) {
        h = key_mix(h, (unsigned long long)row_start[(long)((double)n * i / 256.0)]);
        if (nnz > 0)
            h = key_mix(h, (unsigned long long)col_idx[(long)((double)(nnz - 1) * i / 256.0)]);
    }
    return h;
}

/* Pick the tile order and grid of the two SpMM shapes a solve runs (a basis
 * step: k = m columns of the n x sm basis with the Chebyshev epilogue; A Q:
 * k = sm) by timing a few schedules on this matrix, once per solver.  The
 * work blocks serve as scratch (zero-filled first; the solve overwrites
 * them).  A schedule changes no result, so neither does the choice. */
static int tune_spmm(cacg_solver *sv, int blk)
{
    static const cacg_spmm_sched cand[] = {
        {0, 0, 0},       {4096, 0, 0},    {8192, 0, 0},       {0, 1024, 0},
        {0, 2048, 0},    {0, 4096, 0},    {4096, 2048, 0},    {8192, 4096, 0},
        {16384, 512, 4}, {32768, 512, 4}};
    const int nc = (int)(sizeof(cand) / sizeof(cand[0]));
    sv->tuned = 1;
    if (sv->n < TUNE_MIN_ROWS)
        return 0;
    const unsigned long long key = sv->pattern_key ^ (blk ? 0x9e3779b97f4a7c15ULL : 0ULL);
    pthread_mutex_lock(&g_sched_mu);
    for (int i = 0; i < SCHED_CACHE; i++)
        if (g_sched[i].used && g_sched[i].key == key) {
            sv->sch_basis = g_sched[i].basis;
            sv->sch_aq = g_sched[i].aq;
            pthread_mutex_unlock(&g_sched_mu);
            return 0;
        }
    pthread_mutex_unlock(&g_sched_mu);
    const int m = (int)sv->m, sm = (int)sv->sm;
    const int f32 = sv->opts.precision == CACG_PREC_FP32_REFINE;
    const long nsm = sv->n * sv->sm;
    int rc = cacg_k_fill_f64(sv->d_Q, nsm, 0.0, sv->st);
    rc = rc ? rc : cacg_k_fill_f64(sv->d_AQ, nsm, 0.0, sv->st);
    if (!rc && f32)
        rc = (int)hipMemsetAsync(sv->f_S, 0, sizeof(float) * (size_t)nsm, sv->st);
    else if (!rc)
        rc = cacg_k_fill_f64(sv->d_S, nsm, 0.0, sv->st);
    if (rc)
        return rc;
    /* basis step: x, p1 one block of S, p2 another, y a block of Q (FP64) or of the
     * FP32 basis (the precision form) */
    const void *bx = f32 ? (const void *)sv->f_S : (const void *)sv->d_S;
    const void *bp2 = f32 ? (const void *)(sv->f_S + (sm > m ? m : 0)) : (const void *)sv->d_AQ;
    void *by = f32 ? (void *)(sv->f_S + (sm > 2 * m ? 2 * m : 0)) : (void *)sv->d_Q;
    float best_b = -1.f, best_a = -1.f;
    for (int c = 0; c < nc; c++) {
        /* the basis steps' leading dimension: m for a block-major basis */
        const float tb = time_spmm(sv, f32, m, blk ? m : sm, bx, by, bx, bp2, CACG_EPI_CHEB,
                                   &cand[c], 2);
        if (tb >= 0.f && (best_b < 0.f || tb < best_b)) {
            best_b = tb;
            sv->sch_basis = cand[c];
        }
        const float ta = time_spmm(sv, 0, sm, sm, sv->d_Q, sv->d_AQ, NULL, NULL, CACG_EPI_NONE,
                                   &cand[c], 2);
        if (ta >= 0.f && (best_a < 0.f || ta < best_a)) {
            best_a = ta;
            sv->sch_aq = cand[c];
        }
    }
    if (best_b < 0.f || best_a < 0.f)
        return (int)hipErrorLaunchFailure;
    pthread_mutex_lock(&g_sched_mu);
    g_sched[g_sched_next].key = key;
    g_sched[g_sched_next].used = 1;
    g_sched[g_sched_next].basis = sv->sch_basis;
    g_sched[g_sched_next].aq = sv->sch_aq;
    g_sched_next = (g_sched_next + 1) % SCHED_CACHE;
    pthread_mutex_unlock(&g_sched_mu);
    return 0;
}

/* enqueue what the Chebyshev interval needs beyond Gerschgorin: the Lanczos
 * coefficients, on a solver's first solve with CACG_BOUNDS_LANCZOS */
static int interval_enqueue(cacg_solver *sv)
{
    if (sv->opts.bounds != CACG_BOUNDS_LANCZOS || sv->have_lz)
        return 0;
    int rc = cacg_k_lanczos(&sv->A, sv->lz_steps, sv->d_lz_work, sv->d_lz_ab, sv->st);
    if (rc)
        return rc;
    return (int)hipMemcpyAsync(sv->h_lz, sv->d_lz_ab, sizeof(double) * 2 * (size_t)sv->lz_steps,
                               hipMemcpyDeviceToHost, sv->st);
}

/* after the stream synchronised: the interval of this solve from the
 * Gerschgorin interval g */
static int interval_finish(cacg_solver *sv, const double g[2], double iv[2])
{
    if (sv->opts.bounds == CACG_BOUNDS_GIVEN) {
        iv[0] = sv->opts.bound_lo;
        iv[1] = sv->opts.bound_hi;
    } else if (sv->opts.bounds == CACG_BOUNDS_LANCZOS) {
        if (!sv->have_lz) {
            if (cacg_lanczos_interval(sv->h_lz, sv->h_lz + sv->lz_steps, sv->lz_steps, g,
                                      sv->lz_iv) != CACG_OK)
                return CACG_BREAKDOWN;
            sv->have_lz = 1;
        }
        iv[0] = sv->lz_iv[0];
        iv[1] = sv->lz_iv[1];
    } else {
        iv[0] = g[0];
        iv[1] = g[1];
    }
    return CACG_OK;
}

#define SCHK(call)                                                                               \
    do {                                                                                         \
        int e_ = (int)(call);                                                                    \
        if (e_ != 0) {                                                                           \
            fprintf(stderr, "cacg: %s failed: %s (%s:%d)\n", #call,                              \
                    hipGetErrorString((hipError_t)e_), __FILE__, __LINE__);                      \
            status = CACG_EDEVICE;                                                               \
            goto out;                                                                            \
        }                                                                                        \
    } while (0)

static int solve_refine(cacg_solver *sv, const double *d_B, double *d_X, cacg_report *rep);

int cacg_solver_solve_device(cacg_solver *sv, const double *d_B, double *d_X, cacg_report *rep)
{
    if (!sv || !d_B || !d_X)
        return CACG_EINVAL;
    if (sv->opts.precision == CACG_PREC_FP32_REFINE)
        return solve_refine(sv, d_B, d_X, rep);
    const cacg_opts *o = &sv->opts;
    const int mode = o->order == CACG_ORDER_REFERENCE ? CACG_MODE_REFERENCE : CACG_MODE_FAST;
    const int n = (int)sv->n, m = (int)sv->m, s = (int)sv->s, sm = (int)sv->sm;
    /* fast order: the X update deferred into the next Q update (Q read once) */
    const int split = mode == CACG_MODE_FAST && cacg_k_split_supported(sm, m);
    /* fast order: from iteration 1 on, the basis is stored block-major (S_j as
     * n x m blocks n m elements apart), so each basis step reads and writes
     * whole contiguous rows; iteration 0 writes it row-major into Q (Q = S) */
    const int blk = split && m % 4 == 0 && s >= 2 && !sv->basis_rows &&
                    cacg_gram_blk_supported(n, sm, sm, m);
    const long bs = (long)n * m;
    hipStream_t st = sv->st;
    double *h_norm2 = sv->h_pin, *h_bnorm2 = sv->h_pin + m, *h_bounds = sv->h_pin + 2 * m;
    int *h_flag = (int *)(sv->h_pin + 2 * m + 2);
    double *S = sv->d_S, *Q = sv->d_Q;
    int status = CACG_MAXITER;
    long it = 0;
    if (rep) {
        double *keep = rep->res_hist;
        long cap = rep->res_hist_cap;
        memset(rep, 0, sizeof(*rep));
        rep->res_hist = keep;
        rep->res_hist_cap = cap;
        rep->true_relres = -1.0;
    }

    if (!sv->tuned)
        SCHK(tune_spmm(sv, blk));
    sv->nev = 0;
    mark(sv, CACG_PHASE_OTHER);
    /* setup: bounds, R = B - A X (src/bcbcgKernel.c:106-154) */
    SCHK(cacg_k_fill_i32(sv->d_flag, 1, 0, st));
    SCHK(cacg_k_gerschgorin(&sv->A, sv->d_gwork, sv->d_bounds, st));
    SCHK(interval_enqueue(sv));
    SCHK(cacg_k_spmm(&sv->A, m, d_X, m, sv->d_R, m, CACG_EPI_RESID, 0.0, 0.0, d_B, m, NULL, 0,
                     NULL, 0, st));
    if (o->stop_mode == CACG_STOP_REL_PER_COLUMN) {
        if (mode == CACG_MODE_REFERENCE) {
            SCHK(cacg_k_colnorm2_ref(n, m, d_B, m, sv->d_norm2, st));
            SCHK(hipMemcpyAsync(h_bnorm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        } else {
            /* B^T B, diagonal taken as one strided copy */
            SCHK(cacg_k_gram_fast(n, d_B, m, m, d_B, m, m, NULL, 0, 0, sv->d_QtAS, NULL,
                                  sv->d_gram_work, st));
            SCHK(hipMemcpy2DAsync(h_bnorm2, sizeof(double), sv->d_QtAS, sizeof(double) * (m + 1),
                                  sizeof(double), m, hipMemcpyDeviceToHost, st));
        }
    }
    mark(sv, CACG_PHASE_SETUP);
    SCHK(hipMemcpyAsync(h_bounds, sv->d_bounds, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    harvest(sv, rep);
    double iv[2];
    status = interval_finish(sv, h_bounds, iv);
    if (status != CACG_OK)
        goto out;
    status = CACG_MAXITER;
    if (rep) {
        rep->bounds[0] = iv[0];
        rep->bounds[1] = iv[1];
    }
    /* src/bcbcgKernel.c:113-115 */
    const double sa = 2.0 / (iv[1] - iv[0]);
    const double sb = -(iv[1] + iv[0]) / (iv[1] - iv[0]);
    const double sa2 = 2.0 * sa, sb2 = 2.0 * sb;

    for (it = 0; it < o->maxiter; it++) {
        mark(sv, CACG_PHASE_OTHER);
        /* Chebyshev basis (src/common.c:182-256): block j at SB + j * bo, rows
         * lb apart; the first iteration's basis is its Q */
        const int bm = blk && it > 0;
        double *SB = it == 0 ? Q : S;
        const long bo = bm ? bs : m, lb = bm ? m : sm;
        if (s == 1) {
            SCHK(hipMemcpyAsync(SB, sv->d_R, sizeof(double) * (size_t)n * m, hipMemcpyDeviceToDevice, st));
        } else {
            SCHK(cacg_k_spmm_x(&sv->A, m, sv->d_R, m, SB + bo, lb, CACG_EPI_CHEB1, sa, sb, NULL, 0,
                               NULL, 0, SB, lb, &sv->sch_basis, st));
            mark(sv, CACG_PHASE_BASIS1);
            for (int j = 2; j < s; j++)
                SCHK(cacg_k_spmm_x(&sv->A, m, SB + (j - 1) * bo, lb, SB + j * bo, lb,
                                   CACG_EPI_CHEB, sa2, sb2, SB + (j - 1) * bo, lb,
                                   SB + (j - 2) * bo, lb, NULL, 0, &sv->sch_basis, st));
        }
        if (s != 2)
            mark(sv, s == 1 ? CACG_PHASE_BASIS1 : CACG_PHASE_BASIS);
        if (it > 0) {
            if (bm)
                SCHK(cacg_k_gram_fast_blk(n, sv->d_AQ, sm, sm, S, m, bs, sm, sv->d_QtAS,
                                          sv->d_gram_work, st));
            else
                SCHK(gram(sv, mode, sv->d_AQ, S, sm, sm, NULL, 0, 0, sv->d_QtAS, NULL));
            mark(sv, CACG_PHASE_GRAM_AS);
            /* Q^T A Q of the previous iteration: its factor is still in d_lu_work */
            SCHK(cacg_k_lu_solve(sm, sv->d_QtAQ, sm, sv->d_QtAS, sv->d_beta, sv->d_lu_work,
                                 sv->d_flag, mode, 1, st));
            mark(sv, CACG_PHASE_SOLVE);
            if (split && bm) /* X += Q alpha of the previous iteration, then Q = S - Q beta */
                SCHK(cacg_k_qxupdate_blk(n, sm, m, S, bs, Q, sv->d_beta, sv->d_alpha, d_X, st));
            else if (split)
                SCHK(cacg_k_qxupdate(n, sm, m, S, Q, sv->d_beta, sv->d_alpha, d_X, st));
            else
                SCHK(cacg_k_qupdate(n, sm, S, Q, sv->d_beta, mode, st));
            mark(sv, CACG_PHASE_QUPDATE);
        }
        SCHK(cacg_k_spmm_x(&sv->A, sm, Q, sm, sv->d_AQ, sm, CACG_EPI_NONE, 0.0, 0.0, NULL, 0, NULL,
                           0, NULL, 0, &sv->sch_aq, st));
        mark(sv, CACG_PHASE_AQ);
        SCHK(gram(sv, mode, Q, sv->d_AQ, sm, sm, sv->d_R, m, m, sv->d_QtAQ, sv->d_QtR));
        mark(sv, CACG_PHASE_GRAM);
        SCHK(cacg_k_lu_solve(sm, sv->d_QtAQ, m, sv->d_QtR, sv->d_alpha, sv->d_lu_work, sv->d_flag,
                             mode, 0, st));
        mark(sv, CACG_PHASE_SOLVE);
        if (split) /* R and its norms now; X with the next Q update */
            SCHK(cacg_k_xrupdate_split(n, sm, m, NULL, sv->d_AQ, sv->d_alpha, NULL, sv->d_R,
                                       sv->d_norm2, sv->d_xr_work, 2, st));
        else
            SCHK(cacg_k_xrupdate(n, sm, m, Q, sv->d_AQ, sv->d_alpha, d_X, sv->d_R, sv->d_norm2,
                                 sv->d_xr_work, mode, st));
        mark(sv, CACG_PHASE_XRUPDATE);
        SCHK(hipMemcpyAsync(h_norm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        SCHK(hipMemcpyAsync(h_flag, sv->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        harvest(sv, rep);

        int done = 1, finite = 1;
        for (int c = 0; c < m; c++) {
            const double rn = sqrt(h_norm2[c]);
            if (rep && rep->res_hist && it < rep->res_hist_cap)
                rep->res_hist[it * m + c] = rn;
            if (!isfinite(rn))
                finite = 0;
            if (o->stop_mode == CACG_STOP_REL_PER_COLUMN && !(rn < o->tol * sqrt(h_bnorm2[c])))
                done = 0;
        }
        const double r0 = sqrt(h_norm2[0]);
        if (o->stop_mode == CACG_STOP_REF_COL0_ABS)
            done = r0 < o->eps; /* src/bcbcgKernel.c:445-447, 524-526 */
        if (o->print_period > 0 &&
            (it % o->print_period == 0 || done || it == o->maxiter - 1)) /* :449-452 */
            printf("Loop: %ld, R[%d]_norm=%30.30f\n", it, 1, r0);
        if (done) {
            status = CACG_OK;
            it++;
            break;
        }
        if (!finite || *h_flag) {
            status = CACG_BREAKDOWN;
            it++;
            break;
        }
    }
    if (split && it > 0) { /* the last iteration's X += Q alpha */
        SCHK(cacg_k_xrupdate_split(n, sm, m, Q, NULL, sv->d_alpha, d_X, NULL, NULL, NULL, 1, st));
        SCHK(hipStreamSynchronize(st));
    }
out:
    /* leave the two basis-sized buffers in allocation roles for the next solve */
    sv->d_S = S;
    sv->d_Q = Q;
    if (rep) {
        rep->status = status;
        rep->iters = it;
    }
    return status;
}

/* ---- the entry point that allocates nothing -------------------------------------------- */

#define NOALLOC_ALIGN 256L

static long na_round(long bytes) { return (bytes + NOALLOC_ALIGN - 1) / NOALLOC_ALIGN * NOALLOC_ALIGN; }

/* bytes of scratch a solve of (n, m, s) takes: the work blocks of
 * cacg_solver_create in the same roles (FP64, Gerschgorin interval) */
long cacg_noalloc_bytes(long n, long m, long s)
{
    if (n <= 0 || m <= 0 || s <= 0 || s * m > 4096 || n >= (1L << 31))
        return -1;
    const long sm = s * m, d = (long)sizeof(double);
    long gws = cacg_gram_workspace_n((int)n, (int)sm, (int)(sm + m));
    const long g2 = cacg_gram_workspace_n((int)n, (int)sm, (int)sm);
    const long g3 = cacg_gram_workspace_n((int)n, (int)m, (int)m);
    if (g2 > gws)
        gws = g2;
    if (g3 > gws)
        gws = g3;
    long b = 0;
    b += na_round(n * m * d);                               /* R */
    b += 3 * na_round(n * sm * d);                          /* S, Q, AQ */
    b += 3 * na_round(sm * sm * d);                         /* QtAQ, QtAS, beta */
    b += 2 * na_round(sm * m * d);                          /* QtR, alpha */
    b += na_round(m * d) + na_round(2 * d);                 /* norms, bounds */
    b += na_round(2L * CACG_GERSH_BLOCKS * d);              /* Gerschgorin partials */
    b += na_round(gws * d);                                 /* Gram partials */
    b += na_round(cacg_xrupdate_workspace((int)m) * d);     /* norm partials */
    b += na_round(sm * sm * d);                             /* LU factor */
    b += na_round((long)sizeof(int));                       /* breakdown flag */
    return b;
}

static void *na_take(char **cur, long bytes)
{
    void *p = *cur;
    *cur += na_round(bytes);
    return p;
}

int cacg_bcbcg_device_noalloc(int n, const int *d_rowptr, const int *d_col, const double *d_val,
                              long m, long s, const double *d_B, double *d_X,
                              const cacg_opts *opts, cacg_report *report, void *stream)
{
    if (n <= 0 || !d_rowptr || !d_col || !d_val || !d_B || !d_X || m <= 0 ||
        m > CACG_NOALLOC_MAX_M || s <= 0 || s * m > 4096)
        return CACG_EINVAL;
    cacg_opts o;
    if (opts)
        o = *opts;
    else
        cacg_default_opts(&o);
    if (o.maxiter <= 0 || o.precision != CACG_PREC_FP64 || o.tol < 0.0 ||
        (o.bounds != CACG_BOUNDS_GERSCHGORIN && o.bounds != CACG_BOUNDS_GIVEN) ||
        (o.bounds == CACG_BOUNDS_GIVEN &&
         !(isfinite(o.bound_lo) && isfinite(o.bound_hi) && o.bound_lo < o.bound_hi)))
        return CACG_EINVAL;
    const long need = cacg_noalloc_bytes(n, m, s);
    void *base = NULL;
    long cap = 0;
    int rc = check_device();
    if (rc != CACG_OK)
        return rc;
    if (cacg_noalloc_scratch(&base, &cap) != 0)
        return CACG_EDEVICE;
    if (need < 0 || need > cap) {
        fprintf(stderr,
                "cacg: cacg_bcbcg_device_noalloc: n = %d, m = %ld, s = %ld needs %ld bytes of "
                "scratch, more than the %ld bytes it has\n",
                n, m, s, need, cap);
        return CACG_EINVAL;
    }
    /* the solver of this call: on the stack, its work blocks carved from the
     * scratch (every block is written before it is read), no events, no SpMM
     * tuning (the default schedules), host-side norms in a stack buffer */
    cacg_solver sv;
    double hbuf[2 * CACG_NOALLOC_MAX_M + 4];
    memset(&sv, 0, sizeof(sv));
    sv.n = n;
    sv.m = m;
    sv.s = s;
    sv.sm = s * m;
    sv.opts = o;
    sv.st = (hipStream_t)stream;
    sv.A.n = n;
    sv.A.rowptr = d_rowptr;
    sv.A.col = d_col;
    sv.A.val = d_val;
    sv.h_pin = hbuf;
    sv.tuned = 1;
    sv.noev = 1;
    sv.basis_rows = 0;
    const long sm = sv.sm, d = (long)sizeof(double);
    char *cur = (char *)base;
    sv.d_R = (double *)na_take(&cur, n * m * d);
    sv.d_S = (double *)na_take(&cur, n * sm * d);
    sv.d_Q = (double *)na_take(&cur, n * sm * d);
    sv.d_AQ = (double *)na_take(&cur, n * sm * d);
    sv.d_QtAQ = (double *)na_take(&cur, sm * sm * d);
    sv.d_QtAS = (double *)na_take(&cur, sm * sm * d);
    sv.d_beta = (double *)na_take(&cur, sm * sm * d);
    sv.d_QtR = (double *)na_take(&cur, sm * m * d);
    sv.d_alpha = (double *)na_take(&cur, sm * m * d);
    sv.d_norm2 = (double *)na_take(&cur, m * d);
    sv.d_bounds = (double *)na_take(&cur, 2 * d);
    sv.d_gwork = (double *)na_take(&cur, 2L * CACG_GERSH_BLOCKS * d);
    {
        long gws = cacg_gram_workspace_n(n, (int)sm, (int)(sm + m));
        const long g2 = cacg_gram_workspace_n(n, (int)sm, (int)sm);
        const long g3 = cacg_gram_workspace_n(n, (int)m, (int)m);
        if (g2 > gws)
            gws = g2;
        if (g3 > gws)
            gws = g3;
        sv.d_gram_work = (double *)na_take(&cur, gws * d);
    }
    sv.d_xr_work = (double *)na_take(&cur, cacg_xrupdate_workspace((int)m) * d);
    sv.d_lu_work = (double *)na_take(&cur, sm * sm * d);
    sv.d_flag = (int *)na_take(&cur, (long)sizeof(int));
    if (cur - (char *)base != need)
        return CACG_EDEVICE; /* the layout above and cacg_noalloc_bytes disagree */
    return cacg_solver_solve_device(&sv, d_B, d_X, report);
}

/* ---- block CG (src/bcgKernel.c:22-462) ---------------------------------------------- */

/* O'Leary block CG on the solver's matrix: P, A P in the leading n x m parts of
 * the Q and A Q buffers, P^T A P, R^T R and the coefficients in the m x m
 * leading parts of the small blocks.  Per iteration (src/bcgKernel.c:184-443):
 *   AP = A P; PtAP = P^T AP; alpha = PtAP^{-1} RtR      :192, :213, :239
 *   X += P alpha; R -= AP alpha; ||R(:,c)||^2           :258-291, :375
 *   stop test on the R the iteration started from        :301-365
 *   RnRn = R^T R; beta = RtR^{-1} RnRn; RtR = RnRn        :389-419
 *   P = R + P beta, as R - P (-beta): negation is exact   :428-443
 * The X update comes before the stop test, as in the reference, so the
 * iteration that stops has already moved X once more.  The new residual's
 * norms are computed with the X/R update and read back once per iteration;
 * the test of iteration k reads those of iteration k - 1. */
int cacg_solver_solve_bcg_device(cacg_solver *sv, const double *d_B, double *d_X, cacg_report *rep)
{
    if (!sv || !d_B || !d_X || sv->opts.precision != CACG_PREC_FP64)
        return CACG_EINVAL;
    const cacg_opts *o = &sv->opts;
    const int mode = o->order == CACG_ORDER_REFERENCE ? CACG_MODE_REFERENCE : CACG_MODE_FAST;
    const int n = (int)sv->n, m = (int)sv->m;
    hipStream_t st = sv->st;
    double *h_norm2 = sv->h_pin, *h_bnorm2 = sv->h_pin + m;
    int *h_flag = (int *)(sv->h_pin + 2 * m + 2);
    double *P = sv->d_Q, *AP = sv->d_AQ, *R = sv->d_R;
    double *PtAP = sv->d_QtAQ, *RtR = sv->d_QtR, *RnRn = sv->d_QtAS;
    double *alpha = sv->d_alpha, *beta = sv->d_beta;
    double *prev = (double *)malloc(sizeof(double) * (size_t)m);
    int status = CACG_MAXITER;
    long it = 0;
    if (!prev)
        return CACG_ENOMEM;
    if (rep) {
        double *keep = rep->res_hist;
        long cap = rep->res_hist_cap;
        memset(rep, 0, sizeof(*rep));
        rep->res_hist = keep;
        rep->res_hist_cap = cap;
        rep->true_relres = -1.0;
    }
#define GRAMM(L_, R_, G_)                                                                        \
    (mode == CACG_MODE_REFERENCE                                                                 \
         ? cacg_k_gram_ref(n, L_, m, m, R_, m, m, NULL, 0, 0, G_, NULL, st)                      \
         : cacg_k_gram_fast(n, L_, m, m, R_, m, m, NULL, 0, 0, G_, NULL, sv->d_gram_work, st))

    sv->nev = 0;
    mark(sv, CACG_PHASE_OTHER);
    /* R = B - A X, P = R, RtR = R^T R, ||R(:,c)||^2 (src/bcgKernel.c:89-178) */
    SCHK(cacg_k_fill_i32(sv->d_flag, 1, 0, st));
    SCHK(cacg_k_spmm(&sv->A, m, d_X, m, R, m, CACG_EPI_RESID, 0.0, 0.0, d_B, m, NULL, 0, NULL, 0,
                     st));
    SCHK(hipMemcpyAsync(P, R, sizeof(double) * (size_t)n * m, hipMemcpyDeviceToDevice, st));
    SCHK(GRAMM(R, R, RtR));
    if (mode == CACG_MODE_REFERENCE)
        SCHK(cacg_k_colnorm2_ref(n, m, R, m, sv->d_norm2, st));
    else
        SCHK(cacg_k_colnorm2_fast(n, m, R, sv->d_norm2, sv->d_xr_work, st));
    SCHK(hipMemcpyAsync(prev, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    if (o->stop_mode == CACG_STOP_REL_PER_COLUMN) {
        if (mode == CACG_MODE_REFERENCE)
            SCHK(cacg_k_colnorm2_ref(n, m, d_B, m, sv->d_norm2, st));
        else
            SCHK(cacg_k_colnorm2_fast(n, m, d_B, sv->d_norm2, sv->d_xr_work, st));
        SCHK(hipMemcpyAsync(h_bnorm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    }
    mark(sv, CACG_PHASE_SETUP);
    SCHK(hipStreamSynchronize(st));
    harvest(sv, rep);

    for (it = 0; it < o->maxiter; it++) {
        mark(sv, CACG_PHASE_OTHER);
        SCHK(cacg_k_spmm(&sv->A, m, P, m, AP, m, CACG_EPI_NONE, 0.0, 0.0, NULL, 0, NULL, 0, NULL, 0,
                         st));
        mark(sv, CACG_PHASE_AQ);
        SCHK(GRAMM(P, AP, PtAP));
        mark(sv, CACG_PHASE_GRAM);
        SCHK(cacg_k_lu_solve(m, PtAP, m, RtR, alpha, sv->d_lu_work, sv->d_flag, mode, 0, st));
        mark(sv, CACG_PHASE_SOLVE);
        SCHK(cacg_k_xrupdate(n, m, m, P, AP, alpha, d_X, R, sv->d_norm2, sv->d_xr_work, mode, st));
        mark(sv, CACG_PHASE_XRUPDATE);
        SCHK(hipMemcpyAsync(h_norm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        SCHK(hipMemcpyAsync(h_flag, sv->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        harvest(sv, rep);

        int done = 1, finite = 1;
        for (int c = 0; c < m; c++) {
            const double rn = sqrt(prev[c]);
            if (rep && rep->res_hist && it < rep->res_hist_cap)
                rep->res_hist[it * m + c] = rn;
            if (!isfinite(rn))
                finite = 0;
            if (o->stop_mode == CACG_STOP_REL_PER_COLUMN && !(rn < o->tol * sqrt(h_bnorm2[c])))
                done = 0;
        }
        const double r0 = sqrt(prev[0]);
        if (o->stop_mode == CACG_STOP_REF_COL0_ABS)
            done = r0 < o->eps; /* src/bcgKernel.c:301, 364-366 */
        if (o->print_period > 0 &&
            (it % o->print_period == 0 || done || it == o->maxiter - 1)) /* :302-305 */
            printf("id:%d, Loop: %ld, R[%d]_norm=%30.30f\n", 0, it, 1, r0);
        if (done) {
            status = CACG_OK;
            it++;
            break;
        }
        if (!finite || *h_flag) {
            status = CACG_BREAKDOWN;
            it++;
            break;
        }
        memcpy(prev, h_norm2, sizeof(double) * m);
        mark(sv, CACG_PHASE_OTHER);
        SCHK(GRAMM(R, R, RnRn));
        mark(sv, CACG_PHASE_GRAM_AS);
        SCHK(cacg_k_lu_solve(m, RtR, m, RnRn, beta, sv->d_lu_work, sv->d_flag, mode, 0, st));
        mark(sv, CACG_PHASE_SOLVE);
        double *t = RtR;
        RtR = RnRn;
        RnRn = t;
        SCHK(cacg_k_scale_f64(beta, (long)m * m, -1.0, st));
        SCHK(cacg_k_qupdate(n, m, R, P, beta, mode, st));
        mark(sv, CACG_PHASE_QUPDATE);
    }
#undef GRAMM
out:
    free(prev);
    if (rep) {
        rep->status = status;
        rep->iters = it;
    }
    return status;
}

/* ---- precision form: FP32 basis, FP64 refinement ----------------------------------- */

/* 1 when the column norms (squared) meet the stop criterion */
static int stop_met(const cacg_opts *o, const double *norm2, const double *bnorm2, int m)
{
    if (o->stop_mode == CACG_STOP_REF_COL0_ABS)
        return sqrt(norm2[0]) < o->eps;
    for (int c = 0; c < m; c++)
        if (!(sqrt(norm2[c]) < o->tol * sqrt(bnorm2[c])))
            return 0;
    return 1;
}

static int solve_refine(cacg_solver *sv, const double *d_B, double *d_X, cacg_report *rep)
{
    const cacg_opts *o = &sv->opts;
    const int n = (int)sv->n, m = (int)sv->m, s = (int)sv->s, sm = (int)sv->sm;
    const int restart = o->refine_restart;
    /* defaults measured on the 128^3 Laplacian, 16 RHS, s = 8: replacement
     * every 1e-2 of residual reduction (53 iterations, the FP64 path 45);
     * restarts every 1e-5 (56) */
    const double eta = o->inner_tol > 0.0 ? o->inner_tol : (restart ? 1e-5 : 1e-2);
    const long inner_cap = o->inner_maxiter > 0 ? o->inner_maxiter : o->maxiter;
    hipStream_t st = sv->st;
    double *h_norm2 = sv->h_pin, *h_bnorm2 = sv->h_pin + m, *h_bounds = sv->h_pin + 2 * m;
    int *h_flag = (int *)(sv->h_pin + 2 * m + 2);
    float *S = sv->f_S;
    double *Q = sv->d_Q, *AQ = sv->d_AQ;
    int status = CACG_MAXITER;
    long it = 0, refinements = 0;
    double *tn2 = (double *)malloc(sizeof(double) * (size_t)m);  /* true residual norms^2 */
    double *target = (double *)malloc(sizeof(double) * (size_t)m);
    double *fin = (double *)malloc(sizeof(double) * (size_t)m);
    if (!tn2 || !target || !fin) {
        free(tn2);
        free(target);
        free(fin);
        return CACG_ENOMEM;
    }
    if (rep) {
        double *keep = rep->res_hist;
        long cap = rep->res_hist_cap;
        memset(rep, 0, sizeof(*rep));
        rep->res_hist = keep;
        rep->res_hist_cap = cap;
        rep->true_relres = -1.0;
    }

    if (!sv->tuned)
        SCHK(tune_spmm(sv, 0));
    sv->nev = 0;
    mark(sv, CACG_PHASE_OTHER);
    SCHK(cacg_k_fill_i32(sv->d_flag, 1, 0, st));
    SCHK(cacg_k_gerschgorin(&sv->A, sv->d_gwork, sv->d_bounds, st));
    SCHK(interval_enqueue(sv));
    SCHK(cacg_k_colnorm2_fast(n, m, d_B, sv->d_norm2, sv->d_xr_work, st));
    SCHK(hipMemcpyAsync(h_bnorm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    SCHK(hipMemcpyAsync(h_bounds, sv->d_bounds, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    mark(sv, CACG_PHASE_SETUP);
    SCHK(hipStreamSynchronize(st));
    harvest(sv, rep);
    double iv[2];
    status = interval_finish(sv, h_bounds, iv);
    if (status != CACG_OK)
        goto out;
    status = CACG_MAXITER;
    if (rep) {
        rep->bounds[0] = iv[0];
        rep->bounds[1] = iv[1];
    }
    const double sa = 2.0 / (iv[1] - iv[0]);
    const double sb = -(iv[1] + iv[0]) / (iv[1] - iv[0]);
    const double sa2 = 2.0 * sa, sb2 = 2.0 * sb;
    for (int c = 0; c < m; c++)
        fin[c] = o->stop_mode == CACG_STOP_REF_COL0_ABS ? (c == 0 ? o->eps : 0.0)
                                                        : o->tol * sqrt(h_bnorm2[c]);

    int fresh = 1;   /* next iteration starts the recurrence (Q = S) */
    long since = 0;  /* iterations since the last true residual */
    /* X += Q alpha deferred into the next Q update (split) and flushed wherever X
     * is read: before a true residual and at the end */
    const int split = cacg_k_split_supported(sm, m);
    int xpend = 0;
    /* opts.fp32_basis_below > 0: the basis is built in FP64 (d_S, the FP64
     * matrix) until the true residual is below it in every column, in FP32 from
     * then on */
    int use64 = o->fp32_basis_below > 0.0 && sv->d_S;
    for (;;) {
        if (xpend) {
            SCHK(cacg_k_xrupdate_split(n, sm, m, Q, NULL, sv->d_alpha, d_X, NULL, NULL, NULL, 1,
                                       st));
            xpend = 0;
        }
        /* true residual in FP64 with the FP64 matrix; it replaces the recursive one */
        mark(sv, CACG_PHASE_OTHER);
        SCHK(cacg_k_spmm(&sv->A, m, d_X, m, sv->d_R, m, CACG_EPI_RESID, 0.0, 0.0, d_B, m, NULL, 0,
                         NULL, 0, st));
        SCHK(cacg_k_colnorm2_fast(n, m, sv->d_R, sv->d_norm2, sv->d_xr_work, st));
        mark(sv, CACG_PHASE_REFINE);
        SCHK(hipMemcpyAsync(tn2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        harvest(sv, rep);
        int finite = 1;
        for (int c = 0; c < m; c++)
            if (!isfinite(tn2[c]))
                finite = 0;
        if (!finite) {
            status = CACG_BREAKDOWN;
            break;
        }
        if (stop_met(o, tn2, h_bnorm2, m)) {
            status = CACG_OK;
            break;
        }
        if (it >= o->maxiter) {
            status = CACG_MAXITER;
            break;
        }
        if (use64) {
            int below = 1;
            for (int c = 0; c < m; c++)
                if (!(sqrt(tn2[c]) < o->fp32_basis_below * sqrt(h_bnorm2[c])))
                    below = 0;
            if (below)
                use64 = 0;
        }
        if (it > 0)
            refinements++;
        for (int c = 0; c < m; c++) {
            const double red = eta * sqrt(tn2[c]);
            target[c] = red > fin[c] ? red : fin[c];
        }
        if (restart)
            fresh = 1;
        since = 0;
        int broke = 0;
        while (since < inner_cap && it < o->maxiter) {
            mark(sv, CACG_PHASE_OTHER);
            if (fresh && xpend) { /* Q is about to be replaced: apply its pending X update */
                SCHK(cacg_k_xrupdate_split(n, sm, m, Q, NULL, sv->d_alpha, d_X, NULL, NULL, NULL,
                                           1, st));
                xpend = 0;
            }
            if (use64) {
                /* FP64 basis, row-major; a fresh one is written straight into Q */
                double *SB = fresh ? Q : sv->d_S;
                if (s == 1) {
                    SCHK(hipMemcpyAsync(SB, sv->d_R, sizeof(double) * (size_t)n * m,
                                        hipMemcpyDeviceToDevice, st));
                } else {
                    SCHK(cacg_k_spmm_x(&sv->A, m, sv->d_R, m, SB + m, sm, CACG_EPI_CHEB1, sa, sb,
                                       NULL, 0, NULL, 0, SB, sm, &sv->sch_basis, st));
                    mark(sv, CACG_PHASE_BASIS1);
                    for (int j = 2; j < s; j++)
                        SCHK(cacg_k_spmm_x(&sv->A, m, SB + (long)(j - 1) * m, sm, SB + (long)j * m,
                                           sm, CACG_EPI_CHEB, sa2, sb2, SB + (long)(j - 1) * m, sm,
                                           SB + (long)(j - 2) * m, sm, NULL, 0, &sv->sch_basis,
                                           st));
                }
            } else if (s == 1) {
                SCHK(cacg_k_cvt_f64_f32((long)n * m, sv->d_R, S, st));
            } else {
                SCHK(cacg_k_cheb1_f32_x(&sv->A32, m, sv->d_R, m, S + m, sm, S, sm, sa, sb,
                                        &sv->sch_basis, st));
                mark(sv, CACG_PHASE_BASIS1);
                for (int j = 2; j < s; j++)
                    SCHK(cacg_k_spmm_f32_x(&sv->A32, m, S + (long)(j - 1) * m, sm,
                                           S + (long)j * m, sm, CACG_EPI_CHEB, sa2, sb2,
                                           S + (long)(j - 1) * m, sm, S + (long)(j - 2) * m, sm,
                                           &sv->sch_basis, st));
            }
            if (s != 2)
                mark(sv, s == 1 ? CACG_PHASE_BASIS1 : CACG_PHASE_BASIS);
            if (fresh) {
                if (!use64)
                    SCHK(cacg_k_cvt_f32_f64((long)n * sm, S, Q, st));
                fresh = 0;
            } else if (use64) {
                SCHK(cacg_k_gram_fast(n, AQ, sm, sm, sv->d_S, sm, sm, NULL, 0, 0, sv->d_QtAS, NULL,
                                      sv->d_gram_work, st));
                mark(sv, CACG_PHASE_GRAM_AS);
                SCHK(cacg_k_lu_solve(sm, sv->d_QtAQ, sm, sv->d_QtAS, sv->d_beta, sv->d_lu_work,
                                     sv->d_flag, CACG_MODE_FAST, 1, st));
                mark(sv, CACG_PHASE_SOLVE);
                if (xpend)
                    SCHK(cacg_k_qxupdate(n, sm, m, sv->d_S, Q, sv->d_beta, sv->d_alpha, d_X, st));
                else
                    SCHK(cacg_k_qupdate(n, sm, sv->d_S, Q, sv->d_beta, CACG_MODE_FAST, st));
                xpend = 0;
            } else {
                SCHK(cacg_k_gram_fast_f64f32(n, AQ, sm, sm, S, sm, sm, NULL, 0, 0, sv->d_QtAS, NULL,
                                             sv->d_gram_work, st));
                mark(sv, CACG_PHASE_GRAM_AS);
                SCHK(cacg_k_lu_solve(sm, sv->d_QtAQ, sm, sv->d_QtAS, sv->d_beta, sv->d_lu_work,
                                     sv->d_flag, CACG_MODE_FAST, 1, st));
                mark(sv, CACG_PHASE_SOLVE);
                if (xpend)
                    SCHK(cacg_k_qxupdate_s32(n, sm, m, S, Q, sv->d_beta, sv->d_alpha, d_X, st));
                else
                    SCHK(cacg_k_qupdate_s32(n, sm, S, Q, sv->d_beta, st));
                xpend = 0;
            }
            mark(sv, CACG_PHASE_QUPDATE);
            SCHK(cacg_k_spmm_x(&sv->A, sm, Q, sm, AQ, sm, CACG_EPI_NONE, 0.0, 0.0, NULL, 0, NULL,
                               0, NULL, 0, &sv->sch_aq, st));
            mark(sv, CACG_PHASE_AQ);
            SCHK(cacg_k_gram_fast(n, Q, sm, sm, AQ, sm, sm, sv->d_R, m, m, sv->d_QtAQ, sv->d_QtR,
                                  sv->d_gram_work, st));
            mark(sv, CACG_PHASE_GRAM);
            SCHK(cacg_k_lu_solve(sm, sv->d_QtAQ, m, sv->d_QtR, sv->d_alpha, sv->d_lu_work,
                                 sv->d_flag, CACG_MODE_FAST, 0, st));
            mark(sv, CACG_PHASE_SOLVE);
            if (split) {
                SCHK(cacg_k_xrupdate_split(n, sm, m, NULL, AQ, sv->d_alpha, NULL, sv->d_R,
                                           sv->d_norm2, sv->d_xr_work, 2, st));
                xpend = 1;
            } else {
                SCHK(cacg_k_xrupdate(n, sm, m, Q, AQ, sv->d_alpha, d_X, sv->d_R, sv->d_norm2,
                                     sv->d_xr_work, CACG_MODE_FAST, st));
            }
            mark(sv, CACG_PHASE_XRUPDATE);
            SCHK(hipMemcpyAsync(h_norm2, sv->d_norm2, sizeof(double) * m, hipMemcpyDeviceToHost, st));
            SCHK(hipMemcpyAsync(h_flag, sv->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
            SCHK(hipStreamSynchronize(st));
            harvest(sv, rep);
            int reached = 1;
            finite = 1;
            for (int c = 0; c < m; c++) {
                const double rn = sqrt(h_norm2[c]);
                if (rep && rep->res_hist && it < rep->res_hist_cap)
                    rep->res_hist[it * m + c] = rn;
                if (!isfinite(rn))
                    finite = 0;
                if (!(rn < target[c]))
                    reached = 0;
            }
            if (o->print_period > 0 && (it % o->print_period == 0 || it == o->maxiter - 1))
                printf("Loop: %ld, R[%d]_norm=%30.30f\n", it, 1, sqrt(h_norm2[0]));
            it++;
            since++;
            if (!finite || *h_flag) {
                broke = 1;
                break;
            }
            if (reached)
                break;
        }
        if (broke) {
            status = CACG_BREAKDOWN;
            break;
        }
    }
    if (xpend) {
        SCHK(cacg_k_xrupdate_split(n, sm, m, Q, NULL, sv->d_alpha, d_X, NULL, NULL, NULL, 1, st));
        SCHK(hipStreamSynchronize(st));
    }
    if (rep && o->stop_mode == CACG_STOP_REL_PER_COLUMN) {
        double worst = 0.0;
        for (int c = 0; c < m; c++) {
            const double r = sqrt(tn2[c]) / sqrt(h_bnorm2[c]);
            if (!(r <= worst))
                worst = r;
        }
        rep->true_relres = worst;
    }
out:
    free(tn2);
    free(target);
    free(fin);
    if (rep) {
        rep->status = status;
        rep->iters = it;
        rep->refinements = refinements;
    }
    return status;
}

/* ---- host-buffer API --------------------------------------------------------------- */

int cacg_bcbcg(long n, const long *row_start, const long *col_idx, const double *val, long m,
               long s, const double *B, double *X, const cacg_opts *opts, cacg_report *report)
{
    if (!B || !X)
        return CACG_EINVAL;
    cacg_solver *sv = NULL;
    int rc = cacg_solver_create(&sv, n, row_start, col_idx, val, m, s, opts);
    if (rc != CACG_OK)
        return rc;
    double *dB = NULL, *dX = NULL;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)m;
    if (hipMalloc((void **)&dB, bytes) != hipSuccess || hipMalloc((void **)&dX, bytes) != hipSuccess) {
        hipFree(dB);
        cacg_solver_destroy(sv);
        return CACG_ENOMEM;
    }
    rc = CACG_EDEVICE;
    if (hipMemcpy(dB, B, bytes, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice) == hipSuccess) {
        rc = cacg_solver_solve_device(sv, dB, dX, report);
        if (rc != CACG_EDEVICE && rc != CACG_EINVAL && rc != CACG_ENOMEM &&
            hipMemcpy(X, dX, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = CACG_EDEVICE;
    }
    hipFree(dB);
    hipFree(dX);
    cacg_solver_destroy(sv);
    return rc;
}

int cacg_bcg(long n, const long *row_start, const long *col_idx, const double *val, long m,
             const double *B, double *X, const cacg_opts *opts, cacg_report *report)
{
    if (!B || !X)
        return CACG_EINVAL;
    if (opts && opts->precision != CACG_PREC_FP64)
        return CACG_EINVAL;
    cacg_solver *sv = NULL;
    int rc = cacg_solver_create(&sv, n, row_start, col_idx, val, m, 1, opts);
    if (rc != CACG_OK)
        return rc;
    double *dB = NULL, *dX = NULL;
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)m;
    if (hipMalloc((void **)&dB, bytes) != hipSuccess || hipMalloc((void **)&dX, bytes) != hipSuccess) {
        hipFree(dB);
        cacg_solver_destroy(sv);
        return CACG_ENOMEM;
    }
    rc = CACG_EDEVICE;
    if (hipMemcpy(dB, B, bytes, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(dX, X, bytes, hipMemcpyHostToDevice) == hipSuccess) {
        rc = cacg_solver_solve_bcg_device(sv, dB, dX, report);
        if (rc != CACG_EDEVICE && rc != CACG_EINVAL && rc != CACG_ENOMEM &&
            hipMemcpy(X, dX, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = CACG_EDEVICE;
    }
    hipFree(dB);
    hipFree(dX);
    cacg_solver_destroy(sv);
    return rc;
}

/* ---- drop-in entry points (src/cbcgKernel.h:21, src/bcbcgKernel.h:22, src/bcgKernel.h:21) */

enum { DROPIN_CBCG, DROPIN_BCBCG, DROPIN_BCG };

static void dropin(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
                   int numprocs, long maxiter, int which)
{
    static const char *names[] = {"cbcg_v1", "bcbcg_v1", "bcg_v1"};
    g_last_iters = 0;
    if (numprocs != 1 || myid != 0 || mat.start != 0 || mat.num_rows != B.local_num_row ||
        B.local_num_row != B.global_num_row || X.local_num_row != B.local_num_row ||
        X.local_num_col != B.local_num_col || mat.num_cols != mat.num_rows || !B.data || !X.data) {
        fprintf(stderr, "cacg: %s: the whole system must be local to one process (numprocs == 1)\n",
                names[which]);
        g_last_status = CACG_EINVAL;
        return;
    }
    cacg_opts o;
    cacg_default_opts(&o);
    o.stop_mode = CACG_STOP_REF_COL0_ABS;
    o.eps = epsilon;
    o.maxiter = maxiter;
    const char *ord = getenv("CACG_ORDER");
    o.order = (ord && strcmp(ord, "fast") == 0) ? CACG_ORDER_FAST : CACG_ORDER_REFERENCE;
    const char *quiet = getenv("CACG_QUIET");
    o.print_period = (quiet && quiet[0] == '1') ? 0 : 10;
    cacg_report rep;
    memset(&rep, 0, sizeof(rep));
    if (which == DROPIN_BCG)
        g_last_status = cacg_bcg(mat.num_rows, mat.row_start, mat.col_idx, mat.csrdata,
                                 B.local_num_col, B.data, X.data, &o, &rep);
    else
        g_last_status = cacg_bcbcg(mat.num_rows, mat.row_start, mat.col_idx, mat.csrdata,
                                   B.local_num_col, s, B.data, X.data, &o, &rep);
    g_last_iters = rep.iters;
    fflush(stdout);
}

void cbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
             int numprocs)
{
    /* one right-hand side, 1000-iteration cap (src/cbcgKernel.c:2) */
    if (B.local_num_col != 1) {
        fprintf(stderr, "cacg: cbcg_v1 takes a single right-hand side\n");
        g_last_status = CACG_EINVAL;
        g_last_iters = 0;
        return;
    }
    dropin(mat, B, X, s, epsilon, myid, numprocs, 1000, DROPIN_CBCG);
}

void bcbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
              int numprocs)
{
    dropin(mat, B, X, s, epsilon, myid, numprocs, 100, DROPIN_BCBCG); /* src/bcbcgKernel.c:3 */
}

void bcg_v1(csrType_local mat, denseType B, denseType X, double epsilon, int myid, int numprocs)
{
    dropin(mat, B, X, 1, epsilon, myid, numprocs, 1000, DROPIN_BCG); /* src/bcgKernel.c:2 */
}
