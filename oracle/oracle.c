/*
 * CPU oracle: restatement of sclc/cacg's s-step Chebyshev-basis (block) CG.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Loops keep the reference's
 * iteration and summation orders, because those orders decide the rounding
 * of every sum; a different order would not be the reference's trajectory.
 * Build with -O -ffp-contract=off (no fused multiply-add), as the reference
 * is built with `mpicc -O` for x86-64 without FMA (makefile:8).
 *
 * nranks > 1 emulates the reference's 1-D row-block distribution
 * (src/SparseMatrixDistribution.c:64-87: n/P rows per rank, the last rank
 * takes the remainder) with one thread per rank; each rank copies the whole
 * operand before a sparse product, as MPI_Allgatherv does
 * (src/BLAS3.c:48-50), and per-rank partial Gram blocks are summed in rank
 * order in place of MPI_Allreduce (src/BLAS3.c:721-722).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* row-block ranks                                                       */

typedef void (*rank_fn)(void *ctx, int rank, long r0, long r1);

typedef struct {
    rank_fn fn;
    void *ctx;
    int rank;
    long r0, r1;
} rank_job;

static void *rank_trampoline(void *p)
{
    rank_job *j = (rank_job *)p;
    j->fn(j->ctx, j->rank, j->r0, j->r1);
    return NULL;
}

/* rows [r0, r1) of rank p out of P over n rows (src/SparseMatrixDistribution.c:64-87) */
static void rank_rows(long n, int P, int p, long *r0, long *r1)
{
    long per = n / P;
    *r0 = per * p;
    *r1 = (p == P - 1) ? n : per * (p + 1);
}

static void run_ranks(int P, long n, rank_fn fn, void *ctx)
{
    if (P <= 1) {
        fn(ctx, 0, 0, n);
        return;
    }
    pthread_t th[P];
    rank_job jobs[P];
    for (int p = 0; p < P; p++) {
        jobs[p].fn = fn;
        jobs[p].ctx = ctx;
        jobs[p].rank = p;
        rank_rows(n, P, p, &jobs[p].r0, &jobs[p].r1);
        pthread_create(&th[p], NULL, rank_trampoline, &jobs[p]);
    }
    for (int p = 0; p < P; p++)
        pthread_join(th[p], NULL);
}

/* ------------------------------------------------------------------ */
/* kernels                                                               */

/* src/common.c:53-106.  Per row: off = sum |a_ij| (j != i, CSR order),
 * d = a_ii (0 if absent).  min over d - off, max over d + off.  The running
 * max starts at DBL_MIN (smallest positive normal), as at src/common.c:56. */
void orc_gerschgorin(const orc_csr *A, double bounds[2])
{
    double lo = DBL_MAX, hi = DBL_MIN;
    for (long i = 0; i < A->n; i++) {
        double off = 0.0, d = 0.0;
        for (long p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) {
            if (A->col[p] != i + A->row_offset)
                off += (A->val[p] > 0 ? A->val[p] : -A->val[p]);
            else
                d = A->val[p];
        }
        if (hi < off + d)
            hi = off + d;
        if (lo > d - off)
            lo = d - off;
    }
    bounds[0] = lo;
    bounds[1] = hi;
}

/* rows [r0, r1) of y = A x; each y[i, c] sums a_p * x[col_p, c] from 0 in CSR
 * order (src/BLAS3.c:78-112, 207-222, 950-965). */
static void spmm_rows(const orc_csr *A, const double *x, long k, double *y, long r0, long r1)
{
    for (long i = r0; i < r1; i++) {
        double *yi = y + i * k;
        for (long c = 0; c < k; c++)
            yi[c] = 0.0;
        for (long p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) {
            const double a = A->val[p];
            const double *xr = x + A->col[p] * k;
            for (long c = 0; c < k; c++)
                yi[c] += a * xr[c];
        }
    }
}

typedef struct {
    const orc_csr *A;
    const double *x;
    long k;
    double *y;
    long xrows;
    int gather;
} spmm_ctx;

static void spmm_rank(void *vp, int rank, long r0, long r1)
{
    (void)rank;
    spmm_ctx *c = (spmm_ctx *)vp;
    const double *x = c->x;
    double *g = NULL;
    if (c->gather) {
        /* each rank receives the whole operand (MPI_Allgatherv, src/BLAS3.c:197-200) */
        g = (double *)malloc(sizeof(double) * c->xrows * c->k);
        memcpy(g, c->x, sizeof(double) * c->xrows * c->k);
        x = g;
    }
    spmm_rows(c->A, x, c->k, c->y, r0, r1);
    free(g);
}

static void spmm_ranked(const orc_csr *A, const double *x, long k, double *y, int P)
{
    spmm_ctx c = {A, x, k, y, A->n, 1};
    run_ranks(P, A->n, spmm_rank, &c);
}

void orc_spmm(const orc_csr *A, const double *x, long k, double *y)
{
    spmm_rows(A, x, k, y, 0, A->n);
}

/* src/BLAS3.c:579-612: out[i] = a * x[i] + b * y[i] */
void orc_lincomb2(long len, double a, const double *x, double b, const double *y, double *out)
{
    for (long i = 0; i < len; i++)
        out[i] = a * x[i] + b * y[i];
}

/* src/BLAS3.c:978-994: out[i] = a x[i] + b y[i] + c z[i], evaluated left to right */
static void lincomb3(long len, double a, const double *x, double b, const double *y,
                     double c, const double *z, double *out)
{
    for (long i = 0; i < len; i++)
        out[i] = a * x[i] + b * y[i] + c * z[i];
}

/* Chebyshev basis, src/common.c:182-256 (block) and 109-179 (m = 1):
 *   S_0 = R
 *   S_1 = s_alpha (A R) + s_beta R                        (src/BLAS3.c:828-847)
 *   S_j = 2 s_alpha (A S_{j-1}) + 2 s_beta S_{j-1} - S_{j-2}   (src/BLAS3.c:978-994)
 * then the chunk transpose S[i, j m + c] = S_j[i, c]
 * (src/MatrixOperations.c:70-92).  The transpose moves values only. */
static void cheb_basis_ranked(const orc_csr *A, const double *R, long m, long s,
                              double s_alpha, double s_beta, double *S, int P)
{
    const long n = A->n, chunk = n * m, sm = s * m;
    double *chunks = (double *)calloc((size_t)chunk * s, sizeof(double));
    double *ax = (double *)calloc((size_t)chunk, sizeof(double));
    const double a2 = 2.0 * s_alpha, b2 = 2.0 * s_beta;
    for (long j = 0; j < s; j++) {
        if (j == 0) {
            memcpy(chunks, R, sizeof(double) * chunk);
        } else if (j == 1) {
            spmm_ranked(A, R, m, ax, P);
            orc_lincomb2(chunk, s_alpha, ax, s_beta, R, chunks + chunk);
        } else {
            spmm_ranked(A, chunks + (j - 1) * chunk, m, ax, P);
            lincomb3(chunk, a2, ax, b2, chunks + (j - 1) * chunk, -1.0,
                     chunks + (j - 2) * chunk, chunks + j * chunk);
        }
    }
    for (long i = 0; i < n; i++)
        for (long j = 0; j < s; j++)
            memcpy(S + i * sm + j * m, chunks + j * chunk + i * m, sizeof(double) * m);
    free(chunks);
    free(ax);
}

void orc_cheb_basis(const orc_csr *A, const double *R, long m, long s,
                    double s_alpha, double s_beta, double *S)
{
    cheb_basis_ranked(A, R, m, s, s_alpha, s_beta, S, 1);
}

/* rows [r0, r1) of G += L^T R: row index outer, then the kl rows of G, then
 * the kr columns (src/BLAS3.c:709-717, 741-749). */
static void gram_rows(const double *L, long kl, const double *R, long kr, double *G,
                      long r0, long r1)
{
    for (long i = r0; i < r1; i++) {
        const double *li = L + i * kl;
        const double *ri = R + i * kr;
        for (long a = 0; a < kl; a++) {
            const double t = li[a];
            double *ga = G + a * kr;
            for (long b = 0; b < kr; b++)
                ga[b] += ri[b] * t;
        }
    }
}

typedef struct {
    const double *L, *R;
    long kl, kr;
    double *part; /* [P][kl*kr] */
} gram_ctx;

static void gram_rank(void *vp, int rank, long r0, long r1)
{
    gram_ctx *c = (gram_ctx *)vp;
    gram_rows(c->L, c->kl, c->R, c->kr, c->part + (long)rank * c->kl * c->kr, r0, r1);
}

static void gram_ranked(long n, const double *L, long kl, const double *R, long kr,
                        double *G, int P)
{
    const long sz = kl * kr;
    if (P <= 1) {
        memset(G, 0, sizeof(double) * sz);
        gram_rows(L, kl, R, kr, G, 0, n);
        return;
    }
    gram_ctx c = {L, R, kl, kr, (double *)calloc((size_t)sz * P, sizeof(double))};
    run_ranks(P, n, gram_rank, &c);
    /* stand-in for MPI_Allreduce(SUM): rank order */
    memcpy(G, c.part, sizeof(double) * sz);
    for (int p = 1; p < P; p++)
        for (long e = 0; e < sz; e++)
            G[e] += c.part[(long)p * sz + e];
    free(c.part);
}

void orc_gram(long n, const double *L, long kl, const double *R, long kr, double *G)
{
    gram_ranked(n, L, kl, R, kr, G, 1);
}

/* rows [r0, r1) of Y = L W: row, then k, then output column (src/BLAS3.c:811-819) */
static void gemm_rows(const double *L, long k, const double *W, long j, double *Y,
                      long r0, long r1)
{
    for (long i = r0; i < r1; i++) {
        double *yi = Y + i * j;
        for (long c = 0; c < j; c++)
            yi[c] = 0.0;
        for (long q = 0; q < k; q++) {
            const double t = L[i * k + q];
            const double *wq = W + q * j;
            for (long c = 0; c < j; c++)
                yi[c] += wq[c] * t;
        }
    }
}

typedef struct {
    const double *L, *W;
    long k, j;
    double *Y;
} gemm_ctx;

static void gemm_rank(void *vp, int rank, long r0, long r1)
{
    (void)rank;
    gemm_ctx *c = (gemm_ctx *)vp;
    gemm_rows(c->L, c->k, c->W, c->j, c->Y, r0, r1);
}

static void gemm_ranked(long n, const double *L, long k, const double *W, long j, double *Y, int P)
{
    gemm_ctx c = {L, W, k, j, Y};
    run_ranks(P, n, gemm_rank, &c);
}

void orc_gemm_ts(long n, const double *L, long k, const double *W, long j, double *Y)
{
    gemm_rows(L, k, W, j, Y, 0, n);
}

/* src/DLSsolvers.c:106-208.  Crout factorisation with unit-diagonal U and no
 * pivoting on a copy of lhs (U scaled by the reciprocal pivot, then a
 * right-looking trailing update), forward substitution (multiply by the
 * reciprocal of L_ii), backward substitution with ascending k. */
void orc_lu_solve(long N, const double *lhs, long k, const double *rhs, double *x)
{
    double *a = (double *)malloc(sizeof(double) * N * N);
    double *y = (double *)calloc((size_t)N * k, sizeof(double));
    memcpy(a, lhs, sizeof(double) * N * N);
    for (long p = 0; p < N; p++) {
        const double inv = 1.0 / a[p * N + p];
        for (long c = p + 1; c < N; c++)
            a[p * N + c] *= inv;
        for (long r = p + 1; r < N; r++) {
            const double l = a[r * N + p];
            for (long c = p + 1; c < N; c++)
                a[r * N + c] = a[r * N + c] - l * a[p * N + c];
        }
    }
    for (long i = 0; i < N; i++) {
        for (long c = 0; c < k; c++)
            y[i * k + c] = rhs[i * k + c];
        for (long q = 0; q < i; q++) {
            const double l = a[i * N + q];
            for (long c = 0; c < k; c++)
                y[i * k + c] -= y[q * k + c] * l;
        }
        const double inv = 1.0 / a[i * N + i];
        for (long c = 0; c < k; c++)
            y[i * k + c] *= inv;
    }
    for (long c = 0; c < k; c++)
        x[(N - 1) * k + c] = y[(N - 1) * k + c];
    for (long i = N - 2; i >= 0; i--) {
        for (long c = 0; c < k; c++)
            x[i * k + c] = y[i * k + c];
        for (long q = i + 1; q < N; q++) {
            const double u = a[i * N + q];
            for (long c = 0; c < k; c++)
                x[i * k + c] -= x[q * k + c] * u;
        }
    }
    free(a);
    free(y);
}

/* src/MatrixOperations.c:110-123 */
double orc_col_norm2sq(long n, long m, const double *R, long col)
{
    double acc = 0.0;
    for (long i = 0; i < n; i++) {
        const double t = R[i * m + col];
        acc += t * t;
    }
    return acc;
}

/* ------------------------------------------------------------------ */
/* solver                                                                */

typedef struct {
    double *R, *AX, *S, *Q, *AQ, *QtAQ, *QtAS, *QtR, *alpha, *beta;
    double *Qalpha, *AQalpha, *Qbeta;
} orc_work;

static int work_alloc(orc_work *w, long n, long m, long sm)
{
    memset(w, 0, sizeof(*w));
    w->R = (double *)calloc((size_t)n * m, sizeof(double));
    w->AX = (double *)calloc((size_t)n * m, sizeof(double));
    w->S = (double *)calloc((size_t)n * sm, sizeof(double));
    w->Q = (double *)calloc((size_t)n * sm, sizeof(double));
    w->AQ = (double *)calloc((size_t)n * sm, sizeof(double));
    w->QtAQ = (double *)calloc((size_t)sm * sm, sizeof(double));
    w->QtAS = (double *)calloc((size_t)sm * sm, sizeof(double));
    w->QtR = (double *)calloc((size_t)sm * m, sizeof(double));
    w->alpha = (double *)calloc((size_t)sm * m, sizeof(double));
    w->beta = (double *)calloc((size_t)sm * sm, sizeof(double));
    w->Qalpha = (double *)calloc((size_t)n * m, sizeof(double));
    w->AQalpha = (double *)calloc((size_t)n * m, sizeof(double));
    w->Qbeta = (double *)calloc((size_t)n * sm, sizeof(double));
    return w->R && w->AX && w->S && w->Q && w->AQ && w->QtAQ && w->QtAS && w->QtR &&
           w->alpha && w->beta && w->Qalpha && w->AQalpha && w->Qbeta;
}

static void work_free(orc_work *w)
{
    double **p = (double **)w;
    for (size_t i = 0; i < sizeof(*w) / sizeof(double *); i++)
        free(p[i]);
}

/* one outer iteration (src/bcbcgKernel.c:180-437) */
static void outer_iteration(const orc_csr *A, long m, long s, long it, double sa, double sb,
                            double *X, orc_work *w, int P)
{
    const long n = A->n, sm = s * m;
    cheb_basis_ranked(A, w->R, m, s, sa, sb, w->S, P);
    if (it == 0) {
        /* Q takes the basis buffer (src/bcbcgKernel.c:199-203) */
        double *t = w->Q;
        w->Q = w->S;
        w->S = t;
    } else {
        gram_ranked(n, w->AQ, sm, w->S, sm, w->QtAS, P);        /* :214 */
        orc_lu_solve(sm, w->QtAQ, sm, w->QtAS, w->beta);        /* :229 */
        gemm_ranked(n, w->Q, sm, w->beta, sm, w->Qbeta, P);     /* :249 */
        orc_lincomb2(n * sm, 1.0, w->S, -1.0, w->Qbeta, w->Q);  /* :262 */
    }
    spmm_ranked(A, w->Q, sm, w->AQ, P);                         /* :284 */
    gram_ranked(n, w->Q, sm, w->AQ, sm, w->QtAQ, P);            /* :307 */
    gram_ranked(n, w->Q, sm, w->R, m, w->QtR, P);               /* :330 */
    orc_lu_solve(sm, w->QtAQ, m, w->QtR, w->alpha);             /* :344 */
    gemm_ranked(n, w->Q, sm, w->alpha, m, w->Qalpha, P);        /* :376 */
    gemm_ranked(n, w->AQ, sm, w->alpha, m, w->AQalpha, P);      /* :390 */
    orc_lincomb2(n * m, 1.0, X, 1.0, w->Qalpha, X);             /* :414 */
    orc_lincomb2(n * m, 1.0, w->R, -1.0, w->AQalpha, w->R);     /* :437 */
}

static void setup(const orc_csr *A, long m, const double *B, const double *X, orc_work *w,
                  int P, double bounds[2], double *sa, double *sb)
{
    orc_gerschgorin(A, bounds);                                 /* src/bcbcgKernel.c:106 */
    *sa = 2.0 / (bounds[1] - bounds[0]);                        /* :113-115 */
    *sb = -(bounds[1] + bounds[0]) / (bounds[1] - bounds[0]);
    spmm_ranked(A, X, m, w->AX, P);                             /* :136 */
    orc_lincomb2(A->n * m, 1.0, B, -1.0, w->AX, w->R);          /* :154 */
}

int orc_bcbcg(const orc_csr *A, long m, long s, const double *B, double *X,
              const orc_opts *o, long *iters)
{
    const long n = A->n, sm = s * m;
    const int P = o->nranks > 0 ? o->nranks : 1;
    if (n <= 0 || m <= 0 || s <= 0 || o->maxiter <= 0)
        return ORC_INVAL;
    orc_work w;
    if (!work_alloc(&w, n, m, sm)) {
        work_free(&w);
        return ORC_INVAL;
    }
    double bounds[2], sa, sb;
    setup(A, m, B, X, &w, P, bounds, &sa, &sb);
    if (o->bounds) {
        o->bounds[0] = bounds[0];
        o->bounds[1] = bounds[1];
    }
    double *bnorm = (double *)calloc((size_t)m, sizeof(double));
    for (long c = 0; c < m; c++)
        bnorm[c] = sqrt(orc_col_norm2sq(n, m, B, c));

    int status = ORC_MAXITER;
    long it;
    for (it = 0; it < o->maxiter; it++) {
        outer_iteration(A, m, s, it, sa, sb, X, &w, P);
        int done = 1, finite = 1;
        for (long c = 0; c < m; c++) {
            const double rn = sqrt(orc_col_norm2sq(n, m, w.R, c));
            if (o->res_hist)
                o->res_hist[it * m + c] = rn;
            if (!isfinite(rn))
                finite = 0;
            if (o->stop_mode == ORC_STOP_REL_PER_COLUMN && !(rn < o->tol * bnorm[c]))
                done = 0;
        }
        if (o->x_hist)
            memcpy(o->x_hist + it * n * m, X, sizeof(double) * n * m);
        if (o->stop_mode == ORC_STOP_REF_COL0_ABS)  /* src/bcbcgKernel.c:445-447, 524-526 */
            done = sqrt(orc_col_norm2sq(n, m, w.R, 0)) < o->eps;
        if (done) {
            status = ORC_OK;
            it++;
            break;
        }
        if (!finite) {
            status = ORC_BREAKDOWN;
            it++;
            break;
        }
    }
    *iters = it;
    free(bnorm);
    work_free(&w);
    return status;
}

/* src/bcgKernel.c:22-462 */
int orc_bcg(const orc_csr *A, long m, const double *B, double *X, const orc_opts *o, long *iters)
{
    const long n = A->n;
    const int P = o->nranks > 0 ? o->nranks : 1;
    if (n <= 0 || m <= 0 || o->maxiter <= 0)
        return ORC_INVAL;
    const size_t nm = (size_t)n * m, mm = (size_t)m * m;
    double *R = calloc(nm, sizeof(double)), *AX = calloc(nm, sizeof(double));
    double *Pm = calloc(nm, sizeof(double)), *AP = calloc(nm, sizeof(double));
    double *Pa = calloc(nm, sizeof(double)), *APa = calloc(nm, sizeof(double));
    double *Pb = calloc(nm, sizeof(double));
    double *PtAP = calloc(mm, sizeof(double)), *RtR = calloc(mm, sizeof(double));
    double *RnRn = calloc(mm, sizeof(double)), *alpha = calloc(mm, sizeof(double));
    double *beta = calloc(mm, sizeof(double)), *bnorm = calloc((size_t)m, sizeof(double));
    int status = ORC_INVAL;
    long it = 0;
    if (!R || !AX || !Pm || !AP || !Pa || !APa || !Pb || !PtAP || !RtR || !RnRn || !alpha ||
        !beta || !bnorm)
        goto out;
    for (long c = 0; c < m; c++)
        bnorm[c] = sqrt(orc_col_norm2sq(n, m, B, c));
    spmm_ranked(A, X, m, AX, P);                                /* :89 */
    orc_lincomb2((long)nm, 1.0, B, -1.0, AX, R);                /* :127 */
    memcpy(Pm, R, nm * sizeof(double));                         /* :138 */
    gram_ranked(n, R, m, R, m, RtR, P);                         /* :178 */
    status = ORC_MAXITER;
    for (it = 0; it < o->maxiter; it++) {
        spmm_ranked(A, Pm, m, AP, P);                           /* :192 */
        gram_ranked(n, Pm, m, AP, m, PtAP, P);                  /* :213 */
        orc_lu_solve(m, PtAP, m, RtR, alpha);                   /* :239 */
        gemm_ranked(n, Pm, m, alpha, m, Pa, P);                 /* :258 */
        orc_lincomb2((long)nm, 1.0, X, 1.0, Pa, X);             /* :278 */
        gemm_ranked(n, AP, m, alpha, m, APa, P);                /* :291 */
        int done = 1, finite = 1;
        for (long c = 0; c < m; c++) {
            const double rn = sqrt(orc_col_norm2sq(n, m, R, c));
            if (o->res_hist)
                o->res_hist[it * m + c] = rn;
            if (!isfinite(rn))
                finite = 0;
            if (o->stop_mode == ORC_STOP_REL_PER_COLUMN && !(rn < o->tol * bnorm[c]))
                done = 0;
        }
        if (o->x_hist)
            memcpy(o->x_hist + it * nm, X, nm * sizeof(double));
        if (o->stop_mode == ORC_STOP_REF_COL0_ABS)              /* :301-365 */
            done = sqrt(orc_col_norm2sq(n, m, R, 0)) < o->eps;
        if (done) {
            status = ORC_OK;
            it++;
            break;
        }
        if (!finite) {
            status = ORC_BREAKDOWN;
            it++;
            break;
        }
        orc_lincomb2((long)nm, 1.0, R, -1.0, APa, R);           /* :375 */
        gram_ranked(n, R, m, R, m, RnRn, P);                    /* :389 */
        orc_lu_solve(m, RtR, m, RnRn, beta);                    /* :404 */
        double *t = RtR;                                        /* :419 */
        RtR = RnRn;
        RnRn = t;
        gemm_ranked(n, Pm, m, beta, m, Pb, P);                  /* :428 */
        orc_lincomb2((long)nm, 1.0, R, 1.0, Pb, Pm);            /* :443 */
    }
out:
    *iters = it;
    free(R);
    free(AX);
    free(Pm);
    free(AP);
    free(Pa);
    free(APa);
    free(Pb);
    free(PtAP);
    free(RtR);
    free(RnRn);
    free(alpha);
    free(beta);
    free(bnorm);
    return status;
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int orc_bcbcg_time_sample(const orc_csr *A, long m, long s, const double *B, double *X,
                          int nranks, long iters, double *t_setup, double *t_iters)
{
    const long n = A->n, sm = s * m;
    const int P = nranks > 0 ? nranks : 1;
    orc_work w;
    double t0 = now_s();
    if (!work_alloc(&w, n, m, sm)) {
        work_free(&w);
        return ORC_INVAL;
    }
    double bounds[2], sa, sb;
    setup(A, m, B, X, &w, P, bounds, &sa, &sb);
    double t1 = now_s();
    *t_setup = t1 - t0;
    for (long it = 0; it < iters; it++) {
        outer_iteration(A, m, s, it, sa, sb, X, &w, P);
        const double t2 = now_s();
        t_iters[it] = t2 - t1;
        t1 = t2;
    }
    work_free(&w);
    return ORC_OK;
}
