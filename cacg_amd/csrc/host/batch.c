/*
 * The batching / serving engine: many s-step block-CG systems resident on
 * one GPU, advanced together one outer iteration per round.
 *
 * Memory: one device arena (cacg_batch_opts.hbm_bytes) split first-fit into
 * one block per admitted system (its CSR, FP32 values, B, X, R, S, Q, A Q,
 * small blocks, Gram and norm partials), so admission and retirement never
 * call hipMalloc / hipFree (hipFree would synchronise the device).  Slot
 * descriptors (cacg_bsys) live in device memory; the column norms and
 * breakdown flags of all slots sit in two slot-indexed arrays that come back
 * to the host in one copy per round.
 *
 * A round (per system, the iteration of src/bcbcgKernel.c:172-526):
 *   basis S from R; first iteration Q = S, else QtAS = AQ^T S, beta,
 *   Q = S - Q beta; AQ = A Q; [QtAQ | QtR], alpha; X += Q alpha;
 *   then for every admitted system R = B - A X (FP64) and its column norms.
 * The recomputed R replaces the recursive update R -= AQ alpha, which keeps
 * the stop test on the true residual and needs no extra pass.  Systems of
 * SOLO_ROWS rows or more instead run the single-system kernels on streams of
 * their own, iterate on the recursive residual as the single-system solver
 * does, and recompute the true residual in one extra round when the recursive
 * one meets the tolerance (or the cap): the true residual still decides.  Newly
 * admitted systems join at the residual step.  Work lists of {slot, tile}
 * items are built on the host each round and uploaded in one copy.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "cacg.h"
#include "cacg_batch.h"
#include "cacg_kernels.h"

enum { RQ_QUEUED = 0, RQ_ACTIVE = 1, RQ_DONE = 2 };

typedef struct {
    long id;
    int state, status, slot;
    int joined;          /* 0: admitted this round (residual only), 1: iterating */
    int fresh;           /* the next iteration starts the recurrence (Q = S) */
    int verify;          /* solo: this round recomputes the true residual only */
    int xpend;           /* solo: X += Q alpha of the last iteration still to apply */
    int ran;             /* solo: launched this round (admitted earlier) */
    long n, nnz, m, iters, round_submitted, rounds_queued;
    long bypassed;       /* queued: requests submitted after this one admitted before it */
    double tol, sa, sb, true_relres;
    int *rp, *ci;        /* host copies, int32 */
    double *val, *B, *X; /* X: X0 in, the solution out */
    double *bn2;         /* ||B(:,c)||^2 */
    long off, bytes;     /* arena block */
} breq;

typedef struct {
    long off, size;
} fblk;

/* streams for the solo systems (default priority), so that several of them and
 * the batched launches run at once; with the main stream they use the four
 * hardware queues a process gets */
#define NSOLO 3

/* a round's work lists: offsets and counts into the uploaded {slot, item} array
 * (order: iterating tiles | fresh tiles | Q-update tiles | X-update tiles |
 * Gram chunks by class | Gram0 blocks | LU0 | Gram1 blocks | LU1 | all tiles |
 * all systems), with what else fixes the round's launches */
typedef struct {
    long k;
    long o_it, n_it, o_fr, n_fr;
    long o_qc[9][2], n_qc[9][2], o_xc[9][2][2], n_xc[9][2][2];
    long o_gc[2][25], n_gc[2][25];
    long o_b0, n_b0, o_lu[2][2], n_lu[2][2], o_b1, n_b1, o_all, n_all, o_sys, n_sys;
    int s, f32, max_sm, pad_;
} wlists;

/* Online SpMM schedule tuning of a solo system: its first NCAND iterations each
 * run their basis steps and A Q under one candidate schedule (tile order and
 * grid, cacg_spmm_sched), timed by events on the system's stream, and the
 * fastest of each shape is kept for the rest of the solve.  The timed launches
 * are the solve's own (a schedule changes no bit), so tuning costs no extra
 * work; the candidates are those of the single-system solver's offline tuner. */
static const cacg_spmm_sched CAND[] = {
    {0, 0, 0},       {4096, 0, 0},    {8192, 0, 0},    {0, 1024, 0},   {0, 2048, 0},
    {0, 4096, 0},    {4096, 2048, 0}, {8192, 4096, 0}, {16384, 512, 4}, {32768, 512, 4}};
#define NCAND ((int)(sizeof(CAND) / sizeof(CAND[0])))

typedef struct {
    hipEvent_t ev[4];          /* basis steps start / end, A Q start / end */
    int have_ev;
    int timing;                /* the candidate timed this round, -1: none */
    int next;                  /* the next candidate; NCAND: done */
    float best_b, best_a;
    cacg_spmm_sched sb, sa;    /* in use: basis steps, A Q */
} stune;

struct cacg_batch {
    cacg_batch_opts o;
    hipStream_t st;
    hipStream_t sst[NSOLO];
    hipEvent_t sev[NSOLO + 1];
    char *arena;
    long arena_bytes, arena_used;
    fblk *fl;            /* free list, sorted by offset */
    int nfl, capfl;
    int nslots;
    breq **slot;
    cacg_bsys *h_sys, *d_sys;
    double *d_norms, *h_norms;   /* nslots x CACG_BATCH_MAX_M */
    int *d_flags, *h_flags;      /* nslots */
    int2 *h_items, *d_items;
    long items_cap;
    breq **req;          /* by id; NULL once collected */
    long nreq, capreq, qhead;
    long rounds;
    long *ret;           /* ids retired and not yet reported by cacg_batch_retired */
    long nret, capret;
    /* the last uploaded work lists (the graph key) and the graph of their launches */
    wlists g_wl;
    int2 *g_items, *g_ditems;
    long g_cap;
    int g_valid;
    hipGraphExec_t g_exec;
    long graph_rounds;
    stune *tune;                /* per slot */
    int *act;                   /* the batched active slots of a round */
    /* guards the request table (req, nreq, capreq, rounds) and everything a
     * round touches, so that cacg_batch_submit may run on another thread while
     * a round waits for the GPU (the round releases it for that wait only) */
    pthread_mutex_t mu;
    int mu_init;
};

#define HCHK(call)                                                                               \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "cacg_batch: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                         \
            return CACG_EDEVICE;                                                                 \
        }                                                                                        \
    } while (0)

#define KCHK(call)                                                                               \
    do {                                                                                         \
        int e_ = (call);                                                                         \
        if (e_ != 0) {                                                                           \
            fprintf(stderr, "cacg_batch: %s failed: %s (%s:%d)\n", #call,                        \
                    hipGetErrorString((hipError_t)e_), __FILE__, __LINE__);                      \
            return CACG_EDEVICE;                                                                 \
        }                                                                                        \
    } while (0)

static long al256(long x) { return (x + 255) & ~255L; }

/* Systems of at least this many rows iterate with the single-system kernels
 * (vector-per-row SpMM, wide split-K Gram, MFMA tall-skinny updates) on their
 * arena block instead of joining the batched launches, which are built for
 * many small systems: alone, a 2-million-row system runs 3-4.5 x slower per
 * iteration through them.  The choice depends on the system's shape and the
 * batch's options only, so a result is still the same alone and in any mix of
 * one batch configuration.  The threshold is cacg_batch_opts.solo_rows,
 * SOLO_ROWS by default. */
#define SOLO_ROWS (1L << 17)

static long solo_rows(const cacg_batch_opts *o) { return o->solo_rows > 0 ? o->solo_rows : SOLO_ROWS; }
/* a system of this batch that runs the single-system kernels on a solo stream */
#define IS_SOLO(nrows) ((nrows) >= solo_rows(&b->o))

/* ---- per-system layout in its arena block ----------------------------------------- */

typedef struct {
    long rowptr, col, val, val32, B, X, R, S, Q, AQ, QtAQ, QtAS, beta, lu, QtR, alpha, gpart,
        npart, total;
} layout;

static layout make_layout(long n, long nnz, long m, int s, int precision, long solo_min)
{
    layout L;
    const long sm = s * m, ntile = (n + CACG_BT - 1) / CACG_BT, nch = (n + CACG_GR - 1) / CACG_GR;
    const int f32 = precision == CACG_PREC_FP32_REFINE;
    long o = 0;
#define PUT(field, bytes)                                                                        \
    do {                                                                                         \
        L.field = o;                                                                             \
        o += al256(bytes);                                                                       \
    } while (0)
    PUT(rowptr, (n + 1) * 4);
    PUT(col, nnz * 4);
    PUT(val, nnz * 8);
    PUT(val32, f32 ? nnz * 4 : 0);
    PUT(B, n * m * 8);
    PUT(X, n * m * 8);
    PUT(R, n * m * 8);
    PUT(S, n * sm * (f32 ? 4 : 8));
    PUT(Q, n * sm * 8);
    PUT(AQ, n * sm * 8);
    PUT(QtAQ, sm * sm * 8);
    PUT(QtAS, sm * sm * 8);
    PUT(beta, sm * sm * 8);
    PUT(lu, sm * sm * 8);
    PUT(QtR, sm * m * 8);
    PUT(alpha, sm * m * 8);
    {
        long gp = nch * sm * (sm + m);
        if (n >= solo_min) { /* the single-system Gram's split-K partials */
            const long ws = cacg_gram_workspace((int)sm, (int)(sm + m));
            if (ws > gp)
                gp = ws;
        }
        PUT(gpart, gp * 8);
    }
    {
        /* per-tile norm partials of the batched kernels; a large system's
         * single-system norm kernels write cacg_xrupdate_workspace(m) of them
         * whatever its row count */
        long np = ntile * m;
        if (n >= solo_min && cacg_xrupdate_workspace((int)m) > np)
            np = cacg_xrupdate_workspace((int)m);
        PUT(npart, np * 8);
    }
#undef PUT
    L.total = o;
    return L;
}

long cacg_batch_footprint(long n, long nnz, long m, int s, int precision)
{
    if (n <= 0 || nnz < 0 || m <= 0 || s <= 0)
        return -1;
    return make_layout(n, nnz, m, s, precision, SOLO_ROWS).total;
}

/* ---- arena: first fit over a sorted free list, neighbours merged on free ------------- */

static long arena_alloc(cacg_batch *b, long bytes)
{
    bytes = al256(bytes);
    for (int i = 0; i < b->nfl; i++)
        if (b->fl[i].size >= bytes) {
            const long off = b->fl[i].off;
            b->fl[i].off += bytes;
            b->fl[i].size -= bytes;
            if (b->fl[i].size == 0) {
                memmove(b->fl + i, b->fl + i + 1, sizeof(fblk) * (size_t)(b->nfl - i - 1));
                b->nfl--;
            }
            b->arena_used += bytes;
            return off;
        }
    return -1;
}

static int arena_free(cacg_batch *b, long off, long bytes)
{
    bytes = al256(bytes);
    int i = 0;
    while (i < b->nfl && b->fl[i].off < off)
        i++;
    if (b->nfl == b->capfl) {
        fblk *nf = (fblk *)realloc(b->fl, sizeof(fblk) * (size_t)(2 * b->capfl));
        if (!nf)
            return CACG_ENOMEM;
        b->fl = nf;
        b->capfl *= 2;
    }
    memmove(b->fl + i + 1, b->fl + i, sizeof(fblk) * (size_t)(b->nfl - i));
    b->fl[i].off = off;
    b->fl[i].size = bytes;
    b->nfl++;
    if (i + 1 < b->nfl && b->fl[i].off + b->fl[i].size == b->fl[i + 1].off) {
        b->fl[i].size += b->fl[i + 1].size;
        memmove(b->fl + i + 1, b->fl + i + 2, sizeof(fblk) * (size_t)(b->nfl - i - 2));
        b->nfl--;
    }
    if (i > 0 && b->fl[i - 1].off + b->fl[i - 1].size == b->fl[i].off) {
        b->fl[i - 1].size += b->fl[i].size;
        memmove(b->fl + i, b->fl + i + 1, sizeof(fblk) * (size_t)(b->nfl - i - 1));
        b->nfl--;
    }
    b->arena_used -= bytes;
    return CACG_OK;
}

/* ---- lifetime ------------------------------------------------------------------------ */

void cacg_batch_default_opts(cacg_batch_opts *o)
{
    memset(o, 0, sizeof(*o));
    o->hbm_bytes = 0;
    o->slots = 1024;
    o->s = 4;
    o->precision = CACG_PREC_FP32_REFINE;
    o->maxiter = 1000;
    o->max_large = 0;
    o->graph = 0;
    o->tune = 1;
    o->policy = CACG_BATCH_FIFO;
    o->max_bypass = 0;
}

static void free_req(breq *r)
{
    if (!r)
        return;
    free(r->rp);
    free(r->ci);
    free(r->val);
    free(r->B);
    free(r->X);
    free(r->bn2);
    free(r);
}

int cacg_batch_destroy(cacg_batch *b)
{
    if (!b)
        return CACG_OK;
    if (b->st)
        hipStreamSynchronize(b->st);
    if (b->g_exec)
        hipGraphExecDestroy(b->g_exec);
    free(b->g_items);
    for (long i = 0; i < b->nreq; i++)
        free_req(b->req[i]);
    free(b->req);
    free(b->ret);
    free(b->slot);
    free(b->act);
    if (b->tune) {
        for (int i = 0; i < b->nslots; i++)
            if (b->tune[i].have_ev)
                for (int e = 0; e < 4; e++)
                    hipEventDestroy(b->tune[i].ev[e]);
        free(b->tune);
    }
    free(b->fl);
    if (b->arena)
        hipFree(b->arena);
    if (b->d_sys)
        hipFree(b->d_sys);
    if (b->d_norms)
        hipFree(b->d_norms);
    if (b->d_flags)
        hipFree(b->d_flags);
    if (b->d_items)
        hipFree(b->d_items);
    if (b->h_sys)
        hipHostFree(b->h_sys);
    if (b->h_norms)
        hipHostFree(b->h_norms);
    if (b->h_flags)
        hipHostFree(b->h_flags);
    if (b->h_items)
        hipHostFree(b->h_items);
    for (int i = 0; i < NSOLO; i++)
        if (b->sst[i]) {
            hipStreamSynchronize(b->sst[i]);
            hipStreamDestroy(b->sst[i]);
        }
    for (int i = 0; i <= NSOLO; i++)
        if (b->sev[i])
            hipEventDestroy(b->sev[i]);
    if (b->st)
        hipStreamDestroy(b->st);
    if (b->mu_init)
        pthread_mutex_destroy(&b->mu);
    free(b);
    return CACG_OK;
}

static int grow_items(cacg_batch *b, long need)
{
    if (need <= b->items_cap)
        return CACG_OK;
    long cap = b->items_cap > 0 ? b->items_cap : 65536;
    while (cap < need)
        cap *= 2;
    if (b->h_items)
        hipHostFree(b->h_items);
    if (b->d_items)
        hipFree(b->d_items);
    b->h_items = NULL;
    b->d_items = NULL;
    b->items_cap = 0;
    if (hipHostMalloc((void **)&b->h_items, sizeof(int2) * (size_t)cap, 0) != hipSuccess ||
        hipMalloc((void **)&b->d_items, sizeof(int2) * (size_t)cap) != hipSuccess)
        return CACG_ENOMEM;
    b->items_cap = cap;
    return CACG_OK;
}

int cacg_batch_create(cacg_batch **out, const cacg_batch_opts *opts)
{
    if (!out)
        return CACG_EINVAL;
    *out = NULL;
    cacg_batch_opts o;
    if (opts)
        o = *opts;
    else
        cacg_batch_default_opts(&o);
    if (o.slots == 0)
        o.slots = 1024;
    if (o.slots < 1 || o.slots > 4096 || o.s < 1 || o.maxiter <= 0 || o.hbm_bytes < 0 ||
        o.max_large < 0 || (o.graph != 0 && o.graph != 1) ||
        (o.tune != 0 && o.tune != 1) || o.max_bypass < 0 || o.solo_rows < 0 ||
        (o.policy != CACG_BATCH_FIFO && o.policy != CACG_BATCH_SMALLEST_FIRST) ||
        (o.precision != CACG_PREC_FP64 && o.precision != CACG_PREC_FP32_REFINE))
        return CACG_EINVAL;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
        return CACG_EDEVICE;
    cacg_batch *b = (cacg_batch *)calloc(1, sizeof(*b));
    if (!b)
        return CACG_ENOMEM;
    if (pthread_mutex_init(&b->mu, NULL) != 0) {
        free(b);
        return CACG_ENOMEM;
    }
    b->mu_init = 1;
    b->o = o;
    b->nslots = o.slots;
    if (o.hbm_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            cacg_batch_destroy(b);
            return CACG_EDEVICE;
        }
        b->o.hbm_bytes = (long)(fr / 4) & ~255L;
    }
    b->arena_bytes = al256(b->o.hbm_bytes);
    b->capfl = 64;
    b->fl = (fblk *)malloc(sizeof(fblk) * (size_t)b->capfl);
    b->slot = (breq **)calloc((size_t)b->nslots, sizeof(breq *));
    b->tune = (stune *)calloc((size_t)b->nslots, sizeof(stune));
    b->act = (int *)calloc((size_t)b->nslots, sizeof(int));
    b->capreq = 1024;
    b->req = (breq **)calloc((size_t)b->capreq, sizeof(breq *));
    if (!b->fl || !b->slot || !b->tune || !b->act || !b->req) {
        cacg_batch_destroy(b);
        return CACG_ENOMEM;
    }
    b->fl[0].off = 0;
    b->fl[0].size = b->arena_bytes;
    b->nfl = 1;
    const size_t ns = (size_t)b->nslots;
    if (hipStreamCreate(&b->st) != hipSuccess) {
        cacg_batch_destroy(b);
        return CACG_EDEVICE;
    }
    for (int i = 0; i < NSOLO; i++)
        if (hipStreamCreate(&b->sst[i]) != hipSuccess) {
            cacg_batch_destroy(b);
            return CACG_EDEVICE;
        }
    for (int i = 0; i <= NSOLO; i++)
        if (hipEventCreateWithFlags(&b->sev[i], hipEventDisableTiming) != hipSuccess) {
            cacg_batch_destroy(b);
            return CACG_EDEVICE;
        }
    if (hipMalloc((void **)&b->arena, (size_t)b->arena_bytes) != hipSuccess ||
        hipMalloc((void **)&b->d_sys, sizeof(cacg_bsys) * ns) != hipSuccess ||
        hipMalloc((void **)&b->d_norms, sizeof(double) * ns * CACG_BATCH_MAX_M) != hipSuccess ||
        hipMalloc((void **)&b->d_flags, sizeof(int) * ns) != hipSuccess ||
        hipHostMalloc((void **)&b->h_sys, sizeof(cacg_bsys) * ns, 0) != hipSuccess ||
        hipHostMalloc((void **)&b->h_norms, sizeof(double) * ns * CACG_BATCH_MAX_M, 0) != hipSuccess ||
        hipHostMalloc((void **)&b->h_flags, sizeof(int) * ns, 0) != hipSuccess ||
        grow_items(b, 65536) != CACG_OK) {
        cacg_batch_destroy(b);
        return CACG_ENOMEM;
    }
    memset(b->h_sys, 0, sizeof(cacg_bsys) * ns);
    *out = b;
    return CACG_OK;
}

/* ---- submission -------------------------------------------------------------------- */

long cacg_batch_submit(cacg_batch *b, long n, const long *row_start, const long *col_idx,
                       const double *val, long m, const double *B, const double *X0, double tol)
{
    if (!b || n <= 0 || n >= (1L << 31) || !row_start || !B || m < 1 || m > CACG_BATCH_MAX_M ||
        (long)b->o.s * m > CACG_BATCH_MAX_SM || !(tol >= 0.0) || row_start[0] != 0)
        return -CACG_EINVAL;
    const long nnz = row_start[n];
    if (nnz < 0 || nnz >= (1L << 31) || (nnz > 0 && (!col_idx || !val)))
        return -CACG_EINVAL;
    for (long i = 0; i < n; i++)
        if (row_start[i + 1] < row_start[i])
            return -CACG_EINVAL;
    for (long p = 0; p < nnz; p++)
        if (col_idx[p] < 0 || col_idx[p] >= n)
            return -CACG_EINVAL;
    const long bytes = make_layout(n, nnz, m, b->o.s, b->o.precision, solo_rows(&b->o)).total;
    if (bytes > b->arena_bytes)
        return -CACG_ENOMEM;
    breq *r = (breq *)calloc(1, sizeof(*r));
    if (!r)
        return -CACG_ENOMEM;
    r->n = n;
    r->nnz = nnz;
    r->m = m;
    r->tol = tol;
    r->bytes = bytes;
    r->slot = -1;
    r->rp = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    r->ci = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    r->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    r->B = (double *)malloc(sizeof(double) * (size_t)(n * m));
    /* X0 = 0 (no X0 given) is not stored: admission zeroes X on the device */
    r->X = X0 ? (double *)malloc(sizeof(double) * (size_t)(n * m)) : NULL;
    r->bn2 = (double *)calloc((size_t)m, sizeof(double));
    if (!r->rp || !r->ci || !r->val || !r->B || (X0 && !r->X) || !r->bn2) {
        free_req(r);
        return -CACG_ENOMEM;
    }
    for (long i = 0; i <= n; i++)
        r->rp[i] = (int)row_start[i];
    for (long p = 0; p < nnz; p++)
        r->ci[p] = (int)col_idx[p];
    if (nnz > 0)
        memcpy(r->val, val, sizeof(double) * (size_t)nnz);
    memcpy(r->B, B, sizeof(double) * (size_t)(n * m));
    if (X0)
        memcpy(r->X, X0, sizeof(double) * (size_t)(n * m));
    for (long i = 0; i < n; i++)
        for (long c = 0; c < m; c++)
            r->bn2[c] += B[i * m + c] * B[i * m + c];
    /* Gerschgorin interval, the arithmetic of src/common.c:53-106 */
    double lo = DBL_MAX, hi = DBL_MIN;
    for (long i = 0; i < n; i++) {
        double off = 0.0, d = 0.0;
        for (long p = row_start[i]; p < row_start[i + 1]; p++) {
            const double v = val[p];
            if (col_idx[p] != i)
                off += v > 0 ? v : -v;
            else
                d = v;
        }
        const double up = off + d, dn = d - off;
        if (hi < up)
            hi = up;
        if (lo > dn)
            lo = dn;
    }
    r->sa = 2.0 / (hi - lo);
    r->sb = -(hi + lo) / (hi - lo);
    /* the copies above ran unlocked; only the append takes the lock */
    pthread_mutex_lock(&b->mu);
    if (b->nreq == b->capreq) {
        breq **nr = (breq **)realloc(b->req, sizeof(breq *) * (size_t)(2 * b->capreq));
        if (!nr) {
            pthread_mutex_unlock(&b->mu);
            free_req(r);
            return -CACG_ENOMEM;
        }
        memset(nr + b->capreq, 0, sizeof(breq *) * (size_t)b->capreq);
        b->req = nr;
        b->capreq *= 2;
    }
    r->id = b->nreq;
    r->round_submitted = b->rounds;
    r->state = RQ_QUEUED;
    b->req[b->nreq++] = r;
    const long id = r->id;
    pthread_mutex_unlock(&b->mu);
    return id;
}

/* ---- rounds ------------------------------------------------------------------------ */

/* 1: admitted into slot; 0: the arena has no room now; < 0: -CACG_EDEVICE */
#define ACHK(call)                                                                               \
    do {                                                                                         \
        if ((call) != hipSuccess) {                                                              \
            fprintf(stderr, "cacg_batch: %s failed (%s:%d)\n", #call, __FILE__, __LINE__);       \
            return -CACG_EDEVICE;                                                                \
        }                                                                                        \
    } while (0)
#define AKCHK(call)                                                                              \
    do {                                                                                         \
        if ((call) != 0) {                                                                       \
            fprintf(stderr, "cacg_batch: %s failed (%s:%d)\n", #call, __FILE__, __LINE__);       \
            return -CACG_EDEVICE;                                                                \
        }                                                                                        \
    } while (0)

static int admit(cacg_batch *b, breq *r, int slot)
{
    const long off = arena_alloc(b, r->bytes);
    if (off < 0)
        return 0;
    const int s = b->o.s, f32 = b->o.precision == CACG_PREC_FP32_REFINE;
    const long n = r->n, m = r->m, sm = s * m;
    const layout L = make_layout(n, r->nnz, m, s, b->o.precision, solo_rows(&b->o));
    char *base = b->arena + off;
    cacg_bsys *d = &b->h_sys[slot];
    memset(d, 0, sizeof(*d));
    d->n = (int)n;
    d->m = (int)m;
    d->sm = (int)sm;
    d->ntile = (int)((n + CACG_BT - 1) / CACG_BT);
    d->nchunk = (int)((n + CACG_GR - 1) / CACG_GR);
    d->rowptr = (const int *)(base + L.rowptr);
    d->col = (const int *)(base + L.col);
    d->val = (const double *)(base + L.val);
    d->val32 = f32 ? (const float *)(base + L.val32) : NULL;
    d->B = (const double *)(base + L.B);
    d->X = (double *)(base + L.X);
    d->R = (double *)(base + L.R);
    d->S = base + L.S;
    d->Q = (double *)(base + L.Q);
    d->AQ = (double *)(base + L.AQ);
    d->QtAQ = (double *)(base + L.QtAQ);
    d->QtAS = (double *)(base + L.QtAS);
    d->beta = (double *)(base + L.beta);
    d->lu = (double *)(base + L.lu);
    d->QtR = (double *)(base + L.QtR);
    d->alpha = (double *)(base + L.alpha);
    d->gpart = (double *)(base + L.gpart);
    d->npart = (double *)(base + L.npart);
    d->norm2 = b->d_norms + (long)slot * CACG_BATCH_MAX_M;
    d->flag = b->d_flags + slot;
    d->sa = r->sa;
    d->sb = r->sb;
    hipStream_t st = b->st;
    ACHK(hipMemcpyAsync((void *)d->rowptr, r->rp, sizeof(int) * (size_t)(n + 1),
                        hipMemcpyHostToDevice, st));
    if (r->nnz > 0) {
        ACHK(hipMemcpyAsync((void *)d->col, r->ci, sizeof(int) * (size_t)r->nnz,
                            hipMemcpyHostToDevice, st));
        ACHK(hipMemcpyAsync((void *)d->val, r->val, sizeof(double) * (size_t)r->nnz,
                            hipMemcpyHostToDevice, st));
        if (f32)
            AKCHK(cacg_k_cvt_f64_f32(r->nnz, d->val, (float *)d->val32, st));
    }
    ACHK(hipMemcpyAsync((void *)d->B, r->B, sizeof(double) * (size_t)(n * m),
                        hipMemcpyHostToDevice, st));
    if (r->X)
        ACHK(hipMemcpyAsync(d->X, r->X, sizeof(double) * (size_t)(n * m), hipMemcpyHostToDevice, st));
    else
        AKCHK(cacg_k_fill_f64(d->X, n * m, 0.0, st));
    AKCHK(cacg_k_fill_i32(d->flag, 1, 0, st));
    ACHK(hipMemcpyAsync(b->d_sys + slot, d, sizeof(*d), hipMemcpyHostToDevice, st));
    r->off = off;
    r->slot = slot;
    r->state = RQ_ACTIVE;
    r->joined = 0;
    r->fresh = 1;
    r->iters = 0;
    r->rounds_queued = b->rounds - r->round_submitted;
    b->slot[slot] = r;
    stune *t = &b->tune[slot];
    t->timing = -1;
    t->next = b->o.tune ? 0 : NCAND;
    t->best_b = t->best_a = -1.f;
    t->sb = t->sa = CAND[0];
    return 1;
}

static int retire(cacg_batch *b, breq *r, int status, const double *norm2)
{
    double worst = 0.0;
    for (long c = 0; c < r->m; c++) {
        const double q = sqrt(norm2[c]) / sqrt(r->bn2[c]);
        if (!(q <= worst))
            worst = q;
    }
    if (!r->X) { /* submitted without X0 */
        r->X = (double *)malloc(sizeof(double) * (size_t)(r->n * r->m));
        if (!r->X)
            return CACG_ENOMEM;
    }
    HCHK(hipMemcpyAsync(r->X, b->h_sys[r->slot].X, sizeof(double) * (size_t)(r->n * r->m),
                        hipMemcpyDeviceToHost, b->st));
    HCHK(hipStreamSynchronize(b->st));
    if (b->nret == b->capret) {
        const long cap = b->capret ? 2 * b->capret : 256;
        long *nr = (long *)realloc(b->ret, sizeof(long) * (size_t)cap);
        if (!nr)
            return CACG_ENOMEM;
        b->ret = nr;
        b->capret = cap;
    }
    b->ret[b->nret++] = r->id;
    r->true_relres = worst;
    r->status = status;
    r->state = RQ_DONE;
    b->slot[r->slot] = NULL;
    r->slot = -1;
    /* keep the matrix copies only as long as they can be needed */
    free(r->rp);
    free(r->ci);
    free(r->val);
    free(r->B);
    r->rp = NULL;
    r->ci = NULL;
    r->val = NULL;
    r->B = NULL;
    return arena_free(b, r->off, r->bytes);
}

/* MFMA tile class of a system's Gram (cacg_kb_gram_class): 1 and (tm, tn) when
 * tm = ceil(sm / 16) is 1 .. 8 and tn = ceil(kr / 16) is in tm .. tm + 2 */
static int gram_class(const cacg_bsys *d, int which, int *tm, int *tn)
{
    const int kr = d->sm + (which ? d->m : 0);
    const int a = (d->sm + 15) / 16, c = (kr + 15) / 16;
    if (a < 1 || a > 8 || c < a || c > a + 2)
        return 0;
    *tm = a;
    *tn = c;
    return 1;
}

/* append {slot, 0 .. count - 1} */
static long put_range(int2 *it, long k, int slot, int count)
{
    for (int i = 0; i < count; i++) {
        it[k].x = slot;
        it[k].y = i;
        k++;
    }
    return k;
}

/* the Gram partials of every listed system (one MFMA launch per tile class),
 * then the chunk-ordered sums */
static int gram_phase(cacg_batch *b, const int2 *di, const long *o, const long *cnt,
                      const int2 *blocks, long nblocks, int which, int f32)
{
    hipStream_t st = b->st;
    int rc = cnt[0] > 0 ? (int)hipErrorInvalidValue : 0;   /* no class: not admissible */
    for (int c = 1; c < 25 && rc == 0; c++) {
        const int tm = 1 + (c - 1) / 3, tn = tm + (c - 1) % 3;
        rc = cacg_kb_gram_class(b->d_sys, di + o[c], (int)cnt[c], which, f32, tm, tn, st);
    }
    if (rc == 0)
        rc = cacg_kb_gram_reduce(b->d_sys, blocks, (int)nblocks, which, st);
    return rc;
}

/* one outer iteration of a solo system with the single-system kernels (fast
 * order), the same sequence as the batched round: basis from the true R; Q = S
 * or QtAS, beta (the previous alpha solve's factor), Q = S - Q beta; AQ;
 * [QtAQ | QtR], alpha; X += Q alpha */
/* X += Q alpha of a solo system's last iteration (the deferred X update) */
static int solo_flush_x(const cacg_bsys *d, hipStream_t st)
{
    KCHK(cacg_k_xrupdate_split(d->n, d->sm, d->m, d->Q, NULL, d->alpha, d->X, NULL, NULL, NULL, 1,
                               st));
    return CACG_OK;
}

/* Solo systems iterate as the single-system solver does: the recursive
 * residual R -= AQ alpha drives the next basis and a provisional stop test,
 * and X += Q alpha rides on the next Q update (*xpend) where the widths allow;
 * a round of solo_residual then recomputes the true residual, which decides. */
static int solo_iteration(const cacg_bsys *d, int fresh, int s, int f32, int *xpend,
                          const cacg_spmm_sched *sb, const cacg_spmm_sched *sa, hipEvent_t *ev,
                          hipStream_t st)
{
    const int split = cacg_k_split_supported(d->sm, d->m);
    const int n = d->n, m = d->m, sm = d->sm;
    const cacg_dcsr A = {n, d->rowptr, d->col, d->val};
    const cacg_scsr A32 = {n, d->rowptr, d->col, d->val32};
    const double sa2 = 2.0 * d->sa, sb2 = 2.0 * d->sb;
    if (ev)
        HCHK(hipEventRecord(ev[0], st));
    if (f32) {
        float *S = (float *)d->S;
        if (s == 1) {
            KCHK(cacg_k_cvt_f64_f32((long)n * m, d->R, S, st));
        } else {
            KCHK(cacg_k_cheb1_f32_x(&A32, m, d->R, m, S + m, sm, S, sm, d->sa, d->sb, sb, st));
            for (int j = 2; j < s; j++)
                KCHK(cacg_k_spmm_f32_x(&A32, m, S + (long)(j - 1) * m, sm, S + (long)j * m, sm,
                                       CACG_EPI_CHEB, sa2, sb2, S + (long)(j - 1) * m, sm,
                                       S + (long)(j - 2) * m, sm, sb, st));
        }
        if (ev)
            HCHK(hipEventRecord(ev[1], st));
        if (fresh) {
            if (*xpend && solo_flush_x(d, st) != CACG_OK)
                return CACG_EDEVICE;
            KCHK(cacg_k_cvt_f32_f64((long)n * sm, S, d->Q, st));
        } else {
            KCHK(cacg_k_gram_fast_f64f32(n, d->AQ, sm, sm, S, sm, sm, NULL, 0, 0, d->QtAS, NULL,
                                         d->gpart, st));
            KCHK(cacg_k_lu_solve(sm, d->QtAQ, sm, d->QtAS, d->beta, d->lu, d->flag,
                                 CACG_MODE_FAST, 1, st));
            if (*xpend)
                KCHK(cacg_k_qxupdate_s32(n, sm, m, S, d->Q, d->beta, d->alpha, d->X, st));
            else
                KCHK(cacg_k_qupdate_s32(n, sm, S, d->Q, d->beta, st));
        }
        *xpend = 0;
    } else {
        /* as the single-system solver: a fresh iteration writes its basis
         * row-major straight into Q (Q = S; the pending X update, which reads
         * the old Q, goes first), the later ones keep S block-major (S_j as
         * n x m blocks n m apart) when the deferred X update rides on the Q
         * update; same bits either way */
        if (fresh && *xpend && solo_flush_x(d, st) != CACG_OK)
            return CACG_EDEVICE;
        const int bm = !fresh && *xpend && split && m % 4 == 0 && s >= 2 &&
                       cacg_gram_blk_supported(n, sm, sm, m);
        const long bs = (long)n * m;
        double *SB = fresh ? d->Q : (double *)d->S;
        const long bo = bm ? bs : m, lb = bm ? m : sm;
        if (s == 1) {
            HCHK(hipMemcpyAsync(SB, d->R, sizeof(double) * (size_t)n * m, hipMemcpyDeviceToDevice,
                                st));
        } else {
            KCHK(cacg_k_spmm_x(&A, m, d->R, m, SB + bo, lb, CACG_EPI_CHEB1, d->sa, d->sb, NULL, 0,
                               NULL, 0, SB, lb, sb, st));
            for (int j = 2; j < s; j++)
                KCHK(cacg_k_spmm_x(&A, m, SB + (j - 1) * bo, lb, SB + j * bo, lb, CACG_EPI_CHEB,
                                   sa2, sb2, SB + (j - 1) * bo, lb, SB + (j - 2) * bo, lb, NULL, 0,
                                   sb, st));
        }
        if (ev)
            HCHK(hipEventRecord(ev[1], st));
        if (!fresh) {
            if (bm)
                KCHK(cacg_k_gram_fast_blk(n, d->AQ, sm, sm, SB, m, bs, sm, d->QtAS, d->gpart, st));
            else
                KCHK(cacg_k_gram_fast(n, d->AQ, sm, sm, SB, sm, sm, NULL, 0, 0, d->QtAS, NULL,
                                      d->gpart, st));
            KCHK(cacg_k_lu_solve(sm, d->QtAQ, sm, d->QtAS, d->beta, d->lu, d->flag,
                                 CACG_MODE_FAST, 1, st));
            if (bm)
                KCHK(cacg_k_qxupdate_blk(n, sm, m, SB, bs, d->Q, d->beta, d->alpha, d->X, st));
            else if (*xpend)
                KCHK(cacg_k_qxupdate(n, sm, m, SB, d->Q, d->beta, d->alpha, d->X, st));
            else
                KCHK(cacg_k_qupdate(n, sm, SB, d->Q, d->beta, CACG_MODE_FAST, st));
        }
        *xpend = 0;
    }
    if (ev)
        HCHK(hipEventRecord(ev[2], st));
    KCHK(cacg_k_spmm_x(&A, sm, d->Q, sm, d->AQ, sm, CACG_EPI_NONE, 0.0, 0.0, NULL, 0, NULL, 0, NULL,
                       0, sa, st));
    if (ev)
        HCHK(hipEventRecord(ev[3], st));
    KCHK(cacg_k_gram_fast(n, d->Q, sm, sm, d->AQ, sm, sm, d->R, m, m, d->QtAQ, d->QtR, d->gpart,
                          st));
    KCHK(cacg_k_lu_solve(sm, d->QtAQ, m, d->QtR, d->alpha, d->lu, d->flag, CACG_MODE_FAST, 0, st));
    if (split) { /* R and its norms now, X with the next Q update */
        KCHK(cacg_k_xrupdate_split(n, sm, m, NULL, d->AQ, d->alpha, NULL, d->R, d->norm2, d->npart,
                                   2, st));
        *xpend = 1;
    } else {
        KCHK(cacg_k_xrupdate(n, sm, m, d->Q, d->AQ, d->alpha, d->X, d->R, d->norm2, d->npart,
                             CACG_MODE_FAST, st));
    }
    return CACG_OK;
}

/* the true residual R = B - A X of a solo system and its column norms */
static int solo_residual(const cacg_bsys *d, hipStream_t st)
{
    const cacg_dcsr A = {d->n, d->rowptr, d->col, d->val};
    KCHK(cacg_k_spmm(&A, d->m, d->X, d->m, d->R, d->m, CACG_EPI_RESID, 0.0, 0.0, d->B, d->m, NULL,
                     0, NULL, 0, st));
    KCHK(cacg_k_colnorm2_fast(d->n, d->m, d->R, d->norm2, d->npart, st));
    return CACG_OK;
}

/* the batched launches of a round, in order, on the main stream: basis steps,
 * Q = S for fresh systems, QtAS, beta, Q = S - Q beta (X += Q alpha deferred in),
 * A Q, [QtAQ | QtR], alpha, X += Q alpha; then R = B - A X and the norms */
static int batched_launches(cacg_batch *b, const wlists *L)
{
    hipStream_t st = b->st;
    const int2 *di = b->d_items;
    const int s = L->s, f32 = L->f32;
    if (L->n_it > 0) {
        if (s == 1) {
            KCHK(cacg_kb_basis(b->d_sys, di + L->o_it, (int)L->n_it, 0, f32, st));
        } else {
            KCHK(cacg_kb_basis(b->d_sys, di + L->o_it, (int)L->n_it, 1, f32, st));
            for (int j = 2; j < s; j++)
                KCHK(cacg_kb_basis(b->d_sys, di + L->o_it, (int)L->n_it, j, f32, st));
        }
        KCHK(cacg_kb_q_from_s(b->d_sys, di + L->o_fr, (int)L->n_fr, f32, st));
        KCHK(gram_phase(b, di, L->o_gc[0], L->n_gc[0], di + L->o_b0, L->n_b0, 0, f32));
        KCHK(cacg_k_lu_batch(b->d_sys, di + L->o_lu[0][0], (int)L->n_lu[0][0], 0, 64, st));
        KCHK(cacg_k_lu_batch(b->d_sys, di + L->o_lu[0][1], (int)L->n_lu[0][1], 0, L->max_sm, st));
        for (int c = 1; c <= 8; c++)
            for (int ex = 0; ex < 2; ex++)
                KCHK(cacg_kb_qupdate_class(b->d_sys, di + L->o_qc[c][ex], (int)L->n_qc[c][ex],
                                           16 * c, ex, f32, st));
        KCHK(cacg_kb_aq(b->d_sys, di + L->o_it, (int)L->n_it, st));
        KCHK(gram_phase(b, di, L->o_gc[1], L->n_gc[1], di + L->o_b1, L->n_b1, 1, f32));
        KCHK(cacg_k_lu_batch(b->d_sys, di + L->o_lu[1][0], (int)L->n_lu[1][0], 1, 64, st));
        KCHK(cacg_k_lu_batch(b->d_sys, di + L->o_lu[1][1], (int)L->n_lu[1][1], 1, L->max_sm, st));
        for (int c = 1; c <= 8; c++)
            for (int ex = 0; ex < 2; ex++)
                for (int h = 0; h < 2; h++)
                    KCHK(cacg_kb_xupdate_class(b->d_sys, di + L->o_xc[c][ex][h],
                                               (int)L->n_xc[c][ex][h], 16 * c, ex, h ? 32 : 16,
                                               st));
    }
    KCHK(cacg_kb_resid(b->d_sys, di + L->o_all, (int)L->n_all, di + L->o_sys, (int)L->n_sys, st));
    return CACG_OK;
}

static void drop_graph(cacg_batch *b)
{
    if (b->g_exec)
        (void)hipGraphExecDestroy(b->g_exec);
    b->g_exec = NULL;
    b->g_valid = 0;
}

/* remember the lists just uploaded (the key a later round is compared with) */
static void keep_lists(cacg_batch *b, const wlists *L, const int2 *it)
{
    if (!b->o.graph)
        return;
    if (L->k > b->g_cap) {
        free(b->g_items);
        b->g_cap = 0;
        b->g_items = (int2 *)malloc(sizeof(int2) * (size_t)L->k);
        if (!b->g_items)
            return;   /* no key, no graph: the rounds launch one by one */
        b->g_cap = L->k;
    }
    memcpy(b->g_items, it, sizeof(int2) * (size_t)L->k);
    b->g_wl = *L;
    b->g_ditems = b->d_items;
    b->g_valid = 1;
}

/* capture the batched launches of a round into b->g_exec (nothing runs during
 * the capture).  The launches read the slot descriptors and work lists from
 * device memory, so a replay serves any later round with the same lists.  If
 * capture or instantiation fails, the stream has left capture mode, graphs are
 * switched off for this batch and the caller launches one by one. */
static void capture_round(cacg_batch *b, const wlists *L)
{
    hipGraph_t g = NULL;
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        b->o.graph = 0;
        return;
    }
    const int rc = batched_launches(b, L);
    const hipError_t e = hipStreamEndCapture(b->st, &g);
    if (rc == CACG_OK && e == hipSuccess && g &&
        hipGraphInstantiate(&b->g_exec, g, NULL, NULL, 0) != hipSuccess)
        b->g_exec = NULL;
    if (g)
        (void)hipGraphDestroy(g);
    if (!b->g_exec) {
        (void)hipGetLastError();
        b->o.graph = 0;
    }
}

static int step_locked(cacg_batch *b, cacg_batch_stats *stats)
{
    hipStream_t st = b->st;
    const int s = b->o.s, f32 = b->o.precision == CACG_PREC_FP32_REFINE;
    int admitted = 0, retired = 0;

    for (int i = 0; i < b->nslots; i++)
        if (b->slot[i])
            b->slot[i]->ran = 0;
    /* solo systems admitted in earlier rounds: one iteration (when joined) or
     * the true residual, dealt round-robin over the solo streams, launched
     * before this round's admission so that its uploads (host-blocking copies
     * from pageable memory) overlap their work; systems admitted now start next
     * round */
    int nsolo = 0;
    for (int i = 0; i < b->nslots; i++)
        if (b->slot[i] && IS_SOLO(b->h_sys[i].n))
            nsolo++;
    if (nsolo > 0) {
        HCHK(hipEventRecord(b->sev[NSOLO], st));
        for (int q = 0; q < NSOLO && q < nsolo; q++)
            HCHK(hipStreamWaitEvent(b->sst[q], b->sev[NSOLO], 0));
        int q = 0;
        for (int i = 0; i < b->nslots; i++) {
            const breq *r = b->slot[i];
            if (!r || !IS_SOLO(b->h_sys[i].n))
                continue;
            hipStream_t ss = b->sst[q];
            q = (q + 1) % NSOLO;
            breq *rw = b->slot[i];
            rw->ran = 1;
            int rc;
            if (r->joined && !r->verify) {
                stune *t = &b->tune[i];
                t->timing = -1;
                if (t->next < NCAND && !t->have_ev) {
                    int e = 0;
                    while (e < 4 && hipEventCreate(&t->ev[e]) == hipSuccess)
                        e++;
                    if (e < 4) { /* no events: keep the default schedules */
                        (void)hipGetLastError();
                        while (e > 0)
                            hipEventDestroy(t->ev[--e]);
                        t->next = NCAND;
                    } else {
                        t->have_ev = 1;
                    }
                }
                if (t->next < NCAND)
                    t->timing = t->next;
                const cacg_spmm_sched *sb = t->timing >= 0 ? &CAND[t->timing] : &t->sb;
                const cacg_spmm_sched *sa = t->timing >= 0 ? &CAND[t->timing] : &t->sa;
                rc = solo_iteration(&b->h_sys[i], r->fresh, s, f32, &rw->xpend, sb, sa,
                                    t->timing >= 0 ? t->ev : NULL, ss);
            } else { /* admitted, or verifying: X complete, then the true residual */
                rc = rw->xpend ? solo_flush_x(&b->h_sys[i], ss) : CACG_OK;
                rw->xpend = 0;
                if (rc == CACG_OK)
                    rc = solo_residual(&b->h_sys[i], ss);
            }
            if (rc != CACG_OK)
                return rc;
        }
    }
    /* admission, FIFO: the head waits for a slot and memory (and, for a large
     * system, for the large systems in flight to drop below max_large) */
    int free_slot = 0, nlarge = 0;
    for (int i = 0; i < b->nslots; i++)
        if (b->slot[i] && IS_SOLO(b->slot[i]->n))
            nlarge++;
    /* smallest first: among the queued requests the one of smallest footprint
     * that may be admitted now (a large one only below max_large) goes next,
     * unless the oldest queued request has been overtaken max_bypass times:
     * then it goes next, or waits, and nothing overtakes it meanwhile */
    const long maxb = b->o.max_bypass > 0 ? b->o.max_bypass : 16;
    while (b->o.policy == CACG_BATCH_SMALLEST_FIRST) {
        while (b->qhead < b->nreq &&
               (!b->req[b->qhead] || b->req[b->qhead]->state != RQ_QUEUED))
            b->qhead++;
        if (b->qhead >= b->nreq)
            break;
        breq *head = b->req[b->qhead], *pick = NULL;
        const int large_full = b->o.max_large > 0 && nlarge >= b->o.max_large;
        if (head->bypassed >= maxb) {
            if (IS_SOLO(head->n) && large_full)
                break;
            pick = head;
        } else {
            for (long i = b->qhead; i < b->nreq; i++) {
                breq *r = b->req[i];
                if (!r || r->state != RQ_QUEUED || (IS_SOLO(r->n) && large_full))
                    continue;
                if (!pick || r->bytes < pick->bytes)
                    pick = r;
            }
        }
        if (!pick)
            break;
        while (free_slot < b->nslots && b->slot[free_slot])
            free_slot++;
        if (free_slot == b->nslots)
            break;
        const int ok = admit(b, pick, free_slot);
        if (ok < 0)
            return -ok;
        if (ok == 0)
            break;
        admitted++;
        if (IS_SOLO(pick->n))
            nlarge++;
        for (long i = b->qhead; i < pick->id; i++)
            if (b->req[i] && b->req[i]->state == RQ_QUEUED)
                b->req[i]->bypassed++;
    }
    while (b->o.policy == CACG_BATCH_FIFO && b->qhead < b->nreq) {
        breq *r = b->req[b->qhead];
        if (!r || r->state != RQ_QUEUED) {
            b->qhead++;
            continue;
        }
        if (IS_SOLO(r->n) && b->o.max_large > 0 && nlarge >= b->o.max_large)
            break;
        while (free_slot < b->nslots && b->slot[free_slot])
            free_slot++;
        if (free_slot == b->nslots)
            break;
        const int ok = admit(b, r, free_slot);
        if (ok < 0)
            return -ok;
        if (ok == 0)
            break;
        admitted++;
        if (IS_SOLO(r->n))
            nlarge++;
        b->qhead++;
    }

    /* work lists (solo systems take no part in the batched launches) */
#define BATCHED(i) (b->slot[i] && !IS_SOLO(b->h_sys[i].n))
    long need = 0;
    int nact = 0, nit = 0, max_sm = 1, na = 0;
    /* i over the batched active slots (b->act[0 .. na - 1], ascending): the class
     * lists below scan them, not every slot */
#define FOR_BATCHED(i) \
    for (int q_ = 0, i = na > 0 ? b->act[0] : 0; q_ < na; i = ++q_ < na ? b->act[q_] : 0)
    for (int i = 0; i < b->nslots; i++) {
        const breq *r = b->slot[i];
        if (!r)
            continue;
        const cacg_bsys *d = &b->h_sys[i];
        nact++;
        if (IS_SOLO(d->n))
            continue;
        b->act[na++] = i;
        need += 1 + d->ntile;
        if (r->joined) {
            nit++;
            need += 2L * d->ntile + 4L * d->nchunk + 2 +
                    ((long)d->sm * (d->sm + d->m) + 255) / 256 + ((long)d->sm * d->sm + 255) / 256;
            if (d->sm > max_sm)
                max_sm = d->sm;
        }
    }
    if (grow_items(b, need) != CACG_OK)
        return CACG_ENOMEM;
    int2 *it = b->h_items;
    wlists wl;
    memset(&wl, 0, sizeof(wl));   /* compared bytewise: padding included */
    long k = 0;
    wl.o_it = k;
    FOR_BATCHED(i)
        if (BATCHED(i) && b->slot[i]->joined)
            k = put_range(it, k, i, b->h_sys[i].ntile);
    wl.n_it = k - wl.o_it;
    wl.o_fr = k;
    FOR_BATCHED(i)
        if (BATCHED(i) && b->slot[i]->joined && b->slot[i]->fresh)
            k = put_range(it, k, i, b->h_sys[i].ntile);
    wl.n_fr = k - wl.o_fr;
    /* Q and X updates as MFMA per padded width smp = 16 ceil(sm / 16), split by
     * sm == smp (16-byte loads) and, for X, by m <= 16 */
    for (int c = 1; c <= 8; c++)
        for (int ex = 0; ex < 2; ex++) {
            wl.o_qc[c][ex] = k;
            FOR_BATCHED(i) {
                const breq *r = b->slot[i];
                const int sm = b->h_sys[i].sm;
                if (BATCHED(i) && r->joined && !r->fresh && (sm + 15) / 16 == c &&
                    (sm == 16 * c) == ex)
                    k = put_range(it, k, i, b->h_sys[i].nchunk);
            }
            wl.n_qc[c][ex] = k - wl.o_qc[c][ex];
            for (int h = 0; h < 2; h++) {
                wl.o_xc[c][ex][h] = k;
                FOR_BATCHED(i) {
                    const breq *r = b->slot[i];
                    const int sm = b->h_sys[i].sm;
                    if (BATCHED(i) && r->joined && (sm + 15) / 16 == c && (sm == 16 * c) == ex &&
                        (b->h_sys[i].m > 16) == h)
                        k = put_range(it, k, i, b->h_sys[i].nchunk);
                }
                wl.n_xc[c][ex][h] = k - wl.o_xc[c][ex][h];
            }
        }
    /* Gram chunks per tile class: [which][1 + 3 (tm - 1) + tn - tm]; [which][0] stays empty
     * (every admissible system has a class) */
    for (int w = 0; w < 2; w++)
        for (int c = 0; c < 25; c++) {
            wl.o_gc[w][c] = k;
            FOR_BATCHED(i) {
                const breq *r = b->slot[i];
                if (!BATCHED(i) || !r->joined || (w == 0 && r->fresh))
                    continue;
                int tm = 0, tn = 0, cls = 0;
                if (gram_class(&b->h_sys[i], w, &tm, &tn))
                    cls = 1 + 3 * (tm - 1) + (tn - tm);
                if (cls == c)
                    k = put_range(it, k, i, b->h_sys[i].nchunk);
            }
            wl.n_gc[w][c] = k - wl.o_gc[w][c];
        }
    wl.o_b0 = k;
    FOR_BATCHED(i)
        if (BATCHED(i) && b->slot[i]->joined && !b->slot[i]->fresh) {
            const cacg_bsys *d = &b->h_sys[i];
            k = put_range(it, k, i, (int)(((long)d->sm * d->sm + 255) / 256));
        }
    wl.n_b0 = k - wl.o_b0;
    /* LU solves in two LDS classes: sm <= 64 (several workgroups per CU), larger */
    for (int w = 0; w < 2; w++)
        for (int c = 0; c < 2; c++) {
            wl.o_lu[w][c] = k;
            FOR_BATCHED(i) {
                const breq *r = b->slot[i];
                if (BATCHED(i) && r->joined && (w == 1 || !r->fresh) &&
                    ((b->h_sys[i].sm > 64) == c))
                    k = put_range(it, k, i, 1);
            }
            wl.n_lu[w][c] = k - wl.o_lu[w][c];
        }
    wl.o_b1 = k;
    FOR_BATCHED(i)
        if (BATCHED(i) && b->slot[i]->joined) {
            const cacg_bsys *d = &b->h_sys[i];
            k = put_range(it, k, i, (int)(((long)d->sm * (d->sm + d->m) + 255) / 256));
        }
    wl.n_b1 = k - wl.o_b1;
    wl.o_all = k;
    FOR_BATCHED(i)
        if (BATCHED(i))
            k = put_range(it, k, i, b->h_sys[i].ntile);
    wl.n_all = k - wl.o_all;
    wl.o_sys = k;
    FOR_BATCHED(i)
        if (BATCHED(i))
            k = put_range(it, k, i, 1);
    wl.n_sys = k - wl.o_sys;
    wl.k = k;
    wl.s = s;
    wl.f32 = f32;
    wl.max_sm = max_sm;
    /* the same lists as the last upload (the device copy still holds them): the
     * round's launches repeat the last round's, and run as a graph */
    const int same = k > 0 && b->o.graph && b->g_valid && b->g_ditems == b->d_items &&
                     memcmp(&wl, &b->g_wl, sizeof(wl)) == 0 &&
                     memcmp(it, b->g_items, sizeof(int2) * (size_t)k) == 0;
    if (!same && k > 0) {
        HCHK(hipMemcpyAsync(b->d_items, it, sizeof(int2) * (size_t)k, hipMemcpyHostToDevice, st));
        drop_graph(b);
        keep_lists(b, &wl, it);
    }

#undef BATCHED
#undef FOR_BATCHED

    /* one outer iteration of the joined batched systems and the true residual of
     * every batched one: replayed from the graph when the lists repeat (captured on
     * the first repeat), else launched one by one */
    if (same && !b->g_exec)
        capture_round(b, &wl);
    if (same && b->g_exec) {
        HCHK(hipGraphLaunch(b->g_exec, st));
        b->graph_rounds++;
    } else {
        const int rc = batched_launches(b, &wl);
        if (rc != CACG_OK)
            return rc;
    }
    /* the solo streams' work joins the main stream before the norms come back */
    for (int q2 = 0; q2 < NSOLO && q2 < nsolo; q2++) {
        HCHK(hipEventRecord(b->sev[q2], b->sst[q2]));
        HCHK(hipStreamWaitEvent(st, b->sev[q2], 0));
    }
    int hi_slot = 0;
    for (int i = 0; i < b->nslots; i++)
        if (b->slot[i])
            hi_slot = i + 1;
    if (hi_slot > 0) {
        HCHK(hipMemcpyAsync(b->h_norms, b->d_norms,
                            sizeof(double) * (size_t)hi_slot * CACG_BATCH_MAX_M,
                            hipMemcpyDeviceToHost, st));
        HCHK(hipMemcpyAsync(b->h_flags, b->d_flags, sizeof(int) * (size_t)hi_slot,
                            hipMemcpyDeviceToHost, st));
    }
    {   /* the wait for the GPU runs unlocked: cacg_batch_submit may append meanwhile */
        pthread_mutex_unlock(&b->mu);
        const hipError_t e = hipStreamSynchronize(st);
        pthread_mutex_lock(&b->mu);
        if (e != hipSuccess) {
            fprintf(stderr, "cacg_batch: hipStreamSynchronize failed: %s\n", hipGetErrorString(e));
            return CACG_EDEVICE;
        }
    }
    b->rounds++;
    /* the solo systems' timed candidates (their streams joined the main one) */
    for (int i = 0; i < hi_slot; i++) {
        stune *t = &b->tune[i];
        if (!b->slot[i] || t->timing < 0)
            continue;
        float tb = 0.f, ta = 0.f;
        const hipError_t eb = hipEventElapsedTime(&tb, t->ev[0], t->ev[1]);
        const hipError_t ea = hipEventElapsedTime(&ta, t->ev[2], t->ev[3]);
        if (eb != hipSuccess || ea != hipSuccess)
            (void)hipGetLastError();   /* an untimed candidate is skipped, not an error */
        if (eb == hipSuccess && (t->best_b < 0.f || tb < t->best_b)) {
            t->best_b = tb;
            t->sb = CAND[t->timing];
        }
        if (ea == hipSuccess && (t->best_a < 0.f || ta < t->best_a)) {
            t->best_a = ta;
            t->sa = CAND[t->timing];
        }
        t->next = t->timing + 1;
        t->timing = -1;
    }

    /* stop tests on the true residuals; retire */
    for (int i = 0; i < hi_slot; i++) {
        breq *r = b->slot[i];
        if (!r || (IS_SOLO(r->n) && !r->ran))   /* a solo system admitted this round */
            continue;
        const double *nr = b->h_norms + (long)i * CACG_BATCH_MAX_M;
        const int solo = IS_SOLO(r->n);
        /* a solo system's iteration rounds report its recursive residual */
        const int recursive = solo && r->joined && !r->verify;
        if (r->joined && !(solo && r->verify)) {
            r->iters++;
            r->fresh = 0;
        }
        r->joined = 1;
        r->verify = 0;
        int done = 1, finite = 1;
        for (long c = 0; c < r->m; c++) {
            if (!isfinite(nr[c]))
                finite = 0;
            if (!(sqrt(nr[c]) < r->tol * sqrt(r->bn2[c])))
                done = 0;
        }
        int status = -1;
        if (recursive) { /* provisional: the true residual decides next round */
            if (!finite || b->h_flags[i])
                status = CACG_BREAKDOWN;
            else if (done || r->iters >= b->o.maxiter)
                r->verify = 1;
        } else if (done) {
            status = CACG_OK;
        } else if (!finite || b->h_flags[i]) {
            status = CACG_BREAKDOWN;
        } else if (r->iters >= b->o.maxiter) {
            status = CACG_MAXITER;
        }
        if (status >= 0 && r->xpend) { /* the X update still pending */
            const int rc = solo_flush_x(&b->h_sys[i], st);
            if (rc != CACG_OK)
                return rc;
            r->xpend = 0;
        }
        if (status >= 0) {
            const int rc = retire(b, r, status, nr);
            if (rc != CACG_OK)
                return rc;
            retired++;
            nact--;
        }
    }
    if (stats) {
        int queued = 0, finished = 0;
        for (long i = b->qhead; i < b->nreq; i++)
            if (b->req[i] && b->req[i]->state == RQ_QUEUED)
                queued++;
        for (long i = 0; i < b->nreq; i++)
            if (b->req[i] && b->req[i]->state == RQ_DONE)
                finished++;
        stats->active = nact;
        stats->queued = queued;
        stats->finished = finished;
        stats->admitted = admitted;
        stats->retired = retired;
        stats->rounds = b->rounds;
        stats->arena_used = b->arena_used;
        stats->arena_bytes = b->arena_bytes;
        stats->graph_rounds = b->graph_rounds;
    }
    return CACG_OK;
}

int cacg_batch_step(cacg_batch *b, cacg_batch_stats *stats)
{
    if (!b)
        return CACG_EINVAL;
    pthread_mutex_lock(&b->mu);
    const int rc = step_locked(b, stats);
    pthread_mutex_unlock(&b->mu);
    return rc;
}

int cacg_batch_retired(cacg_batch *b, long *ids, int cap)
{
    if (!b || cap < 0 || (cap > 0 && !ids))
        return -CACG_EINVAL;
    pthread_mutex_lock(&b->mu);
    const int k = b->nret < cap ? (int)b->nret : cap;
    memcpy(ids, b->ret, sizeof(long) * (size_t)k);
    memmove(b->ret, b->ret + k, sizeof(long) * (size_t)(b->nret - k));
    b->nret -= k;
    pthread_mutex_unlock(&b->mu);
    return k;
}

int cacg_batch_collect(cacg_batch *b, long id, double *X, cacg_batch_result *res)
{
    if (!b)
        return -CACG_EINVAL;
    pthread_mutex_lock(&b->mu);
    breq *r = id >= 0 && id < b->nreq ? b->req[id] : NULL;
    if (!r || r->state != RQ_DONE) {
        pthread_mutex_unlock(&b->mu);
        return r ? 0 : -CACG_EINVAL;
    }
    /* a retired request: no round touches it again, so the copy runs unlocked */
    b->req[id] = NULL;
    pthread_mutex_unlock(&b->mu);
    if (X)
        memcpy(X, r->X, sizeof(double) * (size_t)(r->n * r->m));
    if (res) {
        res->status = r->status;
        res->iters = r->iters;
        res->rounds_queued = r->rounds_queued;
        res->true_relres = r->true_relres;
    }
    free_req(r);
    return 1;
}
