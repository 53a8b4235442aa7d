/*
 * cacg_batch.h — C-ABI of the batching / serving form: many independent
 * s-step Chebyshev block-CG solves resident on one GPU at once.
 *
 * No counterpart in the reference (it solves one system per MPI job,
 * src/main.c:140-156); the per-system iteration is the one of
 * src/bcbcgKernel.c:172-526 with the true residual recomputed in FP64 after
 * every outer iteration (it drives the stop test and seeds the next basis).
 *
 * Requests are queued by cacg_batch_submit (the matrix, B and X0 are copied,
 * so the caller's buffers can be reused at once).  Each cacg_batch_step is one
 * round: queued requests are admitted while a slot and their HBM footprint are
 * free, in FIFO order (the head waits, later requests do not overtake it) or
 * smallest footprint first (cacg_batch_opts.policy);
 * then every admitted system advances one outer iteration, in launches shared
 * by all of them; systems that converged (every column's true residual below
 * tol * ||B(:,c)||), hit maxiter or broke down are retired and their memory is
 * reused by the next admission.  Results are fetched with cacg_batch_collect.
 *
 * Every per-system sum runs in an order fixed by that system alone, so a
 * system's result is the same bits whether it runs alone or in any mix.
 *
 * Threads: cacg_batch_submit, cacg_batch_step, cacg_batch_retired and
 * cacg_batch_collect may be called from different threads at once (an
 * internal lock; a round releases it while it waits for the GPU, so a
 * submission's host-side copies overlap the round).  A feeder thread that
 * submits while another thread runs rounds is the intended use.
 * cacg_batch_destroy must not overlap any other call on the same batch.
 */
#ifndef CACG_BATCH_H
#define CACG_BATCH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cacg_batch cacg_batch;

typedef struct {
    long hbm_bytes;  /* device arena for the systems' data; 0: a quarter of the free HBM */
    int slots;       /* systems in flight at most, 1..4096; 0: 1024 */
    int s;           /* basis length of every system, >= 1 (s * m <= 128) */
    int precision;   /* CACG_PREC_FP64, or CACG_PREC_FP32_REFINE: FP32 matrix copy, basis and
                        Chebyshev SpMMs; Q, A Q, Gram, LU, X and R FP64 */
    long maxiter;    /* outer-iteration cap per system */
    int max_large;   /* large systems (solo_rows rows or more: each runs the single-system
                        kernels on its own stream) in flight at most; 0: no limit.  A queued
                        large system at the head waits for one to retire (FIFO: later requests
                        wait too) */
    int graph;       /* 1: when a round's batched launches repeat the previous round's (same
                        work lists), capture them once in a HIP graph and replay it, with no
                        work-list upload; 0 (default): launch them one by one, which measured
                        ~2% faster on the 256-system batch.  Same bits either way */
    int tune;        /* 1 (default): the first iterations of a large system run
                        its SpMMs under candidate schedules (tile order, grid), timed on its
                        stream, and the rest under the fastest; 0: default schedules.  Same
                        bits */
    int policy;      /* admission order: CACG_BATCH_FIFO (default) or
                        CACG_BATCH_SMALLEST_FIRST (smallest HBM footprint first among the
                        queued requests that may be admitted now, with max_bypass bounding how
                        often a request is overtaken).  Results do not depend on it */
    int max_bypass;  /* CACG_BATCH_SMALLEST_FIRST: a queued request overtaken by this many
                        later ones is admitted next (and nothing overtakes it while it waits
                        for memory or a slot); 0: 16 */
    long solo_rows;  /* systems of at least this many rows are large: they run the
                        single-system kernels (recursive residual, the X update deferred into
                        the next Q update, a final round on the true residual) on streams of
                        their own instead of joining the batched launches; 0: 2^17.  Part of
                        the batch's configuration: a result is the same bits alone and in any
                        mix of batches with the same solo_rows */
} cacg_batch_opts;

enum { CACG_BATCH_FIFO = 0, CACG_BATCH_SMALLEST_FIRST = 1 };

typedef struct {
    int status;          /* CACG_OK, CACG_MAXITER or CACG_BREAKDOWN */
    long iters;          /* outer iterations */
    long rounds_queued;  /* rounds spent waiting for admission */
    double true_relres;  /* max over columns of ||B - A X||/||B|| at retirement */
} cacg_batch_result;

typedef struct {
    int active;          /* systems admitted and iterating */
    int queued;          /* submitted, waiting for a slot or memory */
    int finished;        /* retired, not yet collected */
    int admitted;        /* admitted in the last round */
    int retired;         /* retired in the last round */
    long rounds;         /* rounds so far */
    long arena_used, arena_bytes;
    long graph_rounds;   /* rounds so far whose batched launches ran as a graph replay */
} cacg_batch_stats;

void cacg_batch_default_opts(cacg_batch_opts *o);
int cacg_batch_create(cacg_batch **out, const cacg_batch_opts *opts);
int cacg_batch_destroy(cacg_batch *b);

/* bytes of arena one system takes (its matrix, blocks and work space) in a
 * batch with the default solo_rows */
long cacg_batch_footprint(long n, long nnz, long m, int s, int precision);

/* Queue one system: CSR of n rows (row_start n + 1, col_idx / val nnz), B and
 * X0 (NULL: zero) n x m row-major, 1 <= m <= 32, s * m <= 128.  Returns the
 * request id (>= 0), or -CACG_EINVAL for bad arguments and -CACG_ENOMEM when
 * the system can never fit the arena. */
long cacg_batch_submit(cacg_batch *b, long n, const long *row_start, const long *col_idx,
                       const double *val, long m, const double *B, const double *X0, double tol);

/* One round (admit, one outer iteration of every active system, retire).
 * Returns CACG_OK or a CACG_E* error; stats may be NULL. */
int cacg_batch_step(cacg_batch *b, cacg_batch_stats *stats);

/* Ids of the requests retired since the last call (at most cap of them, the
 * rest stay for the next call); returns how many were written. */
int cacg_batch_retired(cacg_batch *b, long *ids, int cap);

/* Fetch a retired request: copies its solution into X (n x m) and fills res;
 * returns 1 and forgets the request, 0 while it is queued or active, or
 * -CACG_EINVAL for an unknown id. */
int cacg_batch_collect(cacg_batch *b, long id, double *X, cacg_batch_result *res);

#ifdef __cplusplus
}
#endif
#endif /* CACG_BATCH_H */
