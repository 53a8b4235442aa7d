/*
 * Golden-data harness for the reference solver (TEST INFRASTRUCTURE ONLY).
 *
 * Links the reference's own solver objects, compiled from
 * /root/reference/src by oracle/ref/Makefile, and calls bcbcg_v1 / cbcg_v1 /
 * bcg_v1 (src/bcbcgKernel.h:22, src/cbcgKernel.h:21, src/bcgKernel.h:21)
 * once, at one rank or under
 * `mpiexec -n P`.  Every rank reads the whole input and keeps its row block
 * the way src/SparseMatrixDistribution.c:64-108 and src/GetRHS.c:124-190 lay
 * it out: n/P rows per rank, the last rank takes the remainder, local
 * row_start rebased to 0, global column ids, csrType_local.start = first
 * global row, dense blocks with local rows and global sizes.
 *
 * The reference never writes X (src/DenseMatIO.c:3-7 is empty), so the link
 * wraps norm2square_dist_denseMat_col_n (-Wl,--wrap): the solver calls it once
 * per outer iteration right after the X and R updates
 * (src/bcbcgKernel.c:447, src/cbcgKernel.c:214) with R as its argument (bcg_v1:
 * src/bcgKernel.c:301, after the X update at :278 and before the R update at
 * :375, so it sees the residual the iteration started from); the
 * wrapper forwards the call unchanged and records the global ||R(:,0)||^2 and,
 * at one rank, R and X.  After the solver returns, X is gathered to rank 0.
 *
 * Input file (little-endian):  "CACGIN01", int64 n, nnz, m, s, solver
 * (0 = bcg_v1, 1 = cbcg_v1, 2 = bcbcg_v1), float64 epsilon, int64 rowptr[n+1],
 * int64 col[nnz], float64 val[nnz], float64 B[n*m], float64 X0[n*m].
 * Output file (rank 0): "CACGOUT2", int64 iterations, int64 P, int64 n,
 * int64 m, float64 r2[iterations], then at P == 1 per iteration float64
 * R[n*m], X[n*m], then the final float64 X[n*m].
 *
 * Large runs (environment, P == 1 only): CACG_REF_NOSNAP=1 keeps no R / X
 * snapshots (the output then holds r2 and the final X only) and writes
 * instead, to <output>.cols, float64 ||R(:,c)||^2 for every column c after
 * every outer iteration, each a sequential row-order sum of squares of the
 * recorded R (the reference's own norm loop, src/MatrixOperations.c:110-142,
 * applied to every column).  CACG_REF_PROGRESS=1 prints one line per outer
 * iteration to stderr (iteration, ||R(:,0)||^2, wall seconds).
 */
#include <mpi.h>
#include <stdint.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "DataTypes.h"

void bcbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
              int numprocs);
void cbcg_v1(csrType_local mat, denseType B, denseType X, long s, double epsilon, int myid,
             int numprocs);
void bcg_v1(csrType_local mat, denseType B, denseType X, double epsilon, int myid, int numprocs);
void __real_norm2square_dist_denseMat_col_n(denseType mat, double *result, long n, int myid,
                                            int numprocs);

#define MAXREC 100000
static double g_r2[MAXREC];
static FILE *g_snap; /* P == 1: per-iteration R, X snapshots (temporary file) */
static double *g_X;
static long g_nm_local;
static long g_iters;
static int g_np;
static FILE *g_cols;      /* CACG_REF_NOSNAP: per-iteration column norms^2 */
static long g_m;
static int g_progress;
static double g_t0;

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void __wrap_norm2square_dist_denseMat_col_n(denseType mat, double *result, long n, int myid,
                                            int numprocs)
{
    __real_norm2square_dist_denseMat_col_n(mat, result, n, myid, numprocs);
    if (g_iters < MAXREC)
        g_r2[g_iters] = *result;
    if (g_np == 1 && g_snap) {
        fwrite(mat.data, sizeof(double), (size_t)g_nm_local, g_snap);
        fwrite(g_X, sizeof(double), (size_t)g_nm_local, g_snap);
    }
    if (g_np == 1 && g_cols) {
        const long rows = g_nm_local / g_m;
        for (long c = 0; c < g_m; c++) {
            double acc = 0.0;
            for (long i = 0; i < rows; i++) {
                const double v = mat.data[i * g_m + c];
                const double sq = v * v;
                acc = acc + sq;
            }
            fwrite(&acc, sizeof(double), 1, g_cols);
        }
        fflush(g_cols);
    }
    if (g_progress && myid == 0) {
        fprintf(stderr, "ref_harness: iteration %ld r2 %.17g at %.6f s\n", g_iters, *result,
                wall() - g_t0);
        fflush(stderr);
    }
    g_iters++;
}

static void rd(void *p, size_t sz, size_t cnt, FILE *f)
{
    if (fread(p, sz, cnt, f) != cnt) {
        fprintf(stderr, "ref_harness: short read\n");
        exit(2);
    }
}

int main(int argc, char **argv)
{
    int myid, np;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    MPI_Comm_size(MPI_COMM_WORLD, &np);
    g_np = np;
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    FILE *in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    char magic[8];
    rd(magic, 1, 8, in);
    if (memcmp(magic, "CACGIN01", 8) != 0) {
        fprintf(stderr, "ref_harness: bad magic\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    int64_t hdr[5];
    double eps;
    rd(hdr, sizeof(int64_t), 5, in);
    rd(&eps, sizeof(double), 1, in);
    const long n = hdr[0], nnz = hdr[1], m = hdr[2], s = hdr[3], solver = hdr[4];
    long *rp = (long *)malloc(sizeof(long) * (n + 1));
    long *ci = (long *)malloc(sizeof(long) * (nnz > 0 ? nnz : 1));
    double *va = (double *)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    double *Bg = (double *)malloc(sizeof(double) * n * m);
    double *Xg = (double *)malloc(sizeof(double) * n * m);
    rd(rp, sizeof(long), n + 1, in);
    rd(ci, sizeof(long), nnz, in);
    rd(va, sizeof(double), nnz, in);
    rd(Bg, sizeof(double), n * m, in);
    rd(Xg, sizeof(double), n * m, in);
    fclose(in);

    /* this rank's row block */
    const long per = n / np;
    const long r0 = per * myid;
    const long nloc = (myid == np - 1) ? n - per * (np - 1) : per;
    csrType_local A;
    A.num_rows = nloc;
    A.num_cols = n;
    A.nnz = rp[r0 + nloc] - rp[r0];
    A.start = r0;
    A.row_start = (long *)malloc(sizeof(long) * (nloc + 1));
    for (long i = 0; i <= nloc; i++)
        A.row_start[i] = rp[r0 + i] - rp[r0];
    A.col_idx = ci + rp[r0];
    A.csrdata = va + rp[r0];

    denseType B, X;
    B.local_num_row = nloc;
    B.local_num_col = m;
    B.global_num_row = n;
    B.global_num_col = m;
    B.start_idx = myid * m * per;
    X = B;
    B.data = (double *)malloc(sizeof(double) * nloc * m);
    X.data = (double *)malloc(sizeof(double) * nloc * m);
    memcpy(B.data, Bg + r0 * m, sizeof(double) * nloc * m);
    memcpy(X.data, Xg + r0 * m, sizeof(double) * nloc * m);

    g_X = X.data;
    g_nm_local = nloc * m;
    g_m = m;
    const char *ns = getenv("CACG_REF_NOSNAP"), *pg = getenv("CACG_REF_PROGRESS");
    const int nosnap = ns && ns[0] == '1';
    g_progress = pg && pg[0] == '1';
    /* the reference's timing convention (src/main.c:154-163): barrier, clock,
     * solve, barrier, clock, the maximum over ranks */
    MPI_Barrier(MPI_COMM_WORLD);
    g_t0 = wall();
    char snap_path[4096];
    if (np == 1 && !nosnap) {
        snprintf(snap_path, sizeof(snap_path), "%s.snap", argv[2]);
        g_snap = fopen(snap_path, "wb");
    }
    if (np == 1 && nosnap) {
        char cols_path[4096];
        snprintf(cols_path, sizeof(cols_path), "%s.cols", argv[2]);
        g_cols = fopen(cols_path, "wb");
    }

    if (solver == 0)
        bcg_v1(A, B, X, eps, myid, np);
    else if (solver == 1)
        cbcg_v1(A, B, X, s, eps, myid, np);
    else
        bcbcg_v1(A, B, X, s, eps, myid, np);
    MPI_Barrier(MPI_COMM_WORLD);
    {
        const double mine = wall() - g_t0;
        double tmax = 0.0;
        MPI_Reduce(&mine, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
        if (g_progress && myid == 0) {
            fprintf(stderr, "ref_harness: solve %.6f s (max over %d ranks)\n", tmax, np);
            fflush(stderr);
        }
    }

    /* gather the final X to rank 0 */
    int *cnt = (int *)malloc(sizeof(int) * np), *dsp = (int *)malloc(sizeof(int) * np);
    for (int p = 0; p < np; p++) {
        const long rows = (p == np - 1) ? n - per * (np - 1) : per;
        cnt[p] = (int)(rows * m);
        dsp[p] = (int)(per * p * m);
    }
    MPI_Gatherv(X.data, (int)(nloc * m), MPI_DOUBLE, Xg, cnt, dsp, MPI_DOUBLE, 0, MPI_COMM_WORLD);

    if (myid == 0) {
        if (g_snap)
            fclose(g_snap);
        FILE *out = fopen(argv[2], "wb");
        if (!out) {
            perror(argv[2]);
            MPI_Abort(MPI_COMM_WORLD, 2);
        }
        int64_t h[4] = {g_iters, np, n, m};
        fwrite("CACGOUT2", 1, 8, out);
        fwrite(h, sizeof(int64_t), 4, out);
        fwrite(g_r2, sizeof(double), (size_t)(g_iters < MAXREC ? g_iters : MAXREC), out);
        if (g_cols)
            fclose(g_cols);
        if (np == 1 && g_snap) {
            FILE *sn = fopen(snap_path, "rb");
            char buf[1 << 16];
            size_t got;
            while (sn && (got = fread(buf, 1, sizeof(buf), sn)) > 0)
                fwrite(buf, 1, got, out);
            if (sn)
                fclose(sn);
            remove(snap_path);
        }
        fwrite(Xg, sizeof(double), (size_t)(n * m), out);
        fclose(out);
    }
    MPI_Finalize();
    return 0;
}
