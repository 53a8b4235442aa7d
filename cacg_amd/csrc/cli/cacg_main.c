/*
 * cacg — the reference's command line program (src/main.c) on libcacg.so.
 *
 *   cacg <path> <matrix.mtx> <rhs.csv> <m> <solver> <s> [epsilon]
 *
 * Same arguments as the reference's executable (src/main.c:37-55; the run
 * script passes them as `$MSTORE/$2/ $2.mtx $MSTORE/$2/$2_$3.csv $3 $4 $5`,
 * run:4): the matrix file is <path><matrix.mtx> (plain concatenation, as
 * concatStr does), the RHS file is used as given, m columns are read from it,
 * solver 0 = block CG (bcg_v1), 1 = CBCG (cbcg_v1), 2 = BCBCG (bcbcg_v1), s is
 * the basis length.  As in the reference, X starts at all ones (src/main.c:93)
 * and the stop threshold is the absolute epsilon = 20000 on ||R(:,0)||
 * (src/main.c:35); an optional seventh argument replaces that epsilon (the
 * reference ignores extra arguments).  One process: the whole matrix is
 * local, the role of MPI rank 0 of 1.
 *
 * The program prints the reference's progress lines (src/main.c:64-65,
 * src/readMTX.c:56,91, src/SparseMatrixDistribution.c:38, src/GetRHS.c:172-174,
 * src/main.c:114-164) and the solvers print their Loop lines.  Environment:
 * CACG_ORDER=fast selects the split-K MFMA order (default: the reference's
 * order), CACG_QUIET=1 silences the Loop lines, CACG_X_OUT=<file> writes the
 * solution X as CSV in the RHS format (the reference writes nothing).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "cacg.h"
#include "cacg_io.h"

static double wall(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        printf("Argument setting is wrong.\n");
        return 0;
    }
    const long m = atoi(argv[4]);
    const long solver = atoi(argv[5]);
    const long s = atoi(argv[6]);
    const double epsilon = argc > 7 ? atof(argv[7]) : 20000.0;
    printf("myid:%d, numprocs: %d\n", 0, 1);
    printf("we can want to generate RHS with size of %ld\n", m);

    char *mtx = (char *)malloc(strlen(argv[1]) + strlen(argv[2]) + 1);
    if (!mtx)
        return 1;
    strcpy(mtx, argv[1]);
    strcat(mtx, argv[2]);
    long nr, nc, nnz, *rs, *ci;
    double *val;
    printf("Reading sparse matrix from file ( %s ):\n", mtx);
    int rc = cacg_mtx_read(mtx, &nr, &nc, &nnz, &rs, &ci, &val);
    if (rc != CACG_OK) {
        printf("Unable to read the Matrix Market file %s: %s\n", mtx, cacg_status_string(rc));
        return 1;
    }
    printf(" done reading\n");
    printf("there are %ld elements in matrix \n", nnz);
    if (m <= 0 || nr != nc) {
        printf("cacg: need a square matrix and m >= 1 (got %ld x %ld, m = %ld)\n", nr, nc, m);
        return 1;
    }

    double *B;
    printf("Reading MRHS data from %s ... ...\n", argv[3]);
    rc = cacg_csv_read(argv[3], nr, m, &B);
    if (rc != CACG_OK) {
        printf("Unable to read %ld x %ld values from %s: %s\n", nr, m, argv[3],
               cacg_status_string(rc));
        return 1;
    }
    printf("Reading MRHS data from %s done.\n", argv[3]);
    double *Xd = (double *)malloc(sizeof(double) * (size_t)(nr * m));
    if (!Xd)
        return 1;
    for (long e = 0; e < nr * m; e++)
        Xd[e] = 1.0;

    csrType_local A = {nr, nc, nnz, 0, rs, ci, val};
    denseType Bd = {nr, m, nr, m, 0, B};
    denseType X = {nr, m, nr, m, 0, Xd};
    const char *name = NULL;
    double t1 = 0.0;
    switch (solver) {
    case 0:
        printf("Going to call BCG solver ... ...\n");
        printf("m value:%ld\n", m);
        name = "BCG";
        fflush(stdout);
        t1 = wall();
        bcg_v1(A, Bd, X, epsilon, 0, 1);
        break;
    case 1:
        printf("Going to call CBCG solver ... ...\n");
        printf("s value:%ld; m value:%ld\n", s, m);
        name = "CBCG";
        fflush(stdout);
        t1 = wall();
        cbcg_v1(A, Bd, X, s, epsilon, 0, 1);
        break;
    case 2:
        printf("Going to call BCBCG solver ... ...\n");
        printf("s value:%ld; m value:%ld\n", s, m);
        name = "BCBCG";
        fflush(stdout);
        t1 = wall();
        bcbcg_v1(A, Bd, X, s, epsilon, 0, 1);
        break;
    default:
        printf("Sorry, wrong solver index\n");
        break;
    }
    int status = 0;
    if (name) {
        printf("Total time for %s solver: %f secs\n", name, wall() - t1);
        const int st = cacg_last_status();
        if (st != CACG_OK && st != CACG_MAXITER) {
            fprintf(stderr, "cacg: %s solver: %s\n", name, cacg_status_string(st));
            status = 1;
        }
        const char *xout = getenv("CACG_X_OUT");
        if (xout && xout[0] && cacg_csv_write(xout, nr, m, Xd) != CACG_OK) {
            fprintf(stderr, "cacg: cannot write %s\n", xout);
            status = 1;
        }
    }
    cacg_io_free(rs);
    cacg_io_free(ci);
    cacg_io_free(val);
    cacg_io_free(B);
    free(Xd);
    free(mtx);
    return status;
}
