/*
 * Matrix Market and CSV input/output with the reference's semantics
 * (include/cacg_io.h states them; src/readMTX.c:8-136, src/mmio.c:93-214,
 * src/matrixType.c:4-52, src/common.c:16-48 are what they follow).
 *
 * The whole file is read into memory once and scanned with strtol / strtod,
 * which are the conversions fscanf("%ld") / fscanf("%lf") and atof perform,
 * so every value is the double the reference reads.
 */
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cacg.h"
#include "cacg_io.h"

void cacg_io_free(void *p) { free(p); }

/* the file as one NUL-terminated buffer */
static int slurp(const char *path, char **buf, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return CACG_EIO;
    size_t cap = 1 << 16, n = 0;
    char *b = (char *)malloc(cap + 1);
    if (!b) {
        fclose(f);
        return CACG_ENOMEM;
    }
    for (;;) {
        if (n == cap) {
            cap *= 2;
            char *nb = (char *)realloc(b, cap + 1);
            if (!nb) {
                free(b);
                fclose(f);
                return CACG_ENOMEM;
            }
            b = nb;
        }
        const size_t got = fread(b + n, 1, cap - n, f);
        n += got;
        if (got == 0)
            break;
    }
    const int err = ferror(f);
    fclose(f);
    if (err) {
        free(b);
        return CACG_EIO;
    }
    b[n] = '\0';
    *buf = b;
    *len = n;
    return CACG_OK;
}

/* ---- Matrix Market ------------------------------------------------------------ */

/* up to `max` whitespace-separated words of the line [p, e), lower-cased copies
 * of words 1.. (word 0, the banner, compared as written) */
static int banner_words(const char *p, const char *e, char w[5][64])
{
    int k = 0;
    while (k < 5) {
        while (p < e && isspace((unsigned char)*p))
            p++;
        if (p >= e)
            break;
        int l = 0;
        while (p < e && !isspace((unsigned char)*p)) {
            if (l < 63)
                w[k][l++] = k == 0 ? *p : (char)tolower((unsigned char)*p);
            p++;
        }
        w[k][l] = '\0';
        k++;
    }
    return k;
}

/* one integer token: the whole token must convert */
static int next_long(char **pp, long *v)
{
    char *p = *pp;
    while (isspace((unsigned char)*p))
        p++;
    if (!*p)
        return 0;
    char *end;
    errno = 0;
    const long x = strtol(p, &end, 10);
    if (end == p || errno == ERANGE || (*end && !isspace((unsigned char)*end)))
        return 0;
    *v = x;
    *pp = end;
    return 1;
}

static int next_double(char **pp, double *v)
{
    char *p = *pp;
    while (isspace((unsigned char)*p))
        p++;
    if (!*p)
        return 0;
    char *end;
    const double x = strtod(p, &end);
    if (end == p || (*end && !isspace((unsigned char)*end)))
        return 0;
    *v = x;
    *pp = end;
    return 1;
}

int cacg_mtx_read(const char *path, long *num_rows, long *num_cols, long *nnz, long **row_start,
                  long **col_idx, double **val)
{
    if (!path || !num_rows || !num_cols || !nnz || !row_start || !col_idx || !val)
        return CACG_EINVAL;
    *row_start = NULL;
    *col_idx = NULL;
    *val = NULL;
    char *buf = NULL;
    size_t len = 0;
    int rc = slurp(path, &buf, &len);
    if (rc != CACG_OK)
        return rc;
    long *I = NULL, *J = NULL, *rs = NULL, *ci = NULL;
    double *V = NULL, *cv = NULL;
    rc = CACG_EINVAL;

    /* banner: five words, src/mmio.c:93-175 */
    char *line_end = strchr(buf, '\n');
    if (!line_end)
        line_end = buf + len;
    char w[5][64];
    if (banner_words(buf, line_end, w) != 5 || strncmp(w[0], "%%MatrixMarket", 14) != 0 ||
        strcmp(w[1], "matrix") != 0)
        goto out;
    const int coordinate = strcmp(w[2], "coordinate") == 0;
    if (!coordinate && strcmp(w[2], "array") != 0)
        goto out;
    const int real = strcmp(w[3], "real") == 0, integer = strcmp(w[3], "integer") == 0;
    const int pattern = strcmp(w[3], "pattern") == 0, complex_ = strcmp(w[3], "complex") == 0;
    if (!real && !integer && !pattern && !complex_)
        goto out;
    const int symmetric = strcmp(w[4], "symmetric") == 0;
    if (!symmetric && strcmp(w[4], "general") != 0 && strcmp(w[4], "skew-symmetric") != 0 &&
        strcmp(w[4], "hermitian") != 0)
        goto out;
    /* what src/readMTX.c:33-38 accepts: real, integer or pattern coordinate files */
    if (!coordinate || !(real || integer || pattern))
        goto out;
    /* src/mmio.c:83-91: real hermitian, pattern hermitian / skew-symmetric are invalid */
    if ((real && strcmp(w[4], "hermitian") == 0) ||
        (pattern && (strcmp(w[4], "hermitian") == 0 || strcmp(w[4], "skew-symmetric") == 0)))
        goto out;

    /* comment lines, then rows, cols, entries (src/mmio.c:186-214) */
    char *p = line_end < buf + len ? line_end + 1 : line_end;
    while (*p == '%') {
        char *nl = strchr(p, '\n');
        p = nl ? nl + 1 : buf + len;
    }
    long M, N, L;
    if (!next_long(&p, &M) || !next_long(&p, &N) || !next_long(&p, &L) || M < 0 || N < 0 ||
        L < 0)
        goto out;

    /* entries (src/readMTX.c:58-87) */
    const long cap = symmetric ? 2 * L : L;
    I = (long *)malloc(sizeof(long) * (size_t)(cap > 0 ? cap : 1));
    J = (long *)malloc(sizeof(long) * (size_t)(cap > 0 ? cap : 1));
    V = (double *)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
    rs = (long *)calloc((size_t)M + 2, sizeof(long));
    if (!I || !J || !V || !rs) {
        rc = CACG_ENOMEM;
        goto out;
    }
    long k = 0;
    for (long e = 0; e < L; e++) {
        long r, c;
        double v = 1.0;
        if (!next_long(&p, &r) || !next_long(&p, &c) || (!pattern && !next_double(&p, &v)))
            goto out;
        if (r < 1 || r > M || c < 1 || c > N)
            goto out;
        r--;
        c--;
        I[k] = r;
        J[k] = c;
        V[k] = v;
        k++;
        if (symmetric && r != c) { /* src/readMTX.c:93-134: (i, j) then (j, i) */
            if (c >= M || r >= N)
                goto out;
            I[k] = c;
            J[k] = r;
            V[k] = v;
            k++;
        }
    }

    /* stable bucket sort by row (src/matrixType.c:4-52) */
    ci = (long *)malloc(sizeof(long) * (size_t)(k > 0 ? k : 1));
    cv = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
    if (!ci || !cv) {
        rc = CACG_ENOMEM;
        goto out;
    }
    for (long e = 0; e < k; e++)
        rs[I[e] + 1]++;
    for (long i = 0; i < M; i++)
        rs[i + 1] += rs[i];
    /* rs[i] is now the next free slot of row i; shift back afterwards */
    for (long e = 0; e < k; e++) {
        const long d = rs[I[e]]++;
        ci[d] = J[e];
        cv[d] = V[e];
    }
    for (long i = M; i > 0; i--)
        rs[i] = rs[i - 1];
    rs[0] = 0;

    *num_rows = M;
    *num_cols = N;
    *nnz = k;
    *row_start = rs;
    *col_idx = ci;
    *val = cv;
    rs = NULL;
    ci = NULL;
    cv = NULL;
    rc = CACG_OK;
out:
    free(buf);
    free(I);
    free(J);
    free(V);
    free(rs);
    free(ci);
    free(cv);
    return rc;
}

int cacg_mtx_write(const char *path, long num_rows, long num_cols, const long *row_start,
                   const long *col_idx, const double *val)
{
    if (!path || num_rows < 0 || num_cols < 0 || !row_start || (row_start[num_rows] > 0 &&
                                                                (!col_idx || !val)))
        return CACG_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f)
        return CACG_EIO;
    int ok = fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%ld %ld %ld\n", num_rows,
                     num_cols, row_start[num_rows]) > 0;
    for (long i = 0; i < num_rows && ok; i++)
        for (long q = row_start[i]; q < row_start[i + 1] && ok; q++)
            ok = fprintf(f, "%ld %ld %.17g\n", i + 1, col_idx[q] + 1, val[q]) > 0;
    if (fclose(f) != 0)
        ok = 0;
    return ok ? CACG_OK : CACG_EIO;
}

/* ---- CSV ------------------------------------------------------------------------ */

/* strtok-style field scan of [p, e): skip delimiters, return the field start
 * and set *fe to its end; NULL when no field is left */
static char *field(char *p, char *e, const char *delims, char **fe)
{
    while (p < e && strchr(delims, *p))
        p++;
    if (p >= e)
        return NULL;
    char *q = p;
    while (q < e && !strchr(delims, *q))
        q++;
    *fe = q;
    return p;
}

int cacg_csv_read(const char *path, long rows, long cols, double **data)
{
    if (!path || rows <= 0 || cols <= 0 || !data)
        return CACG_EINVAL;
    *data = NULL;
    char *buf = NULL;
    size_t len = 0;
    int rc = slurp(path, &buf, &len);
    if (rc != CACG_OK)
        return rc;
    double *out = (double *)calloc((size_t)rows * (size_t)cols, sizeof(double));
    if (!out) {
        free(buf);
        return CACG_ENOMEM;
    }
    char *p = buf, *end = buf + len;
    rc = CACG_OK;
    for (long r = 0; r < rows && p < end; r++) {
        char *nl = memchr(p, '\n', (size_t)(end - p));
        char *le = nl ? nl + 1 : end; /* the line keeps its newline, as fgets does */
        long c = 0;
        char *fe;
        /* src/common.c:33-37: first field split at ',' only, the rest at ',' or '\n' */
        for (char *f = field(p, le, ",", &fe); f; f = field(fe, le, ",\n", &fe)) {
            if (c < cols) {
                const char save = *fe;
                *fe = '\0';
                out[r * cols + c] = atof(f);
                *fe = save;
            }
            c++;
        }
        if (c != cols) { /* src/common.c:40 asserts this */
            rc = CACG_EINVAL;
            break;
        }
        p = le;
    }
    free(buf);
    if (rc != CACG_OK) {
        free(out);
        return rc;
    }
    *data = out;
    return CACG_OK;
}

int cacg_csv_write(const char *path, long rows, long cols, const double *data)
{
    if (!path || rows < 0 || cols <= 0 || (rows > 0 && !data))
        return CACG_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f)
        return CACG_EIO;
    int ok = 1;
    for (long r = 0; r < rows && ok; r++)
        for (long c = 0; c < cols && ok; c++)
            ok = fprintf(f, c + 1 < cols ? "%.18e," : "%.18e\n", data[r * cols + c]) > 0;
    if (fclose(f) != 0)
        ok = 0;
    return ok ? CACG_OK : CACG_EIO;
}
