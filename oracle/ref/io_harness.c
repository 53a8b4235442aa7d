/*
 * Input-format harness for the reference (TEST INFRASTRUCTURE ONLY).
 *
 * Links the reference's own reader objects, compiled from /root/reference/src
 * by oracle/ref/Makefile, and dumps what they produce: the Matrix Market file
 * read by readMtx_info_and_coo (src/readMTX.c:8-136) and turned into CSR by
 * Converter_Coo2Csr (src/matrixType.c:4-52), and optionally a CSV block read
 * by parseCSV (src/common.c:16-48).  One process, no MPI calls.
 *
 *   io_harness file.mtx out.bin [rhs.csv rows cols]
 *
 * Output (little-endian): "CACGIO01", int64 num_rows, num_cols, nnz,
 * row_start[num_rows + 1], col_idx[nnz], float64 val[nnz]; then, with a CSV,
 * int64 rows, cols, float64 data[rows * cols].
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "DataTypes.h"

void readMtx_info_and_coo(char *path, char *name, matInfo *info, cooType *mat);
void Converter_Coo2Csr(cooType src, csrType_local *target, matInfo *mat_info);
void parseCSV(char *filename, double **output, long numRows, long numCols);

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 6) {
        fprintf(stderr, "usage: %s file.mtx out.bin [rhs.csv rows cols]\n", argv[0]);
        return 2;
    }
    matInfo info;
    cooType coo;
    csrType_local csr;
    char empty[1] = "";
    readMtx_info_and_coo(empty, argv[1], &info, &coo);
    Converter_Coo2Csr(coo, &csr, &info);
    FILE *out = fopen(argv[2], "wb");
    if (!out) {
        perror(argv[2]);
        return 2;
    }
    int64_t h[3] = {info.num_rows, info.num_cols, info.nnz};
    fwrite("CACGIO01", 1, 8, out);
    fwrite(h, sizeof(int64_t), 3, out);
    fwrite(csr.row_start, sizeof(long), (size_t)info.num_rows + 1, out);
    fwrite(csr.col_idx, sizeof(long), (size_t)info.nnz, out);
    fwrite(csr.csrdata, sizeof(double), (size_t)info.nnz, out);
    if (argc == 6) {
        const long rows = atol(argv[4]), cols = atol(argv[5]);
        double *data = NULL;
        parseCSV(argv[3], &data, rows, cols);
        int64_t d[2] = {rows, cols};
        fwrite(d, sizeof(int64_t), 2, out);
        fwrite(data, sizeof(double), (size_t)(rows * cols), out);
    }
    fclose(out);
    return 0;
}
