/*
 * cacg_io.h — the reference's input formats, for callers that start from
 * files: Matrix Market coordinate matrices and comma-separated right-hand
 * sides.  Host code only (no GPU needed).
 *
 * cacg_mtx_read replaces the reference's readMtx_info_and_coo
 * (src/readMTX.c:8-136, banner and size line as src/mmio.c:93-214 reads them)
 * followed by Converter_Coo2Csr (src/matrixType.c:4-52), and returns the CSR
 * the reference builds, entry for entry:
 *   - "%%MatrixMarket matrix coordinate {real|integer|pattern}
 *     {general|symmetric|skew-symmetric}" is accepted (keywords in any case),
 *     and "integer hermitian", which the reference lets through and reads as
 *     general; array (dense) files, complex ones, real hermitian and pattern
 *     hermitian / skew-symmetric ones are refused, as there;
 *   - indices are 1-based in the file, 0-based in the CSR; pattern entries get
 *     the value 1.0; integer values are read as doubles;
 *   - symmetric files are expanded as the reference does: an off-diagonal
 *     entry (i, j) becomes (i, j) immediately followed by (j, i);
 *     skew-symmetric files are NOT expanded (the reference expands only
 *     symmetric ones, src/readMTX.c:93) and are read as general;
 *   - the CSR is a stable bucket sort of the entries by row: inside a row the
 *     entries keep their file order (no column sort) and duplicates are kept
 *     (not summed), exactly as src/matrixType.c:31-42 leaves them.
 * Where the reference aborts (assert) or reads out of bounds — a short file,
 * an index outside 1..rows / 1..cols, an unparsable token — this returns
 * CACG_EINVAL.  Comment lines ('%' first) are allowed only before the size
 * line.  Sizes are read as 64-bit (the reference reads them as int).
 *
 * cacg_csv_read replaces parseCSV (src/common.c:16-48) as GenVector_ReadCSV
 * (src/GetRHS.c:124-208) calls it: the first `rows` lines are split at ','
 * (the first field of a line at ',' only, later ones at ',' or newline, empty
 * fields skipped), each field converted as atof does, and every line must give
 * exactly `cols` values; rows beyond the end of the file stay zero.  Lines may
 * be of any length (the reference's 1024-byte line buffer aborts on longer
 * ones).  data is rows x cols row-major.
 *
 * The writers produce what the reference's tools consume: cacg_csv_write in
 * the RHS tools' format (numpy.savetxt fmt='%.18e', delimiter=',',
 * RHSTools/MrhsGenUniform.py:19), which round-trips every double exactly;
 * cacg_mtx_write a "coordinate real general" file with %.17g values, rows in
 * order and each row's entries in CSR order, so that reading it back gives
 * the same CSR bit for bit.
 *
 * Arrays returned by the readers are malloc'ed; release them with
 * cacg_io_free.  Return values: CACG_OK, CACG_EINVAL, CACG_ENOMEM, CACG_EIO
 * (file cannot be opened, read or written).
 */
#ifndef CACG_IO_H
#define CACG_IO_H

#ifdef __cplusplus
extern "C" {
#endif

int cacg_mtx_read(const char *path, long *num_rows, long *num_cols, long *nnz, long **row_start,
                  long **col_idx, double **val);
int cacg_mtx_write(const char *path, long num_rows, long num_cols, const long *row_start,
                   const long *col_idx, const double *val);
int cacg_csv_read(const char *path, long rows, long cols, double **data);
int cacg_csv_write(const char *path, long rows, long cols, const double *data);
void cacg_io_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* CACG_IO_H */
