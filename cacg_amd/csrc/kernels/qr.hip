// Tall-skinny QR of an n x k block (k <= 128): A = Q R, Q n x k with
// orthonormal columns, R k x k upper triangular with a positive diagonal.
// gfx950.
//
// Algorithm: CholeskyQR2.  One pass is G = A^T A (the split-K FP64 MFMA
// Gram, a deterministic fixed-order reduction), R1 = chol(G) and
// W = R1^{-1} in one workgroup with G in LDS, then Q1 = A W (the tall-skinny
// MFMA update); the second pass repeats it on Q1 and R = R2 R1.  The second
// pass brings the loss of orthogonality from O(u cond(A)^2) down to O(u) for
// cond(A) up to about u^{-1/2} (1e8); a Cholesky pivot that is not finite or
// not above 64 u times its original diagonal entry (A numerically rank
// deficient, or cond(A) beyond that range) sets *d_flag.
// No reference counterpart: the reference's TSQR.c is empty.
#include <math.h>

#include "kcommon.hpp"

namespace {

constexpr int QT = 1024;

// G (N x N, row-major, symmetric positive definite) -> R (upper, N x N, the
// lower triangle written as zeros) and W = R^{-1} (upper); R is also kept in
// LDS.  mode 1: additionally Rout = R Rprev (both upper).
__global__ __launch_bounds__(QT) void chol_inv_kernel(int N, const double *__restrict__ G,
                                                      double *__restrict__ R,
                                                      double *__restrict__ W,
                                                      const double *__restrict__ Rprev,
                                                      double *__restrict__ Rout, int *flag)
{
    extern __shared__ double a[];  // N x N
    const int t = threadIdx.x;
    for (int e = t; e < N * N; e += QT)
        a[e] = G[e];
    __syncthreads();
    int bad = 0;
    // right-looking Cholesky on the upper triangle: row p of R is a[p][p..N)
    for (int p = 0; p < N; p++) {
        // a pivot below 64 u of its original diagonal entry has lost every
        // significant digit: column p is (numerically) in the span of 0 .. p-1
        const double piv = a[p * N + p];
        if (!(piv > 0x1p-47 * G[p * N + p]) || !isfinite(piv))
            bad = 1;
        const double d = sqrt(piv > 0.0 ? piv : 1.0);
        __syncthreads();
        if (t == 0)
            a[p * N + p] = d;
        const double inv = 1.0 / d;
        for (int c = p + 1 + t; c < N; c += QT)
            a[p * N + c] *= inv;
        __syncthreads();
        // trailing update of the upper triangle: a[r][c] -= a[p][r] a[p][c], p < r <= c
        const int w = N - p - 1;
        for (int e = t; e < w * w; e += QT) {
            const int r = p + 1 + e / w, c = p + 1 + e % w;
            if (c >= r)
                a[r * N + c] -= a[p * N + r] * a[p * N + c];
        }
        __syncthreads();
    }
    if (t == 0 && bad)
        *flag = 1;
    for (int e = t; e < N * N; e += QT) {
        const int r = e / N, c = e % N;
        R[e] = c >= r ? a[e] : 0.0;
    }
    // W = R^{-1}, one column per lane: W[i][c] for i = c down to 0
    for (int c = t; c < N; c += QT) {
        for (int i = 0; i < N; i++)
            if (i > c)
                W[i * N + c] = 0.0;
        W[c * N + c] = 1.0 / a[c * N + c];
        for (int i = c - 1; i >= 0; i--) {
            double s = 0.0;
            for (int j = i + 1; j <= c; j++)
                s += a[i * N + j] * W[j * N + c];
            W[i * N + c] = -s / a[i * N + i];
        }
    }
    if (Rprev) {
        // Rout = R Rprev (upper x upper)
        for (int e = t; e < N * N; e += QT) {
            const int r = e / N, c = e % N;
            double s = 0.0;
            for (int j = r; j <= c; j++)
                s += a[r * N + j] * Rprev[j * N + c];
            Rout[e] = c >= r ? s : 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void copy_rows_kernel(long n, int k, const double *__restrict__ A,
                                                        long lda, double *__restrict__ Q)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n * k; e += (long)gridDim.x * 256) {
        const long i = e / k;
        const int c = (int)(e - i * k);
        Q[e] = A[i * lda + c];
    }
}

}  // namespace

extern "C" long cacg_tsqr_workspace(int k)
{
    // Gram partials, then G, R1, W, R2
    return cacg_gram_workspace(k, k) + 4L * k * k;
}

extern "C" int cacg_k_tsqr(int n, int k, const double *A, long lda, double *Q, double *R,
                           double *d_work, int *d_flag, hipStream_t st)
{
    if (n < 0 || k <= 0 || k > 128 || lda < k || !A || !Q || !R || !d_work || !d_flag)
        return (int)hipErrorInvalidValue;
    double *gw = d_work;
    double *G = gw + cacg_gram_workspace(k, k);
    double *R1 = G + (long)k * k, *W = R1 + (long)k * k, *R2 = W + (long)k * k;
    const size_t shm = (size_t)k * k * sizeof(double);
    cacg_allow_lds(chol_inv_kernel, shm);
    // Q = A (contiguous n x k)
    hipLaunchKernelGGL(copy_rows_kernel, dim3(cacg_grid((long)n * k, 256, 4096)), dim3(256), 0, st,
                       (long)n, k, A, lda, Q);
    CACG_LAUNCH_CHECK();
    for (int pass = 0; pass < 2; pass++) {
        int rc = cacg_k_gram_fast(n, Q, k, k, Q, k, k, NULL, 0, 0, G, NULL, gw, st);
        if (rc)
            return rc;
        // pass 0: R1 = chol(G); pass 1: R2 = chol(G), R = R2 R1
        hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(QT), shm, st, k, G,
                           pass == 0 ? R1 : R2, W, pass == 0 ? nullptr : R1,
                           pass == 0 ? nullptr : R, d_flag);
        CACG_LAUNCH_CHECK();
        // Q = Q W, as Q = 0 - Q (-W) with the in-place tall-skinny update
        rc = cacg_k_scale_f64(W, (long)k * k, -1.0, st);
        if (rc)
            return rc;
        rc = cacg_k_qupdate_zero(n, k, Q, W, st);
        if (rc)
            return rc;
    }
    return 0;
}
