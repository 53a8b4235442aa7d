// Lanczos tridiagonalisation for the Chebyshev interval.  gfx950.
//
// k steps of the symmetric Lanczos recurrence from a fixed start vector
// (entries in [0.5, 1.5) from a splitmix64 hash of the row index, normalised):
//   w = A v_j;  alpha_j = v_j . w;  w -= alpha_j v_j + beta_{j-1} v_{j-1};
//   beta_j = ||w||;  v_{j+1} = w / beta_j
// with no reorthogonalisation (the extreme Ritz values converge first; a lost
// orthogonality only adds copies of converged ones).  Everything stays on the
// device: the scalars are written to d_ab (alpha_0..k-1, then beta_0..k-1)
// and read by the next step's kernels, so the host reads k-by-2 numbers once.
// Every sum runs in a fixed order (LB block partials, then one block), so the
// coefficients do not depend on the device.  No reference counterpart (the
// reference bounds the spectrum by Gerschgorin discs, src/common.c:53-106).
#include <stdint.h>

#include "kcommon.hpp"

namespace {

constexpr int LB = 512;  // partial sums per reduction (fixed)

__device__ __forceinline__ double start_entry(long i)
{
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 0.5 + (double)(z >> 11) * 0x1p-53;
}

__device__ __forceinline__ void block_sum_store(double acc, double *out)
{
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = red[0];
}

// v = start vector; part[b] = this block's sum of v^2
__global__ __launch_bounds__(256) void lz_init_kernel(long n, double *__restrict__ v,
                                                      double *__restrict__ part)
{
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double u = start_entry(i);
        v[i] = u;
        acc += u * u;
    }
    block_sum_store(acc, part + blockIdx.x);
}

// part[b] = this block's sum of x . y
__global__ __launch_bounds__(256) void lz_dot_kernel(long n, const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     double *__restrict__ part)
{
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        acc += x[i] * y[i];
    block_sum_store(acc, part + blockIdx.x);
}

// *out = sum of the LB partials (root = 1: its square root)
__global__ __launch_bounds__(256) void lz_final_kernel(const double *__restrict__ part,
                                                       double *__restrict__ out, int root)
{
    double acc = 0.0;
    for (int g = threadIdx.x; g < LB; g += 256)
        acc += part[g];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = root ? sqrt(red[0]) : red[0];
}

// y = x / *s (y = 0 when *s is 0: an invariant subspace was found)
__global__ __launch_bounds__(256) void lz_scale_kernel(long n, const double *__restrict__ x,
                                                       const double *__restrict__ s,
                                                       double *__restrict__ y)
{
    const double d = *s;
    const double r = d > 0.0 ? 1.0 / d : 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = x[i] * r;
}

// w -= alpha v + beta vprev (vprev may be null); part[b] = block sum of w^2
__global__ __launch_bounds__(256) void lz_update_kernel(long n, double *__restrict__ w,
                                                        const double *__restrict__ v,
                                                        const double *__restrict__ vprev,
                                                        const double *__restrict__ alpha,
                                                        const double *__restrict__ beta,
                                                        double *__restrict__ part)
{
    const double a = *alpha, b = vprev ? *beta : 0.0;
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double x = w[i] - a * v[i];
        if (vprev)
            x -= b * vprev[i];
        w[i] = x;
        acc += x * x;
    }
    block_sum_store(acc, part + blockIdx.x);
}

}  // namespace

extern "C" long cacg_lanczos_workspace(long n)
{
    return 3 * n + LB + 1;
}

extern "C" int cacg_k_lanczos(const cacg_dcsr *A, int steps, double *d_work, double *d_ab,
                              hipStream_t st)
{
    if (!A || A->n <= 0 || steps <= 0 || !d_work || !d_ab)
        return (int)hipErrorInvalidValue;
    const long n = A->n;
    double *buf[3] = {d_work, d_work + n, d_work + 2 * n};
    double *part = d_work + 3 * n, *norm0 = part + LB;
    double *alpha = d_ab, *beta = d_ab + steps;
    const int G = LB;
    // v_0 = u / ||u||
    hipLaunchKernelGGL(lz_init_kernel, dim3(G), dim3(256), 0, st, n, buf[0], part);
    CACG_LAUNCH_CHECK();
    hipLaunchKernelGGL(lz_final_kernel, dim3(1), dim3(256), 0, st, part, norm0, 1);
    CACG_LAUNCH_CHECK();
    hipLaunchKernelGGL(lz_scale_kernel, dim3(G), dim3(256), 0, st, n, buf[0], norm0, buf[0]);
    CACG_LAUNCH_CHECK();
    int cur = 0, prev = -1, wb = 1;
    for (int j = 0; j < steps; j++) {
        int rc = cacg_k_spmm(A, 1, buf[cur], 1, buf[wb], 1, CACG_EPI_NONE, 0.0, 0.0, nullptr, 0,
                             nullptr, 0, nullptr, 0, st);
        if (rc)
            return rc;
        hipLaunchKernelGGL(lz_dot_kernel, dim3(G), dim3(256), 0, st, n, buf[cur], buf[wb], part);
        CACG_LAUNCH_CHECK();
        hipLaunchKernelGGL(lz_final_kernel, dim3(1), dim3(256), 0, st, part, alpha + j, 0);
        CACG_LAUNCH_CHECK();
        hipLaunchKernelGGL(lz_update_kernel, dim3(G), dim3(256), 0, st, n, buf[wb], buf[cur],
                           prev >= 0 ? buf[prev] : nullptr, alpha + j,
                           j > 0 ? beta + j - 1 : alpha + j, part);
        CACG_LAUNCH_CHECK();
        hipLaunchKernelGGL(lz_final_kernel, dim3(1), dim3(256), 0, st, part, beta + j, 1);
        CACG_LAUNCH_CHECK();
        if (j + 1 == steps)
            break;
        // v_{j+1} = w / beta_j into the buffer v_{j-1} used (or the free one)
        const int nxt = prev >= 0 ? prev : 2;
        hipLaunchKernelGGL(lz_scale_kernel, dim3(G), dim3(256), 0, st, n, buf[wb], beta + j,
                           buf[nxt]);
        CACG_LAUNCH_CHECK();
        prev = cur;
        cur = nxt;
        wb = 3 - prev - cur;
    }
    return 0;
}
