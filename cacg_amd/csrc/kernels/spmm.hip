// Multi-RHS CSR SpMM with fused Chebyshev-recurrence epilogues, and the
// Gerschgorin interval.  gfx950 (CDNA4), wave64.
//
// Scheduling: vector-per-row.  A row is served by LPR consecutive lanes, one
// output column per lane (and CPL columns per lane when k > 64), so the
// gather of x(col, :) is one contiguous k*8-byte segment per nonzero and the
// lanes of a row read the same (col, val) pair (one broadcast request).  Each
// output is summed in CSR order from 0, product and sum rounded separately,
// which is exactly the reference's arithmetic (src/BLAS3.c:78-112, 950-965):
// the basis and A*Q blocks are bit-identical to the reference's.  The nonzero
// loop is unrolled by UNR with all (col, val) and x loads issued before the
// in-order accumulation, for memory-level parallelism.
//
// Row ranges are contiguous per workgroup and contiguous per XCD (hardware
// workgroup ids are dealt round-robin over the 8 XCDs; xcd_remap undoes that)
// so the rows that share gathered x lines run under one L2.
#include <float.h>

#include "kcommon.hpp"

namespace {

__device__ __forceinline__ int xcd_remap(int g, int G)
{
    if ((G & 7) != 0)
        return g;
    return (g & 7) * (G >> 3) + (g >> 3);
}

template <int LPR, int CPL, int UNR, int EPI>
__global__ __launch_bounds__(256) void spmm_kernel(
    int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
    const double *__restrict__ val, int k, const double *__restrict__ x, long ldx,
    double *__restrict__ y, long ldy, double a, double b, const double *__restrict__ p1, long ld1,
    const double *__restrict__ p2, long ld2, double *__restrict__ copy0, long ldc, int chunk)
{
    constexpr int RPB = 256 / LPR;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x % LPR;
    const int r_begin = lb * chunk;
    int r_end = r_begin + chunk;
    if (r_end > n)
        r_end = n;
    for (int i = r_begin + threadIdx.x / LPR; i < r_end; i += RPB) {
        const int pb = rowptr[i], pe = rowptr[i + 1];
        for (int cb = 0; cb < k; cb += LPR * CPL) {
            double acc[CPL];
            int c[CPL];
#pragma unroll
            for (int t = 0; t < CPL; t++) {
                acc[t] = 0.0;
                c[t] = cb + lane + t * LPR;
            }
            for (int p = pb; p < pe; p += UNR) {
                int cc[UNR];
                double vv[UNR];
                double xv[UNR][CPL];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    const bool in = p + u < pe;
                    cc[u] = in ? colidx[p + u] : 0;
                    vv[u] = in ? val[p + u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < UNR; u++)
#pragma unroll
                    for (int t = 0; t < CPL; t++)
                        xv[u][t] = (p + u < pe && c[t] < k) ? x[(long)cc[u] * ldx + c[t]] : 0.0;
#pragma unroll
                for (int u = 0; u < UNR; u++)
                    if (p + u < pe)
#pragma unroll
                        for (int t = 0; t < CPL; t++)
                            acc[t] = add_rn(acc[t], mul_rn(vv[u], xv[u][t]));
            }
#pragma unroll
            for (int t = 0; t < CPL; t++) {
                if (c[t] >= k)
                    continue;
                double out;
                if (EPI == CACG_EPI_NONE) {
                    out = acc[t];
                } else if (EPI == CACG_EPI_RESID) {
                    out = sub_rn(p1[(long)i * ld1 + c[t]], acc[t]);
                } else if (EPI == CACG_EPI_CHEB1) {
                    const double xi = x[(long)i * ldx + c[t]];
                    out = add_rn(mul_rn(a, acc[t]), mul_rn(b, xi));
                    copy0[(long)i * ldc + c[t]] = xi;
                } else {
                    out = sub_rn(add_rn(mul_rn(a, acc[t]), mul_rn(b, p1[(long)i * ld1 + c[t]])),
                                 p2[(long)i * ld2 + c[t]]);
                }
                y[(long)i * ldy + c[t]] = out;
            }
        }
    }
}

template <int LPR, int CPL, int UNR>
int launch_spmm_lpr(const cacg_dcsr *A, int k, const double *x, long ldx, double *y, long ldy,
                    int epi, double a, double b, const double *p1, long ld1, const double *p2,
                    long ld2, double *copy0, long ldc, hipStream_t st)
{
    constexpr int RPB = 256 / LPR;
    // ~8 workgroups per CU over 256 CUs; a multiple of 8 for the XCD remap
    long blocks = ((long)A->n + RPB - 1) / RPB;
    int G = blocks > 2048 ? 2048 : (int)blocks;
    if (G >= 8)
        G &= ~7;
    if (G < 1)
        G = 1;
    long chunk = ((long)A->n + G - 1) / G;
    chunk = (chunk + RPB - 1) / RPB * RPB;
    G = (int)(((long)A->n + chunk - 1) / chunk);
    if (G >= 8 && (G & 7) != 0) {
        // keep the XCD remap a bijection: pad the grid, extra blocks get no rows
        G = (G + 7) & ~7;
    }
#define CACG_SPMM_CASE(E)                                                                        \
    hipLaunchKernelGGL((spmm_kernel<LPR, CPL, UNR, E>), dim3(G), dim3(256), 0, st, A->n,          \
                       A->rowptr, A->col, A->val, k, x, ldx, y, ldy, a, b, p1, ld1, p2, ld2, copy0, \
                       ldc, (int)chunk)
    switch (epi) {
    case CACG_EPI_NONE:
        CACG_SPMM_CASE(CACG_EPI_NONE);
        break;
    case CACG_EPI_RESID:
        CACG_SPMM_CASE(CACG_EPI_RESID);
        break;
    case CACG_EPI_CHEB1:
        CACG_SPMM_CASE(CACG_EPI_CHEB1);
        break;
    case CACG_EPI_CHEB:
        CACG_SPMM_CASE(CACG_EPI_CHEB);
        break;
    default:
        return (int)hipErrorInvalidValue;
    }
#undef CACG_SPMM_CASE
    CACG_LAUNCH_CHECK();
    return 0;
}

// ---- Gerschgorin ------------------------------------------------------------

__global__ __launch_bounds__(256) void gersh_partial(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ colidx,
                                                     const double *__restrict__ val,
                                                     double *__restrict__ work)
{
    double lo = DBL_MAX, hi = DBL_MIN;  // src/common.c:56-57
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double off = 0.0, d = 0.0;
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
            const double v = val[p];
            if (colidx[p] != i)
                off = add_rn(off, v > 0 ? v : -v);
            else
                d = v;
        }
        const double up = add_rn(off, d), dn = sub_rn(d, off);
        if (hi < up)
            hi = up;
        if (lo > dn)
            lo = dn;
    }
    __shared__ double slo[256], shi[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        work[2 * blockIdx.x] = slo[0];
        work[2 * blockIdx.x + 1] = shi[0];
    }
}

__global__ __launch_bounds__(256) void gersh_final(int nblk, const double *__restrict__ work,
                                                   double *__restrict__ out)
{
    __shared__ double slo[256], shi[256];
    double lo = DBL_MAX, hi = DBL_MIN;
    for (int g = threadIdx.x; g < nblk; g += 256) {
        lo = fmin(lo, work[2 * g]);
        hi = fmax(hi, work[2 * g + 1]);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = slo[0];
        out[1] = shi[0];
    }
}

}  // namespace

extern "C" int cacg_k_spmm(const cacg_dcsr *A, int k, const double *x, long ldx, double *y,
                           long ldy, int epi, double a, double b, const double *p1, long ld1,
                           const double *p2, long ld2, double *copy0, long ldc, hipStream_t st)
{
    if (!A || k <= 0 || A->n < 0)
        return (int)hipErrorInvalidValue;
    if (A->n == 0)
        return 0;
    if (k >= 64)
        return launch_spmm_lpr<64, 2, 4>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                         ldc, st);
    if (k > 16)
        return launch_spmm_lpr<32, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                         ldc, st);
    if (k > 8)
        return launch_spmm_lpr<16, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                         ldc, st);
    if (k > 4)
        return launch_spmm_lpr<8, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                        ldc, st);
    if (k > 2)
        return launch_spmm_lpr<4, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                        ldc, st);
    if (k > 1)
        return launch_spmm_lpr<2, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0,
                                        ldc, st);
    return launch_spmm_lpr<1, 1, 8>(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0, ldc,
                                    st);
}

extern "C" int cacg_k_gerschgorin(const cacg_dcsr *A, double *d_work, double *d_out,
                                  hipStream_t st)
{
    if (!A || A->n <= 0)
        return (int)hipErrorInvalidValue;
    int G = cacg_grid(A->n, 256, CACG_GERSH_BLOCKS);
    hipLaunchKernelGGL(gersh_partial, dim3(G), dim3(256), 0, st, A->n, A->rowptr, A->col, A->val,
                       d_work);
    CACG_LAUNCH_CHECK();
    hipLaunchKernelGGL(gersh_final, dim3(1), dim3(256), 0, st, G, d_work, d_out);
    CACG_LAUNCH_CHECK();
    return 0;
}
