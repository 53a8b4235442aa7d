// Small dense solve x = lhs^{-1} rhs on one workgroup (1024 lanes), the
// reference's LUsolver_v1_KeeplhsMat (src/DLSsolvers.c:106-208):
// Crout factorisation with unit-diagonal U and no pivoting, forward then
// backward substitution.  gfx950.
//
// Factorisation: right-looking, one barrier-separated step per pivot; every
// element sees the same operations in the same order as in the reference's
// loops, so the factor is bit-identical.  For N <= 128 the matrix is held in
// registers (4 x 4 elements per lane of a 32 x 32 lane grid) and the factor is
// then kept in LDS (128 KiB of the 160 KiB); larger N work in a global
// scratch block.
//
// Forward substitution runs column-oriented (for q: finalise row q, broadcast
// it through LDS, update the rows below).  Each y[i][c] still receives its
// subtractions in ascending q and the reciprocal scaling last, as in the
// reference's row-oriented loop, so it is bit-identical too.  The right-hand
// side is held in registers, JMAX rows per lane.
//
// Backward substitution: the reference subtracts x[q] u_iq in ascending q,
// which makes each row wait for all later rows; mode REFERENCE keeps that
// order (one sequential chain per column, x in LDS, U from the scratch copy).
// Mode FAST runs it column-oriented (descending q).
//
// With the factor in LDS it is stored column-major with an odd stride
// (element (r, c) at c (N + 1) + r), so the reads of a factor column by
// consecutive rows are consecutive words, and fast-order substitutions run
// wave-local: every right-hand-side column is solved by the lanes of one
// wave, which pass finished rows to each other by readlane or lane shuffle,
// with no barrier (lu_subst_wave).
#include <math.h>

#include "kcommon.hpp"

namespace {

constexpr int LU_T = 1024;
constexpr int JMAX = 16;

// right-hand-side columns per substitution pass: the register block (JMAX
// rows per lane) must cover N, and without the LDS factor the reference pass
// keeps N x kc in LDS; 0 when N is out of range
__host__ __device__ inline int lu_columns(int N, int k, bool in_lds)
{
    int kc = k < N ? k : N;
    if ((long)N * kc > (long)LU_T * JMAX)
        kc = (int)((long)LU_T * JMAX / N);
    if (!in_lds && (long)N * kc > 8192)
        kc = 8192 / N;
    if (kc < 1 || (LU_T / kc) * JMAX < N)
        return 0;
    return kc;
}

// dynamic LDS of a solve: factor (N x (N + 1)), broadcast rows (2 x kc),
// pivot row/column (2 x 2 x 128), 2 reciprocal pivots, N reciprocal diagonal
// entries; without the LDS factor an N x kc block instead of the factor
__host__ __device__ inline size_t lu_lds_bytes(int N, int k, bool in_lds)
{
    const int kc = lu_columns(N, k, in_lds);
    if (in_lds)
        return ((size_t)N * (N + 1) + 2 * (size_t)kc + 512 + 2 + N) * sizeof(double);
    return ((size_t)N * kc + 2 * (size_t)kc + N) * sizeof(double);
}

// position of factor element (r, c): column-major with stride N + 1 in LDS,
// row-major in the global scratch block
template <bool IN_LDS>
__device__ __forceinline__ long fidx(int N, int r, int c)
{
    return IN_LDS ? (long)c * (N + 1) + r : (long)r * N + c;
}

// v of lane src (src the same on every lane of the wave)
__device__ __forceinline__ double readlane_f64(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v of lane src of this lane's group of LPC lanes
template <int LPC>
__device__ __forceinline__ double group_bcast(double v, int src)
{
    if constexpr (LPC == 64)
        return readlane_f64(v, src);
    else
        return __shfl(v, src, LPC);
}

// Fast-order forward and backward substitutions with the factor in LDS
// (a: column-major, stride N + 1; dinv: reciprocal diagonal of L), N <= 128.
// Right-hand-side column `col` is solved by LPC lanes of one wave (64 / LPC
// columns per wave, 1024 / LPC per pass); lane lc holds rows lc + LPC r in
// registers.  Row q is finished by its lane and handed to the others of the
// group (readlane, or a lane shuffle for LPC < 64), which update their rows:
// ascending q for L y = rhs with the reciprocal pivot applied last, then
// descending q for U x = y (unit U).  Each element receives exactly the
// operations of the barrier-per-step form, in the same order, so the result
// is the same bits; nothing is shared between waves, so no barrier is needed.
// Columns cbase + w WC + l / LPC, then + cstride, ... (w = the wave in the
// workgroup); the single-workgroup solve passes 0 and the columns its waves
// cover per pass.
template <int LPC>
__device__ __forceinline__ void lu_subst_wave(int N, const double *__restrict__ a,
                                              const double *__restrict__ dinv, int k,
                                              const double *__restrict__ rhs,
                                              double *__restrict__ x, int cbase = 0,
                                              int cstride = (LU_T / 64) * (64 / LPC))
{
    constexpr int RPL = 128 / LPC;  // rows per lane
    constexpr int WC = 64 / LPC;    // columns per wave
    const int R = N + 1;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int lc = l % LPC;
    for (int c0 = cbase; c0 < k; c0 += cstride) {
        const int col = c0 + w * WC + l / LPC;
        const bool has = col < k;
        double y[RPL];
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lc + LPC * r;
            y[r] = (has && i < N) ? rhs[(long)i * k + col] : 0.0;
        }
#pragma unroll
        for (int rq = 0; rq < RPL; rq++) {
#pragma unroll 1
            for (int lq = 0; lq < LPC; lq++) {
                const int q = lq + LPC * rq;
                if (q >= N)
                    break;
                if (lc == lq)
                    y[rq] = mul_rn(y[rq], dinv[q]);
                const double yq = group_bcast<LPC>(y[rq], lq);
                const double *Lq = a + (long)q * R;
#pragma unroll
                for (int r = rq; r < RPL; r++) {
                    const int i = lc + LPC * r;
                    if (i > q && i < N)
                        y[r] = sub_rn(y[r], mul_rn(yq, Lq[i]));
                }
            }
        }
#pragma unroll
        for (int rq = RPL - 1; rq >= 0; rq--) {
#pragma unroll 1
            for (int lq = LPC - 1; lq >= 0; lq--) {
                const int q = lq + LPC * rq;
                if (q >= N)
                    continue;
                const double xq = group_bcast<LPC>(y[rq], lq);
                const double *Uq = a + (long)q * R;
#pragma unroll
                for (int r = 0; r <= rq; r++) {
                    const int i = lc + LPC * r;
                    if (i < q)
                        y[r] = sub_rn(y[r], mul_rn(xq, Uq[i]));
                }
            }
        }
        if (has) {
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const int i = lc + LPC * r;
                if (i < N)
                    x[(long)i * k + col] = y[r];
            }
        }
    }
}

// The solve of one system by one workgroup of LU_T lanes; lds is the dynamic
// LDS block (lu_lds_bytes).  Shared by the single-system kernel and the
// batched one.
template <bool IN_LDS>
__device__ __forceinline__ void lu_solve_body(int N, const double *__restrict__ lhs, int k,
                                              const double *__restrict__ rhs,
                                              double *__restrict__ x, double *__restrict__ gw,
                                              int *flag, int mode, int kc, int reuse, double *lds)
{
    double *a = IN_LDS ? lds : gw;
    double *rowbuf = IN_LDS ? lds + (long)N * (N + 1) : lds;
    double *xs = IN_LDS ? lds : lds + 2 * kc;
    const int t = threadIdx.x;

    // ---- factorisation (or the factor left in gw by the previous call)
    int bad = 0;
    if (reuse) {
        if (IN_LDS) {
            for (long e = t; e < (long)N * N; e += LU_T)
                a[fidx<true>(N, (int)(e / N), (int)(e % N))] = gw[e];
        }
        __syncthreads();
    } else if (IN_LDS) {
        // N <= 128: a 32 x 32 lane grid, lane (tr, tc) owns a[tr + 32 i][tc + 32 j]
        // in registers.  Step p: the owners of row p and of column p publish them
        // (double-buffered, one barrier per step); every lane scales the row by
        // the reciprocal pivot itself (same bits as the row owner) and updates
        // its elements.  Per element: the reference's operations, in its order.
        const int tr = t >> 5, tc = t & 31;
        double v[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = tr + 32 * i, c = tc + 32 * j;
                v[i][j] = (r < N && c < N) ? lhs[(long)r * N + c] : 0.0;
            }
        double *ub = lds + (long)N * (N + 1) + 2 * kc;  // 2 x 128: unscaled row p
        double *lb = ub + 256;                     // 2 x 128: column p
        double *ib = lb + 256;                     // 2: reciprocal pivot
        // pivot p = 32 ip + pr: ip is a compile-time index into the register block
#pragma unroll
        for (int ip = 0; ip < 4; ip++) {
            for (int pr = 0; pr < 32; pr++) {
                const int p = ip * 32 + pr;
                if (p >= N)
                    break;
                const int off = (p & 1) * 128;
                if (tr == pr && tc == pr)
                    ib[p & 1] = 1.0 / v[ip][ip];
                if (tr == pr) {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int c = tc + 32 * j;
                        if (c >= p && c < N)
                            ub[off + c] = v[ip][j];
                    }
                }
                if (tc == pr) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int r = tr + 32 * i;
                        if (r > p && r < N)
                            lb[off + r] = v[i][ip];
                    }
                }
                __syncthreads();
                const double piv = ub[off + p];
                if (!(piv != 0.0 && isfinite(piv)))
                    bad = 1;
                const double inv = ib[p & 1];
                double u[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int c = tc + 32 * j;
                    u[j] = (c > p && c < N) ? mul_rn(ub[off + c], inv) : 0.0;
                }
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int r = tr + 32 * i;
                    if (i == ip && tr == pr) {
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (tc + 32 * j > p && tc + 32 * j < N)
                                v[i][j] = u[j];
                    } else if (r > p && r < N) {
                        const double l = lb[off + r];
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (tc + 32 * j > p && tc + 32 * j < N)
                                v[i][j] = sub_rn(v[i][j], mul_rn(l, u[j]));
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = tr + 32 * i, c = tc + 32 * j;
                if (r < N && c < N) {
                    a[fidx<true>(N, r, c)] = v[i][j];
                    gw[(long)r * N + c] = v[i][j];
                }
            }
        __syncthreads();
    } else {
        for (long e = t; e < (long)N * N; e += LU_T)
            a[e] = lhs[e];
        __syncthreads();
        for (int p = 0; p < N; p++) {
            const double piv = a[(long)p * N + p];
            if (!(piv != 0.0 && isfinite(piv)))
                bad = 1;
            const double inv = 1.0 / piv;
            for (int c = p + 1 + t; c < N; c += LU_T)
                a[(long)p * N + c] = mul_rn(a[(long)p * N + c], inv);
            __syncthreads();
            const int w = N - p - 1;
            // rows p+1.., one row per lane group of 64, columns strided by 64
            for (int r = p + 1 + (t >> 6); r < N; r += LU_T / 64) {
                const double l = a[(long)r * N + p];
                for (int c = p + 1 + (t & 63); c < N; c += 64)
                    a[(long)r * N + c] = sub_rn(a[(long)r * N + c], mul_rn(l, a[(long)p * N + c]));
            }
            (void)w;
            __syncthreads();
        }
    }
    if (t == 0 && bad)
        *flag = 1;
    // reciprocal diagonal of L for the forward pass (the reference's 1.0 / L_ii)
    double *dinv = IN_LDS ? lds + (long)N * (N + 1) + 2 * kc + 512 + 2
                          : rowbuf + 2 * kc + (long)N * kc;
    for (int q = t; q < N; q += LU_T)
        dinv[q] = 1.0 / a[fidx<IN_LDS>(N, q, q)];
    __syncthreads();

    if constexpr (IN_LDS) {
        if (mode == CACG_MODE_FAST) {
            if (k <= 16)
                lu_subst_wave<64>(N, a, dinv, k, rhs, x);
            else if (k <= 32)
                lu_subst_wave<32>(N, a, dinv, k, rhs, x);
            else if (k <= 64)
                lu_subst_wave<16>(N, a, dinv, k, rhs, x);
            else
                lu_subst_wave<8>(N, a, dinv, k, rhs, x);
            return;
        }
    }

    // ---- substitutions, kc columns at a time
    for (int c0 = 0; c0 < k; c0 += kc) {
        const int kw = (k - c0) < kc ? (k - c0) : kc;
        const int G = LU_T / kw;
        const int cl = t % kw, g = t / kw;
        const bool own = g < G;
        if (IN_LDS && mode == CACG_MODE_REFERENCE && c0 > 0) {
            // the previous chunk's backward pass reused the factor's LDS
            for (long e = t; e < (long)N * N; e += LU_T)
                a[fidx<true>(N, (int)(e / N), (int)(e % N))] = gw[e];
            __syncthreads();
        }
        double y[JMAX];
#pragma unroll
        for (int j = 0; j < JMAX; j++) {
            const int i = g + G * j;
            y[j] = (own && i < N) ? rhs[(long)i * k + c0 + cl] : 0.0;
        }
        // forward: L y = rhs.  Row q = gq + G jq is owned by lane group gq as its
        // register y[jq]; jq is a compile-time index.
#pragma unroll
        for (int jq = 0; jq < JMAX; jq++) {
            for (int gq = 0; gq < G; gq++) {
                const int q = gq + G * jq;
                if (q >= N)
                    break;
                double *rb = rowbuf + (q & 1) * kc;
                if (own && g == gq) {
                    y[jq] = mul_rn(y[jq], dinv[q]);
                    rb[cl] = y[jq];
                }
                __syncthreads();
                if (own) {
                    const double yq = rb[cl];
#pragma unroll
                    for (int j = jq; j < JMAX; j++) {
                        const int i = g + G * j;
                        if ((j > jq || g > gq) && i < N)
                            y[j] = sub_rn(y[j], mul_rn(yq, a[fidx<IN_LDS>(N, i, q)]));
                    }
                }
            }
        }
        if (mode == CACG_MODE_REFERENCE) {
            __syncthreads();  // all reads of the factor's LDS done before xs overwrites it
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int i = g + G * j;
                if (own && i < N)
                    xs[(long)i * kw + cl] = y[j];
            }
            __syncthreads();
            if (t < kw) {
                for (int i = N - 2; i >= 0; i--) {
                    double xi = xs[(long)i * kw + t];
                    const double *u = gw + (long)i * N;
                    int q = i + 1;
                    for (; q + 4 <= N; q += 4) {
                        const double u0 = u[q], u1 = u[q + 1], u2 = u[q + 2], u3 = u[q + 3];
                        xi = sub_rn(xi, mul_rn(xs[(long)q * kw + t], u0));
                        xi = sub_rn(xi, mul_rn(xs[(long)(q + 1) * kw + t], u1));
                        xi = sub_rn(xi, mul_rn(xs[(long)(q + 2) * kw + t], u2));
                        xi = sub_rn(xi, mul_rn(xs[(long)(q + 3) * kw + t], u3));
                    }
                    for (; q < N; q++)
                        xi = sub_rn(xi, mul_rn(xs[(long)q * kw + t], u[q]));
                    xs[(long)i * kw + t] = xi;
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int i = g + G * j;
                if (own && i < N)
                    x[(long)i * k + c0 + cl] = xs[(long)i * kw + cl];
            }
            __syncthreads();
        } else {
            __syncthreads();  // the forward pass's last broadcast row is read before reuse
#pragma unroll
            for (int jq = JMAX - 1; jq >= 0; jq--) {
                for (int gq = G - 1; gq >= 0; gq--) {
                    const int q = gq + G * jq;
                    if (q >= N)
                        continue;
                    double *rb = rowbuf + (q & 1) * kc;
                    if (own && g == gq)
                        rb[cl] = y[jq];
                    __syncthreads();
                    if (own) {
                        const double xq = rb[cl];
#pragma unroll
                        for (int j = 0; j <= jq; j++) {
                            const int i = g + G * j;
                            if (j < jq || g < gq)
                                y[j] = sub_rn(y[j], mul_rn(xq, a[fidx<IN_LDS>(N, i, q)]));
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < JMAX; j++) {
                const int i = g + G * j;
                if (own && i < N)
                    x[(long)i * k + c0 + cl] = y[j];
            }
            __syncthreads();
        }
    }
}

template <bool IN_LDS>
__global__ __launch_bounds__(LU_T) void lu_solve_kernel(int N, const double *__restrict__ lhs, int k,
                                                        const double *__restrict__ rhs,
                                                        double *__restrict__ x,
                                                        double *__restrict__ gw, int *flag,
                                                        int mode, int kc, int reuse)
{
    extern __shared__ double lds[];
    lu_solve_body<IN_LDS>(N, lhs, k, rhs, x, gw, flag, mode, kc, reuse, lds);
}

// The fast-order solve with a kept factor (N <= 128), spread over workgroups of
// LU_SUB_W waves: wave w of workgroup b solves column LU_SUB_W b + w (64 lanes,
// 2 rows each), with its own LDS copy of the factor (gw, N x N row-major).
// The single-workgroup solve ran every column on one CU, whose LDS served the
// factor to all of them step after step (~106 us for 128 columns); per element
// the operations are those of lu_subst_wave in its order, so the bits are the
// same whichever lanes and workgroup solve a column.
constexpr int LU_SUB_W = 4;
__global__ __launch_bounds__(LU_SUB_W * 64) void lu_subst_multi_kernel(int N,
                                                                      const double *__restrict__ gw,
                                                                      int k,
                                                                      const double *__restrict__ rhs,
                                                                      double *__restrict__ x)
{
    extern __shared__ double lds[];
    double *a = lds;                          // N x N, column-major, stride N + 1
    double *dinv = lds + (long)N * (N + 1);   // reciprocal diagonal
    for (long e = threadIdx.x; e < (long)N * N; e += LU_SUB_W * 64)
        a[fidx<true>(N, (int)(e / N), (int)(e % N))] = gw[e];
    __syncthreads();
    for (int q = threadIdx.x; q < N; q += LU_SUB_W * 64)
        dinv[q] = 1.0 / a[fidx<true>(N, q, q)];
    __syncthreads();
    lu_subst_wave<64>(N, a, dinv, k, rhs, x, LU_SUB_W * blockIdx.x, LU_SUB_W * gridDim.x);
}

// batched: workgroup b solves system items[b].x (N = sm <= 128, factor in LDS);
// which 0: beta = QtAQ^{-1} QtAS with the factor kept from the last alpha solve,
// which 1: alpha = QtAQ^{-1} QtR, factoring QtAQ
__global__ __launch_bounds__(LU_T) void lu_batch_kernel(const cacg_bsys *__restrict__ sys,
                                                        const int2 *__restrict__ items, int which)
{
    extern __shared__ double lds[];
    const cacg_bsys &d = sys[items[blockIdx.x].x];
    const int N = d.sm, k = which == 0 ? d.sm : d.m;
    const int kc = lu_columns(N, k, true);
    lu_solve_body<true>(N, d.QtAQ, k, which == 0 ? d.QtAS : d.QtR, which == 0 ? d.beta : d.alpha,
                        d.lu, d.flag, CACG_MODE_FAST, kc, which == 0 ? 1 : 0, lds);
}

}  // namespace

extern "C" int cacg_k_lu_batch(const cacg_bsys *d_sys, const int2 *d_items, int count, int which,
                               int max_sm, hipStream_t st)
{
    if (count <= 0)
        return 0;
    if (max_sm <= 0 || max_sm > 128 || (which != 0 && which != 1))
        return (int)hipErrorInvalidValue;
    const size_t shm = lu_lds_bytes(max_sm, max_sm, true);
    cacg_allow_lds(lu_batch_kernel, shm);
    hipLaunchKernelGGL(lu_batch_kernel, dim3(count), dim3(LU_T), shm, st, d_sys, d_items, which);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_k_lu_solve(int N, const double *lhs, int k, const double *rhs, double *x,
                               double *d_work, int *d_flag, int mode, int reuse_factor,
                               hipStream_t st)
{
    if (N <= 0 || k <= 0)
        return (int)hipErrorInvalidValue;
    const bool in_lds = (long)N * N <= 128L * 128L;
    const int kc = lu_columns(N, k, in_lds);
    if (kc < 1)
        return (int)hipErrorInvalidValue;  // N > 16384 is out of range
    if (in_lds && mode == CACG_MODE_FAST && reuse_factor) {
        // the kept factor: one column per wave over as many workgroups as it takes
        const size_t sub = ((size_t)N * (N + 1) + N) * sizeof(double);
        const int G = (k + LU_SUB_W - 1) / LU_SUB_W;
        cacg_allow_lds(lu_subst_multi_kernel, sub);
        hipLaunchKernelGGL(lu_subst_multi_kernel, dim3(G), dim3(LU_SUB_W * 64), sub, st, N, d_work,
                           k, rhs, x);
        CACG_LAUNCH_CHECK();
        return 0;
    }
    const size_t shm = lu_lds_bytes(N, k, in_lds);
    cacg_allow_lds(in_lds ? lu_solve_kernel<true> : lu_solve_kernel<false>, shm);
    if (in_lds)
        hipLaunchKernelGGL(lu_solve_kernel<true>, dim3(1), dim3(LU_T), shm, st, N, lhs, k, rhs, x,
                           d_work, d_flag, mode, kc, reuse_factor);
    else
        hipLaunchKernelGGL(lu_solve_kernel<false>, dim3(1), dim3(LU_T), shm, st, N, lhs, k, rhs, x,
                           d_work, d_flag, mode, kc, reuse_factor);
    CACG_LAUNCH_CHECK();
    return 0;
}
