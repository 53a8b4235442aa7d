// Batched kernels: one launch advances a phase of the s-step iteration for
// every system in a work list (the batching / serving form).  gfx950, wave64.
//
// Systems of the batch differ in n, nnz, m (<= 32) and sm = s*m (<= 128), so
// the SpMM kernels map lanes to the elements (row, column) of a 64-row tile:
// lane e of the tile handles row e / k, column e % k, which keeps every lane
// busy for any k.  The dense products (Gram blocks, Q = S - Q beta,
// X += Q alpha) run on FP64 MFMA over 2048-row chunks, one launch per tile
// class (the system's sm and m rounded up to 16-column tiles, zero-padded);
// a second kernel sums the Gram chunk partials in chunk order.  Every sum of
// a system runs in an order fixed by that system alone (its rows, tiles,
// chunks and class), so a system's results are the same bits whatever else
// shares the launch.
#include <type_traits>

#include "kcommon.hpp"

namespace {

constexpr int BT = CACG_BT;
constexpr int GR = CACG_GR;
constexpr int GS = 16;  // rows per LDS stage of the Gram
constexpr int MS = CACG_BATCH_MAX_SM, MM = CACG_BATCH_MAX_M;

__device__ __forceinline__ int tile_rows(int n, int tile)
{
    const int r = n - tile * BT;
    return r < BT ? r : BT;
}

// sum over row i of val[p] * x[col[p] * ldx + c] in CSR order, products and sums
// rounded separately (the reference's SpMM arithmetic); the gathers of 8
// nonzeros are issued before they are accumulated, for memory-level parallelism
template <typename TV, typename TX>
__device__ __forceinline__ TV row_dot(const int *__restrict__ rowptr, const int *__restrict__ col,
                                      const TV *__restrict__ val, const TX *__restrict__ x, long ldx,
                                      int c, long i)
{
    const int pb = rowptr[i], pe = rowptr[i + 1];
    TV acc = (TV)0;
    for (int p = pb; p < pe; p += 8) {
        int cc[8];
        TV vv[8];
        TX xv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool in = p + u < pe;
            cc[u] = in ? col[p + u] : 0;
            vv[u] = in ? val[p + u] : (TV)0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            xv[u] = p + u < pe ? x[(long)cc[u] * ldx + c] : (TX)0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (p + u < pe)
                acc = add_rn(acc, mul_rn(vv[u], (TV)xv[u]));
    }
    return acc;
}

// Vector path of the tile SpMMs: the tile's row pointers and (col, val)
// pairs staged in LDS (when at most CAPB), each lane producing 4 adjacent
// columns of one row from 16-byte gathers of x, 4 nonzeros gathered ahead of
// their in-order accumulation.  Each output is the CSR-order sum of separately
// rounded products, the bits of row_dot.
constexpr int CAPB = 1024;

template <typename TV>
struct TileCsr {
    int rp[BT + 1];
    int col[CAPB];
    TV val[CAPB];
};

// stage rows r0 .. r0 + rows - 1; returns whether the pairs fit (else the
// caller uses row_dot); ends with a barrier
template <typename TV>
__device__ __forceinline__ bool stage_tile(TileCsr<TV> &tc, const int *__restrict__ rowptr,
                                           const int *__restrict__ col,
                                           const TV *__restrict__ val, long r0, int rows)
{
    for (int e = threadIdx.x; e <= rows; e += 256)
        tc.rp[e] = rowptr[r0 + e];
    __syncthreads();
    const int pb = tc.rp[0], tn = tc.rp[rows] - pb;
    const bool staged = tn <= CAPB;
    if (staged)
        for (int e = threadIdx.x; e < tn; e += 256) {
            tc.col[e] = col[pb + e];
            tc.val[e] = val[pb + e];
        }
    __syncthreads();
    return staged;
}

// 4 elements of x converted to TV: p[0..3] (FP32, one 16-byte load), or the
// FP64 pairs p[0..1] and p[off..off+1] (16-byte aligned)
template <typename TV, typename TX>
__device__ __forceinline__ void load4(const TX *p, TV *o, int off)
{
    if constexpr (sizeof(TX) == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        o[0] = (TV)v.x;
        o[1] = (TV)v.y;
        o[2] = (TV)v.z;
        o[3] = (TV)v.w;
    } else {
        const double2 v0 = *reinterpret_cast<const double2 *>(p);
        const double2 v1 = *reinterpret_cast<const double2 *>(p + off);
        o[0] = (TV)v0.x;
        o[1] = (TV)v0.y;
        o[2] = (TV)v1.x;
        o[3] = (TV)v1.y;
    }
}

// Columns of a thread's 4 elements in an FP64 row block of width w (w % 4 == 0):
// the pairs c0, c0 + 1 and c0 + w / 2, c0 + w / 2 + 1 with c0 = 2 t for the
// t-th thread of the row, so each 16-byte access instruction of the threads of a
// row covers one contiguous half of it (FP32: 4 contiguous from 4 t).  Which
// thread computes an element changes no bit.
struct cols4 {
    int c0, off;
    __device__ __forceinline__ cols4(int t, int w, bool fp64)
        : c0(fp64 ? 2 * t : 4 * t), off(fp64 ? w / 2 : 2) {}
    __device__ __forceinline__ int operator()(int k) const { return k < 2 ? c0 + k : c0 + off + k - 2; }
};

// acc[0..3] = sums over row r of the staged tile of val * x[col * ldx + cs(0..3)]
template <typename TV, typename TX>
__device__ __forceinline__ void row_dot4(const TileCsr<TV> &tc, int r, const TX *__restrict__ x,
                                         long ldx, const cols4 &cs, TV *acc)
{
    const int pb = tc.rp[0], p0 = tc.rp[r] - pb, p1 = tc.rp[r + 1] - pb;
#pragma unroll
    for (int k = 0; k < 4; k++)
        acc[k] = (TV)0;
    for (int p = p0; p < p1; p += 4) {
        TV xv[4][4], vv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool in = p + u < p1;
            vv[u] = in ? tc.val[p + u] : (TV)0;
            if (in)
                load4<TV, TX>(x + (long)tc.col[p + u] * ldx + cs.c0, xv[u], cs.off);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (p + u < p1)
#pragma unroll
                for (int k = 0; k < 4; k++)
                    acc[k] = add_rn(acc[k], mul_rn(vv[u], xv[u][k]));
    }
}

// basis step j for one tile (see cacg_kb_basis); TB = basis type
template <typename TB>
__global__ __launch_bounds__(256) void bbasis_kernel(const cacg_bsys *__restrict__ sys,
                                                     const int2 *__restrict__ items, int j)
{
    __shared__ TileCsr<TB> tc;
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int n = d.n, m = d.m, sm = d.sm;
    const long r0 = (long)it.y * BT;
    const int rows = tile_rows(n, it.y);
    TB *S = reinterpret_cast<TB *>(d.S);
    const TB *val = sizeof(TB) == 4 ? reinterpret_cast<const TB *>(d.val32)
                                    : reinterpret_cast<const TB *>(d.val);
    const TB a1 = (TB)d.sa, b1 = (TB)d.sb, a2 = (TB)(2.0 * d.sa), b2 = (TB)(2.0 * d.sb);
    if (j == 0) {
        for (int e = threadIdx.x; e < rows * m; e += 256) {
            const int r = e / m, c = e - r * m;
            const long i = r0 + r;
            S[i * sm + c] = (TB)d.R[i * m + c];
        }
        return;
    }
    // the vector path: m a multiple of 4 (16-byte aligned rows and blocks)
    const bool vec = m % 4 == 0 && stage_tile<TB>(tc, d.rowptr, d.col, val, r0, rows);
    if (vec) {
        const int q = m / 4;
        constexpr bool f64 = sizeof(TB) == 8;
        for (int e = threadIdx.x; e < rows * q; e += 256) {
            const int r = e / q;
            const cols4 cs(e - r * q, m, f64);
            const long i = r0 + r;
            TB acc[4];
            if (j == 1) {
                // FP32 basis: the FP64 R read as 4 contiguous columns
                const cols4 cr = f64 ? cs : cols4(e - r * q, m, false);
                row_dot4<TB, double>(tc, r, d.R, m, cr, acc);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const TB e1 = (TB)d.R[i * m + cr(k)];
                    S[i * sm + cr(k)] = e1;
                    S[i * sm + m + cr(k)] = add_rn(mul_rn(a1, acc[k]), mul_rn(b1, e1));
                }
            } else {
                const TB *xs = S + (long)(j - 1) * m;
                row_dot4<TB, TB>(tc, r, xs, sm, cs, acc);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const TB e1 = xs[i * sm + cs(k)], e2 = S[i * sm + (long)(j - 2) * m + cs(k)];
                    S[i * sm + (long)j * m + cs(k)] =
                        sub_rn(add_rn(mul_rn(a2, acc[k]), mul_rn(b2, e1)), e2);
                }
            }
        }
        return;
    }
    for (int e = threadIdx.x; e < rows * m; e += 256) {
        const int r = e / m, c = e - r * m;
        const long i = r0 + r;
        if (j == 1) {
            const TB acc = row_dot<TB, double>(d.rowptr, d.col, val, d.R, m, c, i);
            const TB e1 = (TB)d.R[i * m + c];
            S[i * sm + c] = e1;
            S[i * sm + m + c] = add_rn(mul_rn(a1, acc), mul_rn(b1, e1));
        } else {
            const TB *xs = S + (long)(j - 1) * m;
            const TB acc = row_dot<TB, TB>(d.rowptr, d.col, val, xs, sm, c, i);
            const TB e1 = xs[i * sm + c], e2 = S[i * sm + (long)(j - 2) * m + c];
            S[i * sm + (long)j * m + c] = sub_rn(add_rn(mul_rn(a2, acc), mul_rn(b2, e1)), e2);
        }
    }
}

template <typename TB>
__global__ __launch_bounds__(256) void bq_from_s_kernel(const cacg_bsys *__restrict__ sys,
                                                        const int2 *__restrict__ items)
{
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const long base = (long)it.y * BT * d.sm;
    const int cnt = tile_rows(d.n, it.y) * d.sm;
    const TB *S = reinterpret_cast<const TB *>(d.S);
    for (int e = threadIdx.x; e < cnt; e += 256)
        d.Q[base + e] = (double)S[base + e];
}

__global__ __launch_bounds__(256) void baq_kernel(const cacg_bsys *__restrict__ sys,
                                                  const int2 *__restrict__ items)
{
    __shared__ TileCsr<double> tc;
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int sm = d.sm;
    const long r0 = (long)it.y * BT;
    const int rows = tile_rows(d.n, it.y);
    if (sm % 4 == 0 && stage_tile<double>(tc, d.rowptr, d.col, d.val, r0, rows)) {
        const int q = sm / 4;
        for (int e = threadIdx.x; e < rows * q; e += 256) {
            const int r = e / q;
            const cols4 cs(e - r * q, sm, true);
            double acc[4];
            row_dot4<double, double>(tc, r, d.Q, sm, cs, acc);
            double *y = d.AQ + (r0 + r) * sm;
            *reinterpret_cast<double2 *>(y + cs.c0) = double2{acc[0], acc[1]};
            *reinterpret_cast<double2 *>(y + cs.c0 + cs.off) = double2{acc[2], acc[3]};
        }
        return;
    }
    for (int e = threadIdx.x; e < rows * sm; e += 256) {
        const int r = e / sm, c = e - r * sm;
        const long i = r0 + r;
        d.AQ[i * sm + c] = row_dot<double, double>(d.rowptr, d.col, d.val, d.Q, sm, c, i);
    }
}

// LDS row stride of k doubles: 16 mod 32, so a wave's operand read (4 rows x
// 16 doubles) falls on disjoint halves of the 64 banks
constexpr int bstride(int k)
{
    return k + (48 - k % 32) % 32;
}

// waves of the Gram kernel of tile class TM
constexpr int gram_waves(int TM)
{
    return TM >= 4 ? 8 : 4;
}

// rows per stage of the Gram kernel: the narrow classes do little work per
// row, so their stages are longer (fewer load round trips and barriers; the
// rows are still taken in order)
constexpr int gram_stage(int TM)
{
    return TM <= 2 ? 64 : 16;
}

// MFMA Gram partial of one 2048-row chunk for systems with TM = ceil(sm / 16)
// row tiles and TN = ceil(kr / 16) column tiles (columns past sm and kr are
// zeros in LDS and are not written); WHICH 0: Q^T A S = AQ^T S, 1: Q^T [AQ | R].
// NW waves split the TM x TN tiles into GMG x GNG groups of MT x NT tiles.
// Stages of GS rows of L and of [R1 | R2] are loaded with 16-byte vectors
// where a row's width allows (single elements otherwise) into registers one
// stage ahead, stored to one of two LDS buffers and read as MFMA operands
// (contraction index = row).  Per tile the k-steps run in row order from
// zero (rows past the chunk are zeros), so the partial depends on the system
// alone.
template <typename TB, int TM, int TN, int WHICH>
__global__ __launch_bounds__(gram_waves(TM) * 64) void bgram_mfma_kernel(
    const cacg_bsys *__restrict__ sys, const int2 *__restrict__ items)
{
    constexpr int NW = gram_waves(TM), NTH = NW * 64, GS = gram_stage(TM);
    constexpr int GMG = TM < NW ? TM : NW, GNG = NW / GMG;
    constexpr int MT = (TM + GMG - 1) / GMG, NT = (TN + GNG - 1) / GNG;
    constexpr int KL = TM * 16, KR = TN * 16, LP = bstride(KL), RP = bstride(KR);
    constexpr int BUF = GS * (LP + RP);
    // register slots per lane and stage: L and R1 up to GS x KL elements, R2 up
    // to GS x 32, one element or one 16-byte vector each
    constexpr int SLL = (GS * KL + NTH - 1) / NTH, SL2 = (GS * MM + NTH - 1) / NTH;
    using T1 = typename std::conditional<WHICH == 1, double, TB>::type;
    extern __shared__ double lds[];
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int n = d.n, m = d.m, sm = d.sm;
    const int kr = sm + (WHICH ? m : 0);
    const double *L = WHICH ? d.Q : d.AQ;
    const T1 *R1 = WHICH ? reinterpret_cast<const T1 *>(d.AQ) : reinterpret_cast<const T1 *>(d.S);
    const int c0 = it.y * GR, c1 = (c0 + GR) < n ? (c0 + GR) : n;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, lr = lane >> 4, lc = lane & 15;
    const int gm = wave % GMG, gn = wave / GMG;
    const bool active = gn < GNG;

    // element widths of one load: 2 doubles / 4 floats where the row width
    // (and so every row start) allows 16-byte alignment
    constexpr int V1 = 16 / (int)sizeof(T1);
    const int wl = sm % 2 == 0 ? 2 : 1, w1 = sm % V1 == 0 ? V1 : 1;
    const int w2 = m % 2 == 0 ? 2 : 1;
    const int nl = sm / wl, n1 = sm / w1, n2 = WHICH ? m / w2 : 0;
    int lrow[SLL], lcol[SLL], r1row[SLL], r1col[SLL], r2row[SL2], r2col[SL2];
#pragma unroll
    for (int u = 0; u < SLL; u++) {
        const int v = t + NTH * u;
        lrow[u] = v < GS * nl ? v / nl : -1;
        lcol[u] = v < GS * nl ? (v % nl) * wl : 0;
        r1row[u] = v < GS * n1 ? v / n1 : -1;
        r1col[u] = v < GS * n1 ? (v % n1) * w1 : 0;
    }
#pragma unroll
    for (int u = 0; u < SL2; u++) {
        const int v = t + NTH * u;
        r2row[u] = v < GS * n2 ? v / n2 : -1;
        r2col[u] = v < GS * n2 ? (v % n2) * w2 : 0;
    }
    using V1T = typename std::conditional<sizeof(T1) == 4, float4, double2>::type;
    double2 gl[SLL], g2[SL2];
    V1T g1[SLL];
    auto gload = [&](int rb) {
#pragma unroll
        for (int u = 0; u < SLL; u++) {
            const long i = rb + lrow[u];
            gl[u] = double2{0.0, 0.0};
            if (lrow[u] >= 0 && i < c1) {
                const double *p = L + i * sm + lcol[u];
                if (wl == 2)
                    gl[u] = *reinterpret_cast<const double2 *>(p);
                else
                    gl[u] = double2{p[0], 0.0};
            }
            const long i1 = rb + r1row[u];
            if constexpr (sizeof(T1) == 4)
                g1[u] = float4{0.f, 0.f, 0.f, 0.f};
            else
                g1[u] = double2{0.0, 0.0};
            if (r1row[u] >= 0 && i1 < c1) {
                const T1 *p = R1 + i1 * sm + r1col[u];
                if (w1 == V1)
                    g1[u] = *reinterpret_cast<const V1T *>(p);
                else if constexpr (sizeof(T1) == 4)
                    g1[u] = float4{p[0], 0.f, 0.f, 0.f};
                else
                    g1[u] = double2{p[0], 0.0};
            }
        }
#pragma unroll
        for (int u = 0; u < SL2; u++) {
            const long i = rb + r2row[u];
            g2[u] = double2{0.0, 0.0};
            if (r2row[u] >= 0 && i < c1) {
                const double *p = d.R + i * m + r2col[u];
                if (w2 == 2)
                    g2[u] = *reinterpret_cast<const double2 *>(p);
                else
                    g2[u] = double2{p[0], 0.0};
            }
        }
    };
    auto lstore = [&](int buf) {
        double *Ls = lds + buf * BUF, *Rs = Ls + GS * LP;
#pragma unroll
        for (int u = 0; u < SLL; u++) {
            if (lrow[u] >= 0) {
                Ls[lrow[u] * LP + lcol[u]] = gl[u].x;
                if (wl == 2)
                    Ls[lrow[u] * LP + lcol[u] + 1] = gl[u].y;
            }
            if (r1row[u] >= 0) {
                double *q = Rs + r1row[u] * RP + r1col[u];
                if constexpr (sizeof(T1) == 4) {
                    q[0] = (double)g1[u].x;
                    if (w1 == 4) {
                        q[1] = (double)g1[u].y;
                        q[2] = (double)g1[u].z;
                        q[3] = (double)g1[u].w;
                    }
                } else {
                    q[0] = g1[u].x;
                    if (w1 == 2)
                        q[1] = g1[u].y;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SL2; u++) {
            if (r2row[u] >= 0) {
                double *q = Rs + r2row[u] * RP + sm + r2col[u];
                q[0] = g2[u].x;
                if (w2 == 2)
                    q[1] = g2[u].y;
            }
        }
    };

    // the padding columns (past sm in L, past kr in R) stay zero in both buffers
    for (int e = t; e < 2 * GS; e += NTH) {
        double *Ls = lds + (e / GS) * BUF, *Rs = Ls + GS * LP;
        const int r = e % GS;
        for (int c = sm; c < KL; c++)
            Ls[r * LP + c] = 0.0;
        for (int c = kr; c < KR; c++)
            Rs[r * RP + c] = 0.0;
    }
    dbl4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
            acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int nst = (c1 - c0 + GS - 1) / GS;
    gload(c0);
    lstore(0);
    if (nst > 1)
        gload(c0 + GS);
    __syncthreads();
    for (int stg = 0; stg < nst; stg++) {
        const int cur = stg & 1;
        if (active) {
            const double *Ls = lds + cur * BUF, *Rs = Ls + GS * LP;
#pragma unroll
            for (int ks = 0; ks < GS / 4; ks++) {
                const int row = ks * 4 + lr;
                double av[MT], bv[NT];
#pragma unroll
                for (int i = 0; i < MT; i++) {
                    const int a = (gm * MT + i) * 16 + lc;
                    av[i] = a < KL ? Ls[row * LP + a] : 0.0;
                }
#pragma unroll
                for (int j = 0; j < NT; j++) {
                    const int b = (gn * NT + j) * 16 + lc;
                    bv[j] = b < KR ? Rs[row * RP + b] : 0.0;
                }
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int j = 0; j < NT; j++)
                        acc[i][j] = mfma16x16x4(av[i], bv[j], acc[i][j]);
            }
        }
        if (stg + 1 < nst) {
            lstore(cur ^ 1);
            if (stg + 2 < nst)
                gload(c0 + (stg + 2) * GS);
        }
        __syncthreads();
    }
    if (!active)
        return;
    double *gp = d.gpart + (long)it.y * sm * kr;
#pragma unroll
    for (int i = 0; i < MT; i++) {
        const int tm = gm * MT + i;
        if (tm >= TM)
            continue;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int tn = gn * NT + j;
            if (tn >= TN)
                continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = tm * 16 + lr + 4 * r, b = tn * 16 + lc;
                if (a < sm && b < kr)
                    gp[(long)a * kr + b] = acc[i][j][r];
            }
        }
    }
}

// chunk-ordered sum of the partials, 256 result entries per workgroup
__global__ __launch_bounds__(256) void bgram_reduce_kernel(const cacg_bsys *__restrict__ sys,
                                                           const int2 *__restrict__ items,
                                                           int which)
{
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int sm = d.sm, m = d.m;
    const int kr = sm + (which ? m : 0);
    const int e = it.y * 256 + threadIdx.x;
    if (e >= sm * kr)
        return;
    const long stride = (long)sm * kr;
    double acc = 0.0;
    for (int ch = 0; ch < d.nchunk; ch++)
        acc += d.gpart[ch * stride + e];
    const int a = e / kr, b = e - a * kr;
    if (which == 0)
        d.QtAS[(long)a * sm + b] = acc;
    else if (b < sm)
        d.QtAQ[(long)a * sm + b] = acc;
    else
        d.QtR[(long)a * m + (b - sm)] = acc;
}

// Q = S - Q beta with FP64 MFMA for systems with 16 (SM / 16 - 1) < sm <= SM:
// one 2048-row chunk per 512-lane workgroup, beta in LDS (zero-padded to
// SM x SM), each wave takes 16-row tiles of the chunk; the contraction index
// is permuted across the lane groups (group g supplies k = g SM/4 + t at step
// t), so each lane's A operand is SM/4 contiguous values of one row (16-byte
// loads when sm == SM, EXACT; guarded loads with zeros past sm otherwise),
// read before the tile is written (in place).  Per output the k order is
// fixed by the system alone.
template <typename TB, int SM, bool EXACT>
__global__ __launch_bounds__(512) void bqupdate_mfma_kernel(const cacg_bsys *__restrict__ sys,
                                                            const int2 *__restrict__ items)
{
    constexpr int KQ = SM / 4, NTL = SM / 16;
    extern __shared__ double bl[];  // SM x SM
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int n = d.n, sm = EXACT ? SM : d.sm;
    for (int e = threadIdx.x; e < SM * SM; e += 512) {
        const int q = e / SM, c = e - q * SM;
        bl[e] = (q < sm && c < sm) ? d.beta[(long)q * sm + c] : 0.0;
    }
    __syncthreads();
    const TB *S = reinterpret_cast<const TB *>(d.S);
    double *Q = d.Q;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lg = lane >> 4, lc = lane & 15;
    const int c0 = it.y * GR, c1 = (c0 + GR) < n ? (c0 + GR) : n;
    for (int i0 = c0 + wave * 16; i0 < c1; i0 += 8 * 16) {
        const long arow = i0 + lc;
        double av[KQ];
        if (EXACT && arow < c1) {
            const double2 *src = reinterpret_cast<const double2 *>(Q + arow * SM + lg * KQ);
#pragma unroll
            for (int u = 0; u < KQ / 2; u++) {
                const double2 v = src[u];
                av[2 * u] = v.x;
                av[2 * u + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                const int k = lg * KQ + u;
                av[u] = (arow < c1 && k < sm) ? Q[arow * sm + k] : 0.0;
            }
        }
#pragma unroll 1
        for (int j = 0; j < NTL; j++) {
            const int c = j * 16 + lc;
            double sv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long i = i0 + lg + 4 * r;
                sv[r] = (i < c1 && c < sm) ? (double)S[i * sm + c] : 0.0;
            }
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
            const double *bcol = bl + (lg * KQ) * SM + c;
#pragma unroll
            for (int t = 0; t < KQ; t++)
                acc = mfma16x16x4(av[t], bcol[t * SM], acc);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long i = i0 + lg + 4 * r;
                if (i < c1 && c < sm)
                    Q[i * sm + c] = sv[r] - acc[r];
            }
        }
    }
}

// X += Q alpha with FP64 MFMA for systems with 16 (SM / 16 - 1) < sm <= SM:
// one 2048-row chunk per 256-lane workgroup, alpha (sm x m, m <= 16 NTM) in
// registers in the K-permuted order of the A operand (zeros past sm), 16-row
// tiles per wave; 16-byte Q loads when sm == SM (EXACT).
template <int SM, int NTM, bool EXACT>
__global__ __launch_bounds__(256) void bxupdate_mfma_kernel(const cacg_bsys *__restrict__ sys,
                                                            const int2 *__restrict__ items)
{
    constexpr int KQ = SM / 4;
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int n = d.n, m = d.m, sm = EXACT ? SM : d.sm;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lg = lane >> 4, lc = lane & 15;
    double alf[NTM][KQ];
#pragma unroll
    for (int j = 0; j < NTM; j++) {
        const int c = j * 16 + lc;
#pragma unroll
        for (int t = 0; t < KQ; t++) {
            const int k = lg * KQ + t;
            alf[j][t] = (c < m && k < sm) ? d.alpha[(long)k * m + c] : 0.0;
        }
    }
    const int c0 = it.y * GR, c1 = (c0 + GR) < n ? (c0 + GR) : n;
    for (int i0 = c0 + wave * 16; i0 < c1; i0 += 4 * 16) {
        const long arow = i0 + lc;
        double av[KQ];
        if (EXACT && arow < c1) {
            const double2 *src = reinterpret_cast<const double2 *>(d.Q + arow * SM + lg * KQ);
#pragma unroll
            for (int u = 0; u < KQ / 2; u++) {
                const double2 v = src[u];
                av[2 * u] = v.x;
                av[2 * u + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                const int k = lg * KQ + u;
                av[u] = (arow < c1 && k < sm) ? d.Q[arow * sm + k] : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < NTM; j++) {
            dbl4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < KQ; t++)
                acc = mfma16x16x4(av[t], alf[j][t], acc);
            const int c = j * 16 + lc;
            if (c < m) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const long i = i0 + lg + 4 * r;
                    if (i < c1)
                        d.X[i * m + c] += acc[r];
                }
            }
        }
    }
}

// R = B - A X in the reference's SpMM arithmetic, plus the tile's column sums
// of R^2 (rows in order)
__global__ __launch_bounds__(256) void bresid_kernel(const cacg_bsys *__restrict__ sys,
                                                     const int2 *__restrict__ items)
{
    __shared__ double sq[BT * MM];
    const int2 it = items[blockIdx.x];
    const cacg_bsys &d = sys[it.x];
    const int m = d.m;
    const long r0 = (long)it.y * BT;
    const int rows = tile_rows(d.n, it.y);
    for (int e = threadIdx.x; e < rows * m; e += 256) {
        const int r = e / m, c = e - r * m;
        const long i = r0 + r;
        const double acc = row_dot<double, double>(d.rowptr, d.col, d.val, d.X, m, c, i);
        const double rv = sub_rn(d.B[i * m + c], acc);
        d.R[i * m + c] = rv;
        sq[e] = rv * rv;
    }
    __syncthreads();
    if ((int)threadIdx.x < m) {
        double s = 0.0;
        for (int r = 0; r < rows; r++)
            s += sq[r * m + threadIdx.x];
        d.npart[(long)it.y * m + threadIdx.x] = s;
    }
}

// norm2[c] = sum over the system's tiles, in tile order
__global__ __launch_bounds__(64) void bnorm_kernel(const cacg_bsys *__restrict__ sys,
                                                   const int2 *__restrict__ items)
{
    const cacg_bsys &d = sys[items[blockIdx.x].x];
    const int c = threadIdx.x;
    if (c >= d.m)
        return;
    // tile order, as before; the loads of 16 tiles go out before their adds
    // (one dependent load per tile made a 1024-tile system take ~170 us)
    double s = 0.0;
    int t = 0;
    for (; t + 16 <= d.ntile; t += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
            v[u] = d.npart[(long)(t + u) * d.m + c];
#pragma unroll
        for (int u = 0; u < 16; u++)
            s += v[u];
    }
    for (; t < d.ntile; t++)
        s += d.npart[(long)t * d.m + c];
    d.norm2[c] = s;
}

}  // namespace

extern "C" int cacg_kb_basis(const cacg_bsys *d_sys, const int2 *d_tiles, int count, int j, int f32,
                             hipStream_t st)
{
    if (count <= 0)
        return 0;
    if (j < 0)
        return (int)hipErrorInvalidValue;
    if (f32)
        hipLaunchKernelGGL(bbasis_kernel<float>, dim3(count), dim3(256), 0, st, d_sys, d_tiles, j);
    else
        hipLaunchKernelGGL(bbasis_kernel<double>, dim3(count), dim3(256), 0, st, d_sys, d_tiles, j);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_kb_q_from_s(const cacg_bsys *d_sys, const int2 *d_tiles, int count, int f32,
                                hipStream_t st)
{
    if (count <= 0)
        return 0;
    if (f32)
        hipLaunchKernelGGL(bq_from_s_kernel<float>, dim3(count), dim3(256), 0, st, d_sys, d_tiles);
    else
        hipLaunchKernelGGL(bq_from_s_kernel<double>, dim3(count), dim3(256), 0, st, d_sys, d_tiles);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_kb_aq(const cacg_bsys *d_sys, const int2 *d_tiles, int count, hipStream_t st)
{
    if (count <= 0)
        return 0;
    hipLaunchKernelGGL(baq_kernel, dim3(count), dim3(256), 0, st, d_sys, d_tiles);
    CACG_LAUNCH_CHECK();
    return 0;
}

namespace {
template <typename TB, int TM, int TN, int WHICH>
void launch_bgram(const cacg_bsys *d_sys, const int2 *d_chunks, int nchunks, hipStream_t st)
{
    constexpr int NW = gram_waves(TM);
    const size_t shm =
        2 * (size_t)gram_stage(TM) * (bstride(TM * 16) + bstride(TN * 16)) * sizeof(double);
    cacg_allow_lds(bgram_mfma_kernel<TB, TM, TN, WHICH>, shm);
    hipLaunchKernelGGL((bgram_mfma_kernel<TB, TM, TN, WHICH>), dim3(nchunks), dim3(NW * 64), shm,
                       st, d_sys, d_chunks);
}

template <typename TB, int TM>
int launch_gram_class(const cacg_bsys *d_sys, const int2 *d_chunks, int nchunks, int which, int tn,
                      hipStream_t st)
{
    // which 0: kr = sm (TN = TM); which 1: kr = sm + m, m <= 32 (TN - TM in 0..2)
    if (which == 0 && tn == TM)
        launch_bgram<TB, TM, TM, 0>(d_sys, d_chunks, nchunks, st);
    else if (which == 1 && tn == TM)
        launch_bgram<TB, TM, TM, 1>(d_sys, d_chunks, nchunks, st);
    else if (which == 1 && tn == TM + 1)
        launch_bgram<TB, TM, TM + 1, 1>(d_sys, d_chunks, nchunks, st);
    else if (which == 1 && tn == TM + 2)
        launch_bgram<TB, TM, TM + 2, 1>(d_sys, d_chunks, nchunks, st);
    else
        return (int)hipErrorInvalidValue;
    CACG_LAUNCH_CHECK();
    return 0;
}
}  // namespace

extern "C" int cacg_kb_gram_class(const cacg_bsys *d_sys, const int2 *d_chunks, int nchunks,
                                  int which, int f32, int tm, int tn, hipStream_t st)
{
    if (nchunks <= 0)
        return 0;
    if (which != 0 && which != 1)
        return (int)hipErrorInvalidValue;
#define CACG_GCLASS(TMV)                                                                          \
    case TMV:                                                                                     \
        return f32 ? launch_gram_class<float, TMV>(d_sys, d_chunks, nchunks, which, tn, st)      \
                   : launch_gram_class<double, TMV>(d_sys, d_chunks, nchunks, which, tn, st);
    switch (tm) {
        CACG_GCLASS(1)
        CACG_GCLASS(2)
        CACG_GCLASS(3)
        CACG_GCLASS(4)
        CACG_GCLASS(5)
        CACG_GCLASS(6)
        CACG_GCLASS(7)
        CACG_GCLASS(8)
    default:
        return (int)hipErrorInvalidValue;
    }
#undef CACG_GCLASS
}

extern "C" int cacg_kb_gram_reduce(const cacg_bsys *d_sys, const int2 *d_blocks, int nblocks,
                                   int which, hipStream_t st)
{
    if (nblocks <= 0)
        return 0;
    hipLaunchKernelGGL(bgram_reduce_kernel, dim3(nblocks), dim3(256), 0, st, d_sys, d_blocks,
                       which);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_kb_qupdate_class(const cacg_bsys *d_sys, const int2 *d_chunks, int count,
                                     int smp, int exact, int f32, hipStream_t st)
{
    if (count <= 0)
        return 0;
    const size_t shm = (size_t)smp * smp * sizeof(double);
#define CACG_QLAUNCH(TBV, SMV, EX)                                                                \
    do {                                                                                          \
        cacg_allow_lds(bqupdate_mfma_kernel<TBV, SMV, EX>, shm);                                  \
        hipLaunchKernelGGL((bqupdate_mfma_kernel<TBV, SMV, EX>), dim3(count), dim3(512), shm, st, \
                           d_sys, d_chunks);                                                      \
    } while (0)
#define CACG_QCLASS(SMV)                                                                          \
    case SMV:                                                                                     \
        if (f32) {                                                                                \
            if (exact)                                                                            \
                CACG_QLAUNCH(float, SMV, true);                                                   \
            else                                                                                  \
                CACG_QLAUNCH(float, SMV, false);                                                  \
        } else {                                                                                  \
            if (exact)                                                                            \
                CACG_QLAUNCH(double, SMV, true);                                                  \
            else                                                                                  \
                CACG_QLAUNCH(double, SMV, false);                                                 \
        }                                                                                         \
        break;
    switch (smp) {
        CACG_QCLASS(16)
        CACG_QCLASS(32)
        CACG_QCLASS(48)
        CACG_QCLASS(64)
        CACG_QCLASS(80)
        CACG_QCLASS(96)
        CACG_QCLASS(112)
        CACG_QCLASS(128)
    default:
        return (int)hipErrorInvalidValue;
    }
#undef CACG_QCLASS
#undef CACG_QLAUNCH
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_kb_xupdate_class(const cacg_bsys *d_sys, const int2 *d_chunks, int count,
                                     int smp, int exact, int m, hipStream_t st)
{
    if (count <= 0)
        return 0;
    if (m < 1 || m > 32)
        return (int)hipErrorInvalidValue;
#define CACG_XLAUNCH(SMV, NT, EX)                                                                 \
    hipLaunchKernelGGL((bxupdate_mfma_kernel<SMV, NT, EX>), dim3(count), dim3(256), 0, st, d_sys, \
                       d_chunks)
#define CACG_XCLASS(SMV)                                                                          \
    case SMV:                                                                                     \
        if (m <= 16) {                                                                            \
            if (exact)                                                                            \
                CACG_XLAUNCH(SMV, 1, true);                                                       \
            else                                                                                  \
                CACG_XLAUNCH(SMV, 1, false);                                                      \
        } else {                                                                                  \
            if (exact)                                                                            \
                CACG_XLAUNCH(SMV, 2, true);                                                       \
            else                                                                                  \
                CACG_XLAUNCH(SMV, 2, false);                                                      \
        }                                                                                         \
        break;
    switch (smp) {
        CACG_XCLASS(16)
        CACG_XCLASS(32)
        CACG_XCLASS(48)
        CACG_XCLASS(64)
        CACG_XCLASS(80)
        CACG_XCLASS(96)
        CACG_XCLASS(112)
        CACG_XCLASS(128)
    default:
        return (int)hipErrorInvalidValue;
    }
#undef CACG_XCLASS
#undef CACG_XLAUNCH
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_kb_resid(const cacg_bsys *d_sys, const int2 *d_tiles, int count,
                             const int2 *d_sysitems, int nsys, hipStream_t st)
{
    if (count > 0) {
        hipLaunchKernelGGL(bresid_kernel, dim3(count), dim3(256), 0, st, d_sys, d_tiles);
        CACG_LAUNCH_CHECK();
    }
    if (nsys > 0) {
        hipLaunchKernelGGL(bnorm_kernel, dim3(nsys), dim3(64), 0, st, d_sys, d_sysitems);
        CACG_LAUNCH_CHECK();
    }
    return 0;
}
