// Multi-RHS CSR SpMM with fused Chebyshev-recurrence epilogues, and the
// Gerschgorin interval.  gfx950 (CDNA4), wave64.
//
// Scheduling: vector-per-row.  A row is served by LPR consecutive lanes, each
// lane owning VW contiguous output columns (VW = 4 when k is a multiple of 4
// and x is 16-byte aligned with an even leading dimension, else 1), so the
// gather of x(col, :) is one contiguous k*8-byte segment per nonzero read
// with 16-byte loads, and the lanes of a row read the same (col, val) pair
// (an LDS broadcast).  Each output is summed in CSR order from 0, product and
// sum rounded separately, which is exactly the reference's arithmetic
// (src/BLAS3.c:78-112, 950-965): the basis and A*Q blocks are bit-identical
// to the reference's.  The nonzero loop goes in batches of 4 or 8 with all
// x loads of a batch issued before its in-order accumulation, for
// memory-level parallelism.
//
// Tile order: hardware workgroup ids are dealt round-robin over the 8 XCDs,
// so workgroup h runs on XCD h % 8.  Two orders, chosen per launch:
//   bands (P = 0): each XCD owns a contiguous band of row tiles, and its
//     workgroups take the band's tiles interleaved (workgroup j of the XCD:
//     tiles j, j + Gx, j + 2 Gx, ...), so the XCD sweeps its band as one
//     compact front: rows that gather the same x rows (neighbours i +- 1,
//     i +- a grid line) run close together in time under the same L2;
//   chunks (P > 0): chunks of P tiles are dealt round-robin over the XCDs and
//     each XCD sweeps its chunks in order, so an XCD owns a slice of every
//     grid plane and the x rows a plane away (i +- a grid plane) are gathered
//     again under the same L2 one chunk later, instead of a plane later;
//   chunks in Q passes (P > 0, Q > 1): the XCD's chunks l = 0, 1, ... are swept
//     as Q passes, pass h taking chunks h, h + Q, h + 2 Q, ...  When an XCD owns
//     Q chunks of every grid plane, a pass walks one narrow column of the grid
//     plane by plane, so the three planes' rows a row gathers (its own, and
//     +- a plane) are a Q-th of the chunk-order working set: at k = 128 the
//     XCD's slice of three planes is 6 MB, which its 4 MB L2 cannot hold.
// The order changes no result (each row is summed by one lane group, in CSR
// order); which is faster depends on the matrix, so the solver times both on
// its own matrix (cacg_spmm_sched).  Operands used once (the matrix, S_{j-2},
// the outputs) are loaded and stored non-temporally so they do not evict the
// gathered rows from L2.
#include <float.h>
#include <stdint.h>

#include <type_traits>

#include "kcommon.hpp"

namespace {


// VW consecutive elements of T from a 16-byte aligned address (VW = 4) or
// one element (VW = 1)
template <typename T, int VW>
struct vecld;
template <typename T>
struct vecld<T, 1> {
    T v[1];
    __device__ __forceinline__ void load(const T *p) { v[0] = *p; }
};
template <>
struct vecld<double, 4> {
    double v[4];
    // two 16-byte pieces: columns c, c + 1 and c + OFF, c + OFF + 1
    template <int OFF = 2>
    __device__ __forceinline__ void load(const double *p)
    {
        const double2 a = *reinterpret_cast<const double2 *>(p);
        const double2 b = *reinterpret_cast<const double2 *>(p + OFF);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = b.x;
        v[3] = b.y;
    }
};
// load a lane's VW elements: the FP64 pairs OFF columns apart, else contiguous
template <int OFF, typename T, int VW>
__device__ __forceinline__ void vload(vecld<T, VW> &d, const T *p)
{
    if constexpr (std::is_same<T, double>::value && VW == 4)
        d.template load<OFF>(p);
    else
        d.load(p);
}

template <>
struct vecld<float, 4> {
    float v[4];
    __device__ __forceinline__ void load(const float *p)
    {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
    }
};

// One workgroup = 256 lanes = NG lane groups of LPR lanes; it walks a range
// of row tiles (TR = NG * RPG rows each).  Per tile the row pointers and,
// when they fit (CAP), the tile's (col, val) pairs are staged in LDS by
// coalesced cooperative loads, so each row's gathers wait on one global
// latency instead of three chained ones (row pointer -> column id -> x).
// Lane group g takes rows g, g + NG, ... of the tile; the epilogue operands of
// a row are loaded before its gathers, which go out in batches of
// spmm_unroll<LPR>() nonzeros.  Requires k % VW == 0 and, for VW > 1,
// x and ldx 16-byte aligned.
// Types: TV = matrix values and arithmetic, TX = gathered x, TY = y, p1, p2
// and copy0; x and the epilogue operands are converted to TV on load.  The
// FP64 solver uses <double, double, double>; the FP32 basis <float, float,
// float>, and <float, double, float> for its first step, which reads the FP64
// residual and writes its FP32 copy S_0.
// The lt-th tile of an XCD's chunks of P tiles swept in Q passes: passes
// r .. Q-1 hold F chunks each, then passes 0 .. r-1 hold F + 1, for the XCD's
// L = F Q + r chunks.  Recomputed per tile from the launch constants (a few
// scalar operations per tile) rather than kept live: the extra scalar registers
// made hipcc spill SGPRs in the k = 32 kernel and then fail instruction
// selection.
__device__ __forceinline__ int passes_tile(int lt, int P, int Q, int nx, int xcd, int ntiles)
{
    const int nch = (ntiles + P - 1) / P;
    const int L = nch > xcd ? (nch - 1 - xcd) / nx + 1 : 0;
    const int u = lt / P;
    if (u >= L)
        return -1;
    const int F = L / Q, r = L % Q, u1 = (Q - r) * F;
    int h, i;
    if (u < u1) {
        h = r + u / F;
        i = u % F;
    } else {
        h = (u - u1) / (F + 1);
        i = (u - u1) % (F + 1);
    }
    return ((i * Q + h) * nx + xcd) * P + lt % P;
}

// The XCD's tiles in sweep order: the lt-th tile of its band (P == 0) or of
// its chunks of P tiles (in Q passes when Q > 1); -1 past the end.  Only the
// last chunk can be partial, and every order visits it last (the passes start
// at pass L % Q, so the pass holding chunk L - 1 comes last), so the first -1
// ends a workgroup's sweep.
__device__ __forceinline__ int sweep_tile(int lt, int P, int Q, int nx, int xcd, int band,
                                          int ntiles)
{
    int tile;
    if (P == 0) {
        if (lt >= band)
            return -1;
        tile = xcd * band + lt;
    } else if (Q <= 1) {
        tile = ((lt / P) * nx + xcd) * P + lt % P;
    } else {
        tile = passes_tile(lt, P, Q, nx, xcd, ntiles);
        if (tile < 0)
            return -1;
    }
    return tile < ntiles ? tile : -1;
}

// The rows of one tile (rows r0 .. r0 + rows - 1) whose row pointers are in
// s_rp and, when staged, (col, val) pairs in s_col / s_val (offsets from
// pbase): lane group g takes rows g, g + NG, ..., each lane VW columns.
// Nonzeros gathered per batch: 4 for the shapes where the smaller register
// footprint buys more waves in flight (k = 16: 101 instead of 145 VGPRs, 5
// waves per SIMD instead of 3, 13% faster; k = 128: 2%), 8 elsewhere.
template <int LPR>
constexpr int spmm_unroll()
{
    return (LPR == 4 || LPR == 32) ? 4 : 8;
}

// Columns of a lane: VW = 4 contiguous ones from c0 (IL = false), or with IL
// (FP64, VW = 4, k a multiple of 4 LPR) the pairs c0, c0 + 1 and c0 + 2 LPR,
// c0 + 2 LPR + 1 with c0 = cb + 2 lane, so that each 16-byte load or store
// instruction of a lane group covers one contiguous half of the row's block
// instead of every other 16 bytes of all of it (whole 64-byte runs per
// instruction for the non-temporal stores).  Each output is still summed by one
// lane in CSR order: the layout changes no bit.
template <typename TV, typename TX, typename TY, int LPR, int VW, int RPG, int EPI, bool IL>
__device__ __forceinline__ void tile_rows(const int *s_rp, const int *s_col, const TV *s_val,
                                          bool staged, int pbase, int r0, int rows, int g,
                                          int lane, const int *__restrict__ colidx,
                                          const TV *__restrict__ val, int k,
                                          const TX *__restrict__ x, long ldx, TY *__restrict__ y,
                                          long ldy, TV a, TV b, const TY *__restrict__ p1,
                                          long ld1, const TY *__restrict__ p2, long ld2,
                                          TY *__restrict__ copy0, long ldc, int long_min)
{
    constexpr int NG = 256 / LPR;
    constexpr int UNR = spmm_unroll<LPR>();
    static_assert(!IL || VW == 4, "interleaved columns need 4 per lane");
    constexpr int OFF = IL ? 2 * LPR : 2;   // column distance of a lane's second pair
    for (int cb = 0; cb < k; cb += LPR * VW) {
        const int c0 = cb + lane * (IL ? 2 : VW);
        if (c0 >= k)
            continue;
        // column of a lane's element v
        auto col = [&](int v) { return VW == 4 && v >= 2 ? c0 + OFF + v - 2 : c0 + v; };
        for (int rr = 0; rr < RPG; rr++) {
            const int li = rr * NG + g;
            if (li >= rows)
                continue;
            const long i = r0 + li;
            const int pb = s_rp[li] - pbase, pe0 = s_rp[li + 1] - pbase;
            // a long row is the row-split kernel's: no gathers, no stores
            const bool lng = long_min > 0 && pe0 - pb >= long_min;
            const int pe = lng ? pb : pe0;
            TV acc[VW], e1[VW], e2[VW];
#pragma unroll
            for (int v = 0; v < VW; v++) {
                acc[v] = (TV)0;
                if (EPI == CACG_EPI_RESID) {
                    e1[v] = (TV)__builtin_nontemporal_load(p1 + i * ld1 + col(v));
                } else if (EPI == CACG_EPI_CHEB1) {
                    e1[v] = (TV)x[i * ldx + col(v)];
                } else if (EPI == CACG_EPI_CHEB) {
                    e1[v] = (TV)p1[i * ld1 + col(v)];
                    e2[v] = (TV)__builtin_nontemporal_load(p2 + i * ld2 + col(v));
                }
            }
            for (int p = pb; p < pe; p += UNR) {
                int cc[UNR];
                TV vv[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    const bool in = p + u < pe;
                    cc[u] = in ? (staged ? s_col[p + u] : colidx[pbase + p + u]) : 0;
                    vv[u] = in ? (staged ? s_val[p + u] : val[pbase + p + u]) : (TV)0;
                }
                vecld<TX, VW> xv[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++)
                    if (p + u < pe)
                        vload<OFF>(xv[u], x + (long)cc[u] * ldx + c0);
#pragma unroll
                for (int u = 0; u < UNR; u++)
                    if (p + u < pe) {
#pragma unroll
                        for (int v = 0; v < VW; v++)
                            acc[v] = add_rn(acc[v], mul_rn(vv[u], (TV)xv[u].v[v]));
                    }
            }
#pragma unroll
            for (int v = 0; v < VW; v++) {
                TV out;
                if (EPI == CACG_EPI_NONE) {
                    out = acc[v];
                } else if (EPI == CACG_EPI_RESID) {
                    out = sub_rn(e1[v], acc[v]);
                } else if (EPI == CACG_EPI_CHEB1) {
                    out = add_rn(mul_rn(a, acc[v]), mul_rn(b, e1[v]));
                    if (!lng)
                        __builtin_nontemporal_store((TY)e1[v], copy0 + i * ldc + col(v));
                } else {
                    out = sub_rn(add_rn(mul_rn(a, acc[v]), mul_rn(b, e1[v])), e2[v]);
                }
                if (!lng)
                    __builtin_nontemporal_store((TY)out, y + i * ldy + col(v));
            }
        }
    }
}

// Minimum waves per SIMD asked of the register allocator: 6 for the FP64 A Q
// form (k > 64, no epilogue: 80 VGPRs with 3 spilled instead of 82 at 5 waves,
// 1306 against 1331 us at C2); 1 (no constraint) elsewhere.  Asking 5 of the
// k = 16 basis step (96 VGPRs, 1 spilled) gained nothing.
template <typename TV, int LPR, int EPI>
constexpr int spmm_min_waves()
{
    return (sizeof(TV) == 8 && LPR == 32 && EPI == CACG_EPI_NONE) ? 6 : 1;
}

template <typename TV, typename TX, typename TY, int LPR, int VW, int RPG, int EPI, bool IL>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(spmm_min_waves<TV, LPR, EPI>()))) void spmm_kernel(
    int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
    const TV *__restrict__ val, int k, const TX *__restrict__ x, long ldx, TY *__restrict__ y,
    long ldy, double ad, double bd, const TY *__restrict__ p1, long ld1, const TY *__restrict__ p2,
    long ld2, TY *__restrict__ copy0, long ldc, int P, int Q, int long_min)
{
    constexpr int NG = 256 / LPR, TR = NG * RPG;
    constexpr int CAP = TR * 8 > 256 ? TR * 8 : 256;
    __shared__ int s_rp[TR + 1];
    __shared__ int s_col[CAP];
    __shared__ TV s_val[CAP];
    const int t = threadIdx.x, g = t / LPR, lane = t % LPR;
    const int ntiles = (n + TR - 1) / TR;
    const int nx = gridDim.x >= 8 ? 8 : 1;
    const int xcd = blockIdx.x % nx, gx = gridDim.x / nx, j = blockIdx.x / nx;
    const int band = (ntiles + nx - 1) / nx;
    for (int lt = j;; lt += gx) {
        const int tile = sweep_tile(lt, P, Q, nx, xcd, band, ntiles);
        if (tile < 0)
            break;
        const int r0 = tile * TR;
        const int rows = (n - r0) < TR ? (n - r0) : TR;
        for (int e = t; e <= rows; e += 256)
            s_rp[e] = rowptr[r0 + e];
        __syncthreads();
        const int pbase = s_rp[0];
        const int tnnz = s_rp[rows] - pbase;
        const bool staged = tnnz <= CAP;
        if (staged)
            for (int e = t; e < tnnz; e += 256) {
                s_col[e] = __builtin_nontemporal_load(colidx + pbase + e);
                s_val[e] = __builtin_nontemporal_load(val + pbase + e);
            }
        __syncthreads();
        tile_rows<TV, TX, TY, LPR, VW, RPG, EPI, IL>(s_rp, s_col, s_val, staged, pbase, r0, rows, g,
                                                 lane, colidx, val, k, x, ldx, y, ldy, (TV)ad,
                                                 (TV)bd, p1, ld1, p2, ld2, copy0, ldc, long_min);
        __syncthreads();
    }
}

// The same with the tile staging pipelined (LPR 4 and 8: k = 16, 32): while a
// tile is gathered, the (col, val) pairs of the next tile and the row
// pointers of the one after are already in flight into registers, so the
// chain row pointer -> (col, val) -> x is paid once per workgroup instead of
// once per tile (8% on the 128^3 Laplacian at k = 16; no gain for wide rows,
// which keep spmm_kernel, and not used for LPR <= 2, whose doubled LDS
// staging would halve occupancy).  Two LDS buffers alternate between tiles.
template <typename TV, typename TX, typename TY, int LPR, int VW, int RPG, int EPI, bool IL>
__global__ __launch_bounds__(256) void spmm_pipe_kernel(
    int n, const int *__restrict__ rowptr, const int *__restrict__ colidx,
    const TV *__restrict__ val, int k, const TX *__restrict__ x, long ldx, TY *__restrict__ y,
    long ldy, double ad, double bd, const TY *__restrict__ p1, long ld1, const TY *__restrict__ p2,
    long ld2, TY *__restrict__ copy0, long ldc, int P, int long_min)
{
    constexpr int NG = 256 / LPR, TR = NG * RPG;
    constexpr int CAP = TR * 8 > 256 ? TR * 8 : 256;
    constexpr int PER = CAP / 256;  // staged pairs per thread
    static_assert(CAP % 256 == 0, "CAP must be a multiple of the workgroup size");
    __shared__ int s_rp[2][TR + 1];
    __shared__ int s_col[2][CAP];
    __shared__ TV s_val[2][CAP];
    // one pass: the multi-pass order is not built into the pipelined form (with
    // it, hipcc spills SGPRs in these kernels and fails instruction selection
    // in the FP64 k = 17..32 one; for the same reason the kernel takes no Q
    // argument beside long_min)
    const int Q = 1;
    const int t = threadIdx.x, g = t / LPR, lane = t % LPR;
    const int ntiles = (n + TR - 1) / TR;
    const int nx = gridDim.x >= 8 ? 8 : 1;
    const int xcd = blockIdx.x % nx, gx = gridDim.x / nx, j = blockIdx.x / nx;
    const int band = (ntiles + nx - 1) / nx;
    int lt = j;
    int tile = sweep_tile(lt, P, Q, nx, xcd, band, ntiles);
    if (tile < 0)
        return;
    auto rows_of = [&](int tl) { return (n - tl * TR) < TR ? (n - tl * TR) : TR; };
    // registers in flight: entry t of a tile's row pointers, and PER (col, val) pairs
    int rp_reg = 0;
    int col_reg[PER];
    TV val_reg[PER];
    // prologue: tile 0 staged in buffer 0, the row pointers of tile 1 in flight
    int cur = 0;
    int rows = rows_of(tile);
    if (t <= rows)
        s_rp[0][t] = rowptr[tile * TR + t];
    __syncthreads();
    {
        const int pb = s_rp[0][0], tn = s_rp[0][rows] - pb;
        if (tn <= CAP)
            for (int e = t; e < tn; e += 256) {
                s_col[0][e] = __builtin_nontemporal_load(colidx + pb + e);
                s_val[0][e] = __builtin_nontemporal_load(val + pb + e);
            }
    }
    int tile_n = sweep_tile(lt + gx, P, Q, nx, xcd, band, ntiles);
    if (tile_n >= 0 && t <= rows_of(tile_n))
        rp_reg = rowptr[tile_n * TR + t];
    __syncthreads();
    for (;;) {
        const int nxt = cur ^ 1;
        const int r0 = tile * TR;
        const int pbase = s_rp[cur][0];
        const bool staged = s_rp[cur][rows] - pbase <= CAP;
        // the next tile's row pointers to LDS, then its (col, val) pairs and the
        // row pointers of the tile after it into registers
        int rows_n = 0, pbn = 0, tnn = 0;
        if (tile_n >= 0) {
            rows_n = rows_of(tile_n);
            if (t <= rows_n)
                s_rp[nxt][t] = rp_reg;
        }
        __syncthreads();
        const int tile_nn = tile_n >= 0 ? sweep_tile(lt + 2 * gx, P, Q, nx, xcd, band, ntiles) : -1;
        if (tile_n >= 0) {
            pbn = s_rp[nxt][0];
            tnn = s_rp[nxt][rows_n] - pbn;
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int e = t + 256 * u;
                const bool in = tnn <= CAP && e < tnn;
                col_reg[u] = in ? __builtin_nontemporal_load(colidx + pbn + e) : 0;
                val_reg[u] = in ? __builtin_nontemporal_load(val + pbn + e) : (TV)0;
            }
            if (tile_nn >= 0 && t <= rows_of(tile_nn))
                rp_reg = rowptr[tile_nn * TR + t];
        }
        tile_rows<TV, TX, TY, LPR, VW, RPG, EPI, IL>(s_rp[cur], s_col[cur], s_val[cur], staged, pbase,
                                                 r0, rows, g, lane, colidx, val, k, x, ldx, y, ldy,
                                                 (TV)ad, (TV)bd, p1, ld1, p2, ld2, copy0, ldc,
                                                 long_min);
        if (tile_n < 0)
            break;
        if (tnn <= CAP) {
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int e = t + 256 * u;
                if (e < tnn) {
                    s_col[nxt][e] = col_reg[u];
                    s_val[nxt][e] = val_reg[u];
                }
            }
        }
        __syncthreads();
        cur = nxt;
        lt += gx;
        tile = tile_n;
        rows = rows_n;
        tile_n = tile_nn;
    }
}

// Row split for long rows.  A row with far more nonzeros than the mean holds
// its lane group for one chain of batched gathers per few nonzeros (two
// dependent global latencies per batch of 4 or 8), and the launch ends with
// that workgroup.  Rows with at least long_min nonzeros are left out of the
// kernels above and taken here, one workgroup of LR_T lanes per row: waves
// 1 .. LR_T / 64 - 1 form the products val(p) * x(col(p), c) of the next
// segment of the row into LDS,
// every product rounded as above, while the lanes c < kc of wave 0 add the
// current segment's in CSR order, each keeping the chain of column cb + c.
// Same products, same order: the same bits as the vector-per-row kernels,
// with one LDS read per nonzero on the chain instead of a global gather.
// Columns go 64 at a time (cb = 0, 64, ...), each block one pass over the row.
constexpr int LR_T = 1024;                     // lanes per workgroup: wave 0 sums, the rest produce
constexpr int LR_P = LR_T - 64;                // producer lanes
constexpr int LR_SEGE = 8192;                  // products per LDS buffer (two, in dynamic LDS)
constexpr int LR_J = (LR_SEGE + LR_P - 1) / LR_P;   // products per producer lane and segment

template <typename TV, typename TX, typename TY, int EPI>
__global__ __launch_bounds__(LR_T) void spmm_long_kernel(
    const int *__restrict__ long_rows, const int *__restrict__ rowptr,
    const int *__restrict__ colidx, const TV *__restrict__ val, int k, const TX *__restrict__ x,
    long ldx, TY *__restrict__ y, long ldy, double ad, double bd, const TY *__restrict__ p1,
    long ld1, const TY *__restrict__ p2, long ld2, TY *__restrict__ copy0, long ldc)
{
    extern __shared__ double lr_lds[];
    TV *prod = reinterpret_cast<TV *>(lr_lds);   // two segment buffers of LR_SEGE
    const int t = threadIdx.x;
    const long i = long_rows[blockIdx.x];
    const int pb = rowptr[i], pe = rowptr[i + 1];
    const TV a = (TV)ad, b = (TV)bd;
    for (int cb = 0; cb < k; cb += 64) {
        const int kc = k - cb < 64 ? k - cb : 64;
        const int seg = LR_SEGE / kc;   // nonzeros per segment
        const int nseg = (pe - pb + seg - 1) / seg;
        // products of segment sg into buffer sg & 1 (producer lanes t >= 64):
        // all (col, val) loads of the lane, then all its x gathers, then the
        // products, so each step's loads are in flight together
        auto produce = [&](int sg) {
            TV *buf = prod + (sg & 1) * LR_SEGE;
            const int p0 = pb + sg * seg;
            const int cnt = (pe - p0 < seg ? pe - p0 : seg) * kc;
            const int tp = t - 64;
            int cc[LR_J];
            TV vv[LR_J];
#pragma unroll
            for (int j = 0; j < LR_J; j++) {
                const int e = tp + LR_P * j;
                cc[j] = e < cnt ? colidx[p0 + e / kc] : 0;
                vv[j] = e < cnt ? val[p0 + e / kc] : (TV)0;
            }
            TX xv[LR_J];
#pragma unroll
            for (int j = 0; j < LR_J; j++) {
                const int e = tp + LR_P * j;
                xv[j] = e < cnt ? x[(long)cc[j] * ldx + cb + e % kc] : (TX)0;
            }
#pragma unroll
            for (int j = 0; j < LR_J; j++) {
                const int e = tp + LR_P * j;
                if (e < cnt)
                    buf[e] = mul_rn(vv[j], (TV)xv[j]);
            }
        };
        if (t >= 64 && nseg > 0)
            produce(0);
        __syncthreads();
        TV acc = (TV)0;
        for (int sg = 0; sg < nseg; sg++) {
            if (t >= 64) {
                if (sg + 1 < nseg)
                    produce(sg + 1);
            } else if (t < kc) {
                const TV *buf = prod + (sg & 1) * LR_SEGE;
                const int p0 = pb + sg * seg;
                const int cnt = pe - p0 < seg ? pe - p0 : seg;
#pragma unroll 8
                for (int q = 0; q < cnt; q++)
                    acc = add_rn(acc, buf[q * kc + t]);
            }
            __syncthreads();
        }
        if (t < kc) {
            const int c = cb + t;
            TV out;
            if (EPI == CACG_EPI_NONE) {
                out = acc;
            } else if (EPI == CACG_EPI_RESID) {
                out = sub_rn((TV)p1[i * ld1 + c], acc);
            } else if (EPI == CACG_EPI_CHEB1) {
                const TV e1 = (TV)x[i * ldx + c];
                out = add_rn(mul_rn(a, acc), mul_rn(b, e1));
                copy0[i * ldc + c] = (TY)e1;
            } else {
                out = sub_rn(add_rn(mul_rn(a, acc), mul_rn(b, (TV)p1[i * ld1 + c])),
                             (TV)p2[i * ld2 + c]);
            }
            y[i * ldy + c] = (TY)out;
        }
    }
}

// Launch shape (measured on the 128^3 Laplacian, k = 16 and k = 128): narrow
// rows (LPR <= 8) one row per lane group and 32 workgroups per CU; wide rows
// two rows per lane group and 16 workgroups per CU.
template <typename TV, typename TX, typename TY, int LPR, int VW>
int launch_spmm(int n, const int *rowptr, const int *col, const TV *val, int k, const TX *x,
                long ldx, TY *y, long ldy, int epi, double a, double b, const TY *p1, long ld1,
                const TY *p2, long ld2, TY *copy0, long ldc, const cacg_spmm_sched *sc,
                const int *long_rows, int n_long, int long_min, hipStream_t st)
{
    constexpr int RPG = LPR >= 16 ? 2 : 1;
    constexpr int TR = (256 / LPR) * RPG;
    const int gmax = sc && sc->grid > 0 ? sc->grid : (LPR >= 16 ? 4096 : 8192);
    const int ntiles = (n + TR - 1) / TR;
    // a multiple of 8 so every XCD gets the same number of workgroups
    int G = ntiles < gmax ? ntiles : gmax;
    if (G >= 8)
        G &= ~7;
    // chunk order: chunk_rows rounded to whole tiles
    int P = 0, Q = 1;
    if (sc && sc->chunk_rows > 0) {
        P = sc->chunk_rows >= TR ? sc->chunk_rows / TR : 1;
        Q = sc->passes > 1 ? sc->passes : 1;
    }
    const int lmin = n_long > 0 ? long_min : 0;
    // the row-split kernel's two segment buffers exceed the 64 KiB default
    constexpr size_t lr_bytes = 2 * LR_SEGE * sizeof(TV);
#define CACG_SPMM_LAUNCH(E, IL)                                                                  \
    do {                                                                                         \
        if (lmin > 0 && hipFuncSetAttribute(                                                     \
                            reinterpret_cast<const void *>(&spmm_long_kernel<TV, TX, TY, E>),    \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lr_bytes) != hipSuccess) \
            return (int)hipErrorInvalidConfiguration;                                            \
        if (lmin > 0)                                                                            \
            hipLaunchKernelGGL((spmm_long_kernel<TV, TX, TY, E>), dim3(n_long), dim3(LR_T), lr_bytes, \
                               st, long_rows, rowptr, col, val, k, x, ldx, y, ldy, a, b, p1, ld1, \
                               p2, ld2, copy0, ldc);                                             \
        if constexpr (LPR == 4 || LPR == 8)                                                      \
            hipLaunchKernelGGL((spmm_pipe_kernel<TV, TX, TY, LPR, VW, RPG, E, IL>), dim3(G),     \
                               dim3(256), 0, st, n, rowptr, col, val, k, x, ldx, y, ldy, a, b, p1, \
                               ld1, p2, ld2, copy0, ldc, P, lmin);                               \
        else                                                                                     \
            hipLaunchKernelGGL((spmm_kernel<TV, TX, TY, LPR, VW, RPG, E, IL>), dim3(G), dim3(256), \
                               0, st, n, rowptr, col, val, k, x, ldx, y, ldy, a, b, p1, ld1, p2,    \
                               ld2, copy0, ldc, P, Q, lmin);                                     \
    } while (0)
    // FP64 with 4 columns per lane: interleaved pairs when every lane's block is full
#define CACG_SPMM_CASE(E)                                                                        \
    do {                                                                                         \
        if constexpr (f64 && VW == 4) {                                                          \
            if (k % (LPR * VW) == 0)                                                             \
                CACG_SPMM_LAUNCH(E, true);                                                       \
            else                                                                                 \
                CACG_SPMM_LAUNCH(E, false);                                                      \
        } else {                                                                                 \
            CACG_SPMM_LAUNCH(E, false);                                                          \
        }                                                                                        \
    } while (0)
    constexpr bool f64 = sizeof(TV) == 8 && sizeof(TX) == 8 && sizeof(TY) == 8;
    constexpr bool f32 = sizeof(TV) == 4 && sizeof(TX) == 4 && sizeof(TY) == 4;
    if constexpr (f64) {
        if (epi == CACG_EPI_NONE)
            CACG_SPMM_CASE(CACG_EPI_NONE);
        else if (epi == CACG_EPI_RESID)
            CACG_SPMM_CASE(CACG_EPI_RESID);
        else if (epi == CACG_EPI_CHEB1)
            CACG_SPMM_CASE(CACG_EPI_CHEB1);
        else if (epi == CACG_EPI_CHEB)
            CACG_SPMM_CASE(CACG_EPI_CHEB);
        else
            return (int)hipErrorInvalidValue;
    } else if constexpr (f32) {
        if (epi == CACG_EPI_NONE)
            CACG_SPMM_CASE(CACG_EPI_NONE);
        else if (epi == CACG_EPI_CHEB)
            CACG_SPMM_CASE(CACG_EPI_CHEB);
        else
            return (int)hipErrorInvalidValue;
    } else {
        // FP32 arithmetic on an FP64 x: the first basis step only
        if (epi == CACG_EPI_CHEB1)
            CACG_SPMM_CASE(CACG_EPI_CHEB1);
        else
            return (int)hipErrorInvalidValue;
    }
#undef CACG_SPMM_CASE
#undef CACG_SPMM_LAUNCH
    CACG_LAUNCH_CHECK();
    return 0;
}

// lanes per row and columns per lane from k and the alignment of x
template <typename TV, typename TX, typename TY>
int dispatch_spmm(int n, const int *rowptr, const int *col, const TV *val, int k, const TX *x,
                  long ldx, TY *y, long ldy, int epi, double a, double b, const TY *p1, long ld1,
                  const TY *p2, long ld2, TY *copy0, long ldc, const cacg_spmm_sched *sc,
                  const int *long_rows, int n_long, int long_min, hipStream_t st)
{
#define CACG_SPMM_ARGS n, rowptr, col, val, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0, ldc, sc, long_rows, n_long, long_min, st
    const bool vec4 = k % 4 == 0 && ((uintptr_t)x & 15) == 0 && (ldx * (long)sizeof(TX)) % 16 == 0;
    if (vec4) {
        // LPR = k / 4 rounded up to a power of two, at most 32 (wider k loops)
        const int q = k / 4;
        if (q > 16)
            return launch_spmm<TV, TX, TY, 32, 4>(CACG_SPMM_ARGS);
        if (q > 8)
            return launch_spmm<TV, TX, TY, 16, 4>(CACG_SPMM_ARGS);
        if (q > 4)
            return launch_spmm<TV, TX, TY, 8, 4>(CACG_SPMM_ARGS);
        if (q > 2)
            return launch_spmm<TV, TX, TY, 4, 4>(CACG_SPMM_ARGS);
        if (q > 1)
            return launch_spmm<TV, TX, TY, 2, 4>(CACG_SPMM_ARGS);
        return launch_spmm<TV, TX, TY, 1, 4>(CACG_SPMM_ARGS);
    }
    if (k > 32)
        return launch_spmm<TV, TX, TY, 64, 1>(CACG_SPMM_ARGS);
    if (k > 16)
        return launch_spmm<TV, TX, TY, 32, 1>(CACG_SPMM_ARGS);
    if (k > 8)
        return launch_spmm<TV, TX, TY, 16, 1>(CACG_SPMM_ARGS);
    if (k > 4)
        return launch_spmm<TV, TX, TY, 8, 1>(CACG_SPMM_ARGS);
    if (k > 2)
        return launch_spmm<TV, TX, TY, 4, 1>(CACG_SPMM_ARGS);
    if (k > 1)
        return launch_spmm<TV, TX, TY, 2, 1>(CACG_SPMM_ARGS);
    return launch_spmm<TV, TX, TY, 1, 1>(CACG_SPMM_ARGS);
#undef CACG_SPMM_ARGS
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

// the long-row fields of a CSR: a negative count or threshold, or long rows
// without their list
template <typename CSR>
bool bad_long_rows(const CSR *A)
{
    return A->n_long < 0 || A->long_min < 0 ||
           (A->long_min > 0 && A->n_long > 0 && !A->long_rows) || A->n_long > A->n;
}

}  // namespace

extern "C" int cacg_k_spmm_x(const cacg_dcsr *A, int k, const double *x, long ldx, double *y,
                             long ldy, int epi, double a, double b, const double *p1, long ld1,
                             const double *p2, long ld2, double *copy0, long ldc,
                             const cacg_spmm_sched *sc, hipStream_t st)
{
    if (!A || k <= 0 || A->n < 0 || bad_long_rows(A) ||
        (sc && (sc->grid < 0 || sc->chunk_rows < 0 || sc->passes < 0)))
        return (int)hipErrorInvalidValue;
    if (A->n == 0)
        return 0;
    return dispatch_spmm<double, double, double>(A->n, A->rowptr, A->col, A->val, k, x, ldx, y,
                                                 ldy, epi, a, b, p1, ld1, p2, ld2, copy0, ldc, sc,
                                                 A->long_rows, A->n_long, A->long_min, st);
}

extern "C" int cacg_k_spmm(const cacg_dcsr *A, int k, const double *x, long ldx, double *y,
                           long ldy, int epi, double a, double b, const double *p1, long ld1,
                           const double *p2, long ld2, double *copy0, long ldc, hipStream_t st)
{
    return cacg_k_spmm_x(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, copy0, ldc, nullptr,
                         st);
}

extern "C" int cacg_k_spmm_f32_x(const cacg_scsr *A, int k, const float *x, long ldx, float *y,
                                 long ldy, int epi, double a, double b, const float *p1, long ld1,
                                 const float *p2, long ld2, const cacg_spmm_sched *sc,
                                 hipStream_t st)
{
    if (!A || k <= 0 || A->n < 0 || bad_long_rows(A) || (epi != CACG_EPI_NONE && epi != CACG_EPI_CHEB) ||
        (sc && (sc->grid < 0 || sc->chunk_rows < 0 || sc->passes < 0)))
        return (int)hipErrorInvalidValue;
    if (A->n == 0)
        return 0;
    return dispatch_spmm<float, float, float>(A->n, A->rowptr, A->col, A->val, k, x, ldx, y, ldy,
                                              epi, a, b, p1, ld1, p2, ld2, NULL, 0, sc,
                                              A->long_rows, A->n_long, A->long_min, st);
}

extern "C" int cacg_k_spmm_f32(const cacg_scsr *A, int k, const float *x, long ldx, float *y,
                               long ldy, int epi, double a, double b, const float *p1, long ld1,
                               const float *p2, long ld2, hipStream_t st)
{
    return cacg_k_spmm_f32_x(A, k, x, ldx, y, ldy, epi, a, b, p1, ld1, p2, ld2, nullptr, st);
}

extern "C" int cacg_k_cheb1_f32_x(const cacg_scsr *A, int k, const double *r, long ldr,
                                  float *s1, long ld1, float *s0, long ld0, double a, double b,
                                  const cacg_spmm_sched *sc, hipStream_t st)
{
    if (!A || k <= 0 || A->n < 0 || !s0 || bad_long_rows(A) ||
        (sc && (sc->grid < 0 || sc->chunk_rows < 0 || sc->passes < 0)))
        return (int)hipErrorInvalidValue;
    if (A->n == 0)
        return 0;
    return dispatch_spmm<float, double, float>(A->n, A->rowptr, A->col, A->val, k, r, ldr, s1,
                                               ld1, CACG_EPI_CHEB1, a, b, NULL, 0, NULL, 0, s0,
                                               ld0, sc, A->long_rows, A->n_long, A->long_min, st);
}

extern "C" int cacg_k_cheb1_f32(const cacg_scsr *A, int k, const double *r, long ldr, float *s1,
                                long ld1, float *s0, long ld0, double a, double b, hipStream_t st)
{
    return cacg_k_cheb1_f32_x(A, k, r, ldr, s1, ld1, s0, ld0, a, b, nullptr, st);
}

namespace {
__global__ __launch_bounds__(256) void cvt_f64_f32_kernel(long count, const double *__restrict__ src,
                                                          float *__restrict__ dst)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long)gridDim.x * 256)
        dst[e] = (float)src[e];
}
}  // namespace

namespace {
__global__ __launch_bounds__(256) void cvt_f32_f64_kernel(long count, const float *__restrict__ src,
                                                          double *__restrict__ dst)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long)gridDim.x * 256)
        dst[e] = (double)src[e];
}
}  // namespace

extern "C" int cacg_k_cvt_f32_f64(long count, const float *src, double *dst, hipStream_t st)
{
    if (count < 0)
        return (int)hipErrorInvalidValue;
    if (count == 0)
        return 0;
    hipLaunchKernelGGL(cvt_f32_f64_kernel, dim3(cacg_grid(count, 256, 4096)), dim3(256), 0, st,
                       count, src, dst);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_k_cvt_f64_f32(long count, const double *src, float *dst, hipStream_t st)
{
    if (count < 0)
        return (int)hipErrorInvalidValue;
    if (count == 0)
        return 0;
    hipLaunchKernelGGL(cvt_f64_f32_kernel, dim3(cacg_grid(count, 256, 4096)), dim3(256), 0, st,
                       count, src, dst);
    CACG_LAUNCH_CHECK();
    return 0;
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
