// Tall-skinny block updates of the outer iteration (src/bcbcgKernel.c:249-262
// and 376-437): Q = S - Q beta, and X += Q alpha, R -= AQ alpha with the
// column norms of the new R.  gfx950, wave64.
//
// Reference order (bitwise): one lane per output entry, the k-sum sequential
// from 0 in the reference's order with separately rounded products and sums
// (src/BLAS3.c:811-819), the rows staged through LDS so the in-place update
// of Q reads old values only; norms as one sequential chain per column.
//
// Fast order: FP64 MFMA v_mfma_f64_16x16x4f64 per 16-row tile.  The
// contraction index is permuted across lane groups: lane group g (lanes
// 16g..16g+15) supplies k = g*K/4 + t at MFMA step t, so each lane's A operand
// is K/4 contiguous doubles of one row (vector loads, no LDS staging); the B
// operand (beta or alpha) is permuted the same way.  Any consistent k order
// gives the same product; only rounding differs.
#include "kcommon.hpp"

#include <type_traits>

namespace {

__device__ __forceinline__ int xcd_remap(int g, int G)
{
    if ((G & 7) != 0)
        return g;
    return (g & 7) * (G >> 3) + (g >> 3);
}

// K contiguous values of one row (16-byte aligned, K a multiple of 4) into
// doubles: 16-byte vector loads of 2 doubles or 4 floats
template <typename T, int K>
__device__ __forceinline__ void load_row(const T *p, double *out)
{
    if constexpr (sizeof(T) == 8) {
        const double2 *v = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int u = 0; u < K / 2; u++) {
            const double2 w = v[u];
            out[2 * u] = w.x;
            out[2 * u + 1] = w.y;
        }
    } else {
        const float4 *v = reinterpret_cast<const float4 *>(p);
#pragma unroll
        for (int u = 0; u < K / 4; u++) {
            const float4 w = v[u];
            out[4 * u] = w.x;
            out[4 * u + 1] = w.y;
            out[4 * u + 2] = w.z;
            out[4 * u + 3] = w.w;
        }
    }
}

// ---- reference order ----------------------------------------------------------

// RPT > 1 (sm divides 256): a thread updates RPT rows of one column, rows
// 256 / sm apart, with one read of beta(q, c) for all of them; each entry
// still sums q = 0, 1, ... in order.
template <typename TS, typename T, bool BETA_LDS, int RPT>
__global__ __launch_bounds__(256) void qupdate_ref_kernel(int n, int sm, const TS *__restrict__ S,
                                                          T *Q, const double *__restrict__ beta,
                                                          int rb)
{
    extern __shared__ double lds[];
    double *bl = lds;
    double *qt = lds + (BETA_LDS ? (long)sm * sm : 0);
    const double *B = BETA_LDS ? bl : beta;
    const int t = threadIdx.x;
    if (BETA_LDS) {
        for (long e = t; e < (long)sm * sm; e += 256)
            bl[e] = beta[e];
    }
    __syncthreads();
    const int ntile = (n + rb - 1) / rb;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const long i0 = (long)tile * rb;
        const int rows = (n - i0) < rb ? (int)(n - i0) : rb;
        for (int e = t; e < rows * sm; e += 256)
            qt[e] = (double)Q[i0 * sm + e];
        __syncthreads();
        if (RPT > 1) {
            const int P = 256 / sm, c = t % sm, r0 = t / sm;
            double acc[RPT];
#pragma unroll
            for (int u = 0; u < RPT; u++)
                acc[u] = 0.0;
            for (int q = 0; q < sm; q++) {
                const double bq = B[(long)q * sm + c];
#pragma unroll
                for (int u = 0; u < RPT; u++)
                    acc[u] = add_rn(acc[u], mul_rn(bq, qt[(long)(r0 + u * P) * sm + q]));
            }
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int r = r0 + u * P;
                if (r < rows) {
                    const long e = (long)r * sm + c;
                    Q[i0 * sm + e] = (T)sub_rn(S ? (double)S[i0 * sm + e] : 0.0, acc[u]);
                }
            }
        } else {
            for (int e = t; e < rows * sm; e += 256) {
                const int r = e / sm, c = e % sm;
                const double *qr = qt + (long)r * sm;
                double acc = 0.0;
                for (int q = 0; q < sm; q++)
                    acc = add_rn(acc, mul_rn(B[(long)q * sm + c], qr[q]));
                Q[i0 * sm + e] = (T)sub_rn(S ? (double)S[i0 * sm + e] : 0.0, acc);
            }
        }
        __syncthreads();
    }
}

// 512 lanes (two waves per SIMD), sm even and dividing 512: a thread updates
// QR2_RPT rows of one column c, rows 512 / sm apart.  beta(q, c) and
// beta(q + 1, c) come with one paired LDS read, each row's Q(r, q), Q(r, q + 1)
// with one 16-byte LDS read (all lanes of a wave read the same row), so two
// steps cost 1 + QR2_RPT LDS reads for 4 QR2_RPT FP64 operations.  The next
// tile's Q values (each thread's own four, coalesced) and the tile's S values
// are in flight while the tile is summed.  Each entry sums q = 0, 1, ... in
// order, as the other reference-order kernel does.
constexpr int QR2_T = 512, QR2_RPT = 4;

template <typename TS, typename T>
__global__ __launch_bounds__(QR2_T) void qupdate_ref2_kernel(int n, int sm,
                                                             const TS *__restrict__ S, T *Q,
                                                             const double *__restrict__ beta)
{
    extern __shared__ double lds[];
    double *bl = lds;                     // sm x sm
    double *qt = lds + (long)sm * sm;     // rb x sm
    const int t = threadIdx.x;
    const int P = QR2_T / sm, rb = QR2_RPT * P;
    const int c = t % sm, r0 = t / sm;
    for (int e = t; e < sm * sm; e += QR2_T)
        bl[e] = beta[e];
    const int ntile = (n + rb - 1) / rb;
    double v[QR2_RPT];
    auto load = [&](int tile) {
        const long i0 = (long)tile * rb;
#pragma unroll
        for (int u = 0; u < QR2_RPT; u++) {
            const long i = i0 + r0 + u * P;
            v[u] = i < n ? (double)Q[i * sm + c] : 0.0;
        }
    };
    int tile = blockIdx.x;
    if (tile < ntile)
        load(tile);
    for (; tile < ntile; tile += gridDim.x) {
#pragma unroll
        for (int u = 0; u < QR2_RPT; u++)
            qt[(r0 + u * P) * sm + c] = v[u];
        __syncthreads();
        if (tile + (int)gridDim.x < ntile)
            load(tile + gridDim.x);
        const long i0 = (long)tile * rb;
        double sv[QR2_RPT], acc[QR2_RPT];
#pragma unroll
        for (int u = 0; u < QR2_RPT; u++) {
            const long i = i0 + r0 + u * P;
            sv[u] = (i < n && S) ? (double)S[i * sm + c] : 0.0;
            acc[u] = 0.0;
        }
        const double *bc = bl + c, *qr = qt + (long)r0 * sm;
        for (int q = 0; q < sm; q += 2) {
            const double b0 = bc[(long)q * sm], b1 = bc[(long)(q + 1) * sm];
#pragma unroll
            for (int u = 0; u < QR2_RPT; u++) {
                const double2 x = *(const double2 *)(qr + (long)u * P * sm + q);
                acc[u] = add_rn(acc[u], mul_rn(b0, x.x));
                acc[u] = add_rn(acc[u], mul_rn(b1, x.y));
            }
        }
#pragma unroll
        for (int u = 0; u < QR2_RPT; u++) {
            const long i = i0 + r0 + u * P;
            if (i < n)
                Q[i * sm + c] = (T)sub_rn(sv[u], acc[u]);
        }
        __syncthreads();   // the tile buffer is refilled next
    }
}

// ALPHA_LDS: alpha (sm x m) staged in LDS; otherwise read from global memory
// (wide shapes whose alpha does not fit beside the row tiles)
template <typename T, bool ALPHA_LDS>
__global__ __launch_bounds__(256) void xrupdate_ref_kernel(int n, int sm, int m,
                                                           const T *__restrict__ Q,
                                                           const T *__restrict__ AQ,
                                                           const double *__restrict__ alpha,
                                                           double *__restrict__ X,
                                                           double *__restrict__ R, int rb)
{
    extern __shared__ double lds[];
    double *qt = lds + (ALPHA_LDS ? (long)sm * m : 0);
    double *aqt = qt + (long)rb * sm;
    const double *al = ALPHA_LDS ? lds : alpha;
    const int t = threadIdx.x;
    if (ALPHA_LDS) {
        for (long e = t; e < (long)sm * m; e += 256)
            lds[e] = alpha[e];
    }
    __syncthreads();
    const int ntile = (n + rb - 1) / rb;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const long i0 = (long)tile * rb;
        const int rows = (n - i0) < rb ? (int)(n - i0) : rb;
        for (int e = t; e < rows * sm; e += 256) {
            qt[e] = (double)Q[i0 * sm + e];
            aqt[e] = (double)AQ[i0 * sm + e];
        }
        __syncthreads();
        for (int e = t; e < rows * m; e += 256) {
            const int r = e / m, c = e % m;
            const double *qr = qt + (long)r * sm, *ar = aqt + (long)r * sm;
            double accq = 0.0, acca = 0.0;
            for (int q = 0; q < sm; q++) {
                const double w = al[(long)q * m + c];
                accq = add_rn(accq, mul_rn(w, qr[q]));
                acca = add_rn(acca, mul_rn(w, ar[q]));
            }
            X[i0 * m + e] = add_rn(X[i0 * m + e], accq);
            R[i0 * m + e] = sub_rn(R[i0 * m + e], acca);
        }
        __syncthreads();
    }
}

// 512 lanes (two waves per SIMD), m even and dividing 512, sm even: a thread
// updates XR2_RPT rows of one column c of X and R, rows 512 / m apart.
// alpha(q, c) and alpha(q + 1, c) come with one paired LDS read and each
// row's Q(r, q), Q(r, q + 1) and AQ(r, q), AQ(r, q + 1) with one 16-byte read
// each; the next tile's Q and AQ values are in flight during the sums.  Each
// entry sums q = 0, 1, ... in order and X += (the Q sum), R -= (the AQ sum)
// as the other reference-order kernel does.
constexpr int XR2_T = 512, XR2_RPT = 2;

template <typename T, int PER>
__global__ __launch_bounds__(XR2_T) void xrupdate_ref2_kernel(int n, int sm, int m,
                                                              const T *__restrict__ Q,
                                                              const T *__restrict__ AQ,
                                                              const double *__restrict__ alpha,
                                                              double *__restrict__ X,
                                                              double *__restrict__ R)
{
    extern __shared__ double lds[];
    const int P = XR2_T / m, rb = XR2_RPT * P;
    double *al = lds;                          // sm x m
    double *qt = lds + (long)sm * m;           // rb x sm
    double *aqt = qt + (long)rb * sm;          // rb x sm
    const int t = threadIdx.x;
    const int c = t % m, r0 = t / m;
    for (int e = t; e < sm * m; e += XR2_T)
        al[e] = alpha[e];
    const int ntile = (n + rb - 1) / rb;
    // PER = rb sm / 512 = XR2_RPT sm / m Q (and AQ) values per thread and tile
    double vq[PER], va[PER];
    auto load = [&](int tile) {
        const long base = (long)tile * rb * sm;
        const long lim = (long)n * sm;
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const long e = base + t + (long)j * XR2_T;
            vq[j] = e < lim ? (double)Q[e] : 0.0;
            va[j] = e < lim ? (double)AQ[e] : 0.0;
        }
    };
    int tile = blockIdx.x;
    if (tile < ntile)
        load(tile);
    for (; tile < ntile; tile += gridDim.x) {
#pragma unroll
        for (int j = 0; j < PER; j++) {
            qt[t + j * XR2_T] = vq[j];
            aqt[t + j * XR2_T] = va[j];
        }
        __syncthreads();
        if (tile + (int)gridDim.x < ntile)
            load(tile + gridDim.x);
        const long i0 = (long)tile * rb;
        double xo[XR2_RPT], ro[XR2_RPT], accq[XR2_RPT], acca[XR2_RPT];
#pragma unroll
        for (int u = 0; u < XR2_RPT; u++) {
            const long i = i0 + r0 + u * P;
            xo[u] = i < n ? X[i * m + c] : 0.0;
            ro[u] = i < n ? R[i * m + c] : 0.0;
            accq[u] = 0.0;
            acca[u] = 0.0;
        }
        const double *ac = al + c;
        for (int q = 0; q < sm; q += 2) {
            const double w0 = ac[(long)q * m], w1 = ac[(long)(q + 1) * m];
#pragma unroll
            for (int u = 0; u < XR2_RPT; u++) {
                const long ro_ = (long)(r0 + u * P) * sm + q;
                const double2 x = *(const double2 *)(qt + ro_);
                const double2 y = *(const double2 *)(aqt + ro_);
                accq[u] = add_rn(accq[u], mul_rn(w0, x.x));
                acca[u] = add_rn(acca[u], mul_rn(w0, y.x));
                accq[u] = add_rn(accq[u], mul_rn(w1, x.y));
                acca[u] = add_rn(acca[u], mul_rn(w1, y.y));
            }
        }
#pragma unroll
        for (int u = 0; u < XR2_RPT; u++) {
            const long i = i0 + r0 + u * P;
            if (i < n) {
                X[i * m + c] = add_rn(xo[u], accq[u]);
                R[i * m + c] = sub_rn(ro[u], acca[u]);
            }
        }
        __syncthreads();   // the tile buffers are refilled next
    }
}

// ||R(:, c)||^2, one sequential chain per column in row order, 64 columns (W
// of them) per workgroup.  Lanes c < W of wave 0 keep the chains; waves 1-7
// meanwhile load the next stage of CREF_SEGE / W rows into the other of two
// LDS buffers, column by column (stride S, even), so a chain lane reads two
// rows of its column with one 16-byte LDS read, and reads the next batch of
// rows while it adds the previous one.  The loaders store each value's square
// (the same product the chain would form), so a chain lane issues one FP64
// add per row: a wave issues an FP64 instruction about every 8.5 cycles, and
// the chain's time per row is its instruction count.  The loader lanes walk
// their elements with increments (no division per element): they share the
// chain wave's SIMD.
// VEC = 2: a loader lane loads two adjacent columns of a row with one 16-byte
// load (m and ldr even, R 16-byte aligned).
constexpr int CREF_T = 512;                  // lanes: wave 0 sums, the rest load
constexpr int CREF_P = CREF_T - 64;
constexpr int CREF_SEGE = 8192;              // values per stage
constexpr int CREF_BUF = CREF_SEGE + 256;    // doubles per LDS buffer (two, dynamic LDS)

template <int VEC>
__global__ __launch_bounds__(CREF_T) void colnorm2_ref_kernel(int n, int m,
                                                              const double *__restrict__ R,
                                                              long ldr, double *__restrict__ out)
{
    constexpr int J = (CREF_SEGE / VEC + CREF_P - 1) / CREF_P;   // loads per loader lane and stage
    extern __shared__ double cbuf[];
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * 64;
    const int W = m - c0 < 64 ? m - c0 : 64;
    const int T = CREF_SEGE / W;           // rows per stage
    const int S = ((T + 1) & ~1) + 2;      // LDS column stride; W * S <= CREF_BUF
    const long nst = ((long)n + T - 1) / T;
    // loader lane: first element VEC * (t - 64) of a stage (row r0, column q0),
    // each next one VEC * CREF_P further (dr rows and dq columns on)
    const int e0 = VEC * (t - 64), de = VEC * CREF_P;
    const int r0 = e0 / W, q0 = e0 % W, dr = de / W, dq = de % W;
    auto produce = [&](long st) {
        double *buf = cbuf + (st & 1) * CREF_BUF;
        const long i0 = st * T;
        const int rows = n - i0 < T ? (int)(n - i0) : T;
        using V = typename std::conditional<VEC == 2, double2, double>::type;
        V v[J];
        int r = r0, q = q0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const double *p = R + (i0 + r) * ldr + c0 + q;
            if (r < rows)
                v[j] = *(const V *)p;
            r += dr;
            q += dq;
            if (q >= W) {
                q -= W;
                r++;
            }
        }
        r = r0;
        q = q0;
#pragma unroll
        for (int j = 0; j < J; j++) {
            if (r < rows) {
                if constexpr (VEC == 2) {
                    buf[q * S + r] = mul_rn(v[j].x, v[j].x);
                    buf[(q + 1) * S + r] = mul_rn(v[j].y, v[j].y);
                } else {
                    buf[q * S + r] = mul_rn(v[j], v[j]);
                }
            }
            r += dr;
            q += dq;
            if (q >= W) {
                q -= W;
                r++;
            }
        }
    };
    if (t >= 64 && nst > 0)
        produce(0);
    __syncthreads();
    double acc = 0.0;
    for (long st = 0; st < nst; st++) {
        if (t >= 64) {
            if (st + 1 < nst)
                produce(st + 1);
        } else if (t < W) {
            const double *col = cbuf + (st & 1) * CREF_BUF + t * S;
            const long i0 = st * T;
            const int rows = n - i0 < T ? (int)(n - i0) : T;
            // NB rows a batch (fewer with VEC 1, whose loader holds more registers)
            constexpr int NB = VEC == 2 ? 32 : 16, NV = NB / 2;
            double2 a[NV], b[NV];
            int r = 0;
            if (rows >= NB) {
#pragma unroll
                for (int u = 0; u < NV; u++)
                    a[u] = *(const double2 *)(col + 2 * u);
            }
            for (; r + 2 * NB <= rows; r += 2 * NB) {
#pragma unroll
                for (int u = 0; u < NV; u++)
                    b[u] = *(const double2 *)(col + r + NB + 2 * u);
#pragma unroll
                for (int u = 0; u < NV; u++) {
                    acc = add_rn(acc, a[u].x);
                    acc = add_rn(acc, a[u].y);
                }
                if (r + 3 * NB <= rows) {
#pragma unroll
                    for (int u = 0; u < NV; u++)
                        a[u] = *(const double2 *)(col + r + 2 * NB + 2 * u);
                }
#pragma unroll
                for (int u = 0; u < NV; u++) {
                    acc = add_rn(acc, b[u].x);
                    acc = add_rn(acc, b[u].y);
                }
            }
            if (r + NB <= rows) {
#pragma unroll
                for (int u = 0; u < NV; u++) {
                    acc = add_rn(acc, a[u].x);
                    acc = add_rn(acc, a[u].y);
                }
                r += NB;
            }
            for (; r < rows; r++)
                acc = add_rn(acc, col[r]);
        }
        __syncthreads();
    }
    if (t < W)
        out[c0 + t] = acc;
}

// ---- fast order: MFMA -----------------------------------------------------------

// acc after the MFMAs of a[tt] x b[tt * stride], tt = 0 .. KQ - 1 in order, the
// B values read from LDS one chunk of 4 steps ahead into alternating
// registers: the scheduling barriers keep each chunk's reads ahead of the
// previous chunk's MFMAs (left alone, the compiler reads each pair of steps
// into one register pair just before its MFMAs and waits on it there).
template <int KQ>
__device__ __forceinline__ dbl4 mfma_chain_lds(const double *a, const double *b, int stride,
                                               dbl4 acc)
{
    constexpr int CH = 4;
    static_assert(KQ % CH == 0, "chunks of 4 steps");
    double bq[2][CH];
#pragma unroll
    for (int u = 0; u < CH; u++)
        bq[0][u] = b[u * stride];
#pragma unroll
    for (int c = 0; c < KQ / CH; c++) {
        if (c + 1 < KQ / CH) {
#pragma unroll
            for (int u = 0; u < CH; u++)
                bq[(c + 1) & 1][u] = b[((c + 1) * CH + u) * stride];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < CH; u++)
            acc = mfma16x16x4(a[c * CH + u], bq[c & 1][u], acc);
        __builtin_amdgcn_sched_barrier(0);
    }
    return acc;
}

// rows per 16-row tile handled by one wave; K = SM, KQ = SM / 4 per lane group
// TS, T = storage types of S and Q (double, or float for an FP32 basis); the
// products and sums are FP64 either way
// NTX > 0: also X += Q alpha with the Q rows before their update (the X
// update of the previous iteration, deferred so Q is read once): alpha is
// SM x m, m <= 16 NTX, held in registers; the products run in the same
// operand order as xrupdate_mfma_kernel's, so X gets the same bits.
template <typename TS, typename T, int SM, int NTX>
__global__ __launch_bounds__(512) void qupdate_mfma_kernel(int n, const TS *__restrict__ S, T *Q,
                                                           const double *__restrict__ beta, int m,
                                                           const double *__restrict__ alpha,
                                                           double *__restrict__ X, long sbs,
                                                           int sbw)
{
    constexpr int KQ = SM / 4, NTL = SM / 16;
    constexpr int NA = NTX > 0 ? NTX : 1;
    // beta (SM x SM), then with NTX > 0 alpha (SM x m, zero-padded to 16 NTX
    // columns): both read from LDS as MFMA B operands
    extern __shared__ double bl[];
    double *al = bl + SM * SM;
    for (int e = threadIdx.x; e < SM * SM; e += 512)
        bl[e] = beta[e];
    if (NTX > 0) {
        for (int e = threadIdx.x; e < SM * 16 * NA; e += 512) {
            const int k = e / (16 * NA), c = e % (16 * NA);
            al[e] = c < m ? alpha[(long)k * m + c] : 0.0;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lg = lane >> 4, lc = lane & 15;
    const int nw = gridDim.x * 8;
    const int wid = xcd_remap(blockIdx.x, gridDim.x) * 8 + wave;
    const int ntiles = (n + 15) / 16;
    // contiguous tile ranges per wave
    const int per = (ntiles + nw - 1) / nw;
    const int tb = wid * per;
    int te = tb + per;
    if (te > ntiles)
        te = ntiles;
    for (int tile = tb; tile < te; tile++) {
        const long i0 = (long)tile * 16;
        const long arow = i0 + lc;
        double av[KQ];
        if (arow < n) {
            load_row<T, KQ>(Q + arow * SM + lg * KQ, av);
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++)
                av[u] = 0.0;
        }
        // the X values of the deferred update, loaded with the Q rows so the
        // update after the MFMAs is a plain store (X + Q alpha, same sum)
        double xv[NA][4];
        if (NTX > 0) {
#pragma unroll
            for (int j = 0; j < NA; j++) {
                const int c = j * 16 + lc;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const long i = i0 + lg + 4 * r;
                    xv[j][r] = (c < m && i < n) ? X[i * m + c] : 0.0;
                }
            }
        }
        // S for one 16-column tile: FP64 S by guarded loads (they stay in
        // flight until the subtraction); FP32 S from a clamped row without a
        // branch, loaded a tile ahead into alternating registers and widened at
        // the subtraction (a guarded FP32 load is widened, and so waited for, on
        // the spot).  The FP32 basis is never absent, so it has no null test.
        // S row-major (sbs == 0) or block-major: blocks of sbw columns, sbs
        // elements apart
        auto load_s = [&](int j, TS *dst) {
            const int col = j * 16 + lc;
            const long co = sbs ? (long)(col / sbw) * sbs + col % sbw : (long)col;
            const long rs = sbs ? (long)sbw : (long)SM;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long i = i0 + lg + 4 * r;
                if constexpr (sizeof(TS) == 4)
                    dst[r] = S[(i < n ? i : n - 1) * rs + co];
                else
                    dst[r] = (i < n && S) ? S[i * rs + co] : (TS)0;
            }
        };
        // Q(:, tile j) = S - Q beta for the rows of this tile
        auto update_tile = [&](int j, const TS *sv) {
            const double *bcol = bl + (lg * KQ) * SM + j * 16 + lc;
            const dbl4 acc = mfma_chain_lds<KQ>(av, bcol, SM, dbl4{0.0, 0.0, 0.0, 0.0});
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long i = i0 + lg + 4 * r;
                if (i < n)
                    Q[i * SM + j * 16 + lc] = (T)((double)sv[r] - acc[r]);
            }
        };
        if constexpr (sizeof(TS) == 4 && NTL % 2 == 0) {
            TS sa[4], sb[4];
            load_s(0, sa);
#pragma unroll 1
            for (int j = 0; j < NTL; j += 2) {
                if (j + 1 < NTL)
                    load_s(j + 1, sb);
                update_tile(j, sa);
                if (j + 1 >= NTL)
                    break;
                if (j + 2 < NTL)
                    load_s(j + 2, sa);
                update_tile(j + 1, sb);
            }
        } else if constexpr (sizeof(TS) == 4) {
            // odd tile counts: one tile ahead through a register copy
            TS sv[4], svn[4];
            load_s(0, sv);
#pragma unroll 1
            for (int j = 0; j < NTL; j++) {
                if (j + 1 < NTL)
                    load_s(j + 1, svn);
                update_tile(j, sv);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    sv[r] = svn[r];
            }
        } else {
            // FP64 S: every column tile's S values are loaded with the Q rows,
            // so no tile's subtraction waits on a load issued just before it
            TS sall[NTL][4];
#pragma unroll
            for (int j = 0; j < NTL; j++)
                load_s(j, sall[j]);
#pragma unroll
            for (int j = 0; j < NTL; j++) {
                // keep each tile's beta reads after the previous tile's MFMAs
                asm volatile("" ::: "memory");
                update_tile(j, sall[j]);
            }
        }
        if (NTX > 0) {
#pragma unroll
            for (int j = 0; j < NA; j++) {
                asm volatile("" ::: "memory");
                const double *acol = al + (lg * KQ) * (16 * NA) + j * 16 + lc;
                const dbl4 ax =
                    mfma_chain_lds<KQ>(av, acol, 16 * NA, dbl4{0.0, 0.0, 0.0, 0.0});
                const int c = j * 16 + lc;
                if (c < m) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const long i = i0 + lg + 4 * r;
                        if (i < n)
                            X[i * m + c] = xv[j][r] + ax[r];
                    }
                }
            }
        }
    }
}

// NTM = output column tiles (m <= 16 * NTM); norms: per-block partials
// DOX: X += Q alpha; DOR: R -= AQ alpha and the column norms of the new R
template <typename T, int SM, int NTM, bool DOX, bool DOR>
__global__ __launch_bounds__(256) void xrupdate_mfma_kernel(int n, int m, const T *__restrict__ Q,
                                                            const T *__restrict__ AQ,
                                                            const double *__restrict__ alpha,
                                                            double *__restrict__ X,
                                                            double *__restrict__ R,
                                                            double *__restrict__ part)
{
    constexpr int KQ = SM / 4;
    __shared__ double red[4][64][NTM];
    // alpha (SM x m, zero-padded to 16 NTM columns) in LDS, not in registers: the
    // 32 k-values x 16 columns a lane would hold cost 64 VGPRs and a wave per SIMD
    __shared__ double s_alf[SM][16 * NTM];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lg = lane >> 4, lc = lane & 15;
    for (int e = threadIdx.x; e < SM * 16 * NTM; e += 256) {
        const int k = e / (16 * NTM), c = e % (16 * NTM);
        s_alf[k][c] = c < m ? alpha[(long)k * m + c] : 0.0;
    }
    __syncthreads();
    double nrm[NTM];
#pragma unroll
    for (int j = 0; j < NTM; j++)
        nrm[j] = 0.0;
    const int nw = gridDim.x * 4;
    const int wid = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
    const int ntiles = (n + 15) / 16;
    const int per = (ntiles + nw - 1) / nw;
    const int tb = wid * per;
    int te = tb + per;
    if (te > ntiles)
        te = ntiles;
    for (int tile = tb; tile < te; tile++) {
        // alpha is re-read from LDS for every tile (a compiler barrier keeps the
        // loop-invariant reads from being hoisted back into registers)
        asm volatile("" ::: "memory");
        const long i0 = (long)tile * 16;
        const long arow = i0 + lc;
        double av[KQ], aq[KQ];
        if (arow < n) {
            if (DOX)
                load_row<T, KQ>(Q + arow * SM + lg * KQ, av);
            if (DOR)
                load_row<T, KQ>(AQ + arow * SM + lg * KQ, aq);
        } else {
#pragma unroll
            for (int u = 0; u < KQ; u++) {
                av[u] = 0.0;
                aq[u] = 0.0;
            }
        }
        // the X and R values this lane updates, loaded with the operand rows so
        // the updates after the MFMAs wait on nothing
        double xo[NTM][4], ro[NTM][4];
#pragma unroll
        for (int j = 0; j < NTM; j++) {
            const int c = j * 16 + lc;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long i = i0 + lg + 4 * r;
                const bool in = c < m && i < n;
                xo[j][r] = (DOX && in) ? X[i * m + c] : 0.0;
                ro[j][r] = (DOR && in) ? R[i * m + c] : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < NTM; j++) {
            dbl4 ax = {0.0, 0.0, 0.0, 0.0}, ar = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int tt = 0; tt < KQ; tt++) {
                const double b = s_alf[lg * KQ + tt][j * 16 + lc];
                if (DOX)
                    ax = mfma16x16x4(av[tt], b, ax);
                if (DOR)
                    ar = mfma16x16x4(aq[tt], b, ar);
            }
            const int c = j * 16 + lc;
            if (c < m) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const long i = i0 + lg + 4 * r;
                    if (i < n) {
                        if (DOX)
                            X[i * m + c] = xo[j][r] + ax[r];
                        if (DOR) {
                            const double rn = ro[j][r] - ar[r];
                            R[i * m + c] = rn;
                            nrm[j] += rn * rn;
                        }
                    }
                }
            }
        }
    }
    if (!DOR)
        return;
#pragma unroll
    for (int j = 0; j < NTM; j++)
        red[wave][lane][j] = nrm[j];
    __syncthreads();
    if (threadIdx.x < NTM * 16) {
        const int j = threadIdx.x / 16, l = threadIdx.x % 16;
        const int c = j * 16 + l;
        double acc = 0.0;
        for (int w = 0; w < 4; w++)
            for (int g = 0; g < 4; g++)
                acc += red[w][g * 16 + l][j];
        if (c < m)
            part[(long)blockIdx.x * m + c] = acc;
    }
}

// per-block partial column norms of a row-major n x m block (fast order)
__global__ __launch_bounds__(256) void colnorm2_part_kernel(int n, int m, const double *__restrict__ R,
                                                           double *__restrict__ part)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    // lanes cover (row offset, column) pairs: 256 / m rows per step when m <= 256
    const int cw = m < 256 ? m : 256;
    const int rstep = 256 / cw;
    const int c0 = t % cw, rr = t / cw;
    const long rows = ((long)n + gridDim.x - 1) / gridDim.x;
    const long r0 = (long)blockIdx.x * rows;
    long r1 = r0 + rows;
    if (r1 > n)
        r1 = n;
    for (int cb = 0; cb < m; cb += cw) {
        double acc = 0.0;
        const int c = cb + c0;
        if (rr < rstep && c < m)
            for (long i = r0 + rr; i < r1; i += rstep) {
                const double v = R[i * m + c];
                acc += v * v;
            }
        red[t] = acc;
        __syncthreads();
        if (t < cw && cb + t < m) {
            double s = 0.0;
            for (int q = 0; q < rstep; q++)
                s += red[q * cw + t];
            part[(long)blockIdx.x * m + cb + t] = s;
        }
        __syncthreads();
    }
}

// one block per column: strided partial sums, then a fixed LDS tree
__global__ __launch_bounds__(256) void colnorm2_final_kernel(const double *__restrict__ part, int nb,
                                                            int m, double *__restrict__ out)
{
    __shared__ double red[256];
    const int c = blockIdx.x, t = threadIdx.x;
    double acc = 0.0;
    for (int b = t; b < nb; b += 256)
        acc += part[(long)b * m + c];
    red[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0)
        out[c] = red[0];
}

constexpr int XR_BLOCKS = 1024;  // fixed: the norm partials depend on it

template <typename TS, typename T, int SM, int NTX = 0>
int launch_qupdate_mfma(int n, const TS *S, T *Q, const double *beta, hipStream_t st, int m = 0,
                        const double *alpha = nullptr, double *X = nullptr, long sbs = 0,
                        int sbw = 1)
{
    const int ntiles = (n + 15) / 16;
    int G = (ntiles + 7) / 8;
    if (G > 256)
        G = 256;
    if (G >= 8)
        G &= ~7;
    const size_t shm = ((size_t)SM * SM + (NTX > 0 ? (size_t)SM * 16 * NTX : 0)) * sizeof(double);
    cacg_allow_lds(qupdate_mfma_kernel<TS, T, SM, NTX>, shm);
    hipLaunchKernelGGL((qupdate_mfma_kernel<TS, T, SM, NTX>), dim3(G), dim3(512), shm, st, n, S,
                       Q, beta, m, alpha, X, sbs, sbw);
    CACG_LAUNCH_CHECK();
    return 0;
}

template <typename T, int SM, int NTM, bool DOX = true, bool DOR = true>
int launch_xrupdate_mfma(int n, int m, const T *Q, const T *AQ, const double *alpha, double *X,
                         double *R, double *d_norm2, double *work, hipStream_t st)
{
    hipLaunchKernelGGL((xrupdate_mfma_kernel<T, SM, NTM, DOX, DOR>), dim3(XR_BLOCKS), dim3(256), 0,
                       st, n, m, Q, AQ, alpha, X, R, work);
    CACG_LAUNCH_CHECK();
    if (!DOR)
        return 0;
    hipLaunchKernelGGL(colnorm2_final_kernel, dim3(m), dim3(256), 0, st, work, XR_BLOCKS,
                       m, d_norm2);
    CACG_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" int cacg_k_colnorm2_ref(int n, int m, const double *R, long ldr, double *d_norm2,
                                   hipStream_t st)
{
    if (n < 0 || m <= 0)
        return (int)hipErrorInvalidValue;
    const size_t lds = 2 * CREF_BUF * sizeof(double);
    if (m % 2 == 0 && ldr % 2 == 0 && (uintptr_t)R % 16 == 0) {
        cacg_allow_lds(colnorm2_ref_kernel<2>, lds);
        hipLaunchKernelGGL(colnorm2_ref_kernel<2>, dim3((m + 63) / 64), dim3(CREF_T), lds, st, n,
                           m, R, ldr, d_norm2);
    } else {
        cacg_allow_lds(colnorm2_ref_kernel<1>, lds);
        hipLaunchKernelGGL(colnorm2_ref_kernel<1>, dim3((m + 63) / 64), dim3(CREF_T), lds, st, n,
                           m, R, ldr, d_norm2);
    }
    CACG_LAUNCH_CHECK();
    return 0;
}

namespace {

template <typename TS, typename T>
int qupdate(int n, int sm, const TS *S, T *Q, const double *beta, int mode, hipStream_t st)
{
    if (n < 0 || sm <= 0)
        return (int)hipErrorInvalidValue;
    if (n == 0)
        return 0;
    if (mode == CACG_MODE_FAST) {
        switch (sm) {
        case 128: return launch_qupdate_mfma<TS, T, 128>(n, S, Q, beta, st);
        case 112: return launch_qupdate_mfma<TS, T, 112>(n, S, Q, beta, st);
        case 96: return launch_qupdate_mfma<TS, T, 96>(n, S, Q, beta, st);
        case 80: return launch_qupdate_mfma<TS, T, 80>(n, S, Q, beta, st);
        case 64: return launch_qupdate_mfma<TS, T, 64>(n, S, Q, beta, st);
        case 48: return launch_qupdate_mfma<TS, T, 48>(n, S, Q, beta, st);
        case 32: return launch_qupdate_mfma<TS, T, 32>(n, S, Q, beta, st);
        case 16: return launch_qupdate_mfma<TS, T, 16>(n, S, Q, beta, st);
        default: break;  // other widths: the reference-order kernel
        }
    }
    // 512 lanes when sm is even, divides 512 and beta plus the tile fit in LDS
    if (sm % 2 == 0 && sm <= QR2_T && QR2_T % sm == 0 &&
        (long)sm * sm + (long)QR2_RPT * QR2_T <= 20480 - 256) {
        const size_t shm = ((size_t)sm * sm + (size_t)QR2_RPT * QR2_T) * sizeof(double);
        const int rb = QR2_RPT * (QR2_T / sm);
        const int ntile = (n + rb - 1) / rb;
        const int cap = shm > 80 * 1024 ? 256 : 1024;   // one workgroup per CU above 80 KiB
        const int G = ntile < cap ? ntile : cap;
        cacg_allow_lds(qupdate_ref2_kernel<TS, T>, shm);
        hipLaunchKernelGGL((qupdate_ref2_kernel<TS, T>), dim3(G), dim3(QR2_T), shm, st, n, sm, S,
                           Q, beta);
        CACG_LAUNCH_CHECK();
        return 0;
    }
    // 4 rows per thread when sm divides 256 and beta plus the tile fit in LDS
    const bool rows4 = sm <= 256 && 256 % sm == 0 &&
                       (long)sm * sm + 4L * (256 / sm) * sm <= 20480 - 256;
    const int rb = rows4 ? 4 * (256 / sm) : (sm >= 256 ? 1 : 256 / sm);
    const bool beta_lds = (long)sm * sm + (long)rb * sm <= 20480 - 256;
    const size_t shm = ((beta_lds ? (size_t)sm * sm : 0) + (size_t)rb * sm) * sizeof(double);
    const int ntile = (n + rb - 1) / rb;
    const int G = ntile < 1024 ? ntile : 1024;
#define CACG_QREF(BL, RPT)                                                                        \
    do {                                                                                          \
        cacg_allow_lds(qupdate_ref_kernel<TS, T, BL, RPT>, shm);                                  \
        hipLaunchKernelGGL((qupdate_ref_kernel<TS, T, BL, RPT>), dim3(G), dim3(256), shm, st, n,  \
                           sm, S, Q, beta, rb);                                                   \
    } while (0)
    if (rows4)
        CACG_QREF(true, 4);
    else if (beta_lds)
        CACG_QREF(true, 1);
    else
        CACG_QREF(false, 1);
#undef CACG_QREF
    CACG_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" int cacg_k_colnorm2_fast(int n, int m, const double *R, double *d_norm2,
                                    double *d_work, hipStream_t st)
{
    if (n < 0 || m <= 0)
        return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(colnorm2_part_kernel, dim3(XR_BLOCKS), dim3(256), 0, st, n, m, R, d_work);
    CACG_LAUNCH_CHECK();
    hipLaunchKernelGGL(colnorm2_final_kernel, dim3(m), dim3(256), 0, st, d_work, XR_BLOCKS, m,
                       d_norm2);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" long cacg_xrupdate_workspace(int m)
{
    return (long)XR_BLOCKS * m;
}

namespace {

template <typename T>
int xrupdate(int n, int sm, int m, const T *Q, const T *AQ, const double *alpha, double *X,
             double *R, double *d_norm2, double *d_work, int mode, hipStream_t st)
{
    if (n < 0 || sm <= 0 || m <= 0)
        return (int)hipErrorInvalidValue;
    if (mode == CACG_MODE_FAST && m <= 32) {
#define CACG_XR(SMV)                                                                              \
    case SMV:                                                                                     \
        return m <= 16 ? launch_xrupdate_mfma<T, SMV, 1>(n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st) \
                       : launch_xrupdate_mfma<T, SMV, 2>(n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st);
        switch (sm) {
            CACG_XR(128)
            CACG_XR(112)
            CACG_XR(96)
            CACG_XR(80)
            CACG_XR(64)
            CACG_XR(48)
            CACG_XR(32)
            CACG_XR(16)
        default: break;
        }
#undef CACG_XR
    }
    // 512 lanes when m is even and divides 512, s <= 8 (a thread holds 2 s values
    // of the next tile) and alpha plus the two tiles fit in LDS
    const int per2 = sm % m == 0 ? XR2_RPT * (sm / m) : 0;
    if (n > 0 && m % 2 == 0 && XR2_T % m == 0 && per2 >= 2 && per2 <= 16 &&
        ((long)sm * m + 2L * XR2_RPT * (XR2_T / m) * sm) * (long)sizeof(double) <= 160 * 1024) {
        const int rb = XR2_RPT * (XR2_T / m);
        const size_t shm = ((size_t)sm * m + 2 * (size_t)rb * sm) * sizeof(double);
        const int ntile = (n + rb - 1) / rb;
        const int cap = shm > 80 * 1024 ? 256 : 1024;   // one workgroup per CU above 80 KiB
        const int G = ntile < cap ? ntile : cap;
#define CACG_XR2(PER)                                                                             \
    case PER:                                                                                     \
        cacg_allow_lds(xrupdate_ref2_kernel<T, PER>, shm);                                        \
        hipLaunchKernelGGL((xrupdate_ref2_kernel<T, PER>), dim3(G), dim3(XR2_T), shm, st, n, sm,  \
                           m, Q, AQ, alpha, X, R);                                                \
        break;
        switch (per2) {
            CACG_XR2(2)
            CACG_XR2(4)
            CACG_XR2(6)
            CACG_XR2(8)
            CACG_XR2(10)
            CACG_XR2(12)
            CACG_XR2(14)
            CACG_XR2(16)
        default: return (int)hipErrorInvalidValue;
        }
#undef CACG_XR2
        CACG_LAUNCH_CHECK();
    } else if (n > 0) {
        int rb = m >= 256 ? 1 : 256 / m;
        if ((long)rb * sm > 4096)
            rb = 4096 / sm > 0 ? 4096 / sm : 1;
        const size_t tiles = 2 * (size_t)rb * sm * sizeof(double);
        const size_t al_bytes = (size_t)sm * m * sizeof(double);
        const bool al_lds = al_bytes + tiles <= 160 * 1024;
        const size_t shm = tiles + (al_lds ? al_bytes : 0);
        if (shm > 160 * 1024)
            return (int)hipErrorInvalidValue;
        const int ntile = (n + rb - 1) / rb;
        const int G = ntile < 1024 ? ntile : 1024;
        if (al_lds) {
            cacg_allow_lds(xrupdate_ref_kernel<T, true>, shm);
            hipLaunchKernelGGL((xrupdate_ref_kernel<T, true>), dim3(G), dim3(256), shm, st, n, sm,
                               m, Q, AQ, alpha, X, R, rb);
        } else {
            cacg_allow_lds(xrupdate_ref_kernel<T, false>, shm);
            hipLaunchKernelGGL((xrupdate_ref_kernel<T, false>), dim3(G), dim3(256), shm, st, n,
                               sm, m, Q, AQ, alpha, X, R, rb);
        }
        CACG_LAUNCH_CHECK();
    }
    if (mode == CACG_MODE_REFERENCE)
        return cacg_k_colnorm2_ref(n, m, R, m, d_norm2, st);
    const int nb = XR_BLOCKS;
    hipLaunchKernelGGL(colnorm2_part_kernel, dim3(nb), dim3(256), 0, st, n, m, R, d_work);
    CACG_LAUNCH_CHECK();
    hipLaunchKernelGGL(colnorm2_final_kernel, dim3(m), dim3(256), 0, st, d_work, nb, m,
                       d_norm2);
    CACG_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" int cacg_k_qupdate(int n, int sm, const double *S, double *Q, const double *beta,
                              int mode, hipStream_t st)
{
    return qupdate<double, double>(n, sm, S, Q, beta, mode, st);
}

namespace {
// the widths with MFMA Q / X / R updates: sm a multiple of 16 up to 128, m <= 32
bool split_supported(int sm, int m)
{
    return sm >= 16 && sm <= 128 && sm % 16 == 0 && m >= 1 && m <= 32;
}
}  // namespace

extern "C" int cacg_k_split_supported(int sm, int m) { return split_supported(sm, m) ? 1 : 0; }

namespace {
template <typename TS>
int qxupdate(int n, int sm, int m, const TS *S, double *Q, const double *beta,
             const double *alpha, double *X, hipStream_t st, long sbs = 0)
{
    if (n < 0 || !split_supported(sm, m))
        return (int)hipErrorInvalidValue;
    if (n == 0)
        return 0;
#define CACG_QX(SMV)                                                                              \
    case SMV:                                                                                     \
        return m <= 16                                                                            \
                   ? launch_qupdate_mfma<TS, double, SMV, 1>(n, S, Q, beta, st, m, alpha, X, sbs, m) \
                   : launch_qupdate_mfma<TS, double, SMV, 2>(n, S, Q, beta, st, m, alpha, X, sbs, m);
    switch (sm) {
        CACG_QX(128)
        CACG_QX(112)
        CACG_QX(96)
        CACG_QX(80)
        CACG_QX(64)
        CACG_QX(48)
        CACG_QX(32)
        CACG_QX(16)
    default: break;
    }
#undef CACG_QX
    return (int)hipErrorInvalidValue;
}
}  // namespace

extern "C" int cacg_k_qxupdate(int n, int sm, int m, const double *S, double *Q,
                               const double *beta, const double *alpha, double *X, hipStream_t st)
{
    return qxupdate<double>(n, sm, m, S, Q, beta, alpha, X, st);
}

extern "C" int cacg_k_qxupdate_blk(int n, int sm, int m, const double *S, long sbs, double *Q,
                                   const double *beta, const double *alpha, double *X,
                                   hipStream_t st)
{
    if (sbs < (long)n * m)
        return (int)hipErrorInvalidValue;
    return qxupdate<double>(n, sm, m, S, Q, beta, alpha, X, st, sbs);
}

extern "C" int cacg_k_qxupdate_s32(int n, int sm, int m, const float *S, double *Q,
                                   const double *beta, const double *alpha, double *X,
                                   hipStream_t st)
{
    return qxupdate<float>(n, sm, m, S, Q, beta, alpha, X, st);
}

extern "C" int cacg_k_xrupdate_split(int n, int sm, int m, const double *Q, const double *AQ,
                                     const double *alpha, double *X, double *R, double *d_norm2,
                                     double *d_work, int which, hipStream_t st)
{
    if (n < 0 || !split_supported(sm, m) || (which != 1 && which != 2))
        return (int)hipErrorInvalidValue;
    if (n == 0)
        return 0;
#define CACG_XS(SMV)                                                                              \
    case SMV:                                                                                     \
        if (which == 1)                                                                           \
            return m <= 16 ? launch_xrupdate_mfma<double, SMV, 1, true, false>(                   \
                                 n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st)                   \
                           : launch_xrupdate_mfma<double, SMV, 2, true, false>(                   \
                                 n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st);                  \
        return m <= 16 ? launch_xrupdate_mfma<double, SMV, 1, false, true>(                       \
                             n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st)                       \
                       : launch_xrupdate_mfma<double, SMV, 2, false, true>(                       \
                             n, m, Q, AQ, alpha, X, R, d_norm2, d_work, st);
    switch (sm) {
        CACG_XS(128)
        CACG_XS(112)
        CACG_XS(96)
        CACG_XS(80)
        CACG_XS(64)
        CACG_XS(48)
        CACG_XS(32)
        CACG_XS(16)
    default: break;
    }
#undef CACG_XS
    return (int)hipErrorInvalidValue;
}

extern "C" int cacg_k_qupdate_zero(int n, int sm, double *Q, const double *beta, hipStream_t st)
{
    return qupdate<double, double>(n, sm, (const double *)nullptr, Q, beta, CACG_MODE_FAST, st);
}

extern "C" int cacg_k_qupdate_s32(int n, int sm, const float *S, double *Q, const double *beta,
                                  hipStream_t st)
{
    return qupdate<float, double>(n, sm, S, Q, beta, CACG_MODE_FAST, st);
}

extern "C" int cacg_k_xrupdate(int n, int sm, int m, const double *Q, const double *AQ,
                               const double *alpha, double *X, double *R, double *d_norm2,
                               double *d_work, int mode, hipStream_t st)
{
    return xrupdate<double>(n, sm, m, Q, AQ, alpha, X, R, d_norm2, d_work, mode, st);
}

// ---- fills ---------------------------------------------------------------------

namespace {
template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *d, long count, T v)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long)gridDim.x * 256)
        d[e] = v;
}
}  // namespace

extern "C" int cacg_k_fill_i32(int *d, long count, int v, hipStream_t st)
{
    if (count <= 0)
        return 0;
    hipLaunchKernelGGL(fill_kernel<int>, dim3(cacg_grid(count, 256, 1024)), dim3(256), 0, st, d,
                       count, v);
    CACG_LAUNCH_CHECK();
    return 0;
}

namespace {
__global__ __launch_bounds__(256) void scale_kernel(double *d, long count, double a)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < count; e += (long)gridDim.x * 256)
        d[e] *= a;
}
}  // namespace

extern "C" int cacg_k_scale_f64(double *d, long count, double a, hipStream_t st)
{
    if (count <= 0)
        return 0;
    hipLaunchKernelGGL(scale_kernel, dim3(cacg_grid(count, 256, 1024)), dim3(256), 0, st, d, count,
                       a);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" int cacg_k_fill_f64(double *d, long count, double v, hipStream_t st)
{
    if (count <= 0)
        return 0;
    hipLaunchKernelGGL(fill_kernel<double>, dim3(cacg_grid(count, 256, 1024)), dim3(256), 0, st, d,
                       count, v);
    CACG_LAUNCH_CHECK();
    return 0;
}
