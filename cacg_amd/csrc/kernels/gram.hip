// Gram blocks G = L^T [R1 | R2] over tall-skinny row-major blocks
// (Q^T A S, Q^T A Q, Q^T R of src/BLAS3.c:699-758).  gfx950, wave64.
//
// Reference order: one lane per entry G[a][b], one sequential chain over all
// rows in row order, product and sum rounded separately: the reference's
// summation at one rank, bit for bit.  The chain bounds its time at n
// dependent steps whatever the number of entries.
//
// Fast order: split-K.  Workgroup x owns a contiguous chunk of rows and a
// group of 16x16 output tiles per wave; FP64 MFMA v_mfma_f64_16x16x4f64 with
// the contraction index = row index, so both operands load straight from
// row-major memory: the 16 lanes of a lane group read 16 consecutive doubles
// (128 B) of one row and the 4 lane groups read 4 consecutive rows.  Loads for
// the next 8 rows are issued before the MFMAs of the current 8 (register
// double buffer).  Each chunk writes a kl_pad x kr_pad partial; a second
// kernel sums the partials in chunk order, so the result is deterministic and
// independent of scheduling.
#include "kcommon.hpp"

namespace {

// One workgroup per 8 x 8 block of G: lane l of wave 0 keeps the chain of
// G[a0 + l / 8][b0 + l % 8], sequential in the row index (the reference's
// summation, bit for bit), so the kernel takes n steps of one product and one
// dependent sum whatever the grid.  Waves 1-3 meanwhile stage the block's
// columns of the next GREF_ROWS rows (8 values of L and 8 of [R1 | R2] a row,
// columns past the last clamped to it) into the other of two LDS buffers, so
// a stage costs the chain one barrier instead of a global latency, and the
// chain lanes read their two operands 16 rows ahead of the sums.
constexpr int GREF_ROWS = 256;                          // rows per stage (two 32 KiB buffers)
constexpr int GREF_J = (GREF_ROWS * 16 + 191) / 192;    // values per loader lane and stage

struct GrefSrc {
    const double *L;
    long ldl;
    int kl;
    const double *R1;
    long ld1;
    int kr1;
    const double *R2;
    long ld2;
    int kr2;
};

__global__ __launch_bounds__(256) void gram_ref_kernel(int n, GrefSrc s, int nbb,
                                                       double *__restrict__ G1,
                                                       double *__restrict__ G2)
{
    __shared__ double stg[2][GREF_ROWS * 16];   // row r: L values 0..7, R values 8..15
    const int t = threadIdx.x;
    const int a0 = (blockIdx.x / nbb) * 8, b0 = (blockIdx.x % nbb) * 8;
    const int kr = s.kr1 + s.kr2;
    const long nst = ((long)n + GREF_ROWS - 1) / GREF_ROWS;
    // loader lanes: stage st into buffer st & 1, all loads before the stores
    auto produce = [&](long st) {
        double *buf = stg[st & 1];
        const long i0 = st * GREF_ROWS;
        const int cnt = (n - i0 < GREF_ROWS ? (int)(n - i0) : GREF_ROWS) * 16;
        const int tp = t - 64;
        double v[GREF_J];
#pragma unroll
        for (int j = 0; j < GREF_J; j++) {
            const int e = tp + 192 * j;
            // rows and columns past the last are clamped to it (never summed)
            const long i = i0 + (e >> 4) < n ? i0 + (e >> 4) : (long)n - 1;
            const int q = e & 15, qb = q < 8 ? 0 : q - 8;
            const int a = a0 + q < s.kl ? a0 + q : s.kl - 1;
            const int bb = b0 + qb < kr ? b0 + qb : kr - 1;
            const double *pr = bb < s.kr1 ? s.R1 + i * s.ld1 + bb : s.R2 + i * s.ld2 + (bb - s.kr1);
            const double *p = q < 8 ? s.L + i * s.ldl + a : pr;
            v[j] = e < cnt ? *p : 0.0;
        }
#pragma unroll
        for (int j = 0; j < GREF_J; j++) {
            const int e = tp + 192 * j;
            if (e < cnt)
                buf[e] = v[j];
        }
    };
    if (t >= 64 && nst > 0)
        produce(0);
    __syncthreads();
    const int ai = t >> 3, bi = t & 7;
    double acc = 0.0;
    for (long st = 0; st < nst; st++) {
        if (t >= 64) {
            if (st + 1 < nst)
                produce(st + 1);
        } else {
            const double *buf = stg[st & 1];
            const long i0 = st * GREF_ROWS;
            const int rows = n - i0 < GREF_ROWS ? (int)(n - i0) : GREF_ROWS;
            int r = 0;
            if (rows >= 32) {
                double la[16], ra[16], lb[16], rb[16];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    la[u] = buf[u * 16 + ai];
                    ra[u] = buf[u * 16 + 8 + bi];
                }
                for (; r + 32 <= rows; r += 32) {
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        lb[u] = buf[(r + 16 + u) * 16 + ai];
                        rb[u] = buf[(r + 16 + u) * 16 + 8 + bi];
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++)
                        acc = add_rn(acc, mul_rn(la[u], ra[u]));
                    if (r + 64 <= rows) {
#pragma unroll
                        for (int u = 0; u < 16; u++) {
                            la[u] = buf[(r + 32 + u) * 16 + ai];
                            ra[u] = buf[(r + 32 + u) * 16 + 8 + bi];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 16; u++)
                        acc = add_rn(acc, mul_rn(lb[u], rb[u]));
                }
            }
            for (; r < rows; r++)
                acc = add_rn(acc, mul_rn(buf[r * 16 + ai], buf[r * 16 + 8 + bi]));
        }
        __syncthreads();
    }
    if (t < 64) {
        const int a = a0 + ai, b = b0 + bi;
        if (a < s.kl) {
            if (b < s.kr1)
                G1[(long)a * s.kr1 + b] = acc;
            else if (b < kr)
                G2[(long)a * s.kr2 + (b - s.kr1)] = acc;
        }
    }
}

// ---- fast: split-K FP64 MFMA ------------------------------------------------

struct GramPlan {
    int TM, TN;      // 16x16 tiles in M (kl) and N (kr)
    int MT, NT;      // tiles per wave
    int GN, NG;      // tile groups along N, total groups
    int WGG, KS;     // groups per workgroup, waves sharing one group (split K inside the workgroup)
    int gy;          // workgroups along the tile-group dimension
    int gx;          // row chunks (split K across workgroups)
    int chunk;       // rows per chunk
};

GramPlan plan_gram(int n, int kl, int kr)
{
    GramPlan p;
    p.TM = (kl + 15) / 16;
    p.TN = (kr + 15) / 16;
    p.NT = p.TN < 9 ? p.TN : 9;
    p.GN = (p.TN + p.NT - 1) / p.NT;
    p.MT = p.TM >= 8 ? 2 : 1;
    const int GM = (p.TM + p.MT - 1) / p.MT;
    p.NG = GM * p.GN;
    if (p.NG >= 4) {
        p.WGG = 4;
        p.KS = 1;
    } else {
        p.WGG = p.NG;
        p.KS = 4 / p.NG;
    }
    p.gy = (p.NG + p.WGG - 1) / p.WGG;
    // about 2 workgroups per CU over 256 CUs, at least 64 rows per chunk; a
    // function of n only, so the summation order does not depend on the device
    int gx = 512 / p.gy;
    if (gx < 1)
        gx = 1;
    const int step = 8 * p.KS;
    long maxc = ((long)n + 63) / 64;
    if (gx > maxc)
        gx = (int)(maxc > 0 ? maxc : 1);
    long chunk = ((long)n + gx - 1) / gx;
    chunk = (chunk + step - 1) / step * step;
    if (chunk < step)
        chunk = step;
    p.chunk = (int)chunk;
    p.gx = (int)(((long)n + chunk - 1) / chunk);
    if (p.gx < 1)
        p.gx = 1;
    return p;
}

// L and R1 are FP64 (the FP64 solver) or FP32 (the FP32 basis: Q, AQ, S);
// R2 (the residual R) is always FP64.  Loaded values are widened to FP64 and
// every product and sum is FP64.  R1 may be block-major (bs1 > 0): column c
// of row i at R1[(c / ld1) * bs1 + i * ld1 + c % ld1], blocks of ld1 columns
// bs1 elements apart (the basis S stored as its blocks S_0, S_1, ...); the
// LDS-staged kernels read it (their 16-byte vectors stay inside a block),
// the unstaged one does not.
template <typename TL, typename T1>
struct SrcT {
    const TL *L;
    long ldl;
    int kl;
    const T1 *R1;
    long ld1;
    int kr1;
    const double *R2;
    long ld2;
    int kr2;
    long bs1;
};

// offset of column c of R1 from its row start
template <typename TL, typename T1>
__device__ __forceinline__ long r1_col(const SrcT<TL, T1> &s, int c)
{
    return s.bs1 ? (long)(c / s.ld1) * s.bs1 + c % s.ld1 : (long)c;
}

template <typename TL, typename T1, int MT, int NT>
__global__ __launch_bounds__(256) void gram_mfma_kernel(int n, SrcT<TL, T1> s,
                                                        double *__restrict__ part,
                                                        int kl_pad, int kr_pad, int chunk, int GN,
                                                        int NG, int WGG, int KS)
{
    extern __shared__ double red[];  // (KS - 1) * WGG * MT * NT * 4 * 64 doubles when KS > 1
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane >> 4, lc = lane & 15;
    const int gb = wave % WGG, ks = wave / WGG;
    const int g = blockIdx.y * WGG + gb;
    const bool active = g < NG && ks < KS;
    const int gm = active ? g / GN : 0, gn = active ? g % GN : 0;
    const int kr = s.kr1 + s.kr2;

    // per-lane operand columns: A tile t reads column ca0 + 16 t of L; B tile t
    // reads column cb0 + 16 t of [R1 | R2]
    const int ca0 = gm * MT * 16 + lc;
    const int cb0 = gn * NT * 16 + lc;

    dbl4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
            acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    const int r0 = blockIdx.x * chunk;
    int r1 = r0 + chunk;
    if (r1 > n)
        r1 = n;

    // k-steps of 4 rows; wave ks takes steps ks, ks + KS, ...; two steps per
    // batch, batches double-buffered in registers
    double av[2][2][MT], bv[2][2][NT];
    auto load = [&](int buf, int rb) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int row = rb + u * 4 * KS + lr;
            const bool rok = row < r1;
            const TL *pl = s.L + (long)row * s.ldl;
            const T1 *p1 = s.R1 + (long)row * s.ld1;
            const double *p2 = s.R2 + (long)row * s.ld2 - s.kr1;
#pragma unroll
            for (int t = 0; t < MT; t++) {
                const int a = ca0 + 16 * t;
                av[buf][u][t] = (rok && a < s.kl) ? (double)pl[a] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int b = cb0 + 16 * t;
                bv[buf][u][t] = (rok && b < kr) ? (b < s.kr1 ? (double)p1[b] : p2[b]) : 0.0;
            }
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[i][j] = mfma16x16x4(av[buf][u][i], bv[buf][u][j], acc[i][j]);
    };
    if (active) {
        const int stride = 8 * KS;
        int rb = r0 + ks * 4;
        if (rb < r1)
            load(0, rb);
        for (; rb < r1; rb += 2 * stride) {
            if (rb + stride < r1)
                load(1, rb + stride);
            compute(0);
            if (rb + stride >= r1)
                break;
            if (rb + 2 * stride < r1)
                load(0, rb + 2 * stride);
            compute(1);
        }
    }

    // combine the KS waves of one tile group in wave order
    if (KS > 1) {
        const int per = MT * NT * 4;
        if (active && ks > 0) {
            double *dst = red + ((long)((ks - 1) * WGG + gb) * per) * 64;
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        dst[((i * NT + j) * 4 + r) * 64 + lane] = acc[i][j][r];
        }
        __syncthreads();
        if (active && ks == 0) {
            for (int q = 1; q < KS; q++) {
                const double *src = red + ((long)((q - 1) * WGG + gb) * per) * 64;
#pragma unroll
                for (int i = 0; i < MT; i++)
#pragma unroll
                    for (int j = 0; j < NT; j++)
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            acc[i][j][r] += src[((i * NT + j) * 4 + r) * 64 + lane];
            }
        }
    }
    if (!active || ks != 0)
        return;
    double *out = part + (long)blockIdx.x * kl_pad * kr_pad;
#pragma unroll
    for (int i = 0; i < MT; i++) {
        const int tm = gm * MT + i;
        if (tm * 16 >= kl_pad)
            continue;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int tn = gn * NT + j;
            if (tn * 16 >= kr_pad)
                continue;
#pragma unroll
            for (int r = 0; r < 4; r++)
                out[(long)(tm * 16 + lr + 4 * r) * kr_pad + tn * 16 + lc] = acc[i][j][r];
        }
    }
}

// 64 consecutive outputs per block; wave w sums chunk partials x in
// [w*gx/4, (w+1)*gx/4) in order, then the four wave sums are added in wave order
__global__ __launch_bounds__(256) void gram_reduce_kernel(const double *__restrict__ part, int gx,
                                                          int kl, int kr1, int kr2, int kl_pad,
                                                          int kr_pad, double *__restrict__ G1,
                                                          double *__restrict__ G2)
{
    __shared__ double red[4][64];
    const int kr = kr1 + kr2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long e = (long)blockIdx.x * 64 + lane;
    const bool ok = e < (long)kl * kr;
    const int a = ok ? (int)(e / kr) : 0, b = ok ? (int)(e % kr) : 0;
    const long stride = (long)kl_pad * kr_pad;
    const double *p = part + (long)a * kr_pad + b;
    const int q = (gx + 3) / 4;
    const int x0 = wave * q, x1 = (x0 + q) < gx ? (x0 + q) : gx;
    double acc = 0.0;
    if (ok) {
        int x = x0;
        for (; x + 4 <= x1; x += 4) {
            const double v0 = p[x * stride], v1 = p[(x + 1) * stride];
            const double v2 = p[(x + 2) * stride], v3 = p[(x + 3) * stride];
            acc += v0;
            acc += v1;
            acc += v2;
            acc += v3;
        }
        for (; x < x1; x++)
            acc += p[x * stride];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && ok) {
        const double sum = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
        if (b < kr1)
            G1[(long)a * kr1 + b] = sum;
        else
            G2[(long)a * kr2 + (b - kr1)] = sum;
    }
}

// LDS-staged variant for the large case (even column counts, 16-byte aligned
// rows).  A stage is SR rows of L and of [R1 | R2]; the NW * 64 lanes load it
// cooperatively with 16-byte loads, LDG stages ahead, holding the data in
// registers; the stage is then written to one of two LDS buffers and every
// wave reads its MFMA operands from there.  Each row is read from HBM once per
// workgroup.  Every wave owns one tile group (MT x NT tiles), so all tiles of
// G are covered by one workgroup when NG <= NW.  Each tile accumulates its
// k-steps of 4 rows in row order whatever SR, LDG, NW, MT and NT are, so the
// partials depend on the row chunks only.
//   <4, MT, NT, 8, 3, 5>: general large case (four tile groups per workgroup);
//   <12, 2, 3, 16, 2, 3>, <16, 2, 2, 16, 2, 2>: a 128-row G against 144 and
//   128 columns, as 12 and 16 waves of 2 x 3 and 2 x 2 tiles (measured on
//   gfx950: 58 and 56 TFLOP/s against 42 and 50 for the general shape).
// LDS row stride (doubles) of a staged k-column block: k rounded up to 16
// mod 32.  A wave reads 4 rows x 16 consecutive doubles per MFMA operand;
// with a stride of 16 mod 32 doubles, rows 0, 1 (and 2, 3) fall into the two
// disjoint halves of the 64 four-byte LDS banks, so a 64-lane read takes the
// 2 cycles its 512 bytes need (a stride of k + 2 overlapped all four rows).
__host__ __device__ inline int lds_row_stride(int k)
{
    return k + (48 - k % 32) % 32;
}

template <typename TL, typename T1, int NW, int MT, int NT, int SR, int LDG, int VPT>
__global__ __launch_bounds__(NW * 64) void gram_lds_kernel(int n, SrcT<TL, T1> s,
                                                           double *__restrict__ part, int kl_pad,
                                                           int kr_pad, int chunk, int GN, int NG)
{
    constexpr int NTH = NW * 64;
    extern __shared__ double lds[];
    const int kr = s.kr1 + s.kr2;
    const int LDP = lds_row_stride(s.kl), RDP = lds_row_stride(kr);  // padded LDS row strides
    const int BUF = SR * (LDP + RDP);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane >> 4, lc = lane & 15;
    const int g = blockIdx.y * NW + wave;
    const bool active = g < NG;
    const int gm = active ? g / GN : 0, gn = active ? g % GN : 0;
    const int ca0 = gm * MT * 16 + lc, cb0 = gn * NT * 16 + lc;

    const int r0 = blockIdx.x * chunk;
    int r1 = r0 + chunk;
    if (r1 > n)
        r1 = n;
    const int nstage = (r1 - r0 + SR - 1) / SR;

    // this lane's share of a stage: 16-byte vector u -> (source row-0 address,
    // row stride in bytes, stage row, LDS offset, FP32 source); an FP32 vector
    // carries 4 columns and is widened into 4 doubles of LDS
    constexpr int EL = 16 / sizeof(TL), E1 = 16 / sizeof(T1);
    const int vl = SR * s.kl / EL, v1 = SR * s.kr1 / E1, v2 = SR * s.kr2 / 2;
    const char *src[VPT];
    long ldb[VPT];
    int srow[VPT], loff[VPT];
    bool isf[VPT];
#pragma unroll
    for (int u = 0; u < VPT; u++) {
        int v = threadIdx.x + NTH * u;
        srow[u] = -1;
        src[u] = reinterpret_cast<const char *>(s.L);
        ldb[u] = 0;
        loff[u] = 0;
        isf[u] = false;
        if (v < vl) {
            const int row = v / (s.kl / EL), c = EL * (v % (s.kl / EL));
            src[u] = reinterpret_cast<const char *>(s.L + c);
            ldb[u] = s.ldl * (long)sizeof(TL);
            srow[u] = row;
            loff[u] = row * LDP + c;
            isf[u] = sizeof(TL) == 4;
        } else if ((v -= vl) < v1) {
            const int row = v / (s.kr1 / E1), c = E1 * (v % (s.kr1 / E1));
            src[u] = reinterpret_cast<const char *>(s.R1 + r1_col(s, c));
            ldb[u] = s.ld1 * (long)sizeof(T1);
            srow[u] = row;
            loff[u] = SR * LDP + row * RDP + c;
            isf[u] = sizeof(T1) == 4;
        } else if ((v -= v1) < v2) {
            const int row = v / (s.kr2 / 2), c = 2 * (v % (s.kr2 / 2));
            src[u] = reinterpret_cast<const char *>(s.R2 + c);
            ldb[u] = s.ld2 * (long)sizeof(double);
            srow[u] = row;
            loff[u] = SR * LDP + row * RDP + s.kr1 + c;
        }
    }
    uint4 stg[LDG][VPT];
    auto gload = [&](int slot, int st) {
        const int rb = r0 + st * SR;
#pragma unroll
        for (int u = 0; u < VPT; u++) {
            const int row = rb + srow[u];
            const bool ok = srow[u] >= 0 && st < nstage && row < r1;
            stg[slot][u] = ok ? *reinterpret_cast<const uint4 *>(src[u] + (long)row * ldb[u])
                              : uint4{0u, 0u, 0u, 0u};
        }
    };
    auto lstore = [&](int slot, int buf) {
        double *b = lds + buf * BUF;
#pragma unroll
        for (int u = 0; u < VPT; u++) {
            if (srow[u] < 0)
                continue;
            const uint4 w = stg[slot][u];
            if ((sizeof(TL) == 4 || sizeof(T1) == 4) && isf[u]) {
                float4 f;
                __builtin_memcpy(&f, &w, 16);
                *reinterpret_cast<double2 *>(b + loff[u]) = double2{(double)f.x, (double)f.y};
                *reinterpret_cast<double2 *>(b + loff[u] + 2) = double2{(double)f.z, (double)f.w};
            } else {
                double2 d;
                __builtin_memcpy(&d, &w, 16);
                *reinterpret_cast<double2 *>(b + loff[u]) = d;
            }
        }
    };

    dbl4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
            acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
    // the wide shapes (NW >= 12) read the operands of k-step ks + 1 while the
    // MFMAs of ks run (two register sets; the scheduling barriers keep the
    // reads ahead, which the compiler otherwise places just before their MFMAs
    // and waits on); the general shape keeps one set (a second one would cost
    // its 8-tile form a wave per SIMD)
    constexpr bool PF = NW >= 12;
    auto compute = [&](int buf) {
        const double *bl = lds + buf * BUF;
        const double *br = bl + SR * LDP;
        double av[PF ? 2 : 1][MT], bv[PF ? 2 : 1][NT];
        auto operands = [&](int ks, int p) {
            const int row = ks * 4 + lr;
#pragma unroll
            for (int t = 0; t < MT; t++) {
                const int a = ca0 + 16 * t;
                av[p][t] = a < s.kl ? bl[row * LDP + a] : 0.0;
            }
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int b = cb0 + 16 * t;
                bv[p][t] = b < kr ? br[row * RDP + b] : 0.0;
            }
        };
        if constexpr (PF)
            operands(0, 0);
#pragma unroll
        for (int ks = 0; ks < SR / 4; ks++) {
            const int p = PF ? ks & 1 : 0;
            if constexpr (PF) {
                if (ks + 1 < SR / 4)
                    operands(ks + 1, (ks + 1) & 1);
                __builtin_amdgcn_sched_barrier(0);
            } else {
                operands(ks, 0);
            }
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[i][j] = mfma16x16x4(av[p][i], bv[p][j], acc[i][j]);
            if constexpr (PF)
                __builtin_amdgcn_sched_barrier(0);
        }
    };

    // prologue: stages 0 .. LDG-1 in flight, stage 0 in LDS buffer 0
#pragma unroll
    for (int q = 0; q < LDG; q++)
        gload(q, q);
    lstore(0, 0);
    gload(0, LDG);
    __syncthreads();
    // stage st: computed from buffer st & 1; its successor st + 1 sits in
    // register slot (st + 1) % LDG and is stored to the other buffer
    for (int st0 = 0; st0 < nstage; st0 += LDG) {
#pragma unroll
        for (int q = 0; q < LDG; q++) {
            const int st = st0 + q;
            if (st < nstage) {
                if (active)
                    compute(st & 1);
                const int nslot = (q + 1) % LDG;
                lstore(nslot, (st + 1) & 1);
                gload(nslot, st + 1 + LDG);
            }
            __syncthreads();
        }
    }

    if (!active)
        return;
    double *out = part + (long)blockIdx.x * kl_pad * kr_pad;
#pragma unroll
    for (int i = 0; i < MT; i++) {
        const int tm = gm * MT + i;
        if (tm * 16 >= kl_pad)
            continue;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int tn = gn * NT + j;
            if (tn * 16 >= kr_pad)
                continue;
#pragma unroll
            for (int r = 0; r < 4; r++)
                out[(long)(tm * 16 + lr + 4 * r) * kr_pad + tn * 16 + lc] = acc[i][j][r];
        }
    }
}

// 16-byte vectors of one stage of SR rows of L, R1 and R2
template <typename TL, typename T1>
long stage_vectors(const SrcT<TL, T1> &s, int SR)
{
    return (long)SR * (s.kl / (16 / (int)sizeof(TL)) + s.kr1 / (16 / (int)sizeof(T1)) + s.kr2 / 2);
}

// the 128-row shapes of the wide kernels: returns 1 when launched
template <typename TL, typename T1>
int launch_gram_wide(int n, const SrcT<TL, T1> &s, const GramPlan &p, double *work, bool staged_ok,
                     hipStream_t st)
{
    constexpr int SRW = 16;
    const int kr = s.kr1 + s.kr2;
    const int kl_pad = p.TM * 16, kr_pad = p.TN * 16;
    if (!staged_ok || p.TM != 8 || p.chunk % SRW != 0 || (p.TN != 8 && p.TN != 9))
        return 0;
    const size_t shm = 2 * (size_t)SRW * (lds_row_stride(s.kl) + lds_row_stride(kr)) * sizeof(double);
    if (p.TN == 9) {
        if (stage_vectors(s, SRW) > 12 * 64 * 3)
            return 0;
        cacg_allow_lds(gram_lds_kernel<TL, T1, 12, 2, 3, SRW, 2, 3>, shm);
        hipLaunchKernelGGL((gram_lds_kernel<TL, T1, 12, 2, 3, SRW, 2, 3>), dim3(p.gx, 1),
                           dim3(12 * 64), shm, st, n, s, work, kl_pad, kr_pad, p.chunk, 3, 12);
    } else {
        if (stage_vectors(s, SRW) > 16 * 64 * 2)
            return 0;
        cacg_allow_lds(gram_lds_kernel<TL, T1, 16, 2, 2, SRW, 2, 2>, shm);
        hipLaunchKernelGGL((gram_lds_kernel<TL, T1, 16, 2, 2, SRW, 2, 2>), dim3(p.gx, 1),
                           dim3(16 * 64), shm, st, n, s, work, kl_pad, kr_pad, p.chunk, 4, 16);
    }
    return 1;
}

template <typename TL, typename T1, int MT, int NT>
int launch_gram_mfma(int n, const SrcT<TL, T1> &s, const GramPlan &p, double *G1, double *G2,
                     double *work, hipStream_t st)
{
    constexpr int SR = 8, LDG = 3, VPT = 5;
    constexpr int EL = 16 / sizeof(TL), E1 = 16 / sizeof(T1);
    const int kl_pad = p.TM * 16, kr_pad = p.TN * 16;
    const int kr = s.kr1 + s.kr2;
    auto al16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    // 16-byte vectors: column counts and row strides in whole vectors, aligned rows
    const bool aligned = p.KS == 1 && s.kl % EL == 0 && s.kr1 % E1 == 0 && s.kr2 % 2 == 0 &&
                         s.ldl % EL == 0 && s.ld1 % E1 == 0 && s.bs1 % E1 == 0 &&
                         (s.kr2 == 0 || s.ld2 % 2 == 0) &&
                         al16(s.L) && al16(s.R1) && (s.kr2 == 0 || al16(s.R2));
    const bool staged = aligned && stage_vectors(s, SR) <= 256 * VPT && p.chunk % SR == 0;
    if (launch_gram_wide(n, s, p, work, aligned, st)) {
        // launched
    } else if (s.bs1 && !staged) {
        return (int)hipErrorInvalidValue;  // a block-major R1 needs the staged kernels
    } else if (staged) {
        const size_t shm =
            2 * (size_t)SR * (lds_row_stride(s.kl) + lds_row_stride(kr)) * sizeof(double);
        cacg_allow_lds(gram_lds_kernel<TL, T1, 4, MT, NT, SR, LDG, VPT>, shm);
        hipLaunchKernelGGL((gram_lds_kernel<TL, T1, 4, MT, NT, SR, LDG, VPT>), dim3(p.gx, p.gy),
                           dim3(256), shm, st, n, s, work, kl_pad, kr_pad, p.chunk, p.GN, p.NG);
    } else {
        const size_t shm =
            p.KS > 1 ? (size_t)(p.KS - 1) * p.WGG * MT * NT * 4 * 64 * sizeof(double) : 0;
        cacg_allow_lds(gram_mfma_kernel<TL, T1, MT, NT>, shm);
        hipLaunchKernelGGL((gram_mfma_kernel<TL, T1, MT, NT>), dim3(p.gx, p.gy), dim3(256), shm,
                           st, n, s, work, kl_pad, kr_pad, p.chunk, p.GN, p.NG, p.WGG, p.KS);
    }
    CACG_LAUNCH_CHECK();
    const long outs = (long)s.kl * (s.kr1 + s.kr2);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(cacg_grid(outs, 64, 0)), dim3(256), 0, st, work,
                       p.gx, s.kl, s.kr1, s.kr2, kl_pad, kr_pad, G1, G2);
    CACG_LAUNCH_CHECK();
    return 0;
}

template <typename TL, typename T1>
int gram_fast(int n, const TL *L, long ldl, int kl, const T1 *R1, long ld1, int kr1,
              const double *R2, long ld2, int kr2, double *G1, double *G2, double *d_work,
              hipStream_t st, long bs1 = 0)
{
    if (n < 0 || kl <= 0 || kr1 < 0 || kr2 < 0 || kr1 + kr2 <= 0)
        return (int)hipErrorInvalidValue;
    const GramPlan p = plan_gram(n, kl, kr1 + kr2);
    const SrcT<TL, T1> s = {L, ldl, kl, R1, ld1, kr1, R2, ld2, kr2, bs1};
    // MT in {1, 2}, NT in 1..9
#define CACG_GRAM_NT(MTV)                                                                         \
    switch (p.NT) {                                                                               \
    case 1: return launch_gram_mfma<TL, T1, MTV, 1>(n, s, p, G1, G2, d_work, st);                 \
    case 2: return launch_gram_mfma<TL, T1, MTV, 2>(n, s, p, G1, G2, d_work, st);                 \
    case 3: return launch_gram_mfma<TL, T1, MTV, 3>(n, s, p, G1, G2, d_work, st);                 \
    case 4: return launch_gram_mfma<TL, T1, MTV, 4>(n, s, p, G1, G2, d_work, st);                 \
    case 5: return launch_gram_mfma<TL, T1, MTV, 5>(n, s, p, G1, G2, d_work, st);                 \
    case 6: return launch_gram_mfma<TL, T1, MTV, 6>(n, s, p, G1, G2, d_work, st);                 \
    case 7: return launch_gram_mfma<TL, T1, MTV, 7>(n, s, p, G1, G2, d_work, st);                 \
    case 8: return launch_gram_mfma<TL, T1, MTV, 8>(n, s, p, G1, G2, d_work, st);                 \
    default: return launch_gram_mfma<TL, T1, MTV, 9>(n, s, p, G1, G2, d_work, st);                \
    }
    if (p.MT == 2) {
        CACG_GRAM_NT(2)
    } else {
        CACG_GRAM_NT(1)
    }
#undef CACG_GRAM_NT
}

}  // namespace

extern "C" int cacg_k_gram_ref(int n, const double *L, long ldl, int kl, const double *R1, long ld1,
                               int kr1, const double *R2, long ld2, int kr2, double *G1,
                               double *G2, hipStream_t st)
{
    if (n < 0 || kl <= 0 || kr1 < 0 || kr2 < 0 || kr1 + kr2 <= 0)
        return (int)hipErrorInvalidValue;
    if (kr2 > 0 && !R2)
        return (int)hipErrorInvalidValue;
    if (n == 0) {
        // empty sums: zero blocks
        int rc = kr1 > 0 ? cacg_k_fill_f64(G1, (long)kl * kr1, 0.0, st) : 0;
        if (!rc && kr2 > 0)
            rc = cacg_k_fill_f64(G2, (long)kl * kr2, 0.0, st);
        return rc;
    }
    const int kr = kr1 + kr2;
    const int nab = (kl + 7) / 8, nbb = (kr + 7) / 8;
    const GrefSrc s = {L, ldl, kl, R1, ld1, kr1, R2, ld2, kr2};
    hipLaunchKernelGGL(gram_ref_kernel, dim3((unsigned)(nab * nbb)), dim3(256), 0, st, n, s, nbb,
                       G1, G2);
    CACG_LAUNCH_CHECK();
    return 0;
}

extern "C" long cacg_gram_workspace(int kl, int kr)
{
    // upper bound over n: gx <= 512
    return 512L * ((kl + 15) / 16 * 16) * ((kr + 15) / 16 * 16);
}

extern "C" long cacg_gram_workspace_n(int n, int kl, int kr)
{
    if (n <= 0 || kl <= 0 || kr <= 0)
        return 0;
    const GramPlan p = plan_gram(n, kl, kr);
    return (long)p.gx * (p.TM * 16) * (p.TN * 16);
}

extern "C" int cacg_k_gram_fast(int n, const double *L, long ldl, int kl, const double *R1,
                                long ld1, int kr1, const double *R2, long ld2, int kr2, double *G1,
                                double *G2, double *d_work, hipStream_t st)
{
    return gram_fast<double, double>(n, L, ldl, kl, R1, ld1, kr1, R2, ld2, kr2, G1, G2, d_work,
                                     st);
}

extern "C" int cacg_gram_blk_supported(int n, int kl, int kr1, int bw)
{
    if (n <= 0 || kl <= 0 || bw <= 0 || kr1 <= 0 || kr1 % bw != 0 || bw % 2 != 0 || kl % 2 != 0)
        return 0;
    // the conditions under which launch_gram_mfma stages the operands through
    // LDS (16-byte aligned L, R1 and ldl = kl assumed)
    const GramPlan p = plan_gram(n, kl, kr1);
    const SrcT<double, double> s = {nullptr, kl, kl, nullptr, bw, kr1, nullptr, 1, 0, (long)n * bw};
    if (p.KS != 1)
        return 0;
    const bool wide = p.TM == 8 && p.chunk % 16 == 0 && (p.TN == 8 || p.TN == 9) &&
                      stage_vectors(s, 16) <= (p.TN == 9 ? 12 * 64 * 3 : 16 * 64 * 2);
    const bool staged = stage_vectors(s, 8) <= 256 * 5 && p.chunk % 8 == 0;
    return wide || staged;
}

extern "C" int cacg_k_gram_fast_blk(int n, const double *L, long ldl, int kl, const double *R1,
                                    int bw, long bs, int kr1, double *G1, double *d_work,
                                    hipStream_t st)
{
    if (bw <= 0 || bs < (long)n * bw || kr1 % bw != 0)
        return (int)hipErrorInvalidValue;
    return gram_fast<double, double>(n, L, ldl, kl, R1, bw, kr1, nullptr, 1, 0, G1, nullptr,
                                     d_work, st, bs);
}

extern "C" int cacg_k_gram_fast_f64f32(int n, const double *L, long ldl, int kl, const float *R1,
                                       long ld1, int kr1, const double *R2, long ld2, int kr2,
                                       double *G1, double *G2, double *d_work, hipStream_t st)
{
    return gram_fast<double, float>(n, L, ldl, kl, R1, ld1, kr1, R2, ld2, kr2, G1, G2, d_work,
                                    st);
}
