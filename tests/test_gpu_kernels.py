"""Kernel-level parity of the gfx950 HIP kernels with the CPU oracle.

Reference-order kernels must equal the oracle bit for bit.  Fast-order (MFMA)
kernels are checked exactly on integer-valued data, where every partial sum
is an integer below 2^53 so any summation order gives the exact result (a
wrong lane layout, index or transposed operand cannot pass), and within a
rounding bound on random data:
    |G - L^T R| <= 2 n eps (|L|^T |R|)   (eps = 2^-53)."""

import ctypes

import numpy as np
import pytest

from cacg_amd import problems as P

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def _arrow(L=10, w=-0.1):
    """SPD matrix with one full row/column (ragged rows: 1 row of n entries)."""
    A = P.laplacian_2d(L)
    D = A.to_dense()
    D[0, 1:] += w
    D[1:, 0] += w
    np.fill_diagonal(D, 0.0)
    np.fill_diagonal(D, 1.01 * np.abs(D).sum(axis=1))
    rows, cols = np.nonzero(D)
    rp = np.zeros(A.n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return P.CSR(A.n, np.cumsum(rp), cols.astype(np.int64), D[rows, cols])


MATRICES = {
    "lap2d_16": lambda: P.laplacian_2d(16),
    "lap3d_7": lambda: P.laplacian_3d(7),
    "spd_13": lambda: P.random_spd_like(P.laplacian_2d(13), seed=2),
    "arrow_10": _arrow,
    # a 1600-entry row: the first row tile holds more nonzeros than the LDS
    # stage of any launch shape, so its rows read (col, val) from global memory,
    # and the row is long enough for the row-split kernel
    "arrow_40": lambda: _arrow(40),
    # power-law row lengths: five hub rows of 277..989 entries go to the
    # row-split kernel, the rest stay vector-per-row
    "powerlaw_40": lambda: P.power_law_spd(40, 8, 1500, 5),
}


def test_long_row_rule(gpu_lib):
    """Rows at or above max(256, 16 x mean length) are split; the 3D Laplacian
    has none, the arrow its full row, the power-law matrix its five hubs."""
    lm, ids = gpu_lib.long_rows(MATRICES["lap3d_7"]())
    assert lm == 256 and len(ids) == 0
    lm, ids = gpu_lib.long_rows(MATRICES["arrow_40"]())
    assert lm == 256 and list(ids) == [0]
    A = MATRICES["powerlaw_40"]()
    lm, ids = gpu_lib.long_rows(A)
    d = np.diff(A.rowptr)
    assert list(ids) == list(np.nonzero(d >= lm)[0]) and len(ids) == 5
    assert list(gpu_lib.long_rows(A, 300)[1]) == list(np.nonzero(d >= 300)[0])
    assert len(gpu_lib.long_rows(A, 0)[1]) == 0


@pytest.mark.parametrize("mat", ["arrow_40", "powerlaw_40", "lap3d_7"])
@pytest.mark.parametrize("long_min", [1, 6, 64])
@pytest.mark.parametrize("k", [1, 3, 16, 17, 64, 128, 144])
def test_spmm_row_split_bitwise(gpu_lib, oracle_lib, mat, long_min, k):
    """Row split forced down to short rows (long_min 1: every nonempty row),
    every epilogue: the same bits as the oracle.  k = 128, 144 take column
    blocks of 64 (and 16); the segments of a row are 4096 / min(k, 64)
    nonzeros, so the 1600-entry row crosses several."""
    A = MATRICES[mat]()
    X = P.rhs_normal(A.n, k, 31)
    AX = oracle_lib.spmm(A, X)
    assert np.array_equal(gpu_lib.k_spmm(A, X, long_min=long_min), AX)
    p1, p2 = P.rhs_normal(A.n, k, 32), P.rhs_normal(A.n, k, 33)
    a, b = 0.37, -0.21
    got = gpu_lib.k_spmm(A, X, "cheb", a, b, p1=p1, p2=p2, long_min=long_min)
    assert np.array_equal(got, (a * AX + b * p1) - p2)
    assert np.array_equal(gpu_lib.k_spmm(A, X, "resid", p1=p1, long_min=long_min), p1 - AX)
    S1, S0 = gpu_lib.k_spmm(A, X, "cheb1", a, b, want_copy0=True, long_min=long_min)
    assert np.array_equal(S0, X) and np.array_equal(S1, a * AX + b * X)


@pytest.mark.parametrize("long_min", [1, 64])
def test_spmm_f32_row_split_same_bits(gpu_lib, long_min):
    """FP32 forms: row split against vector-per-row, bit for bit (the same
    FP32 products and sums in the same order)."""
    A = MATRICES["powerlaw_40"]()
    for k in (4, 16, 33):
        X = P.rhs_normal(A.n, k, 41).astype(np.float32)
        p1, p2 = (P.rhs_normal(A.n, k, sd).astype(np.float32) for sd in (42, 43))
        for epi in ("none", "cheb"):
            kw = dict(p1=p1, p2=p2) if epi == "cheb" else {}
            want = gpu_lib.k_spmm_f32(A, X, epi, 0.5, -0.25, long_min=0, **kw)
            got = gpu_lib.k_spmm_f32(A, X, epi, 0.5, -0.25, long_min=long_min, **kw)
            assert np.array_equal(got, want), (k, epi)
        R = P.rhs_normal(A.n, k, 44)
        w1, w0 = gpu_lib.k_cheb1_f32(A, R, 0.5, -0.25, long_min=0)
        g1, g0 = gpu_lib.k_cheb1_f32(A, R, 0.5, -0.25, long_min=long_min)
        assert np.array_equal(g0, w0) and np.array_equal(g1, w1), k


@pytest.mark.parametrize("mat", sorted(MATRICES))
def test_gerschgorin_bitwise(gpu_lib, oracle_lib, mat):
    A = MATRICES[mat]()
    assert np.array_equal(gpu_lib.k_gerschgorin(A), oracle_lib.gerschgorin(A))


@pytest.mark.parametrize("mat", sorted(MATRICES))
@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 16, 17, 33, 64, 128, 144])
def test_spmm_bitwise(gpu_lib, oracle_lib, mat, k):
    A = MATRICES[mat]()
    X = P.rhs_normal(A.n, k, 11)
    assert np.array_equal(gpu_lib.k_spmm(A, X), oracle_lib.spmm(A, X))


@pytest.mark.parametrize("k,off,ld", [(16, 16, 128), (16, 3, 40), (12, 4, 20), (20, 0, 20),
                                      (256, 0, 256), (4, 2, 8), (128, 0, 144), (6, 1, 9)])
@pytest.mark.parametrize("epi", ["none", "cheb"])
@pytest.mark.parametrize("long_min", [0, 5])
def test_spmm_strided_bitwise(gpu_lib, oracle_lib, k, off, ld, epi, long_min):
    """x, p1, p2 and y as column blocks (columns off .. off+k) of wider
    row-major n x ld arrays, the layout of the basis blocks; off and ld select
    the 16-byte gather path (4 | k, 16-byte aligned x, even ld) or the scalar
    one.  long_min 5: the rows of 5 or more entries through the row split."""
    A = P.laplacian_3d(9)
    n = A.n
    Xw, P1w, P2w = (P.rhs_normal(n, ld, sd) for sd in (21, 22, 23))
    a, b = 0.37, -0.21
    dA = gpu_lib.DeviceCSR(A, long_min)
    dX, d1, d2 = (gpu_lib.DeviceBuffer.from_array(M) for M in (Xw, P1w, P2w))
    Yw = np.full((n, ld), 7.0)
    dY = gpu_lib.DeviceBuffer.from_array(Yw)
    epi_id = gpu_lib.EPI[epi]
    rc = gpu_lib.lib().cacg_k_spmm(ctypes.byref(dA.s), k, dX.ptr + 8 * off, ld, dY.ptr + 8 * off,
                                   ld, epi_id, a, b, d1.ptr + 8 * off, ld, d2.ptr + 8 * off, ld,
                                   None, ld, None)
    assert rc == 0
    Y = dY.download((n, ld))
    AX = oracle_lib.spmm(A, np.ascontiguousarray(Xw[:, off:off + k]))
    want = AX if epi == "none" else (a * AX + b * P1w[:, off:off + k]) - P2w[:, off:off + k]
    assert np.array_equal(Y[:, off:off + k], want)
    # columns outside the block are untouched
    assert np.all(Y[:, :off] == 7.0) and np.all(Y[:, off + k:] == 7.0)


@pytest.mark.parametrize("m", [1, 4, 16])
def test_spmm_epilogues_bitwise(gpu_lib, oracle_lib, m):
    A = P.random_spd_like(P.laplacian_2d(12), seed=3)
    R = P.rhs_uniform(A.n, m, 12)
    lo, hi = oracle_lib.gerschgorin(A)
    sa, sb = 2.0 / (hi - lo), -(hi + lo) / (hi - lo)
    S = oracle_lib.cheb_basis(A, R, 3, sa, sb)
    S1, S0 = gpu_lib.k_spmm(A, R, "cheb1", sa, sb, want_copy0=True)
    assert np.array_equal(S0, S[:, :m])
    assert np.array_equal(S1, S[:, m:2 * m])
    S2 = gpu_lib.k_spmm(A, S1, "cheb", 2 * sa, 2 * sb, p1=S1, p2=S0)
    assert np.array_equal(S2, S[:, 2 * m:3 * m])
    X = P.rhs_normal(A.n, m, 13)
    resid = gpu_lib.k_spmm(A, X, "resid", p1=R)
    assert np.array_equal(resid, R - oracle_lib.spmm(A, X))


GRAM_SHAPES = [(1000, 4, 4, 1), (1000, 9, 9, 3), (3000, 16, 16, 4), (2048, 64, 64, 16),
               (5000, 128, 128, 16), (777, 128, 128, 0), (1500, 256, 256, 32), (0, 16, 16, 4)]
# fast-order only: 128-row shapes of the wide-workgroup kernels with many row
# chunks, a 120-row G (padded tiles), and a chunk length that is not a
# multiple of their 16-row stage (general kernel instead)
GRAM_WIDE = [(70000, 128, 128, 16), (131072, 120, 112, 16), (131072, 128, 128, 0),
             (100003, 128, 128, 16)]


@pytest.mark.parametrize("n,kl,kr1,kr2", GRAM_SHAPES)
def test_gram_reference_bitwise(gpu_lib, oracle_lib, n, kl, kr1, kr2):
    rng = np.random.default_rng(n + kl)
    L, R1, R2 = rng.standard_normal((n, kl)), rng.standard_normal((n, kr1)), rng.standard_normal((n, kr2))
    out = gpu_lib.k_gram(L, R1, R2 if kr2 else None, mode="reference")
    G1, G2 = out if kr2 else (out, None)
    assert np.array_equal(G1, oracle_lib.gram(L, R1))
    if kr2:
        assert np.array_equal(G2, oracle_lib.gram(L, R2))


@pytest.mark.parametrize("n,kl,kr1,kr2", GRAM_SHAPES + GRAM_WIDE)
def test_gram_fast_exact_on_integers(gpu_lib, n, kl, kr1, kr2):
    rng = np.random.default_rng(n + kl + 1)
    L = rng.integers(-8, 9, size=(n, kl)).astype(np.float64)
    R1 = rng.integers(-8, 9, size=(n, kr1)).astype(np.float64)
    R2 = rng.integers(-8, 9, size=(n, kr2)).astype(np.float64)
    out = gpu_lib.k_gram(L, R1, R2 if kr2 else None, mode="fast")
    G1, G2 = out if kr2 else (out, None)
    # every partial sum is an integer below 2^53, so the float64 product is exact
    # in any summation order (BLAS included)
    assert np.array_equal(G1, L.T @ R1)
    if kr2:
        assert np.array_equal(G2, L.T @ R2)


@pytest.mark.parametrize("n,kl,kr1,kr2", GRAM_SHAPES[:6])
def test_gram_fast_rounding_bound(gpu_lib, n, kl, kr1, kr2):
    rng = np.random.default_rng(n + kl + 2)
    L, R1 = rng.standard_normal((n, kl)), rng.standard_normal((n, kr1))
    G = gpu_lib.k_gram(L, R1, None, mode="fast")
    exact = (L.astype(np.longdouble).T @ R1.astype(np.longdouble)).astype(np.float64)
    bound = 2 * max(n, 1) * EPS * (np.abs(L).T @ np.abs(R1)) + 1e-300
    assert np.all(np.abs(G - exact) <= bound)


def _spd(N, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((N, N))
    return M @ M.T + N * np.eye(N)


@pytest.mark.parametrize("N,k", [(1, 1), (4, 1), (4, 4), (9, 3), (16, 16), (64, 16), (128, 16),
                                 (128, 128), (160, 32), (256, 32)])
def test_lu_solve(gpu_lib, oracle_lib, N, k):
    A = _spd(N, N + k)
    rhs = np.random.default_rng(k).standard_normal((N, k))
    ref = oracle_lib.lu_solve(A, rhs)
    x, flag = gpu_lib.k_lu_solve(A, rhs, mode="reference")
    assert flag == 0
    assert np.array_equal(x, ref)
    xf, flag = gpu_lib.k_lu_solve(A, rhs, mode="fast")
    assert flag == 0
    assert np.allclose(xf, ref, rtol=1e-12, atol=1e-14)
    # a second solve reusing the stored factor gives the same bits as a fresh one
    rhs2 = np.random.default_rng(k + 1).standard_normal((N, max(1, k // 2)))
    x1, x2, flag = gpu_lib.k_lu_solve(A, rhs, mode="reference", rhs2=rhs2)
    assert flag == 0 and np.array_equal(x1, ref)
    assert np.array_equal(x2, oracle_lib.lu_solve(A, rhs2))


@pytest.mark.parametrize("N,k", [(1, 1), (4, 4), (9, 3), (16, 16), (64, 16), (100, 7), (128, 1),
                                 (128, 16), (128, 17), (128, 32), (128, 33), (128, 64),
                                 (128, 65), (128, 128), (128, 300), (128, 384)])
def test_lu_solve_fast_order_bitwise(gpu_lib, N, k):
    # the fast order (wave-local substitutions with the factor in LDS) against
    # its numpy restatement, bit for bit, fresh and with the kept factor
    from conftest import lu_restated
    A = _spd(N, 3 * N + k)
    rhs = np.random.default_rng(N + k).standard_normal((N, k))
    print(f"lu fast N={N} k={k}", flush=True)
    x, flag = gpu_lib.k_lu_solve(A, rhs, mode="fast")
    assert flag == 0
    assert np.array_equal(x, lu_restated(A, rhs, "fast"))
    rhs2 = np.random.default_rng(N + k + 1).standard_normal((N, max(1, k // 3)))
    x1, x2, flag = gpu_lib.k_lu_solve(A, rhs, mode="fast", rhs2=rhs2)
    assert flag == 0 and np.array_equal(x1, x)
    assert np.array_equal(x2, lu_restated(A, rhs2, "fast"))


def test_lu_zero_pivot_sets_flag(gpu_lib):
    A = _spd(8, 1)
    A[0, 0] = 0.0
    _, flag = gpu_lib.k_lu_solve(A, np.ones((8, 2)), mode="reference")
    assert flag == 1


@pytest.mark.parametrize("n,sm", [(100, 4), (333, 9), (1000, 16), (999, 48), (2000, 128), (50, 256),
                                  (1003, 128), (37, 16), (5001, 64), (777, 32), (70000, 128),
                                  (1003, 8), (4099, 2)])
def test_qupdate(gpu_lib, oracle_lib, n, sm):
    rng = np.random.default_rng(sm)
    S, Q, beta = rng.standard_normal((n, sm)), rng.standard_normal((n, sm)), rng.standard_normal((sm, sm))
    ref = S - oracle_lib.gemm_ts(Q, beta)
    assert np.array_equal(gpu_lib.k_qupdate(S, Q, beta, mode="reference"), ref)
    assert np.allclose(gpu_lib.k_qupdate(S, Q, beta, mode="fast"), ref, rtol=0, atol=1e-12 * sm)
    Si = rng.integers(-8, 9, size=(n, sm)).astype(np.float64)
    Qi = rng.integers(-8, 9, size=(n, sm)).astype(np.float64)
    bi = rng.integers(-8, 9, size=(sm, sm)).astype(np.float64)
    exact = (Si.astype(np.int64) - Qi.astype(np.int64) @ bi.astype(np.int64)).astype(np.float64)
    assert np.array_equal(gpu_lib.k_qupdate(Si, Qi, bi, mode="fast"), exact)


@pytest.mark.parametrize("n,sm,m", [(100, 4, 1), (333, 9, 3), (1000, 16, 4), (1001, 64, 16),
                                    (2003, 128, 16), (700, 128, 32), (300, 96, 24),
                                    (70001, 128, 16), (37, 16, 2), (999, 48, 16), (503, 256, 32)])
def test_xrupdate(gpu_lib, oracle_lib, n, sm, m):
    rng = np.random.default_rng(sm + m)
    Q, AQ = rng.standard_normal((n, sm)), rng.standard_normal((n, sm))
    alpha, X, R = rng.standard_normal((sm, m)), rng.standard_normal((n, m)), rng.standard_normal((n, m))
    Xr = X + oracle_lib.gemm_ts(Q, alpha)
    Rr = R - oracle_lib.gemm_ts(AQ, alpha)
    nr = np.array([oracle_lib.col_norm2sq(Rr, c) for c in range(m)])
    Xg, Rg, ng = gpu_lib.k_xrupdate(Q, AQ, alpha, X, R, mode="reference")
    assert np.array_equal(Xg, Xr) and np.array_equal(Rg, Rr) and np.array_equal(ng, nr)
    Xf, Rf, nf = gpu_lib.k_xrupdate(Q, AQ, alpha, X, R, mode="fast")
    assert np.allclose(Xf, Xr, rtol=0, atol=1e-12 * sm)
    assert np.allclose(Rf, Rr, rtol=0, atol=1e-12 * sm)
    assert np.allclose(nf, nr, rtol=1e-12)
    # integers: exact in any order
    Qi, AQi = (rng.integers(-4, 5, size=(n, sm)).astype(np.float64) for _ in range(2))
    ai = rng.integers(-4, 5, size=(sm, m)).astype(np.float64)
    Xi, Ri = (rng.integers(-4, 5, size=(n, m)).astype(np.float64) for _ in range(2))
    Xe = Xi + (Qi.astype(np.int64) @ ai.astype(np.int64))
    Re = Ri - (AQi.astype(np.int64) @ ai.astype(np.int64))
    Xf, Rf, nf = gpu_lib.k_xrupdate(Qi, AQi, ai, Xi, Ri, mode="fast")
    assert np.array_equal(Xf, Xe) and np.array_equal(Rf, Re)
    assert np.array_equal(nf, (Re.astype(np.int64) ** 2).sum(axis=0).astype(np.float64))


# ---- deferred X update (fast order): the same bits as the undeferred kernels ------

@pytest.mark.parametrize("sm,m", [(128, 16), (64, 8), (96, 32), (48, 20), (16, 1)])
def test_deferred_x_update_kernels_match_undeferred_bitwise(gpu_lib, sm, m):
    rng = np.random.default_rng(sm * 100 + m)
    n = 5003                                       # ragged last tile
    S, Q, AQ = (rng.standard_normal((n, sm)) for _ in range(3))
    beta, alpha = rng.standard_normal((sm, sm)), rng.standard_normal((sm, m))
    X, R = rng.standard_normal((n, m)), rng.standard_normal((n, m))
    print(f"deferred x update sm={sm} m={m}", flush=True)
    assert gpu_lib.lib().cacg_k_split_supported(sm, m) == 1
    Xw, Rw, nw = gpu_lib.k_xrupdate(Q, AQ, alpha, X, R, mode="fast")
    Qw = gpu_lib.k_qupdate(S, Q, beta, mode="fast")
    Qd, Xd = gpu_lib.k_qxupdate(S, Q, beta, alpha, X)
    assert np.array_equal(Qd, Qw) and np.array_equal(Xd, Xw)
    assert np.array_equal(gpu_lib.k_xrupdate_split(Q, AQ, alpha, X, R, 1), Xw)
    Rd, nd = gpu_lib.k_xrupdate_split(Q, AQ, alpha, X, R, 2)
    assert np.array_equal(Rd, Rw) and np.array_equal(nd, nw)
    # against plain float64 products
    assert np.allclose(Xd, X + Q @ alpha, rtol=1e-12, atol=1e-11)
    assert np.allclose(Qd, S - Q @ beta, rtol=1e-12, atol=1e-11)
    # integers: exact in any order
    Si, Qi = (rng.integers(-8, 9, size=(n, sm)).astype(np.float64) for _ in range(2))
    bi = rng.integers(-8, 9, size=(sm, sm)).astype(np.float64)
    ai = rng.integers(-8, 9, size=(sm, m)).astype(np.float64)
    Xi = rng.integers(-8, 9, size=(n, m)).astype(np.float64)
    Qe, Xe = gpu_lib.k_qxupdate(Si, Qi, bi, ai, Xi)
    I64 = np.int64
    assert np.array_equal(Qe, (Si.astype(I64) - Qi.astype(I64) @ bi.astype(I64)).astype(np.float64))
    assert np.array_equal(Xe, (Xi.astype(I64) + Qi.astype(I64) @ ai.astype(I64)).astype(np.float64))


@pytest.mark.parametrize("n,sm,m,gap", [(5003, 128, 16, 0), (4099, 64, 8, 2), (3001, 96, 32, 0),
                                        (2000, 48, 4, 6), (70001, 128, 16, 0)])
def test_block_major_basis_kernels_same_bits(gpu_lib, n, sm, m, gap):
    # the basis stored as its blocks (the fast order from iteration 1 on): the
    # Q^T A S Gram and the deferred Q update give the bits of the row-major forms
    rng = np.random.default_rng(n + sm + m)
    S, Q, AQ = (rng.standard_normal((n, sm)) for _ in range(3))
    beta, alpha = rng.standard_normal((sm, sm)), rng.standard_normal((sm, m))
    X = rng.standard_normal((n, m))
    print(f"block-major basis n={n} sm={sm} m={m}", flush=True)
    Qr, Xr = gpu_lib.k_qxupdate(S, Q, beta, alpha, X)
    Qb, Xb = gpu_lib.k_qxupdate_blk(S, Q, beta, alpha, X, gap)
    assert np.array_equal(Qb, Qr) and np.array_equal(Xb, Xr)
    G = gpu_lib.k_gram(AQ, S, mode="fast")
    assert np.array_equal(gpu_lib.k_gram_blk(AQ, S, m, gap), G)


def test_deferred_forms_refuse_unsupported_widths(gpu_lib):
    L = gpu_lib.lib()
    assert L.cacg_k_split_supported(40, 4) == 0 and L.cacg_k_split_supported(128, 33) == 0
    assert L.cacg_k_qxupdate(10, 40, 4, None, None, None, None, None, None) != 0


@pytest.mark.parametrize("n,m,ld", [(0, 3, 3), (1, 1, 1), (31, 2, 2), (100, 1, 1), (1000, 3, 5),
                                    (5000, 16, 16), (777, 64, 64), (3000, 65, 65),
                                    (20000, 130, 136), (70000, 16, 16), (8193, 1, 1)])
def test_colnorm2_reference_bitwise(gpu_lib, oracle_lib, n, m, ld):
    """Reference-order column norms (one row-order chain per column, columns 64
    per workgroup, rows staged through LDS): the oracle's sums bit for bit, for
    partial stages, several column blocks and a row stride wider than m."""
    R = P.rhs_normal(max(n, 1), m, n + m)[:n]
    got = gpu_lib.k_colnorm2_ref(R, ld)
    want = np.array([oracle_lib.col_norm2sq(R, c) if n else 0.0 for c in range(m)])
    assert np.array_equal(got, want)
