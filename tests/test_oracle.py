"""The CPU oracle, pinned: against the compiled reference's golden fixtures (bit
for bit) and, kernel by kernel, against closed forms, exact integer
arithmetic and textbook algorithms chosen so a dropped term, wrong sign or
index, or transposed operand fails."""

import numpy as np
import pytest

from conftest import golden_case, golden_manifest
from cacg_amd import problems as P

CASES = sorted(k for k, v in golden_manifest().items() if v["solver"] != "bcg")


# ---- golden fixtures: the oracle reproduces the reference bit for bit -------------

@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(oracle_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = oracle_lib.bcbcg(A, B, X0, spec["s"], eps=spec["eps"], maxiter=spec["maxiter"],
                           record_x="x_hist" in g)
    assert out["iters"] == spec["iters_p1"]
    assert np.array_equal(out["res_hist"][:, 0], np.sqrt(g["r2_p1"]))
    assert np.array_equal(out["res_hist"], g["rnorm_cols"])
    assert np.array_equal(out["X"], g["X"])
    if "x_hist" in g:
        assert np.array_equal(out["x_hist"][:2], g["x_hist"])


@pytest.mark.parametrize("name", ["bcbcg_2d16_m4_s4", "cbcg_2d32_s4", "bcbcg_3d12_m16_s8"])
def test_oracle_rank_emulation_tracks_reference_at_4_ranks(oracle_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = oracle_lib.bcbcg(A, B, X0, spec["s"], eps=spec["eps"], maxiter=spec["maxiter"], nranks=4)
    assert out["iters"] == spec["iters_p4"]
    # not bitwise: MPICH's reduction order is its own; same envelope as 4 vs 1 rank
    rel = np.linalg.norm(out["X"] - g["X"]) / np.linalg.norm(g["X"])
    assert rel <= max(1e-12, 10 * float(g["X_p4_reldiff"]))


# ---- kernels against closed forms and exact arithmetic ----------------------------

def test_gerschgorin_closed_forms(oracle_lib):
    # interior rows of the Laplacians: d - sum|off| = 0, d + sum|off| = 8 (2D) / 12 (3D)
    assert np.array_equal(oracle_lib.gerschgorin(P.laplacian_2d(16)), [0.0, 8.0])
    assert np.array_equal(oracle_lib.gerschgorin(P.laplacian_3d(6)), [0.0, 12.0])
    assert np.array_equal(oracle_lib.gerschgorin(P.laplacian_2d(8, shift=0.5)), [0.5, 8.5])
    # 1x1 grid: the only row is diagonal, the max stays above DBL_MIN's seed
    assert np.array_equal(oracle_lib.gerschgorin(P.laplacian_2d(1)), [4.0, 4.0])


def test_gerschgorin_brute_force(oracle_lib):
    A = P.random_spd_like(P.laplacian_2d(7), seed=5)
    D = A.to_dense()
    off = np.abs(D).sum(axis=1) - np.abs(np.diag(D))
    lo, hi = oracle_lib.gerschgorin(A)
    assert lo == pytest.approx(np.min(np.diag(D) - off), rel=1e-15)
    assert hi == pytest.approx(np.max(np.diag(D) + off), rel=1e-15)


def _int_csr(n, seed):
    rng = np.random.default_rng(seed)
    A = P.laplacian_2d(int(np.sqrt(n)))
    val = rng.integers(-4, 5, size=A.nnz).astype(np.float64)
    return P.CSR(A.n, A.rowptr, A.col, val)


def test_spmm_exact_integers(oracle_lib):
    A = _int_csr(64, 1)
    X = np.random.default_rng(2).integers(-9, 10, size=(A.n, 5)).astype(np.float64)
    ref = A.to_dense().astype(np.int64) @ X.astype(np.int64)
    assert np.array_equal(oracle_lib.spmm(A, X), ref.astype(np.float64))


def _eigvec_2d(L, i, j):
    x = np.arange(1, L + 1)
    v = np.outer(np.sin(j * np.pi * x / (L + 1)), np.sin(i * np.pi * x / (L + 1))).ravel()
    lam = 4 - 2 * np.cos(i * np.pi / (L + 1)) - 2 * np.cos(j * np.pi / (L + 1))
    return v, lam


def test_cheb_basis_eigenvector_kat(oracle_lib):
    # R = [v1 | v2], eigenvectors of the 2D Laplacian: S_j = T_j(sa*lam + sb) v, T_j = cos(j acos z)
    L, s = 12, 6
    A = P.laplacian_2d(L)
    (v1, l1), (v2, l2) = _eigvec_2d(L, 1, 2), _eigvec_2d(L, 5, 3)
    R = np.stack([v1, v2], axis=1)
    sa, sb = 0.25, -1.0  # Gerschgorin [0, 8]
    S = oracle_lib.cheb_basis(A, R, s, sa, sb)
    for j in range(s):
        for c, (v, lam) in enumerate([(v1, l1), (v2, l2)]):
            expect = np.cos(j * np.arccos(sa * lam + sb)) * v
            assert np.allclose(S[:, j * 2 + c], expect, rtol=0, atol=1e-12), (j, c)


def test_gram_and_gemm_exact_integers(oracle_lib):
    rng = np.random.default_rng(3)
    Lm = rng.integers(-3, 4, size=(200, 7)).astype(np.float64)
    Rm = rng.integers(-3, 4, size=(200, 5)).astype(np.float64)
    assert np.array_equal(oracle_lib.gram(Lm, Rm), (Lm.T.astype(np.int64) @ Rm.astype(np.int64)))
    W = rng.integers(-3, 4, size=(7, 4)).astype(np.float64)
    assert np.array_equal(oracle_lib.gemm_ts(Lm, W), Lm.astype(np.int64) @ W.astype(np.int64))


def test_lu_solve_against_lapack(oracle_lib):
    rng = np.random.default_rng(4)
    M = rng.standard_normal((40, 40))
    A = M @ M.T + 40 * np.eye(40)  # SPD, well conditioned: no pivoting needed
    rhs = rng.standard_normal((40, 6))
    x = oracle_lib.lu_solve(A, rhs)
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-12, atol=1e-13)
    # a transposed factor or a dropped term would not reproduce rhs
    assert np.allclose(A @ x, rhs, rtol=0, atol=1e-11)


def test_lu_restatement_is_the_reference_solve(oracle_lib):
    # the numpy restatement the fast-order kernel test uses: with the reference's
    # backward order it is the oracle's solve bit for bit (factor, forward pass);
    # with the fast order it is the same solve to rounding
    from conftest import lu_restated
    rng = np.random.default_rng(5)
    for N, k in [(1, 1), (5, 3), (17, 4), (40, 40)]:
        M = rng.standard_normal((N, N))
        A = M @ M.T + N * np.eye(N)
        rhs = rng.standard_normal((N, k))
        assert np.array_equal(lu_restated(A, rhs, "reference"), oracle_lib.lu_solve(A, rhs))
        xf = lu_restated(A, rhs, "fast")
        assert np.allclose(xf, np.linalg.solve(A, rhs), rtol=1e-11, atol=1e-13)
        assert not np.array_equal(xf, lu_restated(A, rhs, "reference")) or N < 5


def test_lu_solve_triangular_exact(oracle_lib):
    # A = L U with integer L (lower, diag 1 or 2) and unit upper U: Crout recovers
    # L and U exactly, and the substitutions solve an integer system exactly
    rng = np.random.default_rng(6)
    n = 8
    Lf = np.tril(rng.integers(-2, 3, size=(n, n))).astype(np.float64)
    np.fill_diagonal(Lf, rng.choice([1.0, 2.0], size=n))
    U = np.triu(rng.integers(-2, 3, size=(n, n)), 1).astype(np.float64) + np.eye(n)
    A = Lf @ U
    x_true = rng.integers(-5, 6, size=(n, 3)).astype(np.float64)
    x = oracle_lib.lu_solve(A, A @ x_true)
    assert np.allclose(x, x_true, rtol=0, atol=1e-9)


# ---- solver special cases -----------------------------------------------------------

def _textbook_cg(D, b, iters):
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    out = []
    for _ in range(iters):
        Ap = D @ p
        a = (r @ r) / (p @ Ap)
        x = x + a * p
        rn = r - a * Ap
        out.append(np.linalg.norm(rn))
        p = rn + ((rn @ rn) / (r @ r)) * p
        r = rn
    return np.array(out), x


def test_s1_single_rhs_is_textbook_cg(oracle_lib):
    A = P.laplacian_2d(16, shift=1.0)
    b = P.rhs_uniform(A.n, 1, 7)[:, 0]
    out = oracle_lib.cbcg(A, b, np.zeros(A.n), 1, eps=0.0, maxiter=20)
    ref, x = _textbook_cg(A.to_dense(), b, 20)
    assert np.allclose(out["res_hist"][:, 0], ref, rtol=1e-9, atol=0)
    assert np.allclose(out["X"][:, 0], x, rtol=1e-10, atol=1e-12)


def test_full_krylov_space_solves_in_one_iteration(oracle_lib):
    # s*m >= n: K_s(A, R) spans R^n, exact after one outer iteration
    A = P.laplacian_2d(4)
    B = P.rhs_uniform(A.n, 4, 8)
    out = oracle_lib.bcbcg(A, B, np.zeros_like(B), 4, eps=0.0, maxiter=1)
    X = np.linalg.solve(A.to_dense(), B)
    assert np.linalg.norm(out["X"] - X) / np.linalg.norm(X) < 1e-9


def test_manufactured_solution_and_true_residual(oracle_lib):
    A = P.laplacian_2d(24)
    Xs = P.rhs_uniform(A.n, 4, 9)
    B = A.to_dense() @ Xs
    out = oracle_lib.bcbcg(A, B, np.zeros_like(B), 4, stop_mode=1, tol=1e-10, maxiter=100)
    assert out["status"] == "ok"
    true_res = np.linalg.norm(B - A.to_dense() @ out["X"], axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(true_res < 1e-9)
    assert np.linalg.norm(out["X"] - Xs) / np.linalg.norm(Xs) < 1e-7


def test_breakdown_on_duplicate_columns(oracle_lib):
    # identical RHS columns make Q^T A Q singular: zero pivot -> non-finite -> breakdown
    A = P.laplacian_2d(8)
    b = P.rhs_uniform(A.n, 1, 1)
    B = np.concatenate([b, b], axis=1)
    out = oracle_lib.bcbcg(A, B, np.zeros_like(B), 2, eps=0.0, maxiter=5)
    assert out["status"] == "breakdown"


def test_fast_order_bars_reject_planted_deviations():
    # the bars of the fast-order parity tests (conftest.check_fast_order) pass
    # the reference's own one-rank data and reject a trajectory that departs
    # early, a residual one iteration off, or a perturbed X
    from conftest import check_fast_order, golden_case
    spec, A, B, X0, g = golden_case("bcbcg_2d32_m4_s4_tol1e-10")
    r1 = np.sqrt(g["r2_p1"])
    check_fast_order("as recorded", spec["iters_p1"], r1, g["X"], spec, g)
    early = r1.copy()
    early[3] *= 1 + 1e-8
    with pytest.raises(AssertionError):
        check_fast_order("early departure", spec["iters_p1"], early, g["X"], spec, g)
    with pytest.raises(AssertionError):
        check_fast_order("shifted", spec["iters_p1"], np.roll(r1, 1), g["X"], spec, g)
    with pytest.raises(AssertionError):
        check_fast_order("X off", spec["iters_p1"], r1, g["X"] * (1 + 1e-9), spec, g)
    with pytest.raises(AssertionError):
        check_fast_order("two iterations more", spec["iters_p1"] + 2, r1, g["X"], spec, g)


def _reference_mpi():
    import os
    import oracle
    return oracle.reference_available() and os.path.exists(oracle.MPIEXEC)


@pytest.mark.skipif(not _reference_mpi(), reason="reference build or MPI not available here")
def test_reference_timing_streamed_and_captured_agree():
    """time_reference(progress=True) passes the harness's per-iteration lines
    through as they come (a long CPU solve stays visibly alive) and parses the
    same stamps as the captured form: same iteration count, one stamp per
    iteration, and the solve time the harness reports."""
    import oracle
    A = P.laplacian_3d(10)
    B = P.rhs_uniform(A.n, 4, 1)
    eps = 1e-10 * float(np.linalg.norm(B[:, 0]))
    streamed = oracle.time_reference(A, B, np.zeros_like(B), 4, eps, "bcbcg", 2, timeout=120,
                                     progress=True)
    captured = oracle.time_reference(A, B, np.zeros_like(B), 4, eps, "bcbcg", 2, timeout=120)
    assert streamed["iters"] == captured["iters"] > 0
    assert len(streamed["t"]) == streamed["iters"]
    assert streamed["solve"] is not None and streamed["solve"] > 0.0
    assert all(b >= a for a, b in zip(streamed["t"], streamed["t"][1:]))
