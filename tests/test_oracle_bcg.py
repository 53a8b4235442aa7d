"""The CPU oracle's block CG (bcg_v1, src/bcgKernel.c:22-462), pinned: bit for
bit against the compiled reference's golden fixtures, and against what the
mathematics fixes (one column is textbook CG; block CG and the s = 1 block
s-step solver span the same spaces, so their iterates agree; n/m iterations
exhaust the Krylov space; a manufactured solution is recovered)."""

import numpy as np
import pytest

from conftest import golden_case, golden_manifest
from cacg_amd import problems as P

BCG = sorted(k for k, v in golden_manifest().items() if v["solver"] == "bcg")


@pytest.mark.parametrize("name", BCG)
def test_oracle_bcg_matches_reference_bitwise(oracle_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = oracle_lib.bcg(A, B, X0, eps=spec["eps"], maxiter=spec["maxiter"],
                         record_x="x_hist" in g)
    assert out["status"] == "ok"
    assert out["iters"] == spec["iters_p1"]
    assert np.array_equal(out["res_hist"][:, 0], np.sqrt(g["r2_p1"]))
    assert np.array_equal(out["res_hist"], g["rnorm_cols"])
    assert np.array_equal(out["X"], g["X"])
    if "x_hist" in g:
        assert np.array_equal(out["x_hist"][:2], g["x_hist"])


@pytest.mark.parametrize("name", ["bcg_2d16_m4", "bcg_2d32_m16_tol1e-10", "bcg_spd16_m4"])
def test_oracle_bcg_rank_emulation_tracks_reference_at_4_ranks(oracle_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = oracle_lib.bcg(A, B, X0, eps=spec["eps"], maxiter=spec["maxiter"], nranks=4)
    assert out["iters"] == spec["iters_p4"]
    rel = np.linalg.norm(out["X"] - g["X"]) / np.linalg.norm(g["X"])
    assert rel <= max(1e-12, 10 * float(g["X_p4_reldiff"]))


def _textbook_cg(D, b, iters):
    """Hestenes-Stiefel CG; returns the residual norm each iteration starts
    from and the iterates after each update."""
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rn, xs = [], []
    for _ in range(iters):
        rn.append(np.linalg.norm(r))
        Ap = D @ p
        a = (r @ r) / (p @ Ap)
        x = x + a * p
        xs.append(x.copy())
        r_new = r - a * Ap
        p = r_new + ((r_new @ r_new) / (r @ r)) * p
        r = r_new
    return np.array(rn), np.array(xs)


def test_one_column_is_textbook_cg(oracle_lib):
    A = P.laplacian_2d(16, shift=1.0)
    b = P.rhs_uniform(A.n, 1, 7)
    out = oracle_lib.bcg(A, b, np.zeros_like(b), eps=0.0, maxiter=20, record_x=True)
    rn, xs = _textbook_cg(A.to_dense(), b[:, 0], 20)
    assert np.allclose(out["res_hist"][:, 0], rn, rtol=1e-10, atol=0)
    assert np.allclose(out["x_hist"][:, :, 0], xs, rtol=1e-10, atol=1e-13)


def test_iterates_equal_block_sstep_with_s1(oracle_lib):
    # P_{k+1} = R_{k+1} + P_k beta spans the same block as the A-orthogonalised
    # S = R of the s = 1 block s-step solver, and both pick X_{k+1} by the same
    # Galerkin condition: the iterates coincide in exact arithmetic
    A = P.laplacian_2d(12, shift=0.5)
    B = P.rhs_uniform(A.n, 4, 11)
    k = 8
    g = oracle_lib.bcg(A, B, np.zeros_like(B), eps=0.0, maxiter=k, record_x=True)
    s = oracle_lib.bcbcg(A, B, np.zeros_like(B), 1, eps=0.0, maxiter=k, record_x=True)
    for i in range(k):
        rel = np.linalg.norm(g["x_hist"][i] - s["x_hist"][i]) / np.linalg.norm(s["x_hist"][i])
        assert rel < 1e-10, (i, rel)


def test_block_krylov_space_exhausted_after_n_over_m_iterations(oracle_lib):
    # n = 16, m = 4: after 4 iterations the block Krylov space is R^16, so the
    # residual that iteration 4 starts from is at the rounding floor
    A = P.laplacian_2d(4)
    B = P.rhs_uniform(A.n, 4, 8)
    out = oracle_lib.bcg(A, B, np.zeros_like(B), eps=0.0, maxiter=5)
    assert np.all(out["res_hist"][4] < 1e-10 * np.linalg.norm(B, axis=0))
    X = np.linalg.solve(A.to_dense(), B)
    assert np.linalg.norm(out["X"] - X) / np.linalg.norm(X) < 1e-10


def test_manufactured_solution(oracle_lib):
    A = P.laplacian_2d(24)
    Xs = P.rhs_uniform(A.n, 4, 9)
    B = A.to_dense() @ Xs
    out = oracle_lib.bcg(A, B, np.zeros_like(B), stop_mode=1, tol=1e-10)
    assert out["status"] == "ok"
    true_res = np.linalg.norm(B - A.to_dense() @ out["X"], axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(true_res < 1e-9)
    assert np.linalg.norm(out["X"] - Xs) / np.linalg.norm(Xs) < 1e-7


def test_stop_test_reads_the_residual_the_iteration_started_from(oracle_lib):
    # eps just above the first norm j >= 1 below every earlier one: the stop
    # test of iteration j fires, after X has been updated j + 1 times
    A = P.laplacian_2d(8)
    B = P.rhs_uniform(A.n, 2, 5)
    full = oracle_lib.bcg(A, B, np.zeros_like(B), eps=0.0, maxiter=10, record_x=True)
    r = full["res_hist"][:, 0]
    j = next(i for i in range(1, 10) if r[i] < r[:i].min())
    out = oracle_lib.bcg(A, B, np.zeros_like(B), eps=0.5 * (r[:j].min() + r[j]), maxiter=10)
    assert out["status"] == "ok" and out["iters"] == j + 1
    assert np.array_equal(out["X"], full["x_hist"][j])


def test_breakdown_on_duplicate_columns(oracle_lib):
    A = P.laplacian_2d(8)
    b = P.rhs_uniform(A.n, 1, 1)
    out = oracle_lib.bcg(A, np.concatenate([b, b], axis=1), np.zeros((A.n, 2)), eps=0.0,
                         maxiter=5)
    assert out["status"] == "breakdown"
