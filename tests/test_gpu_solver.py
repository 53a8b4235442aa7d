"""Solver-level parity of the gfx950 path.

Reference order: bit for bit against the compiled reference's golden
fixtures (iteration count, ||R(:,c)|| after every outer iteration, final X),
through both the extended API and the drop-in entry points.

Fast order (split-K MFMA): the accuracy targets where the reference itself
meets them (conftest.check_fast_order): iterations within +-1, the
per-iteration residual within 1e-9 relative until the reference's own 4-rank
run departs from its 1-rank run, within 2 x that departure after it, and X
within max(1e-10, 2 x the reference's 4-rank X difference)."""

import ctypes

import numpy as np
import pytest

from conftest import check_fast_order, golden_case, golden_manifest
from cacg_amd import problems as P

pytestmark = pytest.mark.gpu
CASES = sorted(k for k, v in golden_manifest().items() if v["solver"] != "bcg")


def _run(gpu_lib, spec, A, B, X0, order):
    return gpu_lib.bcbcg(A, B, X0, spec["s"], stop_mode=0, eps=spec["eps"],
                         maxiter=spec["maxiter"], order=order)


@pytest.mark.parametrize("name", CASES)
def test_reference_order_bitwise_vs_golden(gpu_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = _run(gpu_lib, spec, A, B, X0, "reference")
    assert out["status"] == "ok"
    assert out["iters"] == spec["iters_p1"]
    assert np.array_equal(out["res_hist"][:, 0], np.sqrt(g["r2_p1"]))
    assert np.array_equal(out["res_hist"], g["rnorm_cols"])
    assert np.array_equal(out["X"], g["X"])


class CsrType(ctypes.Structure):
    _fields_ = [("num_rows", ctypes.c_long), ("num_cols", ctypes.c_long), ("nnz", ctypes.c_long),
                ("start", ctypes.c_long), ("row_start", ctypes.c_void_p),
                ("col_idx", ctypes.c_void_p), ("csrdata", ctypes.c_void_p)]


class DenseType(ctypes.Structure):
    _fields_ = [("local_num_row", ctypes.c_long), ("local_num_col", ctypes.c_long),
                ("global_num_row", ctypes.c_long), ("global_num_col", ctypes.c_long),
                ("start_idx", ctypes.c_long), ("data", ctypes.c_void_p)]


@pytest.mark.parametrize("name", CASES)
def test_dropin_entry_points_bitwise(gpu_lib, name, monkeypatch, capfd):
    monkeypatch.setenv("CACG_QUIET", "1")
    spec, A, B, X0, g = golden_case(name)
    L = gpu_lib.lib()
    fn = L.cbcg_v1 if spec["solver"] == "cbcg" else L.bcbcg_v1
    fn.argtypes = [CsrType, DenseType, DenseType, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                   ctypes.c_int]
    fn.restype = None
    rp, ci = np.ascontiguousarray(A.rowptr), np.ascontiguousarray(A.col)
    va = np.ascontiguousarray(A.val)
    Bc, X = np.ascontiguousarray(B), np.ascontiguousarray(X0).copy()
    n, m = B.shape
    mat = CsrType(n, n, A.nnz, 0, rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    dB = DenseType(n, m, n, m, 0, Bc.ctypes.data)
    dX = DenseType(n, m, n, m, 0, X.ctypes.data)
    fn(mat, dB, dX, spec["s"], spec["eps"], 0, 1)
    assert L.cacg_last_status() == 0
    assert L.cacg_last_iterations() == spec["iters_p1"]
    assert np.array_equal(X, g["X"])
    assert np.array_equal(Bc, B)  # B is read-only


def test_dropin_prints_reference_loop_lines(gpu_lib, capfd):
    spec, A, B, X0, g = golden_case("cbcg_2d16_s1")
    L = gpu_lib.lib()
    L.cbcg_v1.argtypes = [CsrType, DenseType, DenseType, ctypes.c_long, ctypes.c_double,
                          ctypes.c_int, ctypes.c_int]
    rp, ci, va = (np.ascontiguousarray(a) for a in (A.rowptr, A.col, A.val))
    X = X0.copy()
    L.cbcg_v1(CsrType(A.n, A.n, A.nnz, 0, rp.ctypes.data, ci.ctypes.data, va.ctypes.data),
              DenseType(A.n, 1, A.n, 1, 0, np.ascontiguousarray(B).ctypes.data),
              DenseType(A.n, 1, A.n, 1, 0, X.ctypes.data), spec["s"], spec["eps"], 0, 1)
    lines = [l for l in capfd.readouterr().out.splitlines() if l.startswith("Loop:")]
    k = spec["iters_p1"]
    want = [i for i in range(k) if i % 10 == 0 or i == k - 1]
    assert [int(l.split(",")[0].split()[1]) for l in lines] == want
    r = np.sqrt(g["r2_p1"])
    assert lines[-1] == "Loop: %d, R[1]_norm=%30.30f" % (k - 1, r[k - 1])


@pytest.mark.parametrize("name", CASES)
def test_fast_order_against_reference(gpu_lib, name):
    spec, A, B, X0, g = golden_case(name)
    out = _run(gpu_lib, spec, A, B, X0, "fast")
    assert out["status"] == "ok"
    check_fast_order(name, out["iters"], out["res_hist"][:, 0], out["X"], spec, g)


# ---- the benchmark configurations at sizes the oracle finishes in seconds ------------

def _config_case(which):
    if which == "C1":        # 2D 128^2, 1 RHS U[0,1) seed 0, s = 4, 1e-8 relres
        A = P.laplacian_2d(128)
        return A, P.rhs_uniform(A.n, 1, 0), 4, 1e-8, 1000
    if which == "C2_16cube":  # C2's shape (3D 7-point, 16 RHS U[0,1) seed 1, s = 8) on 16^3
        A = P.laplacian_3d(16)
        return A, P.rhs_uniform(A.n, 16, 1), 8, 1e-8, 100
    if which == "C3_spd12":  # C3's random SPD variant (seed 2), N(0,1) RHS, on 12^3
        A = P.random_spd_like(P.laplacian_3d(12), seed=2)
        return A, P.rhs_normal(A.n, 16, 2), 8, 1e-10, 100
    raise ValueError(which)


@pytest.mark.parametrize("which", ["C1", "C2_16cube", "C3_spd12"])
def test_configs_reference_bitwise_and_fast_bars(gpu_lib, oracle_lib, which):
    A, B, s, tol, maxiter = _config_case(which)
    X0 = np.zeros_like(B)
    orc = oracle_lib.bcbcg(A, B, X0, s, stop_mode=1, tol=tol, maxiter=maxiter)
    orc4 = oracle_lib.bcbcg(A, B, X0, s, stop_mode=1, tol=tol, maxiter=maxiter, nranks=4)
    assert orc["status"] == "ok"
    ref = gpu_lib.bcbcg(A, B, X0, s, stop_mode=1, tol=tol, maxiter=maxiter, order="reference")
    assert ref["iters"] == orc["iters"]
    assert np.array_equal(ref["res_hist"], orc["res_hist"])
    assert np.array_equal(ref["X"], orc["X"])
    fast = gpu_lib.bcbcg(A, B, X0, s, stop_mode=1, tol=tol, maxiter=maxiter, order="fast")
    assert fast["status"] == "ok"
    assert abs(fast["iters"] - orc["iters"]) <= 1
    # the bars of conftest.check_fast_order, every column, against the oracle
    # at 1 rank and its 4-rank emulation (another summation order)
    x4 = np.linalg.norm(orc4["X"] - orc["X"]) / np.linalg.norm(orc["X"])
    xrel = np.linalg.norm(fast["X"] - orc["X"]) / np.linalg.norm(orc["X"])
    k = min(fast["iters"], orc["iters"], orc4["iters"])
    r1, r4, rf = orc["res_hist"][:k], orc4["res_hist"][:k], fast["res_hist"][:k]
    dev, dev4 = np.abs(rf - r1) / r1, np.abs(r4 - r1) / r1
    over = np.nonzero(dev4.max(axis=1) > 1e-10)[0]
    pre = int(over[0]) if len(over) else k
    print(f"{which}: iterations {fast['iters']} (oracle {orc['iters']}); residual deviation max "
          f"{dev.max():.2e} (4-rank {dev4.max():.2e}), before iteration {pre} "
          f"{dev[:pre].max() if pre else 0.0:.2e}; X {xrel:.2e} (4-rank {x4:.2e})", flush=True)
    assert np.all(dev[:pre] <= 1e-9)
    assert dev.max() <= max(1e-9, 2 * dev4.max())
    assert xrel <= max(1e-10, 2 * x4), (xrel, x4)
    # converged means every column's true residual is at the tolerance too
    D = A.to_dense() if A.n <= 5000 else None
    if D is not None:
        tr = np.linalg.norm(B - D @ fast["X"], axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(tr < 10 * tol)


def test_long_rows_solver_reference_bitwise_and_fast(gpu_lib, oracle_lib):
    """A power-law matrix whose five hub rows (277..989 entries) go through
    the row-split SpMM inside the solver: reference order the same bits as
    the oracle.  On this matrix (cond 1.3e4, Gerschgorin interval [0.02,
    1991] around a spectrum in [0.08, 1001], s = 4) the reference's own
    iteration count moves with its summation order, 37 at one rank, 51 at two,
    60 at four (the oracle's rank emulation), and its recursive residual parts
    from the true one: ||B - A X|| / ||B|| of its solutions is 1e-6, 2e-7 and
    4e-6.  The fast order must stop within that iteration range, with a true
    residual within 2 x the reference's worst."""
    A = P.power_law_spd(40, 8, 1500, 5)
    assert len(gpu_lib.long_rows(A)[1]) == 5
    B = P.rhs_uniform(A.n, 4, 3)
    X0 = np.zeros_like(B)
    eps = 1e-10 * float(np.sqrt(oracle_lib.col_norm2sq(B, 0)))
    orc = oracle_lib.bcbcg(A, B, X0, 4, eps=eps, maxiter=200)
    assert orc["status"] == "ok"
    ref = gpu_lib.bcbcg(A, B, X0, 4, stop_mode=0, eps=eps, maxiter=200, order="reference")
    assert ref["iters"] == orc["iters"]
    assert np.array_equal(ref["res_hist"], orc["res_hist"])
    assert np.array_equal(ref["X"], orc["X"])
    D = A.to_dense()

    def relres(X):
        return float((np.linalg.norm(B - D @ X, axis=0) / np.linalg.norm(B, axis=0)).max())

    runs = [orc] + [oracle_lib.bcbcg(A, B, X0, 4, eps=eps, maxiter=200, nranks=r) for r in (2, 4)]
    spread = [o["iters"] for o in runs]
    worst = max(relres(o["X"]) for o in runs)
    fast = gpu_lib.bcbcg(A, B, X0, 4, stop_mode=0, eps=eps, maxiter=200, order="fast")
    tr = relres(fast["X"])
    print(f"power-law: fast order {fast['iters']} iterations, reference at 1/2/4 ranks {spread}; "
          f"true relres {tr:.2e}, the reference's worst {worst:.2e}", flush=True)
    assert fast["status"] == "ok" and min(spread) - 1 <= fast["iters"] <= max(spread) + 1
    assert tr <= 2 * worst


# ---- statuses and edge cases ------------------------------------------------------

def test_maxiter_status(gpu_lib):
    A = P.laplacian_2d(32)
    B = P.rhs_uniform(A.n, 4, 1)
    out = gpu_lib.bcbcg(A, B, np.zeros_like(B), 2, tol=1e-12, maxiter=2, order="fast")
    assert out["status"] == "maxiter" and out["iters"] == 2


@pytest.mark.parametrize("order", ["reference", "fast"])
def test_breakdown_on_duplicate_columns(gpu_lib, order):
    A = P.laplacian_2d(8)
    b = P.rhs_uniform(A.n, 1, 1)
    B = np.concatenate([b, b], axis=1)
    out = gpu_lib.bcbcg(A, B, np.zeros_like(B), 2, stop_mode=0, eps=0.0, maxiter=5, order=order)
    assert out["status"] == "breakdown"


@pytest.mark.parametrize("order", ["reference", "fast"])
def test_one_by_one_system(gpu_lib, order):
    A = P.laplacian_2d(1)           # [4]
    B = np.array([[2.0]])
    out = gpu_lib.bcbcg(A, B, np.zeros_like(B), 1, stop_mode=0, eps=1e-300, maxiter=3,
                        order=order)
    assert out["status"] == "ok" and out["iters"] == 1
    assert np.array_equal(out["X"], B / 4.0)


def test_solver_reuse_is_deterministic(gpu_lib):
    A = P.laplacian_3d(20)
    B = P.rhs_uniform(A.n, 16, 1)
    sv = gpu_lib.Solver(A, 16, 8, tol=1e-8, maxiter=100, order="fast")
    dB = gpu_lib.DeviceBuffer.from_array(B)
    xs = []
    for _ in range(2):
        dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
        rep = sv.solve(dB, dX)
        assert rep["status"] == "ok"
        xs.append(dX.download(B.shape))
    assert np.array_equal(xs[0], xs[1])
    sv.close()


@pytest.mark.parametrize("grid,m,s", [(20, 16, 8), (24, 8, 4), (16, 4, 3), (48, 16, 8)])
def test_block_major_basis_gives_the_same_bits(gpu_lib, monkeypatch, grid, m, s):
    # the fast order keeps its basis block-major from iteration 1 on; with
    # CACG_BASIS_ROWS=1 (read at solver creation) it stays row-major: every
    # iterate, the residual history and X are the same bits
    A = P.laplacian_3d(grid)
    B = P.rhs_uniform(A.n, m, 2)
    out = {}
    for rows in ("0", "1"):
        monkeypatch.setenv("CACG_BASIS_ROWS", rows)
        print(f"basis rows={rows} grid={grid} m={m} s={s}", flush=True)
        sv = gpu_lib.Solver(A, m, s, tol=1e-10, maxiter=200, order="fast")
        dB = gpu_lib.DeviceBuffer.from_array(B)
        dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
        rep = sv.solve(dB, dX)
        assert rep["status"] == "ok"
        out[rows] = (dX.download(B.shape), rep)
        sv.close()
    (xb, rb), (xr, rr) = out["0"], out["1"]
    assert rb["iters"] == rr["iters"] > 1
    assert np.array_equal(rb["res_hist"], rr["res_hist"])
    assert np.array_equal(xb, xr)


@pytest.mark.parametrize("m,s", [(32, 16), (64, 8)])
def test_wide_shapes_reference_bitwise_and_fast(gpu_lib, oracle_lib, m, s):
    # s m = 512: the X / R update's alpha (s m x m) no longer fits in LDS beside
    # the row tiles and is read from global memory; m = 64 takes the
    # non-MFMA update kernels in the fast order too
    A = P.laplacian_2d(32)
    B = P.rhs_uniform(A.n, m, 1)
    X0 = np.zeros_like(B)
    eps = 1e-8 * float(np.sqrt(oracle_lib.col_norm2sq(B, 0)))
    print(f"wide shape m={m} s={s} n={A.n}", flush=True)
    orc = oracle_lib.bcbcg(A, B, X0, s, eps=eps, maxiter=100)
    assert orc["status"] == "ok"
    ref = gpu_lib.bcbcg(A, B, X0, s, stop_mode=0, eps=eps, maxiter=100, order="reference")
    assert ref["status"] == "ok" and ref["iters"] == orc["iters"]
    assert np.array_equal(ref["res_hist"], orc["res_hist"])
    assert np.array_equal(ref["X"], orc["X"])
    fast = gpu_lib.bcbcg(A, B, X0, s, stop_mode=0, eps=eps, maxiter=100, order="fast")
    assert fast["status"] == "ok" and abs(fast["iters"] - orc["iters"]) <= 1
    D = A.to_dense()
    tr = np.linalg.norm(B[:, 0] - D @ fast["X"][:, 0]) / np.linalg.norm(B[:, 0])
    assert tr < 1e-7
