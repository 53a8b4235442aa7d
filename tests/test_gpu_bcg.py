"""Block CG (bcg_v1, src/bcgKernel.c:22-462) on the gfx950 path.

Reference order: bit for bit against the compiled reference's golden fixtures
(iteration count, ||R(:,c)|| read by every stop test, final X), through the
host API and the drop-in entry point.  Fast order (split-K MFMA Gram, MFMA
block products): iterations within +-1 and X within max(1e-10, 10 x the
reference's own 4-rank vs 1-rank difference), as for the s-step solvers."""

import ctypes

import numpy as np
import pytest

from conftest import check_fast_order, golden_case, golden_manifest
from cacg_amd import problems as P

pytestmark = pytest.mark.gpu
BCG = sorted(k for k, v in golden_manifest().items() if v["solver"] == "bcg")


@pytest.mark.parametrize("name", BCG)
def test_bcg_reference_order_bitwise_vs_golden(gpu_lib, name):
    spec, A, B, X0, g = golden_case(name)
    print(f"bcg reference order {name}", flush=True)
    out = gpu_lib.bcg(A, B, X0, stop_mode=0, eps=spec["eps"], maxiter=spec["maxiter"],
                      order="reference")
    assert out["status"] == "ok"
    assert out["iters"] == spec["iters_p1"]
    assert np.array_equal(out["res_hist"][:, 0], np.sqrt(g["r2_p1"]))
    assert np.array_equal(out["res_hist"], g["rnorm_cols"])
    assert np.array_equal(out["X"], g["X"])


@pytest.mark.parametrize("name", BCG)
def test_bcg_fast_order_against_reference(gpu_lib, name):
    spec, A, B, X0, g = golden_case(name)
    print(f"bcg fast order {name}", flush=True)
    out = gpu_lib.bcg(A, B, X0, stop_mode=0, eps=spec["eps"], maxiter=spec["maxiter"],
                      order="fast")
    assert out["status"] == "ok"
    check_fast_order(name, out["iters"], out["res_hist"][:, 0], out["X"], spec, g)


class CsrType(ctypes.Structure):
    _fields_ = [("num_rows", ctypes.c_long), ("num_cols", ctypes.c_long), ("nnz", ctypes.c_long),
                ("start", ctypes.c_long), ("row_start", ctypes.c_void_p),
                ("col_idx", ctypes.c_void_p), ("csrdata", ctypes.c_void_p)]


class DenseType(ctypes.Structure):
    _fields_ = [("local_num_row", ctypes.c_long), ("local_num_col", ctypes.c_long),
                ("global_num_row", ctypes.c_long), ("global_num_col", ctypes.c_long),
                ("start_idx", ctypes.c_long), ("data", ctypes.c_void_p)]


def _dropin(gpu_lib, A, B, X0, eps):
    L = gpu_lib.lib()
    L.bcg_v1.argtypes = [CsrType, DenseType, DenseType, ctypes.c_double, ctypes.c_int,
                         ctypes.c_int]
    L.bcg_v1.restype = None
    rp, ci, va = (np.ascontiguousarray(a) for a in (A.rowptr, A.col, A.val))
    Bc, X = np.ascontiguousarray(B), np.ascontiguousarray(X0).copy()
    n, m = B.shape
    L.bcg_v1(CsrType(n, n, A.nnz, 0, rp.ctypes.data, ci.ctypes.data, va.ctypes.data),
             DenseType(n, m, n, m, 0, Bc.ctypes.data), DenseType(n, m, n, m, 0, X.ctypes.data),
             eps, 0, 1)
    return L, X, Bc


@pytest.mark.parametrize("name", ["bcg_2d16_m4", "bcg_2d32_m1", "bcg_2d16_m4_x1"])
def test_bcg_dropin_bitwise(gpu_lib, name, monkeypatch):
    monkeypatch.setenv("CACG_QUIET", "1")
    spec, A, B, X0, g = golden_case(name)
    print(f"bcg_v1 drop-in {name}", flush=True)
    L, X, Bc = _dropin(gpu_lib, A, B, X0, spec["eps"])
    assert L.cacg_last_status() == 0
    assert L.cacg_last_iterations() == spec["iters_p1"]
    assert np.array_equal(X, g["X"])
    assert np.array_equal(Bc, B)


def test_bcg_dropin_prints_reference_loop_lines(gpu_lib, capfd):
    spec, A, B, X0, g = golden_case("bcg_2d32_m1")
    print("bcg_v1 loop lines", flush=True)
    capfd.readouterr()
    _dropin(gpu_lib, A, B, X0, spec["eps"])
    lines = [l for l in capfd.readouterr().out.splitlines() if l.startswith("id:0, Loop:")]
    k = spec["iters_p1"]
    want = [i for i in range(k) if i % 10 == 0 or i == k - 1]
    assert [int(l.split(",")[1].split()[1]) for l in lines] == want
    r = np.sqrt(g["r2_p1"])
    assert lines[-1] == "id:0, Loop: %d, R[1]_norm=%30.30f" % (k - 1, r[k - 1])


@pytest.mark.parametrize("order", ["reference", "fast"])
def test_bcg_larger_case_against_oracle(gpu_lib, oracle_lib, order):
    # 3D 16^3, 16 RHS: fast order takes the MFMA Gram and block-update kernels
    A = P.laplacian_3d(16)
    B = P.rhs_uniform(A.n, 16, 1)
    X0 = np.zeros_like(B)
    print(f"bcg 3d16 m16 {order}", flush=True)
    orc = oracle_lib.bcg(A, B, X0, stop_mode=1, tol=1e-8)
    out = gpu_lib.bcg(A, B, X0, stop_mode=1, tol=1e-8, order=order)
    assert out["status"] == "ok"
    if order == "reference":
        assert out["iters"] == orc["iters"]
        assert np.array_equal(out["res_hist"], orc["res_hist"])
        assert np.array_equal(out["X"], orc["X"])
    else:
        assert abs(out["iters"] - orc["iters"]) <= 1
        assert np.linalg.norm(out["X"] - orc["X"]) / np.linalg.norm(orc["X"]) < 1e-9
    D = A.to_dense()
    true_res = np.linalg.norm(B - D @ out["X"], axis=0) / np.linalg.norm(B, axis=0)
    assert np.all(true_res < 1e-8)


def test_bcg_device_solver_deterministic_and_phases(gpu_lib):
    A = P.laplacian_2d(64)
    B = P.rhs_uniform(A.n, 8, 2)
    print("bcg device solver", flush=True)
    sv = gpu_lib.Solver(A, 8, 4, tol=1e-10, maxiter=1000, order="fast")
    dB = gpu_lib.DeviceBuffer.from_array(B)
    outs = []
    for _ in range(2):
        dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
        r = sv.solve_bcg(dX=dX, dB=dB)
        outs.append((r, dX.download(B.shape)))
    sv.close()
    (r1, x1), (r2, x2) = outs
    assert r1["status"] == "ok" and r1["iters"] == r2["iters"]
    assert np.array_equal(x1, x2)
    assert r1["phase_calls"]["aq"] == r1["iters"]
    assert r1["phase_calls"]["gram_as"] == r1["iters"] - 1
    ref = gpu_lib.bcg(A, B, np.zeros_like(B), tol=1e-10, order="fast")
    assert ref["iters"] == r1["iters"] and np.array_equal(ref["X"], x1)


@pytest.mark.parametrize("order", ["reference", "fast"])
def test_bcg_breakdown_and_maxiter(gpu_lib, order):
    A = P.laplacian_2d(8)
    b = P.rhs_uniform(A.n, 1, 1)
    print(f"bcg breakdown {order}", flush=True)
    out = gpu_lib.bcg(A, np.concatenate([b, b], axis=1), np.zeros((A.n, 2)), stop_mode=0,
                      eps=0.0, maxiter=5, order=order)
    assert out["status"] == "breakdown"
    out = gpu_lib.bcg(A, b, np.zeros_like(b), stop_mode=0, eps=0.0, maxiter=3, order=order)
    assert out["status"] == "maxiter" and out["iters"] == 3
