"""The entry point that allocates nothing (cacg_bcbcg_device_noalloc) against
the device-resident solver it is a variant of (cacg_solver_create +
cacg_solver_solve_device): the same kernels on the same inputs, so the same
bits, iteration count and residual history, on two calls in a row, the second
on a smaller system (the scratch keeps the first call's contents)."""

import ctypes

import numpy as np
import pytest

from cacg_amd import problems as P

pytestmark = pytest.mark.gpu


def _bind(gpu_lib):
    L = gpu_lib.lib()
    f = L.cacg_bcbcg_device_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                  ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(gpu_lib.Opts),
                  ctypes.POINTER(gpu_lib.Report), ctypes.c_void_p]
    f.restype = ctypes.c_int
    L.cacg_noalloc_bytes.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_long]
    L.cacg_noalloc_bytes.restype = ctypes.c_long
    return L


def _noalloc_solve(gpu_lib, A, B, s, **kw):
    L = _bind(gpu_lib)
    m = B.shape[1]
    dr = gpu_lib.DeviceBuffer.from_array(np.ascontiguousarray(A.rowptr, dtype=np.int32))
    dc = gpu_lib.DeviceBuffer.from_array(np.ascontiguousarray(A.col, dtype=np.int32))
    dv = gpu_lib.DeviceBuffer.from_array(np.ascontiguousarray(A.val, dtype=np.float64))
    dB = gpu_lib.DeviceBuffer.from_array(B)
    dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
    o = gpu_lib.make_opts(**kw)
    hist = np.zeros((o.maxiter, m))
    rep = gpu_lib.Report()
    rep.res_hist = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rep.res_hist_cap = o.maxiter
    rc = L.cacg_bcbcg_device_noalloc(A.n, dr.ptr, dc.ptr, dv.ptr, m, s, dB.ptr, dX.ptr,
                                     ctypes.byref(o), ctypes.byref(rep), None)
    assert rc in (0, 1), f"cacg_bcbcg_device_noalloc returned {rc}"
    out = {"status": rc, "iters": int(rep.iters), "res_hist": hist[:rep.iters].copy(),
           "X": dX.download(B.shape)}
    return out


def _solver_solve(gpu_lib, A, B, s, **kw):
    sv = gpu_lib.Solver(A, B.shape[1], s, **kw)
    dB = gpu_lib.DeviceBuffer.from_array(B)
    dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
    rep = sv.solve(dB, dX)
    X = dX.download(B.shape)
    sv.close()
    return {"status": {"ok": 0, "maxiter": 1}[rep["status"]], "iters": rep["iters"],
            "res_hist": rep["res_hist"], "X": X}


# (problem, m, s): the second of each pair is smaller than the first
PAIRS = [
    (("lap3d", 24), 16, 8, ("lap2d", 40), 4, 4),
    (("lap2d", 96), 8, 4, ("lap2d", 20), 1, 3),
]


def _problem(kind, size):
    return P.laplacian_3d(size) if kind == "lap3d" else P.laplacian_2d(size)


@pytest.mark.parametrize("order", ["fast", "reference"])
@pytest.mark.parametrize("pair", range(len(PAIRS)))
def test_noalloc_matches_solver_two_calls(gpu_lib, order, pair):
    p1, m1, s1, p2, m2, s2 = PAIRS[pair]
    for (kind, size), m, s in ((p1, m1, s1), (p2, m2, s2)):
        A = _problem(kind, size)
        B = P.rhs_uniform(A.n, m, 1)
        print(f"noalloc {order}: {kind} {size} n={A.n} m={m} s={s}", flush=True)
        kw = dict(stop_mode=1, tol=1e-10, maxiter=300, order=order)
        want = _solver_solve(gpu_lib, A, B, s, **kw)
        got = _noalloc_solve(gpu_lib, A, B, s, **kw)
        assert want["status"] == 0
        assert got["status"] == want["status"]
        assert got["iters"] == want["iters"]
        assert np.array_equal(got["res_hist"], want["res_hist"])
        assert np.array_equal(got["X"], want["X"])


def test_noalloc_rejects_oversize_and_unsupported(gpu_lib):
    L = _bind(gpu_lib)
    cap = 128 << 20
    # the size check is on the host: no kernel runs for a rejected call
    assert L.cacg_noalloc_bytes(2_097_152, 16, 8) > cap
    assert 0 < L.cacg_noalloc_bytes(4096, 4, 4) <= cap
    A = P.laplacian_3d(128)   # n = 2,097,152 with 16 RHS and s = 8 does not fit
    o = gpu_lib.make_opts(tol=1e-8)
    rep = gpu_lib.Report()
    rc = L.cacg_bcbcg_device_noalloc(A.n, 1, 1, 1, 16, 8, 1, 1, ctypes.byref(o),
                                     ctypes.byref(rep), None)
    assert rc == 3  # CACG_EINVAL
    o = gpu_lib.make_opts(tol=1e-8, precision="fp32")
    rc = L.cacg_bcbcg_device_noalloc(1024, 1, 1, 1, 4, 4, 1, 1, ctypes.byref(o),
                                     ctypes.byref(rep), None)
    assert rc == 3
