"""MI355X (gfx950) s-step Chebyshev-basis CG: Python binding of libcacg.so.

The product is the C-ABI library cacg_amd/libcacg.so (include/cacg.h,
include/cacg_kernels.h): HIP kernels for gfx950 behind a C11 host driver.
This module only loads it with ctypes and moves numpy arrays across; there is
no Python or CPU compute path.  If the library is missing, every call raises.

    import cacg_amd, numpy as np
    from cacg_amd import problems
    A = problems.laplacian_2d(128)
    B = problems.rhs_uniform(A.n, 4, seed=0)
    out = cacg_amd.bcbcg(A, B, np.zeros_like(B), s=4, tol=1e-8)
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcacg.so")

STATUS = {0: "ok", 1: "maxiter", 2: "breakdown", 3: "einval", 4: "enomem", 5: "edevice",
          6: "eio"}
STOP_REF_COL0_ABS, STOP_REL_PER_COLUMN = 0, 1
ORDER_REFERENCE, ORDER_FAST = 0, 1
MODE = {"reference": 0, "fast": 1}
EPI = {"none": 0, "resid": 1, "cheb1": 2, "cheb": 3}
PHASES = ["setup", "basis", "gram", "solve", "qupdate", "aq", "xrupdate", "other", "basis1",
          "gram_as", "refine", "unused11"]

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class Opts(ctypes.Structure):
    _fields_ = [("stop_mode", ctypes.c_int), ("eps", ctypes.c_double), ("tol", ctypes.c_double),
                ("maxiter", ctypes.c_long), ("order", ctypes.c_int), ("precision", ctypes.c_int),
                ("print_period", ctypes.c_int), ("inner_tol", ctypes.c_double),
                ("inner_maxiter", ctypes.c_long), ("refine_restart", ctypes.c_int),
                ("bounds", ctypes.c_int), ("lanczos_steps", ctypes.c_int),
                ("bound_lo", ctypes.c_double), ("bound_hi", ctypes.c_double),
                ("fp32_basis_below", ctypes.c_double)]


BOUNDS = {"gerschgorin": 0, "lanczos": 1, "given": 2}


class Report(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_long),
                ("bounds", ctypes.c_double * 2), ("res_hist", _dp), ("res_hist_cap", ctypes.c_long),
                ("seconds", ctypes.c_double), ("phase_ms", ctypes.c_double * 12),
                ("phase_calls", ctypes.c_long * 12), ("refinements", ctypes.c_long),
                ("true_relres", ctypes.c_double)]


class DCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("rowptr", _vp), ("col", _vp), ("val", _vp),
                ("long_rows", _vp), ("n_long", ctypes.c_int), ("long_min", ctypes.c_int)]


# cacg_scsr: the same layout with FP32 values
SCsr = DCsr


_lib = None


def lib() -> ctypes.CDLL:
    """Load libcacg.so (built by cacg_amd.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libcacg.so not built: {LIB_PATH} (run python -m cacg_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
    sig = {
        "cacg_last_status": ([], c_int),
        "cacg_last_iterations": ([], c_long),
        "cacg_status_string": ([c_int], ctypes.c_char_p),
        "cacg_default_opts": ([ctypes.POINTER(Opts)], None),
        "cacg_bcbcg": ([c_long, _lp, _lp, _dp, c_long, c_long, _dp, _dp, ctypes.POINTER(Opts),
                        ctypes.POINTER(Report)], c_int),
        "cacg_solver_create": ([ctypes.POINTER(_vp), c_long, _lp, _lp, _dp, c_long, c_long,
                                ctypes.POINTER(Opts)], c_int),
        "cacg_solver_solve_device": ([_vp, _vp, _vp, ctypes.POINTER(Report)], c_int),
        "cacg_bcg": ([c_long, _lp, _lp, _dp, c_long, _dp, _dp, ctypes.POINTER(Opts),
                      ctypes.POINTER(Report)], c_int),
        "cacg_solver_solve_bcg_device": ([_vp, _vp, _vp, ctypes.POINTER(Report)], c_int),
        "cacg_solver_destroy": ([_vp], c_int),
        "cacg_solver_footprint": ([_vp], c_long),
        "cacg_device_count": ([], c_int),
        "cacg_dev_alloc": ([ctypes.POINTER(_vp), c_long], c_int),
        "cacg_dev_free": ([_vp], c_int),
        "cacg_h2d": ([_vp, _vp, c_long], c_int),
        "cacg_d2h": ([_vp, _vp, c_long], c_int),
        "cacg_d2d": ([_vp, _vp, c_long], c_int),
        "cacg_dev_sync": ([], c_int),
        "cacg_k_gerschgorin": ([ctypes.POINTER(DCsr), _vp, _vp, _vp], c_int),
        "cacg_k_spmm": ([ctypes.POINTER(DCsr), c_int, _vp, c_long, _vp, c_long, c_int, c_double,
                         c_double, _vp, c_long, _vp, c_long, _vp, c_long, _vp], c_int),
        "cacg_k_spmm_x": ([ctypes.POINTER(DCsr), c_int, _vp, c_long, _vp, c_long, c_int, c_double,
                           c_double, _vp, c_long, _vp, c_long, _vp, c_long, _vp, _vp], c_int),
        "cacg_k_gram_ref": ([c_int, _vp, c_long, c_int, _vp, c_long, c_int, _vp, c_long, c_int,
                             _vp, _vp, _vp], c_int),
        "cacg_gram_workspace": ([c_int, c_int], c_long),
        "cacg_k_gram_fast": ([c_int, _vp, c_long, c_int, _vp, c_long, c_int, _vp, c_long, c_int,
                              _vp, _vp, _vp, _vp], c_int),
        "cacg_k_lu_solve": ([c_int, _vp, c_int, _vp, _vp, _vp, _vp, c_int, c_int, _vp], c_int),
        "cacg_k_qupdate": ([c_int, c_int, _vp, _vp, _vp, c_int, _vp], c_int),
        "cacg_xrupdate_workspace": ([c_int], c_long),
        "cacg_k_xrupdate": ([c_int, c_int, c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, c_int, _vp],
                            c_int),
        "cacg_k_colnorm2_ref": ([c_int, c_int, _vp, c_long, _vp, _vp], c_int),
        "cacg_k_colnorm2_fast": ([c_int, c_int, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_spmm_f32": ([ctypes.POINTER(DCsr), c_int, _vp, c_long, _vp, c_long, c_int,
                             c_double, c_double, _vp, c_long, _vp, c_long, _vp], c_int),
        "cacg_k_cheb1_f32": ([ctypes.POINTER(DCsr), c_int, _vp, c_long, _vp, c_long, _vp, c_long,
                              c_double, c_double, _vp], c_int),
        "cacg_k_cvt_f64_f32": ([c_long, _vp, _vp, _vp], c_int),
        "cacg_csr_long_min": ([c_long, _vp], c_int),
        "cacg_csr_long_rows": ([c_long, _vp, c_int, _vp], c_long),
        "cacg_k_qupdate_s32": ([c_int, c_int, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_cvt_f32_f64": ([c_long, _vp, _vp, _vp], c_int),
        "cacg_k_qupdate_zero": ([c_int, c_int, _vp, _vp, _vp], c_int),
        "cacg_tsqr_workspace": ([c_int], c_long),
        "cacg_k_tsqr": ([c_int, c_int, _vp, c_long, _vp, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_gram_fast_f64f32": ([c_int, _vp, c_long, c_int, _vp, c_long, c_int, _vp, c_long,
                                     c_int, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_fill_i32": ([_vp, c_long, c_int, _vp], c_int),
        "cacg_k_fill_f64": ([_vp, c_long, c_double, _vp], c_int),
        "cacg_k_scale_f64": ([_vp, c_long, c_double, _vp], c_int),
        "cacg_k_split_supported": ([c_int, c_int], c_int),
        "cacg_k_qxupdate": ([c_int, c_int, c_int, _vp, _vp, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_qxupdate_blk": ([c_int, c_int, c_int, _vp, c_long, _vp, _vp, _vp, _vp, _vp], c_int),
        "cacg_k_gram_fast_blk": ([c_int, _vp, c_long, c_int, _vp, c_int, c_long, c_int, _vp, _vp,
                                  _vp], c_int),
        "cacg_k_xrupdate_split": ([c_int, c_int, c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, c_int,
                                   _vp], c_int),
        "cacg_mtx_read": ([ctypes.c_char_p, _lp, _lp, _lp, ctypes.POINTER(_lp),
                           ctypes.POINTER(_lp), ctypes.POINTER(_dp)], c_int),
        "cacg_mtx_write": ([ctypes.c_char_p, c_long, c_long, _lp, _lp, _dp], c_int),
        "cacg_csv_read": ([ctypes.c_char_p, c_long, c_long, ctypes.POINTER(_dp)], c_int),
        "cacg_csv_write": ([ctypes.c_char_p, c_long, c_long, _dp], c_int),
        "cacg_io_free": ([_vp], None),
        "cacg_lanczos_workspace": ([c_long], c_long),
        "cacg_k_lanczos": ([ctypes.POINTER(DCsr), c_int, _vp, _vp, _vp], c_int),
        "cacg_lanczos_interval": ([_dp, _dp, c_int, _dp, _dp], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


class DeviceBuffer:
    """A block of device memory (HBM) allocated through the library."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib().cacg_dev_alloc(ctypes.byref(p), max(self.nbytes, 8)), "cacg_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().cacg_h2d(self.ptr, a.ctypes.data, a.nbytes), "cacg_h2d")

    def download(self, shape, dtype=np.float64) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().cacg_d2h(out.ctypes.data, self.ptr, out.nbytes), "cacg_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().cacg_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def long_rows(A, long_min=None):
    """(long_min, ids): the rows the row-split SpMM takes, by the library's rule
    (cacg_csr_long_min) or from a given threshold (0: none)."""
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    L = lib()
    lm = L.cacg_csr_long_min(A.n, rp.ctypes.data) if long_min is None else int(long_min)
    cnt = L.cacg_csr_long_rows(A.n, rp.ctypes.data, lm, None)
    ids = np.zeros(max(cnt, 1), dtype=np.int32)
    L.cacg_csr_long_rows(A.n, rp.ctypes.data, lm, ids.ctypes.data)
    return lm, ids[:cnt]


class DeviceCSR:
    """The int32-index device copy of a problems.CSR (as the solver stores it),
    with its long-row list (long_min: None = the library's rule, 0 = none)."""

    def __init__(self, A, long_min=None):
        self.n = A.n
        self.rowptr = DeviceBuffer.from_array(np.asarray(A.rowptr, dtype=np.int32))
        self.col = DeviceBuffer.from_array(np.asarray(A.col, dtype=np.int32))
        self.val = DeviceBuffer.from_array(np.asarray(A.val, dtype=np.float64))
        lm, ids = long_rows(A, long_min)
        self.n_long = len(ids)
        self.long_min = lm if self.n_long else 0
        self.lrows = DeviceBuffer.from_array(ids) if self.n_long else None
        lp = self.lrows.ptr if self.lrows else None
        self.s = DCsr(A.n, self.rowptr.ptr, self.col.ptr, self.val.ptr, lp, self.n_long,
                      self.long_min)
        self._s32 = None

    @property
    def s32(self):
        """cacg_scsr view: the same indices with an FP32 copy of the values
        (converted on the device, as the solver does)."""
        if self._s32 is None:
            nnz = int(self.val.nbytes // 8)
            self.val32 = DeviceBuffer(4 * max(nnz, 1))
            _check(lib().cacg_k_cvt_f64_f32(nnz, self.val.ptr, self.val32.ptr, None), "cvt")
            self._s32 = SCsr(self.n, self.rowptr.ptr, self.col.ptr, self.val32.ptr, self.s.long_rows,
                             self.n_long, self.long_min)
        return self._s32


# ---- kernel-level wrappers (host arrays in, host arrays out; for tests) -------------


def k_gerschgorin(A) -> np.ndarray:
    dA = DeviceCSR(A)
    work, out = DeviceBuffer(8 * 2 * 1024), DeviceBuffer(16)
    _check(lib().cacg_k_gerschgorin(ctypes.byref(dA.s), work.ptr, out.ptr, None), "gerschgorin")
    return out.download(2)


class SpmmSched(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int), ("chunk_rows", ctypes.c_int), ("passes", ctypes.c_int)]


def k_spmm(A, X, epi="none", a=0.0, b=0.0, p1=None, p2=None, want_copy0=False, sched=None,
           long_min=None):
    """Y = epilogue(A X) for X n x k; returns Y (and copy0 for 'cheb1').
    sched = (grid, chunk_rows[, passes]) selects a launch schedule (cacg_k_spmm_x);
    long_min the row-split threshold (None: the library's rule, 0: no row split)."""
    dA = DeviceCSR(A, long_min)
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(A.n, -1)
    k = X.shape[1]
    dX = DeviceBuffer.from_array(X)
    dY = DeviceBuffer(X.nbytes)
    d1 = DeviceBuffer.from_array(np.ascontiguousarray(p1, dtype=np.float64)) if p1 is not None else None
    d2 = DeviceBuffer.from_array(np.ascontiguousarray(p2, dtype=np.float64)) if p2 is not None else None
    dC = DeviceBuffer(X.nbytes) if epi == "cheb1" else None
    sc = ctypes.byref(SpmmSched(*sched)) if sched is not None else None
    rc = lib().cacg_k_spmm_x(ctypes.byref(dA.s), k, dX.ptr, k, dY.ptr, k, EPI[epi], a, b,
                             d1.ptr if d1 else None, k, d2.ptr if d2 else None, k,
                             dC.ptr if dC else None, k, sc, None)
    _check(rc, "spmm")
    Y = dY.download((A.n, k))
    if want_copy0:
        return Y, dC.download((A.n, k))
    return Y


def k_gram(L, R1, R2=None, mode="reference"):
    L = np.ascontiguousarray(L, dtype=np.float64)
    R1 = np.ascontiguousarray(R1, dtype=np.float64)
    n, kl = L.shape
    kr1 = R1.shape[1]
    kr2 = 0 if R2 is None else R2.shape[1]
    dL, d1 = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R1)
    d2 = DeviceBuffer.from_array(np.ascontiguousarray(R2, dtype=np.float64)) if R2 is not None else None
    g1, g2 = DeviceBuffer(8 * kl * kr1), DeviceBuffer(8 * kl * max(kr2, 1))
    if mode == "reference":
        rc = lib().cacg_k_gram_ref(n, dL.ptr, kl, kl, d1.ptr, kr1, kr1, d2.ptr if d2 else None,
                                   max(kr2, 1), kr2, g1.ptr, g2.ptr, None)
    else:
        ws = DeviceBuffer(8 * lib().cacg_gram_workspace(kl, kr1 + kr2))
        rc = lib().cacg_k_gram_fast(n, dL.ptr, kl, kl, d1.ptr, kr1, kr1, d2.ptr if d2 else None,
                                    max(kr2, 1), kr2, g1.ptr, g2.ptr, ws.ptr, None)
    _check(rc, "gram")
    G1 = g1.download((kl, kr1))
    if R2 is None:
        return G1
    return G1, g2.download((kl, kr2))


def k_lu_solve(lhs, rhs, mode="reference", rhs2=None):
    """x = lhs^{-1} rhs; with rhs2, a second call reuses the factor left by the
    first and returns (x, x2, flag)."""
    lhs = np.ascontiguousarray(lhs, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    N, k = rhs.shape
    dl, dr, dx = DeviceBuffer.from_array(lhs), DeviceBuffer.from_array(rhs), DeviceBuffer(rhs.nbytes)
    work = DeviceBuffer(8 * N * N)
    flag = DeviceBuffer.from_array(np.zeros(1, dtype=np.int32))
    _check(lib().cacg_k_lu_solve(N, dl.ptr, k, dr.ptr, dx.ptr, work.ptr, flag.ptr, MODE[mode], 0,
                                 None), "lu_solve")
    x = dx.download((N, k))
    if rhs2 is None:
        return x, int(flag.download(1, np.int32)[0])
    rhs2 = np.ascontiguousarray(rhs2, dtype=np.float64)
    k2 = rhs2.shape[1]
    dr2, dx2 = DeviceBuffer.from_array(rhs2), DeviceBuffer(rhs2.nbytes)
    _check(lib().cacg_k_lu_solve(N, dl.ptr, k2, dr2.ptr, dx2.ptr, work.ptr, flag.ptr, MODE[mode], 1,
                                 None), "lu_solve")
    return x, dx2.download((N, k2)), int(flag.download(1, np.int32)[0])


def k_qupdate(S, Q, beta, mode="reference"):
    S = np.ascontiguousarray(S, dtype=np.float64)
    n, sm = S.shape
    dS, dQ, db = DeviceBuffer.from_array(S), DeviceBuffer.from_array(Q), DeviceBuffer.from_array(beta)
    _check(lib().cacg_k_qupdate(n, sm, dS.ptr, dQ.ptr, db.ptr, MODE[mode], None), "qupdate")
    return dQ.download((n, sm))


def k_xrupdate(Q, AQ, alpha, X, R, mode="reference"):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    n, sm = Q.shape
    m = alpha.shape[1]
    dQ, dAQ, da = DeviceBuffer.from_array(Q), DeviceBuffer.from_array(AQ), DeviceBuffer.from_array(alpha)
    dX, dR = DeviceBuffer.from_array(X), DeviceBuffer.from_array(R)
    dn = DeviceBuffer(8 * m)
    ws = DeviceBuffer(8 * lib().cacg_xrupdate_workspace(m))
    _check(lib().cacg_k_xrupdate(n, sm, m, dQ.ptr, dAQ.ptr, da.ptr, dX.ptr, dR.ptr, dn.ptr, ws.ptr,
                                 MODE[mode], None), "xrupdate")
    return dX.download((n, m)), dR.download((n, m)), dn.download(m)


def k_qxupdate(S, Q, beta, alpha, X):
    """Deferred fast form: X += Q alpha (old Q), then Q = S - Q beta; returns (Q, X)."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    n, sm = S.shape
    m = alpha.shape[1]
    dS, dQ, db = DeviceBuffer.from_array(S), DeviceBuffer.from_array(Q), DeviceBuffer.from_array(beta)
    da, dX = DeviceBuffer.from_array(alpha), DeviceBuffer.from_array(X)
    _check(lib().cacg_k_qxupdate(n, sm, m, dS.ptr, dQ.ptr, db.ptr, da.ptr, dX.ptr, None), "qxupdate")
    return dQ.download((n, sm)), dX.download((n, m))


def to_blocks(S, m, gap=0):
    """S (n x s m, row-major) as its blocks S_j (n x m each) one after the
    other, block j at j * (n m + gap) elements; returns the flat array."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    n, sm = S.shape
    bs = n * m + gap
    out = np.full(bs * (sm // m), np.nan)
    for j in range(sm // m):
        out[j * bs:j * bs + n * m] = S[:, j * m:(j + 1) * m].reshape(-1)
    return out, bs


def k_qxupdate_blk(S, Q, beta, alpha, X, gap=0):
    """k_qxupdate with S passed block-major (blocks of m columns, gap elements
    of padding between blocks); returns (Q, X)."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    n, sm = S.shape
    m = alpha.shape[1]
    Sb, bs = to_blocks(S, m, gap)
    dS, dQ, db = DeviceBuffer.from_array(Sb), DeviceBuffer.from_array(Q), DeviceBuffer.from_array(beta)
    da, dX = DeviceBuffer.from_array(alpha), DeviceBuffer.from_array(X)
    _check(lib().cacg_k_qxupdate_blk(n, sm, m, dS.ptr, bs, dQ.ptr, db.ptr, da.ptr, dX.ptr, None),
           "qxupdate_blk")
    return dQ.download((n, sm)), dX.download((n, m))


def k_gram_blk(L, R1, bw, gap=0):
    """Fast-order L^T R1 with R1 passed block-major (blocks of bw columns)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    n, kl = L.shape
    kr1 = R1.shape[1]
    Rb, bs = to_blocks(R1, bw, gap)
    dL, dR = DeviceBuffer.from_array(L), DeviceBuffer.from_array(Rb)
    g1 = DeviceBuffer(8 * kl * kr1)
    ws = DeviceBuffer(8 * lib().cacg_gram_workspace(kl, kr1))
    _check(lib().cacg_k_gram_fast_blk(n, dL.ptr, kl, kl, dR.ptr, bw, bs, kr1, g1.ptr, ws.ptr, None),
           "gram_fast_blk")
    return g1.download((kl, kr1))


def k_xrupdate_split(Q, AQ, alpha, X, R, which):
    """which 1: X += Q alpha, returns X; which 2: R -= AQ alpha, returns (R, norms^2)."""
    AQ = np.ascontiguousarray(AQ, dtype=np.float64)
    n, sm = AQ.shape
    m = alpha.shape[1]
    dQ, dAQ, da = DeviceBuffer.from_array(Q), DeviceBuffer.from_array(AQ), DeviceBuffer.from_array(alpha)
    dX, dR = DeviceBuffer.from_array(X), DeviceBuffer.from_array(R)
    dn = DeviceBuffer(8 * m)
    ws = DeviceBuffer(8 * lib().cacg_xrupdate_workspace(m))
    _check(lib().cacg_k_xrupdate_split(n, sm, m, dQ.ptr, dAQ.ptr, da.ptr, dX.ptr, dR.ptr, dn.ptr,
                                       ws.ptr, which, None), "xrupdate_split")
    if which == 1:
        return dX.download((n, m))
    return dR.download((n, m)), dn.download(m)


def k_spmm_f32(A, X, epi="none", a=0.0, b=0.0, p1=None, p2=None, long_min=None):
    """FP32 Y = epilogue(A32 X) for X n x k float32 ('none' or 'cheb')."""
    dA = DeviceCSR(A, long_min)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(A.n, -1)
    k = X.shape[1]
    dX = DeviceBuffer.from_array(X)
    dY = DeviceBuffer(X.nbytes)
    d1 = DeviceBuffer.from_array(np.ascontiguousarray(p1, dtype=np.float32)) if p1 is not None else None
    d2 = DeviceBuffer.from_array(np.ascontiguousarray(p2, dtype=np.float32)) if p2 is not None else None
    rc = lib().cacg_k_spmm_f32(ctypes.byref(dA.s32), k, dX.ptr, k, dY.ptr, k, EPI[epi], a, b,
                               d1.ptr if d1 else None, k, d2.ptr if d2 else None, k, None)
    _check(rc, "spmm_f32")
    return dY.download((A.n, k), np.float32)


def k_cheb1_f32(A, R, a, b, long_min=None):
    """(S_1, S_0) of the FP32 basis from the FP64 residual R (n x k)."""
    dA = DeviceCSR(A, long_min)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(A.n, -1)
    k = R.shape[1]
    dR = DeviceBuffer.from_array(R)
    d1, d0 = DeviceBuffer(4 * A.n * k), DeviceBuffer(4 * A.n * k)
    _check(lib().cacg_k_cheb1_f32(ctypes.byref(dA.s32), k, dR.ptr, k, d1.ptr, k, d0.ptr, k, a, b,
                                  None), "cheb1_f32")
    return d1.download((A.n, k), np.float32), d0.download((A.n, k), np.float32)


def k_gram_f64f32(L, R1):
    """L^T R1 with FP64 L and FP32 R1 (fast order)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    R1 = np.ascontiguousarray(R1, dtype=np.float32)
    n, kl = L.shape
    kr1 = R1.shape[1]
    dL, d1 = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R1)
    g1 = DeviceBuffer(8 * kl * kr1)
    ws = DeviceBuffer(8 * lib().cacg_gram_workspace(kl, kr1))
    _check(lib().cacg_k_gram_fast_f64f32(n, dL.ptr, kl, kl, d1.ptr, kr1, kr1, None, 1, 0, g1.ptr,
                                         None, ws.ptr, None), "gram_f64f32")
    return g1.download((kl, kr1))


def k_qupdate_s32(S, Q, beta):
    """Q (FP64) = S (FP32) - Q beta."""
    S = np.ascontiguousarray(S, dtype=np.float32)
    n, sm = S.shape
    dS, dQ = DeviceBuffer.from_array(S), DeviceBuffer.from_array(np.asarray(Q, dtype=np.float64))
    db = DeviceBuffer.from_array(np.ascontiguousarray(beta, dtype=np.float64))
    _check(lib().cacg_k_qupdate_s32(n, sm, dS.ptr, dQ.ptr, db.ptr, None), "qupdate_s32")
    return dQ.download((n, sm))


def k_tsqr(A):
    """(Q, R, flag) of the tall-skinny QR of A (n x k, k <= 128)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n, k = A.shape
    dA, dQ, dR = DeviceBuffer.from_array(A), DeviceBuffer(A.nbytes), DeviceBuffer(8 * k * k)
    ws = DeviceBuffer(8 * lib().cacg_tsqr_workspace(k))
    flag = DeviceBuffer.from_array(np.zeros(1, dtype=np.int32))
    _check(lib().cacg_k_tsqr(n, k, dA.ptr, k, dQ.ptr, dR.ptr, ws.ptr, flag.ptr, None), "tsqr")
    return dQ.download((n, k)), dR.download((k, k)), int(flag.download(1, np.int32)[0])


def k_lanczos(A, steps):
    """(alpha, beta) of `steps` device Lanczos steps on A (cacg_k_lanczos)."""
    dA = DeviceCSR(A)
    ws = DeviceBuffer(8 * lib().cacg_lanczos_workspace(A.n))
    ab = DeviceBuffer(16 * steps)
    _check(lib().cacg_k_lanczos(ctypes.byref(dA.s), steps, ws.ptr, ab.ptr, None), "lanczos")
    out = ab.download(2 * steps)
    return out[:steps].copy(), out[steps:].copy()


def lanczos_interval(alpha, beta, gersh):
    """The interval CACG_BOUNDS_LANCZOS uses (cacg_lanczos_interval)."""
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    g = np.ascontiguousarray(gersh, dtype=np.float64)
    out = np.zeros(2)
    rc = lib().cacg_lanczos_interval(a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), len(a),
                                     g.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError(f"cacg_lanczos_interval: {STATUS.get(rc, rc)}")
    return out


def k_cvt_f32_f64(S):
    S = np.ascontiguousarray(S, dtype=np.float32)
    dS, dQ = DeviceBuffer.from_array(S), DeviceBuffer(8 * S.size)
    _check(lib().cacg_k_cvt_f32_f64(S.size, dS.ptr, dQ.ptr, None), "cvt_f32_f64")
    return dQ.download(S.shape)


def k_colnorm2_ref(R, ld=None):
    """||R(:, c)||^2 of every column, each a row-order chain (reference order);
    R n x m, stored with row stride ld (default m)."""
    R = np.ascontiguousarray(R, dtype=np.float64)
    n, m = R.shape
    ld = m if ld is None else ld
    W = np.zeros((max(n, 1), ld))
    W[:n, :m] = R
    dR = DeviceBuffer.from_array(W)
    dn = DeviceBuffer(8 * m)
    _check(lib().cacg_k_colnorm2_ref(n, m, dR.ptr, ld, dn.ptr, None), "colnorm2_ref")
    return dn.download(m)


def k_colnorm2_fast(R):
    R = np.ascontiguousarray(R, dtype=np.float64)
    n, m = R.shape
    dR, dn = DeviceBuffer.from_array(R), DeviceBuffer(8 * m)
    ws = DeviceBuffer(8 * lib().cacg_xrupdate_workspace(m))
    _check(lib().cacg_k_colnorm2_fast(n, m, dR.ptr, dn.ptr, ws.ptr, None), "colnorm2_fast")
    return dn.download(m)


# ---- solver ------------------------------------------------------------------------------


PRECISION = {"fp64": 0, "fp32": 1}


def make_opts(stop_mode=STOP_REL_PER_COLUMN, eps=0.0, tol=1e-8, maxiter=100, order="fast",
              print_period=0, precision="fp64", inner_tol=0.0, inner_maxiter=0,
              refine_restart=0, bounds="gerschgorin", lanczos_steps=20, bound_lo=0.0,
              bound_hi=0.0, fp32_basis_below=None) -> Opts:
    """precision 'fp32' is the FP32-basis form with FP64 refinement; bounds
    'gerschgorin' (the reference's), 'lanczos' or 'given' ([bound_lo, bound_hi])."""
    o = Opts()
    lib().cacg_default_opts(ctypes.byref(o))
    o.stop_mode = stop_mode
    o.eps = eps
    o.tol = tol
    o.maxiter = maxiter
    o.order = MODE[order] if isinstance(order, str) else int(order)
    o.print_period = print_period
    o.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    o.inner_tol = inner_tol
    o.inner_maxiter = inner_maxiter
    o.refine_restart = int(refine_restart)
    o.bounds = BOUNDS[bounds] if isinstance(bounds, str) else int(bounds)
    o.lanczos_steps = int(lanczos_steps)
    o.bound_lo = bound_lo
    o.bound_hi = bound_hi
    if fp32_basis_below is not None:   # else the library default
        o.fp32_basis_below = fp32_basis_below
    return o


def _report_dict(rep: Report, hist: np.ndarray, m: int) -> dict:
    k = int(rep.iters)
    return {
        "status": STATUS.get(rep.status, str(rep.status)),
        "iters": k,
        "bounds": np.array(rep.bounds[:]),
        "res_hist": hist[:min(k, hist.shape[0])].copy(),
        "seconds": rep.seconds,
        "phase_ms": dict(zip(PHASES, rep.phase_ms[:])),
        "phase_calls": dict(zip(PHASES, rep.phase_calls[:])),
        "refinements": int(rep.refinements),
        "true_relres": float(rep.true_relres),
    }


def bcbcg(A, B, X0, s, *, stop_mode=STOP_REL_PER_COLUMN, eps=0.0, tol=1e-8, maxiter=100,
          order="fast", print_period=0, precision="fp64", inner_tol=0.0, inner_maxiter=0,
          refine_restart=0, bounds="gerschgorin", lanczos_steps=20, bound_lo=0.0,
          bound_hi=0.0, fp32_basis_below=None) -> dict:
    """Host-buffer solve through cacg_bcbcg().  Returns dict with X and the report."""
    B = np.ascontiguousarray(B, dtype=np.float64).reshape(A.n, -1)
    m = B.shape[1]
    X = np.ascontiguousarray(X0, dtype=np.float64).reshape(A.n, m).copy()
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.col, dtype=np.int64)
    va = np.ascontiguousarray(A.val, dtype=np.float64)
    o = make_opts(stop_mode, eps, tol, maxiter, order, print_period, precision, inner_tol,
                  inner_maxiter, refine_restart, bounds, lanczos_steps, bound_lo, bound_hi,
                  fp32_basis_below)
    hist = np.zeros((maxiter, m))
    rep = Report()
    rep.res_hist = hist.ctypes.data_as(_dp)
    rep.res_hist_cap = maxiter
    rc = lib().cacg_bcbcg(A.n, rp.ctypes.data_as(_lp), ci.ctypes.data_as(_lp), va.ctypes.data_as(_dp),
                          m, s, B.ctypes.data_as(_dp), X.ctypes.data_as(_dp), ctypes.byref(o),
                          ctypes.byref(rep))
    if rc not in (0, 1, 2):
        raise RuntimeError(f"cacg_bcbcg failed: {STATUS.get(rc, rc)}")
    out = _report_dict(rep, hist, m)
    out["X"] = X
    return out


# ---- input formats (include/cacg_io.h) ---------------------------------------------

def read_mtx(path) -> dict:
    """A Matrix Market coordinate file as the reference's reader and COO->CSR
    leave it (cacg_mtx_read): dict(num_rows, num_cols, rowptr, col, val); for a
    square matrix also 'csr', a problems.CSR."""
    nr, nc, nnz = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
    rp, ci, va = _lp(), _lp(), _dp()
    rc = lib().cacg_mtx_read(os.fsencode(path), ctypes.byref(nr), ctypes.byref(nc),
                             ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va))
    if rc != 0:
        raise (OSError if rc == 6 else ValueError)(f"cacg_mtx_read({path}): {STATUS.get(rc, rc)}")
    try:
        n, k = nr.value, nnz.value
        out = {"num_rows": n, "num_cols": nc.value,
               "rowptr": np.ctypeslib.as_array(rp, (n + 1,)).copy(),
               "col": np.ctypeslib.as_array(ci, (k,)).copy() if k else np.zeros(0, np.int64),
               "val": np.ctypeslib.as_array(va, (k,)).copy() if k else np.zeros(0)}
    finally:
        for p in (rp, ci, va):
            lib().cacg_io_free(ctypes.cast(p, _vp))
    if out["num_rows"] == out["num_cols"]:
        from .problems import CSR
        out["csr"] = CSR(out["num_rows"], out["rowptr"], out["col"], out["val"])
    return out


def write_mtx(path, num_rows, num_cols, rowptr, col, val) -> None:
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(col, dtype=np.int64)
    va = np.ascontiguousarray(val, dtype=np.float64)
    rc = lib().cacg_mtx_write(os.fsencode(path), num_rows, num_cols, rp.ctypes.data_as(_lp),
                              ci.ctypes.data_as(_lp), va.ctypes.data_as(_dp))
    if rc != 0:
        raise OSError(f"cacg_mtx_write({path}): {STATUS.get(rc, rc)}")


def read_csv(path, rows, cols) -> np.ndarray:
    """rows x cols values as the reference's parseCSV reads them (cacg_csv_read)."""
    d = _dp()
    rc = lib().cacg_csv_read(os.fsencode(path), rows, cols, ctypes.byref(d))
    if rc != 0:
        raise (OSError if rc == 6 else ValueError)(f"cacg_csv_read({path}): {STATUS.get(rc, rc)}")
    try:
        return np.ctypeslib.as_array(d, (rows, cols)).copy()
    finally:
        lib().cacg_io_free(ctypes.cast(d, _vp))


def write_csv(path, data) -> None:
    a = np.ascontiguousarray(data, dtype=np.float64)
    a = a.reshape(a.shape[0], -1)
    rc = lib().cacg_csv_write(os.fsencode(path), a.shape[0], a.shape[1], a.ctypes.data_as(_dp))
    if rc != 0:
        raise OSError(f"cacg_csv_write({path}): {STATUS.get(rc, rc)}")


def bcg(A, B, X0, *, stop_mode=STOP_REL_PER_COLUMN, eps=0.0, tol=1e-8, maxiter=1000,
        order="fast", print_period=0) -> dict:
    """Host-buffer block CG through cacg_bcg() (the reference's bcg_v1,
    src/bcgKernel.c:22-462).  res_hist[k] holds ||R(:,c)|| of the residual
    iteration k started from (what its stop test reads); X has been updated
    by the iteration that stops."""
    B = np.ascontiguousarray(B, dtype=np.float64).reshape(A.n, -1)
    m = B.shape[1]
    X = np.ascontiguousarray(X0, dtype=np.float64).reshape(A.n, m).copy()
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.col, dtype=np.int64)
    va = np.ascontiguousarray(A.val, dtype=np.float64)
    o = make_opts(stop_mode, eps, tol, maxiter, order, print_period)
    hist = np.zeros((maxiter, m))
    rep = Report()
    rep.res_hist = hist.ctypes.data_as(_dp)
    rep.res_hist_cap = maxiter
    rc = lib().cacg_bcg(A.n, rp.ctypes.data_as(_lp), ci.ctypes.data_as(_lp), va.ctypes.data_as(_dp),
                        m, B.ctypes.data_as(_dp), X.ctypes.data_as(_dp), ctypes.byref(o),
                        ctypes.byref(rep))
    if rc not in (0, 1, 2):
        raise RuntimeError(f"cacg_bcg failed: {STATUS.get(rc, rc)}")
    out = _report_dict(rep, hist, m)
    out["X"] = X
    return out


class Solver:
    """Device-resident solver: matrix uploaded once, work blocks allocated once."""

    def __init__(self, A, m, s, **opt_kw):
        self.n, self.m, self.s = A.n, m, s
        self.maxiter = opt_kw.get("maxiter", 100)
        self._keep = (np.ascontiguousarray(A.rowptr, dtype=np.int64),
                      np.ascontiguousarray(A.col, dtype=np.int64),
                      np.ascontiguousarray(A.val, dtype=np.float64))
        o = make_opts(**opt_kw)
        h = _vp()
        rc = lib().cacg_solver_create(ctypes.byref(h), A.n, self._keep[0].ctypes.data_as(_lp),
                                      self._keep[1].ctypes.data_as(_lp),
                                      self._keep[2].ctypes.data_as(_dp), m, s, ctypes.byref(o))
        _check(rc, "cacg_solver_create")
        self.h = h.value
        self.hist = np.zeros((self.maxiter, m))

    def footprint(self) -> int:
        return lib().cacg_solver_footprint(self.h)

    def solve(self, dB: DeviceBuffer, dX: DeviceBuffer) -> dict:
        rep = Report()
        rep.res_hist = self.hist.ctypes.data_as(_dp)
        rep.res_hist_cap = self.maxiter
        rc = lib().cacg_solver_solve_device(self.h, dB.ptr, dX.ptr, ctypes.byref(rep))
        if rc not in (0, 1, 2):
            raise RuntimeError(f"solve failed: {STATUS.get(rc, rc)}")
        return _report_dict(rep, self.hist, self.m)

    def solve_bcg(self, dB: DeviceBuffer, dX: DeviceBuffer) -> dict:
        """Block CG (bcg_v1) on this solver's matrix and buffers."""
        rep = Report()
        rep.res_hist = self.hist.ctypes.data_as(_dp)
        rep.res_hist_cap = self.maxiter
        rc = lib().cacg_solver_solve_bcg_device(self.h, dB.ptr, dX.ptr, ctypes.byref(rep))
        if rc not in (0, 1, 2):
            raise RuntimeError(f"solve failed: {STATUS.get(rc, rc)}")
        return _report_dict(rep, self.hist, self.m)

    def close(self):
        if self.h:
            lib().cacg_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
