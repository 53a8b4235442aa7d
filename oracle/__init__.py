"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C restatement of sclc/cacg's s-step Chebyshev (block)
CG (oracle/oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, to check results or as the timed CPU
baseline.  The product library never does.

`run_reference` drives the reference's own solver objects, compiled from
/root/reference/src by oracle/ref/Makefile into oracle/_ref/, through
oracle/ref/ref_harness.c.  It is used only where that build exists (the
golden-fixture script and the restatement check in tests/).
"""

from __future__ import annotations

import ctypes
import os
import struct
import sys
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
MPIEXEC = "/opt/conda/bin/mpiexec"

STOP_REF_COL0_ABS = 0
STOP_REL_PER_COLUMN = 1
STATUS = {0: "ok", 1: "maxiter", 2: "breakdown", 3: "inval"}

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


class _Csr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_long), ("row_offset", ctypes.c_long),
                ("rowptr", _lp), ("col", _lp), ("val", _dp)]


class _Opts(ctypes.Structure):
    _fields_ = [("stop_mode", ctypes.c_int), ("eps", ctypes.c_double),
                ("tol", ctypes.c_double), ("maxiter", ctypes.c_long),
                ("nranks", ctypes.c_int), ("res_hist", _dp), ("x_hist", _dp),
                ("bounds", _dp)]


def build() -> None:
    """Compile oracle/liboracle.so (gcc) and, where the reference sources are
    present, oracle/_ref/ref_harness."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/src") and os.path.exists("/opt/conda/include/mpi.h"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_gerschgorin.argtypes = [ctypes.POINTER(_Csr), _dp]
        L.orc_spmm.argtypes = [ctypes.POINTER(_Csr), _dp, ctypes.c_long, _dp]
        L.orc_cheb_basis.argtypes = [ctypes.POINTER(_Csr), _dp, ctypes.c_long, ctypes.c_long,
                                     ctypes.c_double, ctypes.c_double, _dp]
        L.orc_gram.argtypes = [ctypes.c_long, _dp, ctypes.c_long, _dp, ctypes.c_long, _dp]
        L.orc_gemm_ts.argtypes = [ctypes.c_long, _dp, ctypes.c_long, _dp, ctypes.c_long, _dp]
        L.orc_lu_solve.argtypes = [ctypes.c_long, _dp, ctypes.c_long, _dp, _dp]
        L.orc_col_norm2sq.argtypes = [ctypes.c_long, ctypes.c_long, _dp, ctypes.c_long]
        L.orc_col_norm2sq.restype = ctypes.c_double
        L.orc_bcbcg.argtypes = [ctypes.POINTER(_Csr), ctypes.c_long, ctypes.c_long, _dp, _dp,
                                ctypes.POINTER(_Opts), _lp]
        L.orc_bcbcg.restype = ctypes.c_int
        L.orc_bcg.argtypes = [ctypes.POINTER(_Csr), ctypes.c_long, _dp, _dp,
                              ctypes.POINTER(_Opts), _lp]
        L.orc_bcg.restype = ctypes.c_int
        L.orc_bcbcg_time_sample.argtypes = [ctypes.POINTER(_Csr), ctypes.c_long, ctypes.c_long,
                                            _dp, _dp, ctypes.c_int, ctypes.c_long, _dp, _dp]
        L.orc_bcbcg_time_sample.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _CsrHold:
    """Keeps the arrays of a CSR alive while C holds pointers to them."""

    def __init__(self, A, row_offset=0):
        self.rowptr = np.ascontiguousarray(A.rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(A.col, dtype=np.int64)
        self.val = _c64(A.val)
        self.s = _Csr(A.n, row_offset, self.rowptr.ctypes.data_as(_lp),
                      self.col.ctypes.data_as(_lp), _d(self.val))


def gerschgorin(A):
    h = _CsrHold(A)
    out = np.zeros(2)
    lib().orc_gerschgorin(ctypes.byref(h.s), _d(out))
    return out


def spmm(A, X):
    h = _CsrHold(A)
    X = _c64(X).reshape(A.n, -1)
    Y = np.zeros_like(X)
    lib().orc_spmm(ctypes.byref(h.s), _d(X), X.shape[1], _d(Y))
    return Y


def cheb_basis(A, R, s, s_alpha, s_beta):
    h = _CsrHold(A)
    R = _c64(R).reshape(A.n, -1)
    m = R.shape[1]
    S = np.zeros((A.n, s * m))
    lib().orc_cheb_basis(ctypes.byref(h.s), _d(R), m, s, s_alpha, s_beta, _d(S))
    return S


def gram(L, R):
    L, R = _c64(L), _c64(R)
    G = np.zeros((L.shape[1], R.shape[1]))
    lib().orc_gram(L.shape[0], _d(L), L.shape[1], _d(R), R.shape[1], _d(G))
    return G


def gemm_ts(L, W):
    L, W = _c64(L), _c64(W)
    Y = np.zeros((L.shape[0], W.shape[1]))
    lib().orc_gemm_ts(L.shape[0], _d(L), L.shape[1], _d(W), W.shape[1], _d(Y))
    return Y


def lu_solve(lhs, rhs):
    lhs, rhs = _c64(lhs), _c64(rhs)
    x = np.zeros_like(rhs)
    lib().orc_lu_solve(lhs.shape[0], _d(lhs), rhs.shape[1], _d(rhs), _d(x))
    return x


def col_norm2sq(R, col):
    R = _c64(R)
    return lib().orc_col_norm2sq(R.shape[0], R.shape[1], _d(R), col)


def bcbcg(A, B, X0, s, *, stop_mode=STOP_REF_COL0_ABS, eps=0.0, tol=0.0, maxiter=100,
          nranks=1, record_x=False):
    """Run the restated solver.  Returns a dict with X, iters, status,
    res_hist (iters x m, ||R(:,j)||_2 after each outer iteration), bounds and,
    with record_x, x_hist (iters x n x m)."""
    h = _CsrHold(A)
    B = _c64(B).reshape(A.n, -1)
    m = B.shape[1]
    X = _c64(X0).reshape(A.n, m).copy()
    res = np.zeros((maxiter, m))
    xh = np.zeros((maxiter, A.n, m)) if record_x else None
    bounds = np.zeros(2)
    o = _Opts(stop_mode, eps, tol, maxiter, nranks, _d(res),
              _d(xh) if xh is not None else None, _d(bounds))
    it = ctypes.c_long(0)
    st = lib().orc_bcbcg(ctypes.byref(h.s), m, s, _d(B), _d(X), ctypes.byref(o), ctypes.byref(it))
    k = it.value
    out = {"X": X, "iters": k, "status": STATUS[st], "res_hist": res[:k], "bounds": bounds}
    if xh is not None:
        out["x_hist"] = xh[:k]
    return out


def cbcg(A, b, x0, s, **kw):
    """cbcg_v1 (src/cbcgKernel.c:23-250) is bcbcg_v1 with one column and a
    1000-iteration cap (src/cbcgKernel.c:2)."""
    kw.setdefault("maxiter", 1000)
    return bcbcg(A, np.reshape(b, (A.n, 1)), np.reshape(x0, (A.n, 1)), s, **kw)


def bcg(A, B, X0, *, stop_mode=STOP_REF_COL0_ABS, eps=0.0, tol=0.0, maxiter=1000, nranks=1,
        record_x=False):
    """bcg_v1 (src/bcgKernel.c:22-462), O'Leary block CG with the reference's
    1000-iteration cap (src/bcgKernel.c:2).  res_hist[k] is ||R(:,c)|| of the
    residual iteration k started from (the one its stop test reads), x_hist[k]
    the X after iteration k's update."""
    h = _CsrHold(A)
    B = _c64(B).reshape(A.n, -1)
    m = B.shape[1]
    X = _c64(X0).reshape(A.n, m).copy()
    res = np.zeros((maxiter, m))
    xh = np.zeros((maxiter, A.n, m)) if record_x else None
    o = _Opts(stop_mode, eps, tol, maxiter, nranks, _d(res),
              _d(xh) if xh is not None else None, None)
    it = ctypes.c_long(0)
    st = lib().orc_bcg(ctypes.byref(h.s), m, _d(B), _d(X), ctypes.byref(o), ctypes.byref(it))
    k = it.value
    out = {"X": X, "iters": k, "status": STATUS[st], "res_hist": res[:k]}
    if xh is not None:
        out["x_hist"] = xh[:k]
    return out


def time_sample(A, B, X0, s, nranks, iters):
    """(t_setup, [t_iteration_0, ...]) wall seconds of a bounded sample of the
    solve at `nranks` emulated ranks (one thread each)."""
    h = _CsrHold(A)
    B = _c64(B).reshape(A.n, -1)
    X = _c64(X0).reshape(B.shape).copy()
    ts = ctypes.c_double()
    ti = np.zeros(max(iters, 1))
    lib().orc_bcbcg_time_sample(ctypes.byref(h.s), B.shape[1], s, _d(B), _d(X), nranks, iters,
                                ctypes.byref(ts), _d(ti))
    return ts.value, ti[:iters].tolist()


def reference_available() -> bool:
    return os.path.exists(REF_HARNESS)


def time_reference(A, B, X0, s, eps, solver="bcbcg", nranks=1, timeout=600, workdir=None,
                   progress=False):
    """Wall-clock profile of the reference's own solver (the compiled
    oracle/_ref/ref_harness) at `nranks` MPI ranks: the harness stamps every
    outer iteration with the seconds since the solver was called (rank 0,
    after MPI_Init and its input was read), so t[k] = setup + iterations
    0..k; and it times the whole solve as the reference's main does (barrier,
    clock, solve, barrier, clock, the maximum over ranks; src/main.c:154-163).
    Returns dict(iters, t (list of seconds), solve (that time), wall (seconds
    of the whole run, input reading and output writing included)).  The run
    keeps no snapshots.  progress passes the harness's per-iteration lines
    through to stderr as they come."""
    import re
    import time as _time
    if not reference_available():
        raise RuntimeError("reference build missing (make -C oracle ref)")
    B = _c64(B).reshape(A.n, -1)
    n, m = B.shape
    X0 = _c64(X0).reshape(n, m)
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(b"CACGIN01")
            f.write(struct.pack("<5q", n, A.nnz, m, s, {"bcg": 0, "cbcg": 1, "bcbcg": 2}[solver]))
            f.write(struct.pack("<d", eps))
            f.write(np.ascontiguousarray(A.rowptr, dtype="<i8").tobytes())
            f.write(np.ascontiguousarray(A.col, dtype="<i8").tobytes())
            f.write(_c64(A.val).tobytes())
            f.write(B.tobytes())
            f.write(X0.tobytes())
        cmd = [REF_HARNESS, fin, fout]
        if nranks > 1:
            cmd = [MPIEXEC, "-n", str(nranks)] + cmd
        env = dict(os.environ, CACG_REF_NOSNAP="1", CACG_REF_PROGRESS="1")
        t0 = _time.perf_counter()
        if progress:
            lines = []
            with subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                                  text=True, cwd=td, env=env) as proc:
                for line in proc.stderr:
                    lines.append(line)
                    sys.stderr.write(line)
                    sys.stderr.flush()
                    if _time.perf_counter() - t0 > timeout:
                        proc.kill()
                        raise subprocess.TimeoutExpired(cmd, timeout)
            if proc.returncode != 0:
                raise subprocess.CalledProcessError(proc.returncode, cmd)
            err = "".join(lines)
        else:
            err = subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL,
                                 stderr=subprocess.PIPE, text=True, cwd=td, timeout=timeout,
                                 env=env).stderr
        wall = _time.perf_counter() - t0
    stamps = [float(x) for x in re.findall(r"iteration \d+ r2 \S+ at (\S+) s", err)]
    solve = re.findall(r"ref_harness: solve (\S+) s", err)
    return {"iters": len(stamps), "t": stamps, "solve": float(solve[-1]) if solve else None,
            "wall": wall}


def run_reference(A, B, X0, s, eps, solver="bcbcg", nranks=1, snapshots=True, timeout=600,
                  progress=False, workdir=None):
    """Run the reference's own bcbcg_v1 / cbcg_v1 / bcg_v1 at `nranks` MPI
    ranks (mpiexec from the image's MPICH when nranks > 1).

    Returns dict(iters, r2 (iters,), X (final, n x m)) and, at one rank,
    R and x_hist (iters x n x m) recorded at every outer iteration's stop test
    (bcbcg / cbcg: after the R update; bcg: the R the iteration started from,
    with X already updated).  With snapshots=False (one rank, large systems)
    no R / X history is kept; instead `col_r2` (iters x m) holds every
    column's ||R(:,c)||^2 at each stop test, summed in row order.  progress
    passes the harness's per-iteration lines through to stderr; workdir
    holds the temporary input / output files (default: a fresh temp dir).
    """
    if not reference_available():
        raise RuntimeError("reference build missing (make -C oracle ref)")
    B = _c64(B).reshape(A.n, -1)
    n, m = B.shape
    X0 = _c64(X0).reshape(n, m)
    if not snapshots and nranks != 1:
        raise ValueError("snapshots=False is a one-rank mode")
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(b"CACGIN01")
            f.write(struct.pack("<5q", n, A.nnz, m, s, {"bcg": 0, "cbcg": 1, "bcbcg": 2}[solver]))
            f.write(struct.pack("<d", eps))
            f.write(np.ascontiguousarray(A.rowptr, dtype="<i8").tobytes())
            f.write(np.ascontiguousarray(A.col, dtype="<i8").tobytes())
            f.write(_c64(A.val).tobytes())
            f.write(B.tobytes())
            f.write(X0.tobytes())
        cmd = [REF_HARNESS, fin, fout]
        if nranks > 1:
            cmd = [MPIEXEC, "-n", str(nranks)] + cmd
        env = dict(os.environ)
        if not snapshots:
            env["CACG_REF_NOSNAP"] = "1"
        if progress:
            env["CACG_REF_PROGRESS"] = "1"
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL,
                       stderr=None if progress else subprocess.DEVNULL, cwd=td, timeout=timeout,
                       env=env)
        raw = np.fromfile(fout, dtype=np.uint8)
        cols = np.fromfile(fout + ".cols", dtype="<f8") if not snapshots else None
    assert raw[:8].tobytes() == b"CACGOUT2"
    iters, P, n_, m_ = (int(v) for v in raw[8:40].view("<i8"))
    assert (n_, m_, P) == (n, m, nranks)
    body = raw[40:].view("<f8")
    out = {"iters": iters, "r2": body[:iters].copy()}
    pos = iters
    if cols is not None:
        out["col_r2"] = cols.reshape(iters, m).copy()
    elif P == 1:
        rec = body[pos:pos + iters * 2 * n * m].reshape(iters, 2, n, m)
        out["R"] = rec[:, 0].copy()
        out["x_hist"] = rec[:, 1].copy()
        pos += iters * 2 * n * m
    out["X"] = body[pos:pos + n * m].reshape(n, m).copy()
    return out
