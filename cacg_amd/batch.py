"""Python binding of the batching / serving engine (include/cacg_batch.h)
and the request generators and drivers of the batching and serving forms.

    b = Batch(s=4, precision="fp32")
    ids = [b.submit(A, B, tol=1e-10) for A, B in systems]
    results = b.run()                     # rounds until every request retired

Continuous serving: `serve_stream` submits each request when its arrival
time has passed (wall clock), runs one round, collects what retired and
records submit-to-retire latency; `serve_sequential` is the comparison
server that solves the same stream one request at a time (FIFO) with the
single-system FP64 path.
"""

from __future__ import annotations

import ctypes
import threading
import time

import numpy as np

from . import STATUS, lib as _lib
from . import problems as P

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
PRECISION = {"fp64": 0, "fp32": 1}
POLICY = {"fifo": 0, "smallest": 1}


class BatchOpts(ctypes.Structure):
    _fields_ = [("hbm_bytes", ctypes.c_long), ("slots", ctypes.c_int), ("s", ctypes.c_int),
                ("precision", ctypes.c_int), ("maxiter", ctypes.c_long),
                ("max_large", ctypes.c_int), ("graph", ctypes.c_int), ("tune", ctypes.c_int),
                ("policy", ctypes.c_int), ("max_bypass", ctypes.c_int),
                ("solo_rows", ctypes.c_long)]


class BatchResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_long),
                ("rounds_queued", ctypes.c_long), ("true_relres", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("active", ctypes.c_int), ("queued", ctypes.c_int), ("finished", ctypes.c_int),
                ("admitted", ctypes.c_int), ("retired", ctypes.c_int), ("rounds", ctypes.c_long),
                ("arena_used", ctypes.c_long), ("arena_bytes", ctypes.c_long),
                ("graph_rounds", ctypes.c_long)]


def lib():
    L = _lib()
    if not getattr(L, "_batch_sigs", False):
        c_int, c_long, c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
        sig = {
            "cacg_batch_default_opts": ([ctypes.POINTER(BatchOpts)], None),
            "cacg_batch_create": ([ctypes.POINTER(_vp), ctypes.POINTER(BatchOpts)], c_int),
            "cacg_batch_destroy": ([_vp], c_int),
            "cacg_batch_footprint": ([c_long, c_long, c_long, c_int, c_int], c_long),
            "cacg_batch_submit": ([_vp, c_long, _lp, _lp, _dp, c_long, _dp, _dp, c_double], c_long),
            "cacg_batch_step": ([_vp, ctypes.POINTER(BatchStats)], c_int),
            "cacg_batch_collect": ([_vp, c_long, _dp, ctypes.POINTER(BatchResult)], c_int),
            "cacg_batch_retired": ([_vp, _lp, c_int], c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._batch_sigs = True
    return L


def footprint(n, nnz, m, s, precision="fp32") -> int:
    return lib().cacg_batch_footprint(n, nnz, m, s, PRECISION[precision])


class Batch:
    """A batch of device-resident systems advanced together (cacg_batch_*)."""

    def __init__(self, s=4, precision="fp32", slots=1024, hbm_bytes=0, maxiter=1000,
                 max_large=0, graph=False, tune=True, policy="fifo", max_bypass=0,
                 solo_rows=0):
        o = BatchOpts()
        lib().cacg_batch_default_opts(ctypes.byref(o))
        o.hbm_bytes, o.slots, o.s = int(hbm_bytes), int(slots), int(s)
        o.precision, o.maxiter = PRECISION[precision], int(maxiter)
        o.max_large, o.graph, o.tune = int(max_large), int(bool(graph)), int(bool(tune))
        o.policy, o.max_bypass = POLICY[policy], int(max_bypass)
        o.solo_rows = int(solo_rows)
        h = _vp()
        rc = lib().cacg_batch_create(ctypes.byref(h), ctypes.byref(o))
        if rc != 0:
            raise RuntimeError(f"cacg_batch_create failed: {STATUS.get(rc, rc)}")
        self.h = h.value
        self.shapes = {}
        self.stats = BatchStats()

    def submit(self, A, B, X0=None, tol=1e-10) -> int:
        B = np.ascontiguousarray(B, dtype=np.float64).reshape(A.n, -1)
        m = B.shape[1]
        rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
        ci = np.ascontiguousarray(A.col, dtype=np.int64)
        va = np.ascontiguousarray(A.val, dtype=np.float64)
        x0 = None if X0 is None else np.ascontiguousarray(X0, dtype=np.float64).reshape(A.n, m)
        rid = lib().cacg_batch_submit(self.h, A.n, rp.ctypes.data_as(_lp), ci.ctypes.data_as(_lp),
                                      va.ctypes.data_as(_dp), m, B.ctypes.data_as(_dp),
                                      x0.ctypes.data_as(_dp) if x0 is not None else None, tol)
        if rid < 0:
            raise ValueError(f"cacg_batch_submit refused the system: {STATUS.get(-rid, -rid)}")
        self.shapes[rid] = (A.n, m)
        return rid

    def step(self) -> dict:
        rc = lib().cacg_batch_step(self.h, ctypes.byref(self.stats))
        if rc != 0:
            raise RuntimeError(f"cacg_batch_step failed: {STATUS.get(rc, rc)}")
        st = self.stats
        return {"active": st.active, "queued": st.queued, "finished": st.finished,
                "admitted": st.admitted, "retired": st.retired, "rounds": st.rounds,
                "arena_used": st.arena_used, "arena_bytes": st.arena_bytes,
                "graph_rounds": st.graph_rounds}

    def retired(self) -> list:
        """Ids retired since the last call."""
        out = []
        buf = np.empty(1024, dtype=np.int64)
        while True:
            k = lib().cacg_batch_retired(self.h, buf.ctypes.data_as(_lp), len(buf))
            out.extend(int(v) for v in buf[:k])
            if k < len(buf):
                return out

    def collect(self, rid):
        n, m = self.shapes[rid]
        X = np.empty((n, m))
        res = BatchResult()
        rc = lib().cacg_batch_collect(self.h, rid, X.ctypes.data_as(_dp), ctypes.byref(res))
        if rc < 0:
            raise KeyError(rid)
        if rc == 0:
            return None
        del self.shapes[rid]
        return {"X": X, "status": STATUS.get(res.status, str(res.status)), "iters": res.iters,
                "rounds_queued": res.rounds_queued, "true_relres": res.true_relres}

    def run(self, max_rounds=100000) -> dict:
        """Rounds until every submitted request retired; returns results by id."""
        out = {}
        for _ in range(max_rounds):
            st = self.step()
            for rid in self.retired():
                out[rid] = self.collect(rid)
            if st["active"] == 0 and st["queued"] == 0 and not self.shapes:
                return out
        raise RuntimeError("batch did not drain")

    def close(self):
        if self.h:
            lib().cacg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- workloads ---------------------------------------------------------------------------


def batch_systems(count=256, seed=3, lmin=32, lmax=256, ms=(1, 2, 4, 8, 16, 32)):
    """The batching form: shifted 2D Laplacians, L log-uniform in [lmin, lmax],
    shift U(0, 0.1), m uniform over ms (problems.sample_batch_systems), B
    entries U[0,1) drawn from the same stream after the 3*count parameters,
    system after system, row by row."""
    out = []
    off = 3 * count
    for L, shift, m in P.sample_batch_systems(count, seed, lmin, lmax, ms):
        A = P.laplacian_2d(L, shift)
        B = P.uniform(seed, A.n * m, off).reshape(A.n, m)
        off += A.n * m
        out.append((A, B))
    return out


def serving_stream(count, rate, seed=4, l2=(64, 512), e3=(32, 128), mrange=(1, 32)):
    """The serving form: `count` requests with Poisson arrivals at `rate`
    per second; each is a 2D Laplacian (edge uniform in l2) or a 3D 7-point
    Laplacian (edge uniform in e3) with equal probability and m uniform in
    mrange, B entries U[0,1).  Five stream outputs per request (gap, kind,
    edge, m, spare), then the B entries request after request.
    Returns [(t_arrival, A, B)]."""
    u = P.uniform(seed, 5 * count).reshape(count, 5)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    out = []
    off = 5 * count
    for i in range(count):
        if u[i, 1] < 0.5:
            L = int(l2[0] + np.floor(u[i, 2] * (l2[1] - l2[0] + 1)))
            A = P.laplacian_2d(min(L, l2[1]))
        else:
            E = int(e3[0] + np.floor(u[i, 2] * (e3[1] - e3[0] + 1)))
            A = P.laplacian_3d(min(E, e3[1]))
        m = int(mrange[0] + np.floor(u[i, 3] * (mrange[1] - mrange[0] + 1)))
        m = min(m, mrange[1])
        B = P.uniform(seed, A.n * m, off).reshape(A.n, m)
        off += A.n * m
        out.append((float(t[i]), A, B))
    return out


# ---- drivers -----------------------------------------------------------------------------


def serve_stream(stream, batch: Batch, tol=1e-10, timeout=600.0, feeder=True, feeders=1):
    """Continuous batching of a stream of (t_arrival, A, B) requests.

    feeder=True: a feeder thread submits each request once its arrival time has
    passed, so its host-side copies (cacg_batch_submit converts and copies the
    matrix and B, and computes its Gerschgorin interval) run while a round waits
    for the GPU; the calling thread runs rounds while any submitted request is
    unfinished, collects what retired, and waits for the feeder when idle.
    feeders > 1: requests are dealt round-robin to that many feeder threads, so
    several submissions' host-side copies run at once (requests arriving
    together may then be submitted in another order; every result is the same
    bits whatever the order).
    feeder=False: submissions happen between rounds on the calling thread.
    Returns (results in stream order with 'latency' from arrival to the end of
    the retiring round, wall seconds from the first arrival to the last
    retirement).  On an error or the timeout the feeders are stopped and joined
    before the exception propagates, so the caller may close the batch."""
    count = len(stream)
    ids, res = {}, {}
    cv = threading.Condition()
    err = []
    stop = threading.Event()
    t0 = time.perf_counter()

    def submit_due(now):
        while pending and stream[pending[0]][0] <= now:
            i = pending.pop(0)
            _, A, B = stream[i]
            ids[batch.submit(A, B, tol=tol)] = i

    def feed(first=0, step=1):
        try:
            for i in range(first, count, step):
                t_arr, A, B = stream[i]
                d = t_arr - (time.perf_counter() - t0)
                if d > 0 and stop.wait(d):
                    return
                if stop.is_set():
                    return
                rid = batch.submit(A, B, tol=tol)
                with cv:
                    ids[rid] = i
                    cv.notify_all()
        except Exception as e:  # surfaced on the calling thread
            with cv:
                err.append(e)
                cv.notify_all()

    pending = list(range(count))
    ths = []
    if feeder:
        pending = []
        k = max(1, int(feeders))
        ths = [threading.Thread(target=feed, args=(j, k), daemon=True) for j in range(k)]
        for th in ths:
            th.start()
    try:
        _serve_rounds(stream, batch, feeder, count, ids, res, cv, err, t0, timeout, pending,
                      submit_due)
    finally:
        stop.set()
        with cv:
            cv.notify_all()
        for th in ths:
            th.join()
    wall = (time.perf_counter() - t0) - stream[0][0]
    return [res[i] for i in range(count)], wall


def _serve_rounds(stream, batch, feeder, count, ids, res, cv, err, t0, timeout, pending,
                  submit_due):
    """The round loop of serve_stream on the calling thread."""
    while len(res) < count:
        if time.perf_counter() - t0 > timeout:
            raise RuntimeError("serving stream timed out")
        if not feeder:
            submit_due(time.perf_counter() - t0)
            if len(ids) == len(res):   # idle until the next arrival
                time.sleep(max(0.0, stream[pending[0]][0] - (time.perf_counter() - t0)))
                continue
        else:
            with cv:
                while len(ids) == len(res) and not err:
                    cv.wait(timeout=1.0)
                    if time.perf_counter() - t0 > timeout:
                        raise RuntimeError("serving stream timed out")
                if err:
                    raise err[0]
        batch.step()
        t_now = time.perf_counter() - t0
        for rid in batch.retired():
            r = batch.collect(rid)
            with cv:   # the feeder may not have recorded the id yet
                while rid not in ids:
                    cv.wait(timeout=1.0)
                i = ids[rid]
            r["latency"] = t_now - stream[i][0]
            res[i] = r


def serve_sequential(stream, tol=1e-10, s=4, maxiter=1000, precision="fp64"):
    """One request at a time, FIFO, with the single-system path
    (cacg_amd.bcbcg: create, upload, solve, download per request).  The
    server waits for arrivals when idle.  Returns (results, wall seconds)."""
    from . import bcbcg
    res = []
    t0 = time.perf_counter()
    for t_arr, A, B in stream:
        now = time.perf_counter() - t0
        if now < t_arr:
            time.sleep(t_arr - now)
        out = bcbcg(A, B, np.zeros_like(B), s, tol=tol, maxiter=maxiter, precision=precision)
        out["latency"] = (time.perf_counter() - t0) - t_arr
        res.append(out)
    wall = (time.perf_counter() - t0) - stream[0][0]
    return res, wall
