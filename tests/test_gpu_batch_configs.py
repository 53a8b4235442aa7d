"""The batching and serving forms at their configured sizes.

Batching (C4): 256 independent shifted 2D Laplacians, L log-uniform in
[32, 256], shift U(0, 0.1), m from {1, 2, 4, 8, 16, 32} (seed 3), s = 4, FP32
basis with FP64 refinement, tol 1e-10, all submitted at once.
Serving (C5): a Poisson stream (seed 4) of 2D (L 64..512) and 3D (edge
32..128) Laplacians with 1..32 right-hand sides, plus the largest request of
each kind (2D 512^2 and 3D 128^3), continuous batching, tol 1e-10, FP64.

Every system's true relative residual ||B - A X|| / ||B|| is recomputed here
in FP64 (scipy) and must be below the tolerance; a result must be the same
bits alone as in the mix; the smallest systems agree with a direct sparse
solve within cond(A) * tol.  No reference counterpart exists for batching."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cacg_amd import batch as BT
from cacg_amd import problems as P

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _csr(A):
    return sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))


def _relres(A, B, X):
    return np.linalg.norm(B - _csr(A) @ X, axis=0) / np.linalg.norm(B, axis=0)


def _alone(A, B, **kw):
    b = BT.Batch(**kw)
    rid = b.submit(A, B, tol=TOL)
    out = b.run()[rid]
    b.close()
    return out


def test_batching_form_full_size(gpu_lib):
    systems = BT.batch_systems(256, 3)
    ns = [A.n for A, _ in systems]
    assert min(ns) >= 32 * 32 and max(ns) <= 256 * 256
    print(f"batching form: 256 systems, n {min(ns)}..{max(ns)}", flush=True)
    kw = dict(s=4, precision="fp32", hbm_bytes=48 << 30, maxiter=2000)
    b = BT.Batch(**kw)
    ids = [b.submit(A, B, tol=TOL) for A, B in systems]
    res = b.run()
    b.close()
    worst = 0.0
    for (A, B), rid in zip(systems, ids):
        r = res[rid]
        assert r["status"] == "ok", (rid, r["status"])
        rr = _relres(A, B, r["X"])
        worst = max(worst, float(rr.max()))
        assert np.all(rr < TOL), (rid, rr.max())
    print(f"batching form: all converged, worst true relres {worst:.2e}", flush=True)
    # alone = mix, bit for bit: the smallest, the largest, the widest and one more
    pick = {int(np.argmin(ns)), int(np.argmax(ns)),
            int(np.argmax([B.shape[1] for _, B in systems])), 100}
    for k in sorted(pick):
        A, B = systems[k]
        a = _alone(A, B, **dict(kw, hbm_bytes=4 << 30))
        assert a["iters"] == res[ids[k]]["iters"]
        assert np.array_equal(a["X"], res[ids[k]]["X"]), k
    # the two smallest systems against a direct solve: error within cond(A) tol
    for k in np.argsort(ns)[:2]:
        A, B = systems[k]
        As = _csr(A)
        Xd = spla.spsolve(As.tocsc(), B)
        Xd = Xd.reshape(B.shape)
        cond = np.linalg.cond(As.toarray())
        err = np.linalg.norm(res[ids[k]]["X"] - Xd) / np.linalg.norm(Xd)
        assert err <= 2 * cond * TOL, (k, err, cond)


def test_serving_form_stream_with_the_largest_requests(gpu_lib):
    stream = BT.serving_stream(10, 400.0, seed=4)
    big2 = P.laplacian_2d(512)
    big3 = P.laplacian_3d(128)
    t_last = stream[-1][0]
    stream.append((t_last + 0.001, big2, P.uniform(41, big2.n * 32).reshape(big2.n, 32)))
    stream.append((t_last + 0.002, big3, P.uniform(42, big3.n * 4).reshape(big3.n, 4)))
    print("serving form: " + ", ".join(f"n={A.n} m={B.shape[1]}" for _, A, B in stream),
          flush=True)
    kw = dict(s=4, precision="fp64", hbm_bytes=96 << 30, maxiter=3000)
    b = BT.Batch(**kw)
    res, wall = BT.serve_stream(stream, b, tol=TOL)
    b.close()
    for (_, A, B), r in zip(stream, res):
        assert r["status"] == "ok"
        rr = _relres(A, B, r["X"])
        assert np.all(rr < TOL), (A.n, rr.max())
    lat = np.array([r["latency"] for r in res])
    print(f"serving form: {len(res)} requests in {wall:.2f} s, latency p50 "
          f"{np.percentile(lat, 50):.3f} s, max {lat.max():.3f} s", flush=True)
    # alone = in the stream, bit for bit: the largest 2D request and the smallest one
    small = int(np.argmin([A.n * B.shape[1] for _, A, B in stream]))
    for k in (len(stream) - 2, small):
        _, A, B = stream[k]
        a = _alone(A, B, **kw)
        assert a["iters"] == res[k]["iters"]
        assert np.array_equal(a["X"], res[k]["X"]), k
    # the smallest request against a direct solve
    _, A, B = stream[small]
    As = _csr(A)
    Xd = spla.spsolve(As.tocsc(), B).reshape(B.shape)
    err = np.linalg.norm(res[small]["X"] - Xd) / np.linalg.norm(Xd)
    cond = np.linalg.cond(As.toarray()) if A.n <= 5000 else 4 / (np.pi / (A.n ** 0.5 + 1)) ** 2
    assert err <= 2 * cond * TOL, (err, cond)
