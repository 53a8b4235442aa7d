"""The batching / serving engine on the GPU.

A system's result must be the same bits alone and in any mix (every
per-system sum runs in an order fixed by the system), including when a small
HBM arena makes requests wait for admission; every retired system must meet
its stop test on the true residual; and the solutions must agree with the
single-system solver within cond(A) * tol.  No reference counterpart exists
for batching (the reference solves one system per MPI job)."""

import numpy as np
import pytest
import scipy.sparse as sp

from cacg_amd import problems as P
from cacg_amd import batch as BT

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _mix():
    """Eight small systems of differing n, nnz, m and shift (2D and 3D)."""
    specs = [(P.laplacian_2d(40, 0.05), 1), (P.laplacian_2d(33, 0.01), 4),
             (P.laplacian_3d(9, 0.02), 16), (P.laplacian_2d(64, 0.08), 32),
             (P.random_spd_like(P.laplacian_2d(20), seed=2), 8), (P.laplacian_2d(50, 0.0), 2),
             (P.laplacian_3d(12), 3), (P.laplacian_2d(17, 0.1), 5)]
    out = []
    for k, (A, m) in enumerate(specs):
        out.append((A, P.uniform(100 + k, A.n * m).reshape(A.n, m)))
    return out


def _relres(A, B, X):
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
    return (np.linalg.norm(B - As @ X, axis=0) / np.linalg.norm(B, axis=0)).max()


def _run(systems, **kw):
    b = BT.Batch(**kw)
    ids = [b.submit(A, B, tol=TOL) for A, B in systems]
    res = b.run()
    b.close()
    return [res[i] for i in ids]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_alone_equals_mix_bitwise(gpu_lib, precision):
    systems = _mix()
    mixed = _run(systems, s=4, precision=precision, hbm_bytes=1 << 30)
    for (A, B), r in zip(systems, mixed):
        assert r["status"] == "ok"
        rr = _relres(A, B, r["X"])
        assert rr < TOL and abs(rr - r["true_relres"]) <= 1e-3 * rr
    for k, sysk in enumerate(systems):
        alone = _run([sysk], s=4, precision=precision, hbm_bytes=1 << 28)[0]
        assert alone["iters"] == mixed[k]["iters"]
        assert np.array_equal(alone["X"], mixed[k]["X"]), k


def test_admission_by_footprint_keeps_results(gpu_lib):
    """An arena that holds only a few systems: the rest wait in FIFO order,
    every request still completes, with the same bits as an unconstrained run."""
    systems = _mix()
    fps = [BT.footprint(A.n, A.nnz, B.shape[1], 4) for A, B in systems]
    small = int(max(fps) * 1.05) + 4096   # the largest fits alone: the rest must wait
    free = _run(systems, s=4, hbm_bytes=1 << 30)
    b = BT.Batch(s=4, hbm_bytes=small)
    ids = [b.submit(A, B, tol=TOL) for A, B in systems]
    seen_queue, res = 0, {}
    while len(res) < len(ids):
        st = b.step()
        seen_queue = max(seen_queue, st["queued"])
        assert st["arena_used"] <= st["arena_bytes"]
        for rid in b.retired():
            res[rid] = b.collect(rid)
    b.close()
    assert seen_queue > 0
    assert max(res[i]["rounds_queued"] for i in ids) > 0
    for k, rid in enumerate(ids):
        assert res[rid]["status"] == "ok"
        assert np.array_equal(res[rid]["X"], free[k]["X"])


def test_slots_limit_admission(gpu_lib):
    systems = _mix()[:5]
    a = _run(systems, s=4, slots=2, hbm_bytes=1 << 30)
    c = _run(systems, s=4, slots=64, hbm_bytes=1 << 30)
    for x, y in zip(a, c):
        assert np.array_equal(x["X"], y["X"])


def test_agrees_with_single_system_solver(gpu_lib):
    for A, B in _mix()[:4]:
        r = _run([(A, B)], s=4, precision="fp64", hbm_bytes=1 << 28)[0]
        ref = gpu_lib.bcbcg(A, B, np.zeros_like(B), 4, tol=TOL, maxiter=2000)
        As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n)).toarray()
        cond = np.linalg.cond(As)
        err = np.linalg.norm(r["X"] - ref["X"]) / np.linalg.norm(ref["X"])
        assert err <= 2 * cond * TOL


def test_maxiter_and_x0(gpu_lib):
    A, B = _mix()[3]
    r = _run([(A, B)], s=4, maxiter=2, hbm_bytes=1 << 28)[0]
    assert r["status"] == "maxiter" and r["iters"] == 2
    # a converged X0 retires in the admission round without iterating
    full = _run([(A, B)], s=4, hbm_bytes=1 << 28)[0]
    b = BT.Batch(s=4, hbm_bytes=1 << 28)
    rid = b.submit(A, B, X0=full["X"], tol=TOL)
    out = b.run()[rid]
    b.close()
    assert out["status"] == "ok" and out["iters"] == 0


def test_submit_rejects_bad_systems(gpu_lib):
    b = BT.Batch(s=4, hbm_bytes=1 << 24)
    A = P.laplacian_2d(8)
    with pytest.raises(ValueError):
        b.submit(A, np.ones((A.n, 33)))          # m > 32
    with pytest.raises(ValueError):
        b.submit(P.laplacian_2d(400), np.ones((160000, 16)))   # larger than the arena
    b.close()


def test_batching_form_sample(gpu_lib):
    """A slice of the batching form (shifted 2D Laplacians, FP32 basis)."""
    systems = BT.batch_systems(count=24, seed=3, lmax=96)
    res = _run(systems, s=4, precision="fp32", hbm_bytes=4 << 30)
    for (A, B), r in zip(systems, res):
        assert r["status"] == "ok"
        assert _relres(A, B, r["X"]) < TOL


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_large_systems_alone_equal_mix_and_meet_tol(gpu_lib, precision):
    # two systems above the row count that runs the single-system kernels,
    # mixed with small batched ones: same bits alone and in the mix, and with the
    # online schedule tuning off (each run times its own candidates, under
    # different concurrent work, so the kept schedules may differ)
    big = [(P.laplacian_2d(400, 0.01), 4), (P.laplacian_3d(52, 0.02), 8)]
    systems = [(A, P.uniform(200 + k, A.n * m).reshape(A.n, m)) for k, (A, m) in enumerate(big)]
    systems += _mix()[:3]
    print(f"batch large+small {precision}", flush=True)
    mixed = _run(systems, s=4, precision=precision, hbm_bytes=4 << 30)
    for (A, B), r in zip(systems, mixed):
        assert r["status"] == "ok"
        rr = _relres(A, B, r["X"])
        assert rr < TOL and abs(rr - r["true_relres"]) <= 1e-3 * rr
    for k in range(2):
        alone = _run([systems[k]], s=4, precision=precision, hbm_bytes=4 << 30)[0]
        assert alone["iters"] == mixed[k]["iters"]
        assert np.array_equal(alone["X"], mixed[k]["X"]), k
    untuned = _run(systems[:2], s=4, precision=precision, hbm_bytes=4 << 30, tune=False)
    for k in range(2):
        assert untuned[k]["iters"] == mixed[k]["iters"]
        assert np.array_equal(untuned[k]["X"], mixed[k]["X"]), k


def test_max_large_caps_concurrent_large_systems(gpu_lib):
    # max_large = 1: the second large system waits at the head of the queue
    # (and the small one behind it too) until the first retires; the results
    # are the bits of the uncapped run
    big = [(P.laplacian_2d(400, 0.02), 2), (P.laplacian_3d(52, 0.05), 4)]
    systems = [(A, P.uniform(300 + k, A.n * m).reshape(A.n, m)) for k, (A, m) in enumerate(big)]
    systems += _mix()[:1]
    print("batch max_large", flush=True)
    b = BT.Batch(s=4, precision="fp64", hbm_bytes=4 << 30, max_large=1)
    ids = [b.submit(A, B, tol=TOL) for A, B in systems]
    st = b.step()
    assert st["active"] == 1 and st["queued"] == 2
    res = b.run()
    b.close()
    capped = [res[i] for i in ids]
    free = _run(systems, s=4, precision="fp64", hbm_bytes=4 << 30)
    assert capped[1]["rounds_queued"] > 0 and free[1]["rounds_queued"] == 0
    for x, y in zip(capped, free):
        assert x["status"] == "ok" and np.array_equal(x["X"], y["X"])
    with pytest.raises(RuntimeError):
        BT.Batch(s=4, max_large=-1)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_graph_replays_give_the_launch_by_launch_bits(gpu_lib, precision):
    # Two batches driven in lockstep, one replaying captured rounds (graph) and one
    # launching every kernel on its own; submissions arrive mid-run (the lists change:
    # the graph is dropped and captured again) and a large system iterates on its own
    # streams beside the replays.  Every round retires the same systems, and every
    # result has the same bits and iteration count.
    systems = _mix()
    A0 = P.laplacian_2d(400, 0.02)
    systems.insert(2, (A0, P.uniform(400, A0.n * 2).reshape(A0.n, 2)))
    print(f"batch graph {precision}", flush=True)
    runs = [BT.Batch(s=4, precision=precision, hbm_bytes=2 << 30, graph=g) for g in (False, True)]
    ids = [[], []]
    res = [{}, {}]
    for rnd in range(100000):
        if rnd == 0 or rnd == 4:
            part = systems[:6] if rnd == 0 else systems[6:]
            for k in range(2):
                ids[k] += [runs[k].submit(A, B, tol=TOL) for A, B in part]
        st = [b.step() for b in runs]
        got = [sorted(b.retired()) for b in runs]
        assert got[0] == got[1], rnd
        for k in range(2):
            for rid in got[k]:
                res[k][rid] = runs[k].collect(rid)
        keys = ("active", "queued", "finished", "retired", "rounds")
        assert [st[0][q] for q in keys] == [st[1][q] for q in keys], rnd
        if rnd > 4 and st[0]["active"] == 0 and st[0]["queued"] == 0:
            break
    assert st[0]["graph_rounds"] == 0 and st[1]["graph_rounds"] > 10, st[1]
    for b in runs:
        b.close()
    assert ids[0] == ids[1] and len(res[0]) == len(systems)
    for i in ids[0]:
        x, y = res[0][i], res[1][i]
        assert x["status"] == "ok" and y["status"] == "ok"
        assert x["iters"] == y["iters"] and np.array_equal(x["X"], y["X"]), i


def _sjf_order(fps, max_bypass):
    """Admission order of the smallest-first policy with one slot, every
    request queued before the first round (a restatement of the rule)."""
    queued = list(range(len(fps)))
    byp = [0] * len(fps)
    order = []
    while queued:
        head = queued[0]
        pick = head if byp[head] >= max_bypass else min(queued, key=lambda i: (fps[i], i))
        for i in queued:
            if i < pick:
                byp[i] += 1
        queued.remove(pick)
        order.append(pick)
    return order


@pytest.mark.parametrize("max_bypass", [2, 100])
def test_smallest_first_admission(gpu_lib, max_bypass):
    """One slot, six systems queued largest first: they are admitted (so they
    retire) smallest footprint first, a request overtaken max_bypass times goes
    next, and every result has the bits of the FIFO run."""
    systems = [(P.laplacian_2d(L, 0.02), P.uniform(200 + L, L * L * 2).reshape(L * L, 2))
               for L in (60, 52, 44, 36, 28, 20)]
    fps = [BT.footprint(A.n, A.nnz, B.shape[1], 4) for A, B in systems]
    fifo = _run(systems, s=4, slots=1, hbm_bytes=1 << 30)
    b = BT.Batch(s=4, slots=1, hbm_bytes=1 << 30, policy="smallest", max_bypass=max_bypass)
    ids = [b.submit(A, B, tol=TOL) for A, B in systems]
    order, res = [], {}
    while len(res) < len(ids):
        b.step()
        for rid in b.retired():
            order.append(ids.index(rid))
            res[rid] = b.collect(rid)
    b.close()
    assert order == _sjf_order(fps, max_bypass)
    for k, rid in enumerate(ids):
        assert res[rid]["status"] == "ok"
        assert np.array_equal(res[rid]["X"], fifo[k]["X"])


@pytest.mark.parametrize("feeders", [1, 3])
def test_feeder_thread_submissions_give_the_same_results(gpu_lib, feeders):
    """A stream served with submissions from feeder threads (cacg_batch_submit
    beside cacg_batch_step, and with three feeders beside one another) and with
    submissions between rounds: every request retires with the same iterations
    and the same bits, and converges.  Mixed sizes, one large system (2^17 rows
    or more) among them."""
    stream = []
    for k, (L, m) in enumerate([(40, 4), (96, 8), (64, 16), (368, 8), (48, 1), (80, 32),
                                (72, 2), (56, 12)]):
        A = P.laplacian_2d(L, 0.01 * (k + 1))
        stream.append((0.002 * k, A, P.uniform(70 + k, A.n * m).reshape(A.n, m)))
    assert any(A.n >= 1 << 17 for _, A, _ in stream)
    out = {}
    for fd in (False, True):
        b = BT.Batch(s=4, precision="fp64", hbm_bytes=8 << 30, maxiter=2000, policy="smallest",
                     max_large=1)
        out[fd], _ = BT.serve_stream(stream, b, tol=1e-10, feeder=fd, feeders=feeders)
        b.close()
    for (_, A, B), r0, r1 in zip(stream, out[False], out[True]):
        assert r0["status"] == r1["status"] == "ok"
        assert r0["iters"] == r1["iters"]
        assert np.array_equal(r0["X"], r1["X"])
        D = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
        rr = np.linalg.norm(B - D @ r1["X"], axis=0) / np.linalg.norm(B, axis=0)
        assert np.all(rr < 1e-10)


def test_low_solo_threshold_small_systems(gpu_lib):
    """solo_rows = 1024: systems of a few thousand rows run the single-system
    kernels on solo streams (their norm partials are sized for those kernels,
    not by their row tiles).  Alone and in a mix the same bits; converged."""
    systems = []
    for k, (L, m) in enumerate([(40, 4), (64, 16), (200, 8), (24, 2)]):
        A = P.laplacian_2d(L, 0.02)
        systems.append((A, P.uniform(90 + k, A.n * m).reshape(A.n, m)))
    kw = dict(s=4, precision="fp64", hbm_bytes=4 << 30, maxiter=2000, solo_rows=1024)
    b = BT.Batch(**kw)
    ids = [b.submit(A, B, tol=1e-10) for A, B in systems]
    mix = b.run()
    b.close()
    for (A, B), rid in zip(systems, ids):
        r = mix[rid]
        assert r["status"] == "ok"
        D = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
        assert np.all(np.linalg.norm(B - D @ r["X"], axis=0) / np.linalg.norm(B, axis=0) < 1e-10)
        a = BT.Batch(**kw)
        rid1 = a.submit(A, B, tol=1e-10)
        alone = a.run()[rid1]
        a.close()
        assert alone["iters"] == r["iters"] and np.array_equal(alone["X"], r["X"])
