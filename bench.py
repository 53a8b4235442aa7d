#!/usr/bin/env python3
"""Benchmark: BCBCG solves/s, p50/p99 latency and basis-SpMM HBM GB/s on one
MI355X.

Workload (a "step" = one full solve): 3D 7-point Dirichlet Laplacian on a
128^3 grid (n = 2,097,152, nnz = 14,581,760, FP64 CSR, int32 indices), 16
right-hand sides U[0,1) from seed 1 (n x 16 row-major), x0 = 0, Chebyshev
basis s = 8, stop when every column's recursive residual is below 1e-10
relative.  Inputs are resident in HBM before the timed region; each step
resets X to 0 on the device and solves.  Synthetic inputs (no datasets).

    python bench.py [--steps K] [--warmup W] [--tol 1e-10] [--order fast]

Prints one JSON line.  Per-kernel device times come from HIP events the
library records on its own stream around every phase of every solve in the
timed region; the dominant phase is priced against its roofline with the
algorithmic bytes / flops of cacg_amd.metrics.  The CPU baseline times the
reference's own solver (compiled from its sources by oracle/ref/Makefile, run
with MPICH) on this host: setup plus two outer iterations of the same
workload at 16 ranks, projected to the reference's own iteration count, and
full CBCG solves of the 2D 128^2 Laplacian (1 RHS, s = 4, 1e-8) at 1 rank and
16; without the reference build or MPI, the C restatement (oracle/).

Beside the headline line (unless --no-forms), the same JSON object carries
the two further forms, each measured in this run:
  precision_form : the same workload with the FP32 basis and FP64 refinement,
                   its solves/s against the FP64 path and its solution's
                   error against the FP64 solution; then the same Laplacian
                   and its random SPD variant (seed 2) with 16 RHS N(0,1)
                   from seed 2, each against its own FP64 run;
  batching_form  : 256 shifted 2D Laplacians (L log-uniform in [32, 256],
                   shift U(0, 0.1), m from {1, 2, 4, 8, 16, 32}, seed 3,
                   s = 4, 1e-10) in one batch against the same systems solved
                   one at a time through the single-system FP64 path;
  serving_form   : a Poisson stream of 2D (L 64..512) and 3D (edge 32..128)
                   Laplacians with 1..32 RHS (seed 4), continuous batching
                   against one request at a time with the FP64 path.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BCBCG solves/s + p50/p99 latency at 1e-10 relres; basis-SpMM HBM GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--m", type=int, default=16)
    ap.add_argument("--s", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxiter", type=int, default=100)
    ap.add_argument("--order", default="fast", choices=["fast", "reference"])
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu count)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the precision / batching / serving measurements")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="only time one full solve of the benchmark workload by the reference's "
                         "own solver on the host (minutes) and print its JSON")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device-to-device copies of 2 GiB (HBM counter calibration)")
    ap.add_argument("--form", choices=["batching", "serving", "long-rows", "reference-order"],
                    help="only run this form and print its JSON (for profiling it alone)")
    return ap.parse_args()


def copy_probe():
    """Three plain device-to-device copies of 2 GiB: each reads and writes
    2147483648 bytes, the yardstick for FETCH_SIZE / WRITE_SIZE readings."""
    import cacg_amd
    L = cacg_amd.lib()
    nbytes = 2 << 30
    src, dst = cacg_amd.DeviceBuffer(nbytes), cacg_amd.DeviceBuffer(nbytes)
    L.cacg_k_fill_f64(src.ptr, nbytes // 8, 1.0, None)
    L.cacg_dev_sync()
    for _ in range(3):
        if L.cacg_d2d(dst.ptr, src.ptr, nbytes) != 0:
            raise RuntimeError("cacg_d2d failed")
    L.cacg_dev_sync()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))


def cpu_baseline(A, B, s, iters_gpu, threads):
    """The reference's own solver (oracle/_ref/ref_harness, compiled from the
    reference's sources, MPICH) on this host's cores: a bounded sample of the
    benchmark solve at `threads` ranks plus full solves of the CBCG
    configuration; the C restatement with threads as ranks where the
    reference build or MPI is not available."""
    import oracle
    if oracle.reference_available() and os.path.exists(oracle.MPIEXEC):
        try:
            return reference_baseline(A, B, s, threads)
        except Exception as e:  # no MPI launch possible on this host: the restatement
            print(f"bench: reference baseline unavailable ({e}); timing the C restatement",
                  file=sys.stderr, flush=True)
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    t_setup, t_it = oracle.time_sample(A, B, np.zeros_like(B), s, threads, 2)
    proj = t_setup + t_it[0] + max(iters_gpu - 1, 0) * t_it[1]
    return {
        "value": 1.0 / proj,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"C restatement of bcbcg_v1 (oracle/, -O, {threads} row-block ranks as threads) "
                   f"on this host: setup {t_setup:.2f}s + outer iterations 0 and 1 "
                   f"({t_it[0]:.2f}s, {t_it[1]:.2f}s) of the same workload, projected to "
                   f"{iters_gpu} iterations = {proj:.1f}s per solve"),
    }


def reference_full_solve(grid, m, s, tol, ranks):
    """One full solve of the benchmark workload by the reference's own
    bcbcg_v1 (oracle/_ref/ref_harness, compiled from its sources, MPICH) at
    `ranks` MPI ranks on this host, stopped as the reference stops
    (||R(:,0)|| < tol ||B(:,0)||), timed as its main times a solve; the
    harness's per-iteration lines are passed through as they come."""
    import oracle
    from cacg_amd import problems as P
    if not (oracle.reference_available() and os.path.exists(oracle.MPIEXEC)):
        raise SystemExit("bench.py --cpu-full: the reference build or MPI is missing")
    A = P.laplacian_3d(grid)
    B = P.rhs_uniform(A.n, m, 1)
    eps = tol * float(np.linalg.norm(B[:, 0]))
    print(f"bench: reference solver, full solve of the benchmark workload at {ranks} ranks",
          file=sys.stderr, flush=True)
    r = oracle.time_reference(A, B, np.zeros_like(B), s, eps, "bcbcg", ranks, timeout=1500,
                              progress=True)
    secs = r["solve"] if r["solve"] is not None else r["t"][-1]
    return {"workload": f"3D Laplacian {grid}^3, m={m}, s={s}, stop at {tol} relres of "
                        f"column 0", "ranks": ranks, "iters": r["iters"], "seconds": secs,
            "solves_per_s": 1.0 / secs, "setup_plus_iteration0_s": r["t"][0] if r["t"] else None,
            "timing": "barrier, clock, solve, barrier, clock, max over ranks"}


def reference_baseline(A, B, s, ranks):
    """Benchmark solve at `ranks` MPI ranks, stopped after outer iterations 0
    and 1 (epsilon just above the reference's recorded ||R(:,0)|| after
    iteration 1), projected to the reference's own iteration count at 4 ranks
    (51, tests/golden/large); and full CBCG solves of the 2D 128^2 Laplacian
    (1 RHS U[0,1) seed 0, s = 4, 1e-8) at 1 rank and at `ranks`."""
    import oracle
    from cacg_amd import problems as P
    gdir = os.path.join(ROOT, "tests", "golden", "large")
    name = "bcbcg_3d128_m16_s8_tol1e-10"
    with np.load(os.path.join(gdir, f"{name}.npz")) as z:
        r2 = z["r2_p1"]
    with np.load(os.path.join(gdir, f"{name}_p4.npz")) as z:
        iters_ref = len(z["r2_p4"])
    eps = float(np.sqrt(r2[1])) * (1 + 1e-6)
    print(f"bench: reference solver, benchmark workload at {ranks} ranks, 2 iterations",
          file=sys.stderr, flush=True)
    r = oracle.time_reference(A, B, np.zeros_like(B), s, eps, "bcbcg", ranks, timeout=900)
    if r["iters"] != 2:
        raise RuntimeError(f"sample ran {r['iters']} iterations, not 2")
    t0, t1 = r["t"][0], r["t"][1] - r["t"][0]
    proj = t0 + (iters_ref - 1) * t1
    A1 = P.laplacian_2d(128)
    b1 = P.rhs_uniform(A1.n, 1, 0)
    eps1 = 1e-8 * float(np.linalg.norm(b1[:, 0]))
    c1 = {}
    for nr in (1, ranks):
        q = oracle.time_reference(A1, b1, np.zeros_like(b1), 4, eps1, "cbcg", nr, timeout=300)
        c1[f"ranks_{nr}"] = {"seconds": q["solve"] if q["solve"] is not None else q["t"][-1],
                             "iters": q["iters"],
                             "timing": "barrier, clock, solve, barrier, clock, max over ranks"}
    return {
        "value": 1.0 / proj,
        "unit": "solves/s",
        "cores": ranks,
        "kind": "reference",
        "sample": (f"the reference's bcbcg_v1 (compiled from its sources, -O, MPICH) at {ranks} MPI "
                   f"ranks on this host, same workload: setup + outer iteration 0 {t0:.2f}s, "
                   f"iteration 1 {t1:.2f}s, projected to its own {iters_ref} iterations at 1e-10 "
                   f"(4-rank count) = {proj:.1f}s per solve"),
        "c1_full_solve": c1,
    }


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_full:
        print(json.dumps(reference_full_solve(args.grid, args.m, args.s, args.tol,
                                              args.cpu_threads or min(16, os.cpu_count() or 1))))
        return
    if args.form:
        run = {"batching": batching_form, "long-rows": spmm_long_rows,
               "reference-order": reference_order_form,
               "serving": lambda: serving_form(args.serve_requests, args.serve_rate)}[args.form]
        print(json.dumps(run()))
        return
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import cacg_amd
    from cacg_amd import metrics as M
    from cacg_amd import problems as P

    L = cacg_amd.lib()
    if L.cacg_device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    n_grid, m, s = args.grid, args.m, args.s
    A = P.laplacian_3d(n_grid)
    B = P.rhs_uniform(A.n, m, 1)
    n, nnz, sm = A.n, A.nnz, s * m
    print(f"bench: 3D Laplacian {n_grid}^3 n={n} nnz={nnz} m={m} s={s} tol={args.tol} "
          f"order={args.order}", file=sys.stderr, flush=True)

    sv = cacg_amd.Solver(A, m, s, tol=args.tol, maxiter=args.maxiter, order=args.order)
    dB = cacg_amd.DeviceBuffer.from_array(B)
    dX = cacg_amd.DeviceBuffer(B.nbytes)

    def one_solve():
        L.cacg_k_fill_f64(dX.ptr, n * m, 0.0, None)
        L.cacg_dev_sync()
        return sv.solve(dB, dX)

    for w in range(args.warmup):
        rep = one_solve()
        print(f"bench: warmup {w}: {rep['status']} in {rep['iters']} iterations", file=sys.stderr,
              flush=True)

    lat, reps = [], []
    L.cacg_dev_sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        rep = one_solve()
        lat.append(time.perf_counter() - ts)
        reps.append(rep)
        print(f"bench: step {k}: {rep['status']} {rep['iters']} iterations {lat[-1] * 1e3:.1f} ms",
              file=sys.stderr, flush=True)
    L.cacg_dev_sync()
    wall = time.perf_counter() - t0

    X = dX.download(B.shape)
    if any(r["status"] != "ok" for r in reps):
        print("bench: a timed solve did not converge", file=sys.stderr)
    iters = reps[-1]["iters"]
    ph = {k: sum(r["phase_ms"][k] for r in reps) for k in reps[0]["phase_ms"]}
    calls = {k: sum(r["phase_calls"][k] for r in reps) for k in reps[0]["phase_calls"]}
    iters_total = sum(r["iters"] for r in reps)

    # one entry per kernel type: (phase, bound, algorithmic work per launch, launches per phase call)
    nsolve = len(reps)
    it_rest = iters_total - nsolve                       # iterations with Q^T A S and Q beta
    cheb_b = nnz * 12 + 4 * (n + 1) + 3 * n * m * 8      # reads S_{j-1} (gathered) and S_{j-2}, writes S_j
    cheb1_b = nnz * 12 + 4 * (n + 1) + 3 * n * m * 8     # reads R, writes S_0 and S_1
    table = {
        "basis_spmm_cheb_step": ("basis", "hbm", cheb_b, max(s - 2, 0)),
        "basis_spmm_step1": ("basis1", "hbm", cheb1_b, 1),
        "aq_spmm": ("aq", "hbm", M.aq_bytes(n, nnz, sm), 1),
        "gram_QtAQ_QtR": ("gram", "mfma", M.gram_flops(n, sm, sm + m), 1),
        "gram_QtAS": ("gram_as", "mfma", M.gram_flops(n, sm, sm), 1),
        "qupdate": ("qupdate", "mfma", M.qupdate_flops(n, sm), 1),
        "xrupdate": ("xrupdate", "hbm", M.xrupdate_bytes(n, sm, m), 1),
    }
    if args.order == "fast" and cacg_amd.lib().cacg_k_split_supported(sm, m):
        # the X update rides on the next Q update; the X/R update updates R only
        table["qupdate"] = ("qupdate", "mfma", M.qxupdate_flops(n, sm, m), 1)
        table["xrupdate"] = ("xrupdate", "hbm", M.rupdate_bytes(n, sm, m), 1)
    kernel_rates = {}
    for name, (phase, bound, work, per_call) in table.items():
        launches = calls[phase] * per_call
        if launches == 0 or ph[phase] <= 0:
            continue
        avg_ms = ph[phase] / launches
        ach = work / (avg_ms * 1e-3) / (1e9 if bound == "hbm" else 1e12)
        peak = M.HBM_PEAK_GBPS if bound == "hbm" else M.FP64_PEAK_TFLOPS
        kernel_rates[name] = {"bound": bound, "achieved": ach, "peak": peak,
                              "unit": "GB/s" if bound == "hbm" else "TFLOP/s", "frac": ach / peak,
                              "avg_launch_ms": avg_ms, "launches": launches,
                              "total_ms": ph[phase]}
    dom = max(kernel_rates, key=lambda k: kernel_rates[k]["total_ms"])
    roofline = {k: kernel_rates[dom][k] for k in ("bound", "achieved", "peak", "unit", "frac")}
    roofline["kernel"] = dom
    roofline["avg_launch_ms"] = kernel_rates[dom]["avg_launch_ms"]
    roofline["traffic"] = measured_traffic(dom)
    basis_ms = ph["basis"] + ph["basis1"]
    basis_gbps = M.basis_bytes(n, nnz, m, s) * iters_total / (basis_ms * 1e-3) / 1e9

    tr = np.linalg.norm(B - (A_dot(A, X)), axis=0) / np.linalg.norm(B, axis=0)
    flops = M.solve_flops(n, nnz, m, s, iters)
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"BCBCG FP64, 3D 7-point Laplacian {n_grid}^3 (n={n}, nnz={nnz}, int32 CSR), "
                         f"{m} RHS U[0,1) seed 1, x0=0, s={s}, stop at {args.tol:g} relres per column"),
            "n": n, "nnz": nnz, "m": m, "s": s, "tol": args.tol, "order": args.order,
            "parallelism": "single-gpu",
        },
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "iters": iters,
        "status": reps[-1]["status"],
        "true_relres_max": float(tr.max()),
        "basis_spmm_gbps": basis_gbps,
        "gflops": flops / (wall / args.steps) / 1e9,
        "kernel_rates": kernel_rates,
        "phase_ms_per_solve": {k: v / len(reps) for k, v in ph.items()},
        "roofline": roofline,
        "hbm_footprint_gb": sv.footprint() / 1e9,
    }
    sv.close()
    if not args.no_forms:
        out["dropin_host_arrays"] = dropin_host_arrays(A, B, m, s, args, out["ms_per_step"])
        out["c1_cbcg"] = c1_cbcg()
        out["spmm_long_rows"] = spmm_long_rows()
        out["precision_form"] = precision_form(A, B, X, dB, dX, m, s, args, out["ms_per_step"],
                                               out["phase_ms_per_solve"])
        out["batching_form"] = batching_form()
        out["serving_form"] = serving_form(args.serve_requests, args.serve_rate)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"bench: CPU baseline, {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(A, B, s, iters, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def dropin_host_arrays(A, B, m, s, args, ms_device):
    """The same solve through the host-array API (cacg_bcbcg): solver
    creation (HBM allocation, int32 CSR upload, SpMM autotuning), upload of B
    and X, the solve, download of X, teardown, timed end to end (PCIe
    included); then one call of the drop-in entry point bcbcg_v1 itself as a
    caller of the reference links it (csrType_local / denseType by value, its
    stop test ||R(:,0)|| < 1e-10 ||B(:,0)||), in its default reference
    summation order and with CACG_ORDER=fast."""
    import cacg_amd
    lat = []
    for k in range(2):
        t = time.perf_counter()
        r = cacg_amd.bcbcg(A, B, np.zeros_like(B), s, tol=args.tol, maxiter=args.maxiter,
                           order=args.order)
        if k > 0:
            lat.append(time.perf_counter() - t)
    ms = float(np.mean(lat) * 1e3)
    print(f"bench: host-array drop-in {ms:.1f} ms per solve", file=sys.stderr, flush=True)
    out = {"workload": "same solve through cacg_bcbcg with host arrays: create, upload A/B/X, "
                       "solve, download X, destroy",
           "ms_per_solve": ms, "solves_per_s": 1e3 / ms, "device_only_ms": ms_device,
           "status": r["status"], "iters": r["iters"]}
    eps = args.tol * float(np.linalg.norm(B[:, 0]))
    for order in ("reference", "fast"):
        t, it, st = dropin_call(A, B, s, eps, order)
        out[f"bcbcg_v1_{order}_order"] = {"ms_per_solve": t * 1e3, "iters": it, "status": st}
        print(f"bench: bcbcg_v1 ({order} order) {t * 1e3:.1f} ms, {it} iterations",
              file=sys.stderr, flush=True)
    return out


def dropin_call(A, B, s, eps, order):
    """One call of bcbcg_v1 (the reference's signature, by-value structs) with
    host arrays; returns (seconds, iterations, status)."""
    import ctypes
    import cacg_amd

    class CsrType(ctypes.Structure):
        _fields_ = [("num_rows", ctypes.c_long), ("num_cols", ctypes.c_long),
                    ("nnz", ctypes.c_long), ("start", ctypes.c_long),
                    ("row_start", ctypes.c_void_p), ("col_idx", ctypes.c_void_p),
                    ("csrdata", ctypes.c_void_p)]

    class DenseType(ctypes.Structure):
        _fields_ = [("local_num_row", ctypes.c_long), ("local_num_col", ctypes.c_long),
                    ("global_num_row", ctypes.c_long), ("global_num_col", ctypes.c_long),
                    ("start_idx", ctypes.c_long), ("data", ctypes.c_void_p)]

    L = cacg_amd.lib()
    fn = L.bcbcg_v1
    fn.argtypes = [CsrType, DenseType, DenseType, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                   ctypes.c_int]
    fn.restype = None
    rp = np.ascontiguousarray(A.rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(A.col, dtype=np.int64)
    va = np.ascontiguousarray(A.val, dtype=np.float64)
    Bc = np.ascontiguousarray(B)
    X = np.zeros_like(Bc)
    n, m = Bc.shape
    old = {k: os.environ.get(k) for k in ("CACG_ORDER", "CACG_QUIET")}
    os.environ["CACG_QUIET"] = "1"
    os.environ["CACG_ORDER"] = order
    try:
        t = time.perf_counter()
        fn(CsrType(n, n, A.nnz, 0, rp.ctypes.data, ci.ctypes.data, va.ctypes.data),
           DenseType(n, m, n, m, 0, Bc.ctypes.data), DenseType(n, m, n, m, 0, X.ctypes.data),
           s, eps, 0, 1)
        t = time.perf_counter() - t
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return t, int(L.cacg_last_iterations()), cacg_amd.STATUS.get(L.cacg_last_status(), "?")


def reference_order_form(iters=3):
    """A few outer iterations of the reference summation order (the drop-in's
    default: every Gram entry and column norm is one chain over all n rows)
    at the benchmark workload, 128^3 Laplacian, 16 right-hand sides, s = 8:
    one untimed solve of `iters` iterations (a solver's first solve tunes its
    SpMM schedule), then the same solve timed; wall time and per-phase ms per
    iteration, for profiling the reference-order kernels on their own."""
    import cacg_amd
    from cacg_amd import problems as P
    L = cacg_amd.lib()
    A = P.laplacian_3d(128)
    B = P.rhs_uniform(A.n, 16, 1)
    eps = 1e-10 * float(np.sqrt(np.sum(B[:, 0] ** 2)))
    dB = cacg_amd.DeviceBuffer.from_array(B)
    dX = cacg_amd.DeviceBuffer(B.nbytes)
    out = {"workload": "3D Laplacian 128^3, m=16, s=8, reference order"}
    sv = cacg_amd.Solver(A, 16, 8, stop_mode=0, eps=eps, maxiter=iters, order="reference")
    for _ in range(2):   # the second solve is the timed one
        L.cacg_k_fill_f64(dX.ptr, A.n * 16, 0.0, None)
        L.cacg_dev_sync()
        t = time.perf_counter()
        rep = sv.solve(dB, dX)
        L.cacg_dev_sync()
        dt = time.perf_counter() - t
    it = max(rep["iters"], 1)
    out.update({"iters": rep["iters"], "ms_per_iteration": dt * 1e3 / it,
                "phase_ms_per_iteration": {k: round(v / it, 3) for k, v in rep["phase_ms"].items()
                                           if v}})
    print(f"bench: reference order {out['ms_per_iteration']:.1f} ms per iteration",
          file=sys.stderr, flush=True)
    return out


def spmm_long_rows(reps=20):
    """SpMM on an irregular matrix: a 1024^2-grid 5-point pattern plus 64 hub
    rows of power-law length (up to 19,824 entries; problems.power_law_spd),
    k = 16 (a basis step's width) and k = 128 (A Q's), every row vector-per-row
    against the hub rows through the row-split kernel; the same bits asserted.
    Time per launch over `reps` back-to-back launches; GB/s from the
    algorithmic bytes 12 nnz + 4 (n + 1) + 16 n k."""
    import ctypes
    import cacg_amd
    from cacg_amd import problems as P
    L = cacg_amd.lib()
    A = P.power_law_spd(1024, 64, 20000, 7)
    nnz = len(A.col)
    d = np.diff(A.rowptr)
    out = {"matrix": f"power-law SPD, n={A.n}, nnz={nnz}, rows {int(d.min())}..{int(d.max())}, "
                     f"mean {d.mean():.2f}"}
    for k in (16, 128):
        X = ((np.arange(A.n * k) % 1013) * 1e-3 - 0.5).reshape(A.n, k)
        dX = cacg_amd.DeviceBuffer.from_array(X)
        res, Y = {}, {}
        for label, lm in (("vector_per_row", 0), ("row_split", None)):
            dA = cacg_amd.DeviceCSR(A, lm)
            dY = cacg_amd.DeviceBuffer(X.nbytes)
            args = (ctypes.byref(dA.s), k, dX.ptr, k, dY.ptr, k, 0, 0.0, 0.0, None, k, None, k,
                    None, k, None, None)
            if L.cacg_k_spmm_x(*args) != 0:
                raise RuntimeError("spmm launch failed")
            L.cacg_dev_sync()
            t = time.perf_counter()
            for _ in range(reps):
                L.cacg_k_spmm_x(*args)
            L.cacg_dev_sync()
            ms = (time.perf_counter() - t) / reps * 1e3
            by = 12 * nnz + 4 * (A.n + 1) + 16 * A.n * k
            res[label] = {"ms": ms, "gbps": by / (ms * 1e-3) / 1e9, "long_rows": dA.n_long}
            Y[label] = dY.download((A.n, k))
        res["same_bits"] = bool(np.array_equal(Y["vector_per_row"], Y["row_split"]))
        res["speedup"] = res["vector_per_row"]["ms"] / res["row_split"]["ms"]
        out[f"k{k}"] = res
    print(f"bench: long rows k=16 {out['k16']['speedup']:.1f}x, k=128 {out['k128']['speedup']:.1f}x",
          file=sys.stderr, flush=True)
    return out


def c1_cbcg():
    """The CBCG configuration on the GPU: 2D 5-point Laplacian 128^2, 1 RHS
    U[0,1) seed 0, x0 = 0, s = 4, Chebyshev interval from 20 Lanczos steps,
    stop at 1e-8 relative residual; mean of 5 device-resident solves after one."""
    import cacg_amd
    from cacg_amd import problems as P
    L = cacg_amd.lib()
    A = P.laplacian_2d(128)
    b = P.rhs_uniform(A.n, 1, 0)
    sv = cacg_amd.Solver(A, 1, 4, tol=1e-8, maxiter=1000, order="fast", bounds="lanczos",
                         lanczos_steps=20)
    dB, dX = cacg_amd.DeviceBuffer.from_array(b), cacg_amd.DeviceBuffer(b.nbytes)
    lat = []
    for k in range(6):
        L.cacg_k_fill_f64(dX.ptr, A.n, 0.0, None)
        L.cacg_dev_sync()
        t = time.perf_counter()
        rep = sv.solve(dB, dX)
        if k > 0:
            lat.append(time.perf_counter() - t)
    sv.close()
    return {"workload": "CBCG, 2D 5-point Laplacian 128^2, 1 RHS U[0,1) seed 0, s=4, "
                        "Lanczos interval (20 steps), 1e-8 relres",
            "ms_per_solve": float(np.mean(lat) * 1e3), "iters": rep["iters"],
            "status": rep["status"]}


def _reference_order_solution(A, B, s, args):
    """X of the same solve in the reference summation order (bit-identical to
    the reference's bcbcg_v1 at one rank under its own stop test), with the
    bench's stop test: the FP64 oracle the precision form is measured against."""
    import cacg_amd
    r = cacg_amd.bcbcg(A, B, np.zeros_like(B), s, tol=args.tol, maxiter=args.maxiter,
                       order="reference")
    return r["X"]


def _col_err(X, Xr):
    return float((np.linalg.norm(X - Xr, axis=0) / np.linalg.norm(Xr, axis=0)).max())


def precision_form(A, B, X64, dB, dX, m, s, args, ms64, phase64):
    """The FP32-basis / FP64-refined solve of the same workload (the library
    default: FP64 basis until every column's true relative residual is below
    1e-3, FP32 basis after); errors per column (relative 2-norm, maximum over
    columns) against the FP64 fast path and against the reference-order
    solution; the FP32 basis throughout beside it; device time per phase and
    solve of both forms against the FP64 path's (phase64)."""
    import cacg_amd
    L = cacg_amd.lib()
    sv = cacg_amd.Solver(A, m, s, tol=args.tol, maxiter=3 * args.maxiter, order="fast",
                         precision="fp32")
    lat, reps = [], []
    for k in range(1 + 3):
        L.cacg_k_fill_f64(dX.ptr, A.n * m, 0.0, None)
        L.cacg_dev_sync()
        t = time.perf_counter()
        rep = sv.solve(dB, dX)
        L.cacg_dev_sync()
        if k > 0:
            lat.append(time.perf_counter() - t)
            reps.append(rep)
    X = dX.download(B.shape)
    sv.close()
    err = np.linalg.norm(X - X64, axis=0) / np.linalg.norm(X64, axis=0)
    ms = float(np.mean(lat) * 1e3)
    print(f"bench: precision form {ms:.1f} ms per solve, {reps[-1]['iters']} iterations",
          file=sys.stderr, flush=True)
    Xr = _reference_order_solution(A, B, s, args)
    thr = _timed_solves(A, B, m, s, args, "fp32", fp32_basis_below=0.0)
    out = {"workload": "same system, FP32 matrix copy / basis / Chebyshev SpMM once every column's "
                       "true relres is below 1e-3 (FP64 basis before), FP64 Q, A Q, Gram, LU, X, "
                       "R; true residual recomputed in FP64 (residual replacement)",
           "solves_per_s": 1e3 / ms, "ms_per_solve": ms, "vs_fp64_path": ms64 / ms,
           "iters": reps[-1]["iters"], "refinements": reps[-1]["refinements"],
           "status": reps[-1]["status"], "true_relres_max": reps[-1]["true_relres"],
           "err_vs_fp64_solution": float(err.max()),
           "err_vs_reference_order": _col_err(X, Xr),
           "fp64_fast_err_vs_reference_order": _col_err(X64, Xr),
           "fp32_basis_throughout": {
               "ms_per_solve": thr["ms"], "iters": thr["rep"]["iters"],
               "vs_fp64_path": ms64 / thr["ms"], "err_vs_reference_order": _col_err(thr["X"], Xr)},
           "phase_ms_per_solve": {
               "fp64": phase64,
               "precision_form": {k: v for k, v in reps[-1]["phase_ms"].items() if v},
               "fp32_basis_throughout": {k: v for k, v in thr["rep"]["phase_ms"].items() if v}}}
    # the 128^3 Laplacian and its random SPD variant (same pattern, off-diagonals
    # -U(0.5, 1.5) seed 2, diagonal 1.01 x absolute row sum), 16 RHS N(0, 1) seed 2,
    # each against its own FP64 run
    from cacg_amd import problems as P
    out["normal_rhs"] = {}
    for name, mk in (("laplacian", lambda: A), ("random_spd", lambda: P.random_spd_like(A, 2))):
        Am = mk()
        Bn = P.rhs_normal(Am.n, m, 2)
        r64 = _timed_solves(Am, Bn, m, s, args, "fp64")
        r32 = _timed_solves(Am, Bn, m, s, args, "fp32")
        e = np.linalg.norm(r32["X"] - r64["X"], axis=0) / np.linalg.norm(r64["X"], axis=0)
        Xr = _reference_order_solution(Am, Bn, s, args)
        out["normal_rhs"][name] = {
            "fp64_ms_per_solve": r64["ms"], "fp64_iters": r64["rep"]["iters"],
            "ms_per_solve": r32["ms"], "iters": r32["rep"]["iters"],
            "refinements": r32["rep"]["refinements"], "vs_fp64_path": r64["ms"] / r32["ms"],
            "status": r32["rep"]["status"], "true_relres_max": r32["rep"]["true_relres"],
            "err_vs_fp64_solution": float(e.max()),
            "err_vs_reference_order": _col_err(r32["X"], Xr),
            "fp64_fast_err_vs_reference_order": _col_err(r64["X"], Xr)}
        print(f"bench: precision form, {name} N(0,1) RHS: FP64 {r64['ms']:.1f} ms "
              f"({r64['rep']['iters']} it), FP32/refined {r32['ms']:.1f} ms "
              f"({r32['rep']['iters']} it)", file=sys.stderr, flush=True)
    return out


def _timed_solves(A, B, m, s, args, precision, reps=3, **kw):
    """Mean wall time of `reps` solves from x0 = 0 after one untimed solve."""
    import cacg_amd
    L = cacg_amd.lib()
    maxiter = args.maxiter if precision == "fp64" else 3 * args.maxiter
    sv = cacg_amd.Solver(A, m, s, tol=args.tol, maxiter=maxiter, order="fast",
                         precision=precision, **kw)
    dB = cacg_amd.DeviceBuffer.from_array(B)
    dX = cacg_amd.DeviceBuffer(B.nbytes)
    lat = []
    for k in range(1 + reps):
        L.cacg_k_fill_f64(dX.ptr, A.n * m, 0.0, None)
        L.cacg_dev_sync()
        t = time.perf_counter()
        rep = sv.solve(dB, dX)
        L.cacg_dev_sync()
        if k > 0:
            lat.append(time.perf_counter() - t)
    X = dX.download(B.shape)
    sv.close()
    return {"ms": float(np.mean(lat) * 1e3), "rep": rep, "X": X}


def batching_form(count=256):
    """256 shifted 2D Laplacians in one batch vs one at a time (FP64 path).
    All arrive at t = 0; four feeder threads submit them (host-side copies of
    ~0.75 GB in all) while the first rounds already run."""
    import cacg_amd
    from cacg_amd import batch as BT
    systems = BT.batch_systems(count, 3)
    b = BT.Batch(s=4, precision="fp32", hbm_bytes=48 << 30, maxiter=2000)
    out, wall = BT.serve_stream([(0.0, A, Bm) for A, Bm in systems], b, tol=1e-10, feeders=4)
    rounds = b.stats.rounds
    b.close()
    ids = list(range(count))
    res = dict(enumerate(out))
    lat = {i: r["latency"] for i, r in res.items()}
    t1 = time.perf_counter()
    for A, Bm in systems:
        cacg_amd.bcbcg(A, Bm, np.zeros_like(Bm), 4, tol=1e-10, maxiter=2000)
    seq = time.perf_counter() - t1
    L = np.array([lat[i] for i in ids]) * 1e3
    print(f"bench: batching form {count / wall:.1f} solves/s vs {count / seq:.1f} one at a time",
          file=sys.stderr, flush=True)
    return {"workload": f"{count} shifted 2D Laplacians (L log-uniform 32..256, shift U(0,0.1), "
                        "m in {1,2,4,8,16,32}, seed 3), s=4, 1e-10, FP32 basis, all arriving at t=0, "
                        "submitted by four feeder threads",
            "solves_per_s": count / wall, "p50_ms": float(np.percentile(L, 50)),
            "p99_ms": float(np.percentile(L, 99)), "rounds": rounds,
            "converged": int(sum(res[i]["status"] == "ok" for i in ids)),
            "max_true_relres": float(max(res[i]["true_relres"] for i in ids)),
            "one_at_a_time_fp64_solves_per_s": count / seq, "vs_one_at_a_time": seq / wall}


# the serving form's large-system threshold: its stream's systems of 32^3 rows
# and more run the single-system kernels (measured: the batched launches, built
# for many small systems, take twice as long per iteration for a lone system of
# 40,000-130,000 rows); the batching form keeps the library default (2^17)
SERVE_SOLO_ROWS = 32768
# feeder threads submitting the serving stream's requests as they arrive
SERVE_FEEDERS = 4


def serving_form(count, rate):
    """Poisson streams at several offered loads: continuous batching (smallest
    HBM footprint first, at most two large systems in flight, systems of
    SERVE_SOLO_ROWS rows or more on the single-system kernels, requests
    submitted by SERVE_FEEDERS feeder threads as they arrive) against the same stream
    solved one request at a time (FIFO) with the FP64 path.  The
    lowest rate is below the capacity of either server (a few solves/s on
    this mix), so latency there is service time; the higher ones build a
    backlog.  `rate` is the rate also reported at the top level."""
    from cacg_amd import batch as BT
    rates = sorted({2.0, 50.0, float(rate), 2000.0})
    sweep = []
    for r in rates:
        cnt = 16 if r < 10 else count
        stream = BT.serving_stream(cnt, r, seed=4)
        b = BT.Batch(s=4, precision="fp64", hbm_bytes=96 << 30, maxiter=3000, policy="smallest",
                     max_large=2, solo_rows=SERVE_SOLO_ROWS)
        res, wall = BT.serve_stream(stream, b, tol=1e-10, feeders=SERVE_FEEDERS)
        b.close()
        seq, wall_seq = BT.serve_sequential(stream, tol=1e-10, s=4, maxiter=3000)
        lb = np.array([q["latency"] for q in res]) * 1e3
        ls = np.array([q["latency"] for q in seq]) * 1e3
        row = {"rate_per_s": r, "requests": cnt, "solves_per_s": cnt / wall,
               "p50_ms": float(np.percentile(lb, 50)), "p99_ms": float(np.percentile(lb, 99)),
               "converged": int(sum(q["status"] == "ok" for q in res)),
               "sequential_fp64": {"solves_per_s": cnt / wall_seq,
                                   "p50_ms": float(np.percentile(ls, 50)),
                                   "p99_ms": float(np.percentile(ls, 99))}}
        sweep.append(row)
        print(f"bench: serving form at {r:g}/s: {row['solves_per_s']:.2f} solves/s p50 "
              f"{row['p50_ms']:.0f} ms vs sequential {row['sequential_fp64']['solves_per_s']:.2f} "
              f"p50 {row['sequential_fp64']['p50_ms']:.0f} ms", file=sys.stderr, flush=True)
    top = next(x for x in sweep if x["rate_per_s"] == float(rate))
    return dict(top, workload=(f"Poisson streams (seed 4) of 2D L 64..512 and 3D edge 32..128 "
                               "Laplacians, 1..32 RHS, s=4, 1e-10, FP64 basis; continuous batching, "
                               "smallest footprint first, at most 2 large systems in flight, "
                               f"systems of {SERVE_SOLO_ROWS} rows or more on the single-system "
                               f"kernels, submissions from {SERVE_FEEDERS} feeder threads"),
                sweep=sweep)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/<kernel>_c2_traffic.txt: FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, and the copy-kernel scale factors), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_c2_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            key, _, val = line.partition(":")
            kv[key.strip()] = val.strip().split()[0]
    try:
        fetch = float(kv["FETCH_SIZE_KiB"]) * float(kv["fetch_scale"])
        write = float(kv["WRITE_SIZE_KiB"]) * float(kv["write_scale"])
    except (KeyError, ValueError):
        return None
    return (fetch + write) * 1024.0


def A_dot(A, X):
    """A @ X on the host (scipy CSR), for the true-residual check only."""
    import scipy.sparse as sp
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
    return As @ X


if __name__ == "__main__":
    main()
