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
C restatement of the reference solver (oracle/) on this host: setup plus two
outer iterations of the same workload at one row-block rank per thread,
projected to the GPU run's iteration count.
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
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device-to-device copies of 2 GiB (HBM counter calibration)")
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
    import oracle
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


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
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

    # algorithmic work per phase over the timed region
    basis_b = M.basis_bytes(n, nnz, m, s) * iters_total
    aq_b = M.aq_bytes(n, nnz, sm) * iters_total
    gram_f = (M.gram_flops(n, sm, sm + m) * iters_total
              + M.gram_flops(n, sm, sm) * (iters_total - len(reps)))
    qup_f = M.qupdate_flops(n, sm) * (iters_total - len(reps))
    xr_b = M.xrupdate_bytes(n, sm, m) * iters_total
    kernels = {
        "basis_spmm": ("hbm", basis_b / (ph["basis"] * 1e-3) / 1e9, calls["basis"] * max(s - 1, 1)),
        "aq_spmm": ("hbm", aq_b / (ph["aq"] * 1e-3) / 1e9, calls["aq"]),
        "gram_mfma": ("mfma", gram_f / (ph["gram"] * 1e-3) / 1e12, calls["gram"]),
        "qupdate_mfma": ("mfma", qup_f / max(ph["qupdate"] * 1e-3, 1e-12) / 1e12, calls["qupdate"]),
        "xrupdate": ("hbm", xr_b / (ph["xrupdate"] * 1e-3) / 1e9, calls["xrupdate"]),
    }
    phase_of = {"basis_spmm": "basis", "aq_spmm": "aq", "gram_mfma": "gram",
                "qupdate_mfma": "qupdate", "xrupdate": "xrupdate"}
    dom = max(kernels, key=lambda k: ph[phase_of[k]])
    bound, ach, launches = kernels[dom]
    peak = M.HBM_PEAK_GBPS if bound == "hbm" else M.FP64_PEAK_TFLOPS
    roofline = {"kernel": dom, "bound": bound, "achieved": ach, "peak": peak,
                "unit": "GB/s" if bound == "hbm" else "TFLOP/s", "frac": ach / peak,
                "traffic": None,
                "avg_launch_ms": ph[phase_of[dom]] / max(launches, 1)}

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
        "basis_spmm_gbps": kernels["basis_spmm"][1],
        "gflops": flops / (wall / args.steps) / 1e9,
        "kernel_rates": {k: {"bound": v[0], "achieved": v[1], "unit": "GB/s" if v[0] == "hbm" else "TFLOP/s",
                             "frac": v[1] / (M.HBM_PEAK_GBPS if v[0] == "hbm" else M.FP64_PEAK_TFLOPS),
                             "avg_launch_ms": ph[phase_of[k]] / max(v[2], 1)}
                         for k, v in kernels.items()},
        "phase_ms_per_solve": {k: v / len(reps) for k, v in ph.items()},
        "roofline": roofline,
        "hbm_footprint_gb": sv.footprint() / 1e9,
    }
    sv.close()
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        print(f"bench: CPU baseline, {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(A, B, s, iters, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def A_dot(A, X):
    """A @ X on the host (scipy CSR), for the true-residual check only."""
    import scipy.sparse as sp
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
    return As @ X


if __name__ == "__main__":
    main()
