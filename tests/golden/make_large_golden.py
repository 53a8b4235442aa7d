"""Write the compact golden record of the benchmark-size BCBCG solve.

Runs the reference's own bcbcg_v1 (compiled from /root/reference/src by
oracle/ref/Makefile, driven by oracle/ref/ref_harness.c) at one MPI rank on
the 3D 7-point Laplacian of a 128^3 grid (n = 2,097,152), 16 right-hand sides
U[0,1) from seed 1, x0 = 0, s = 8, and the reference's stop test
||R(:,0)|| < eps with eps = 1e-10 ||B(:,0)||.  A full solve takes hours of
one CPU core; the run keeps no per-iteration snapshots (CACG_REF_NOSNAP).

The whole solution (268 MB) is not committed.  Stored, as data only:
  r2_p1       ||R(:,0)||^2 after every outer iteration (the reference's own)
  col_r2      ||R(:,c)||^2 of every column after every outer iteration
              (row-order sums of the reference's R, computed by the harness)
  X_sha256    SHA-256 of the final X as little-endian float64, n x m row-major
  X_colnorm2  ||X(:,c)||^2 of the final X (row-order sums)
  X_rows      the final X at rows `rows` (a fixed spread of 64 rows)

With --nranks 4 the same solve runs under `mpiexec -n 4` (the reference's
row-block distribution; its Allreduce sums in another order) and writes
<name>_p4.npz: r2_p4 (||R(:,0)||^2 after every outer iteration; its length is
the 4-rank iteration count), X_rows_p4 and X_colnorm2_p4.

    python tests/golden/make_large_golden.py [--workdir DIR] [--nranks 4]   # this container only
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cacg_amd import problems as P  # noqa: E402
import oracle  # noqa: E402

OUT_DIR = os.path.join(HERE, "large")
NAME = "bcbcg_3d128_m16_s8_tol1e-10"
GRID, M, S, TOL, RHS_SEED = 128, 16, 8, 1e-10, 1


def sample_rows(n: int) -> np.ndarray:
    """64 rows: both ends, the middle, and a fixed stride through the rest."""
    rows = set([0, 1, n // 2, n - 2, n - 1])
    rows.update(int(r) for r in np.linspace(0, n - 1, 59))
    return np.array(sorted(rows)[:64], dtype=np.int64)


def x_digest(X: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(X, dtype="<f8").tobytes()).hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workdir", default=None, help="where the harness's temporary files go")
    ap.add_argument("--nranks", type=int, default=1, choices=[1, 4])
    args = ap.parse_args()
    if not oracle.reference_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    A = P.laplacian_3d(GRID)
    B = P.rhs_uniform(A.n, M, RHS_SEED)
    X0 = np.zeros_like(B)
    eps = TOL * float(np.sqrt(oracle.col_norm2sq(B, 0)))
    t0 = time.time()
    r = oracle.run_reference(A, B, X0, S, eps, solver="bcbcg", nranks=args.nranks,
                             snapshots=args.nranks > 1, timeout=24 * 3600, progress=True,
                             workdir=args.workdir)
    wall = time.time() - t0
    X = r["X"]
    rows = sample_rows(A.n)
    os.makedirs(OUT_DIR, exist_ok=True)
    mpath = os.path.join(OUT_DIR, "manifest.json")
    if args.nranks > 1:
        np.savez_compressed(
            os.path.join(OUT_DIR, f"{NAME}_p4.npz"), r2_p4=r["r2"], rows=rows, X_rows_p4=X[rows].copy(),
            X_colnorm2_p4=np.array([oracle.col_norm2sq(X, c) for c in range(M)]))
        print(f"{NAME} at 4 ranks: {r['iters']} iterations, {wall:.0f} s")
        return
    np.savez_compressed(
        os.path.join(OUT_DIR, f"{NAME}.npz"), r2_p1=r["r2"], col_r2=r["col_r2"],
        X_colnorm2=np.array([oracle.col_norm2sq(X, c) for c in range(M)]), rows=rows,
        X_rows=X[rows].copy())
    meta = {
        NAME: {
            "matrix": "lap3d", "size": GRID, "m": M, "s": S, "solver": "bcbcg", "tol": TOL,
            "eps": eps, "rhs": "uniform", "rhs_seed": RHS_SEED, "x0": "zero", "n": A.n,
            "nnz": A.nnz, "iters_p1": r["iters"], "maxiter": 100, "X_sha256": x_digest(X),
            "reference_wall_s": round(wall, 1),
        }
    }
    with open(mpath, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"{NAME}: {r['iters']} iterations, {wall:.0f} s, X sha256 {meta[NAME]['X_sha256']}")


if __name__ == "__main__":
    main()
