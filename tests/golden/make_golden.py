"""Write the golden fixtures of the s-step Chebyshev (block) CG path.

Runs the reference's own bcbcg_v1 / cbcg_v1 / bcg_v1 (compiled from
/root/reference/src by oracle/ref/Makefile and driven by
oracle/ref/ref_harness.c) at one MPI rank and at four, on seeded inputs from
cacg_amd.problems, and stores its outputs as small .npz fixtures plus a
manifest.  Only data are committed: inputs are regenerated from the manifest
(their SHA-256 is stored to catch drift), outputs are the reference's.

    python tests/golden/make_golden.py [name ...]   # needs oracle/_ref (this container)

With names, only those fixtures are (re)written and the manifest keeps the
other entries.

Per fixture:
  r2_p1        ||R(:,0)||^2 after each outer iteration, reference at 1 rank
  rnorm_cols   ||R(:,c)|| at each iteration's stop test (sequential row-order
               sum of the reference's own R snapshots, the reference's norm
               loop applied to every column); bcbcg / cbcg: the R after the
               iteration's update, bcg: the R the iteration started from
  X            final X at 1 rank
  x_hist       X after iterations 1 and 2 (small cases only)
  r2_p4, X_p4_reldiff   the reference at 4 ranks: its per-iteration ||R(:,0)||^2
               and the relative 2-norm difference of its final X from the 1-rank X
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cacg_amd import problems as P  # noqa: E402
import oracle  # noqa: E402

# name: (matrix kind, size, matrix param, m, s, solver, tol, rhs kind, rhs seed, x0)
CASES = {
    "cbcg_2d8_s4": ("lap2d", 8, 0.0, 1, 4, "cbcg", 1e-8, "uniform", 0, "zero"),
    "cbcg_2d16_s1": ("lap2d", 16, 0.0, 1, 1, "cbcg", 1e-8, "uniform", 0, "zero"),
    "cbcg_2d16_s2": ("lap2d", 16, 0.0, 1, 2, "cbcg", 1e-8, "uniform", 0, "zero"),
    "cbcg_2d32_s4": ("lap2d", 32, 0.0, 1, 4, "cbcg", 1e-8, "uniform", 0, "zero"),
    "cbcg_2d64_s4": ("lap2d", 64, 0.0, 1, 4, "cbcg", 1e-8, "uniform", 0, "zero"),
    "cbcg_2d32_s8_tol1e-10": ("lap2d", 32, 0.0, 1, 8, "cbcg", 1e-10, "uniform", 0, "zero"),
    "bcbcg_2d8_m2_s4": ("lap2d", 8, 0.0, 2, 4, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_2d16_m4_s4": ("lap2d", 16, 0.0, 4, 4, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_2d16_m4_s2": ("lap2d", 16, 0.0, 4, 2, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_2d16_m2_s1": ("lap2d", 16, 0.0, 2, 1, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_2d16_m4_s4_x1": ("lap2d", 16, 0.0, 4, 4, "bcbcg", 1e-8, "uniform", 1, "ones"),
    "bcbcg_2d32_m4_s4_tol1e-10": ("lap2d", 32, 0.0, 4, 4, "bcbcg", 1e-10, "uniform", 1, "zero"),
    "bcbcg_2d32_m16_s8": ("lap2d", 32, 0.0, 16, 8, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_2d20_m3_s3": ("lap2d", 20, 0.0, 3, 3, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_3d8_m4_s4": ("lap3d", 8, 0.0, 4, 4, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_3d12_m16_s8": ("lap3d", 12, 0.0, 16, 8, "bcbcg", 1e-8, "uniform", 1, "zero"),
    "bcbcg_spd16_m4_s4": ("spd2d", 16, 2.0, 4, 4, "bcbcg", 1e-8, "normal", 2, "zero"),
    "bcbcg_shift16_m8_s4": ("lap2d", 16, 0.05, 8, 4, "bcbcg", 1e-10, "uniform", 3, "zero"),
    # bcg_v1 (block CG, no basis: s is unused and recorded as 1)
    "bcg_2d16_m4": ("lap2d", 16, 0.0, 4, 1, "bcg", 1e-8, "uniform", 1, "zero"),
    "bcg_2d32_m1": ("lap2d", 32, 0.0, 1, 1, "bcg", 1e-8, "uniform", 0, "zero"),
    "bcg_2d16_m4_x1": ("lap2d", 16, 0.0, 4, 1, "bcg", 1e-8, "uniform", 1, "ones"),
    "bcg_3d8_m4": ("lap3d", 8, 0.0, 4, 1, "bcg", 1e-8, "uniform", 1, "zero"),
    "bcg_spd16_m4": ("spd2d", 16, 2.0, 4, 1, "bcg", 1e-8, "normal", 2, "zero"),
    "bcg_2d32_m16_tol1e-10": ("lap2d", 32, 0.0, 16, 1, "bcg", 1e-10, "uniform", 1, "zero"),
    "bcg_shift16_m8": ("lap2d", 16, 0.05, 8, 1, "bcg", 1e-10, "uniform", 3, "zero"),
}


def make_problem(kind, size, param, m, rhs_kind, rhs_seed, x0):
    if kind == "lap2d":
        A = P.laplacian_2d(size, shift=param)
    elif kind == "lap3d":
        A = P.laplacian_3d(size, shift=param)
    elif kind == "spd2d":
        A = P.random_spd_like(P.laplacian_2d(size), seed=int(param))
    else:
        raise ValueError(kind)
    B = P.rhs_uniform(A.n, m, rhs_seed) if rhs_kind == "uniform" else P.rhs_normal(A.n, m, rhs_seed)
    X0 = np.zeros_like(B) if x0 == "zero" else np.ones_like(B)
    return A, B, X0


def input_digest(A, B, X0) -> str:
    h = hashlib.sha256()
    for a in (np.asarray(A.rowptr, "<i8"), np.asarray(A.col, "<i8"), np.asarray(A.val, "<f8"),
              np.asarray(B, "<f8"), np.asarray(X0, "<f8")):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    if not oracle.reference_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    oracle.build()
    names = sys.argv[1:] or list(CASES)
    unknown = [n for n in names if n not in CASES]
    if unknown:
        raise SystemExit(f"unknown fixtures: {unknown}")
    mpath = os.path.join(HERE, "manifest.json")
    manifest = {}
    if sys.argv[1:] and os.path.exists(mpath):
        with open(mpath) as f:
            manifest = json.load(f)
    for name in names:
        spec = CASES[name]
        kind, size, param, m, s, solver, tol, rhs_kind, rhs_seed, x0 = spec
        A, B, X0 = make_problem(kind, size, param, m, rhs_kind, rhs_seed, x0)
        eps = tol * float(np.sqrt(oracle.col_norm2sq(B, 0)))
        r1 = oracle.run_reference(A, B, X0, s, eps, solver=solver, nranks=1)
        r4 = oracle.run_reference(A, B, X0, s, eps, solver=solver, nranks=4)
        k = r1["iters"]
        rn = np.array([[np.sqrt(oracle.col_norm2sq(r1["R"][i], c)) for c in range(m)]
                       for i in range(k)])
        data = {
            "r2_p1": r1["r2"], "rnorm_cols": rn, "X": r1["X"],
            "r2_p4": r4["r2"],
            "X_p4_reldiff": np.array(np.linalg.norm(r4["X"] - r1["X"]) / np.linalg.norm(r1["X"])),
        }
        if A.n * m <= 8192:
            data["x_hist"] = r1["x_hist"][:2]
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
        manifest[name] = {
            "matrix": kind, "size": size, "param": param, "m": m, "s": s, "solver": solver,
            "tol": tol, "eps": eps, "rhs": rhs_kind, "rhs_seed": rhs_seed, "x0": x0,
            "n": A.n, "nnz": A.nnz, "iters_p1": k, "iters_p4": r4["iters"],
            "maxiter": 100 if solver == "bcbcg" else 1000,
            "input_sha256": input_digest(A, B, X0),
        }
        print(f"{name}: n={A.n} m={m} s={s} iters p1={k} p4={r4['iters']} "
              f"X p4 reldiff={float(data['X_p4_reldiff']):.1e}")
    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
