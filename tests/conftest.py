import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libcacg.so")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    oracle.lib()
    return oracle


def golden_manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden_case(name):
    """(spec, A, B, X0, data) of one committed fixture; inputs regenerated and
    checked against the stored SHA-256."""
    sys.path.insert(0, GOLDEN)
    from make_golden import input_digest, make_problem
    spec = golden_manifest()[name]
    A, B, X0 = make_problem(spec["matrix"], spec["size"], spec["param"], spec["m"], spec["rhs"],
                            spec["rhs_seed"], spec["x0"])
    assert input_digest(A, B, X0) == spec["input_sha256"], f"{name}: regenerated inputs drifted"
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    return spec, A, B, X0, data


def check_fast_order(name, iters, res_col0, X, spec, g):
    """The fast (split-K MFMA) order against the one-rank reference of a golden
    fixture.  Its sums run in another order, as the reference's own run at 4
    MPI ranks does (golden r2_p4, X_p4_reldiff), and an s-step recurrence
    amplifies such differences once its residual nears the rounding floor.
    Bars (the solver's accuracy targets, where the reference itself meets them):
      iterations within +-1 of the reference's;
      ||R(:,0)|| within 1e-9 relative on every iteration before the
      reference's 4-rank run departs from its 1-rank run by more than 1e-10;
      on every iteration within max(1e-9, 2 x the largest 4-rank departure);
      X within max(1e-10, 2 x the 4-rank X difference), relative 2-norm.
    Prints the measured deviations."""
    assert abs(iters - spec["iters_p1"]) <= 1, (iters, spec["iters_p1"])
    r1, r4 = np.sqrt(g["r2_p1"]), np.sqrt(g["r2_p4"])
    k = min(len(r1), iters)
    dev = np.abs(np.asarray(res_col0)[:k] - r1[:k]) / r1[:k]
    k4 = min(len(r1), len(r4))
    dev4 = np.abs(r4[:k4] - r1[:k4]) / r1[:k4]
    over = np.nonzero(dev4 > 1e-10)[0]
    pre = int(over[0]) if len(over) else k4
    xrel = float(np.linalg.norm(X - g["X"]) / np.linalg.norm(g["X"]))
    x4 = float(g["X_p4_reldiff"])
    print(f"{name}: iterations {iters} (reference {spec['iters_p1']}); ||R(:,0)|| deviation "
          f"max {dev.max():.2e} (4-rank reference {dev4.max():.2e}), before iteration {pre} "
          f"{dev[:pre].max() if pre else 0.0:.2e}; X {xrel:.2e} (4-rank reference {x4:.2e})",
          flush=True)
    assert np.all(dev[:min(pre, k)] <= 1e-9), dev[:pre]
    assert dev.max() <= max(1e-9, 2 * dev4.max()), (dev.max(), dev4.max())
    assert xrel <= max(1e-10, 2 * x4), (xrel, x4)


@pytest.fixture(scope="session")
def gpu_lib():
    import cacg_amd
    L = cacg_amd.lib()          # raises if libcacg.so is missing: no fallback
    if L.cacg_device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return cacg_amd


def lu_restated(lhs, rhs, backward="reference"):
    """The small solve in numpy, operation for operation: Crout factorisation
    (unit-diagonal U, no pivoting; src/DLSsolvers.c:106-208), forward
    substitution with ascending subtractions and the reciprocal pivot applied
    last, then the backward substitution either in the reference's order
    (row by row from the bottom, ascending subtractions) or in the fast order
    (column by column, descending q).  numpy rounds every product and every
    difference separately, as the kernels do."""
    a = np.array(lhs, dtype=np.float64)
    N = a.shape[0]
    for p in range(N):
        inv = 1.0 / a[p, p]
        a[p, p + 1:] = a[p, p + 1:] * inv
        a[p + 1:, p + 1:] = a[p + 1:, p + 1:] - a[p + 1:, p:p + 1] * a[p:p + 1, p + 1:]
    y = np.array(rhs, dtype=np.float64)
    for i in range(N):
        for q in range(i):
            y[i] = y[i] - y[q] * a[i, q]
        y[i] = y[i] * (1.0 / a[i, i])
    x = y.copy()
    if backward == "reference":
        for i in range(N - 2, -1, -1):
            for q in range(i + 1, N):
                x[i] = x[i] - x[q] * a[i, q]
    else:
        for q in range(N - 1, 0, -1):
            x[:q] = x[:q] - a[:q, q:q + 1] * x[q:q + 1]
    return x
