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
