"""SpMM launch schedules (tile order and grid) change no result: every
candidate the solver's autotuner tries gives the same bits as the default, on
a matrix large enough to be tuned; and a tuned solver stays bit-identical to
the oracle in the reference order."""

import numpy as np
import pytest

from cacg_amd import problems as P

pytestmark = pytest.mark.gpu

SCHEDS = [(0, 0), (4096, 0), (8192, 0), (0, 1024), (0, 2048), (0, 4096), (4096, 2048),
          (8192, 4096), (16384, 512, 4), (32768, 512, 4), (64, 64), (2048, 100),
          # passes: uneven chunk counts per XCD, a partial last chunk, more passes than chunks
          (4096, 1000, 3), (64, 64, 5), (8192, 48, 7), (256, 100000, 4)]


@pytest.mark.parametrize("k", [1, 4, 16, 20, 64, 128])
def test_every_schedule_gives_the_same_bits(gpu_lib, k):
    A = P.laplacian_3d(64)                       # 262144 rows: above the tuning threshold
    X = P.rhs_uniform(A.n, k, 3)
    p2 = P.rhs_uniform(A.n, k, 4)
    print(f"spmm schedules k={k}", flush=True)
    base = gpu_lib.k_spmm(A, X, "cheb", 0.7, -0.3, p1=X, p2=p2)
    for sc in SCHEDS:
        Y = gpu_lib.k_spmm(A, X, "cheb", 0.7, -0.3, p1=X, p2=p2, sched=sc)
        assert np.array_equal(Y, base), sc


def test_tuned_solver_reference_order_matches_oracle(gpu_lib, oracle_lib):
    A = P.laplacian_2d(512)                      # 262144 rows: the solver tunes its SpMMs
    B = P.rhs_uniform(A.n, 4, 6)
    X0 = np.zeros_like(B)
    eps = 1e-6 * float(np.linalg.norm(B[:, 0]))
    print("tuned solver vs oracle 2d512", flush=True)
    out = gpu_lib.bcbcg(A, B, X0, 4, stop_mode=0, eps=eps, maxiter=60, order="reference")
    orc = oracle_lib.bcbcg(A, B, X0, 4, eps=eps, maxiter=60)
    assert out["iters"] == orc["iters"]
    assert np.array_equal(out["res_hist"], orc["res_hist"])
    assert np.array_equal(out["X"], orc["X"])
