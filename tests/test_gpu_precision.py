"""The FP32-basis precision form: its FP32 kernels and the refinement solver.

Kernels.  FP32 SpMM steps are checked bit for bit against a numpy float32
restatement of the same arithmetic (CSR order, every product and sum rounded
to float32 separately).  The mixed Gram (FP64 A Q against the FP32 basis) and
the update Q = S - Q beta (FP32 S, FP64 Q) multiply and add in FP64: on
small-integer data every product and partial sum is exact, so their results
must equal the exact integer products whatever the summation order (a wrong
lane layout, index or transposed operand cannot pass).

Solver.  The stop test is on the true residual ||B - A X|| computed in FP64,
so the recomputed relative residual must be below tol; the solution's error
against the FP64 solution is bounded by cond(A) * tol (checked with the
Laplacian's exact condition number).  No reference counterpart exists (the
reference is FP64 only), so the FP64 GPU path and exact host arithmetic are
the comparison."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cacg_amd import problems as P

pytestmark = pytest.mark.gpu
f32 = np.float32


def _ell(A):
    """Rows padded to the longest row: (col, val, mask), slots in CSR order."""
    lens = np.diff(A.rowptr)
    w = int(lens.max())
    col = np.zeros((A.n, w), dtype=np.int64)
    val = np.zeros((A.n, w))
    mask = np.zeros((A.n, w), dtype=bool)
    for t in range(w):
        has = lens > t
        idx = A.rowptr[:-1][has] + t
        col[has, t] = A.col[idx]
        val[has, t] = A.val[idx]
        mask[has, t] = True
    return col, val, mask


def spmm_f32_ref(A, X):
    """float32 A X, each row summed in CSR order, products and sums rounded
    to float32 separately."""
    col, val, mask = _ell(A)
    v32 = val.astype(f32)
    X = X.astype(f32)
    acc = np.zeros((A.n, X.shape[1]), dtype=f32)
    for t in range(col.shape[1]):
        prod = (v32[:, t:t + 1] * X[col[:, t]]).astype(f32)
        acc = np.where(mask[:, t:t + 1], (acc + prod).astype(f32), acc)
    return acc


MATS = {
    "lap3d_7": lambda: P.laplacian_3d(7),
    "spd_13": lambda: P.random_spd_like(P.laplacian_2d(13), seed=2),
    "lap2d_40": lambda: P.laplacian_2d(40),
}


@pytest.mark.parametrize("mat", sorted(MATS))
@pytest.mark.parametrize("k", [1, 3, 4, 16, 128])
def test_spmm_f32_bitwise(gpu_lib, mat, k):
    A = MATS[mat]()
    rng = np.random.default_rng(k)
    X = rng.standard_normal((A.n, k)).astype(f32)
    assert np.array_equal(gpu_lib.k_spmm_f32(A, X), spmm_f32_ref(A, X))
    P1, P2 = (rng.standard_normal((A.n, k)).astype(f32) for _ in range(2))
    a, b = 0.37, -0.21
    Y = gpu_lib.k_spmm_f32(A, X, "cheb", a, b, p1=P1, p2=P2)
    AX = spmm_f32_ref(A, X)
    want = ((f32(a) * AX).astype(f32) + (f32(b) * P1).astype(f32)).astype(f32) - P2
    assert np.array_equal(Y, want.astype(f32))


@pytest.mark.parametrize("k", [1, 4, 16])
def test_cheb1_f32_reads_fp64_residual(gpu_lib, k):
    A = P.random_spd_like(P.laplacian_2d(13), seed=2)
    R = np.random.default_rng(k).standard_normal((A.n, k))
    a, b = 0.25, -1.0
    S1, S0 = gpu_lib.k_cheb1_f32(A, R, a, b)
    assert np.array_equal(S0, R.astype(f32))
    AS0 = spmm_f32_ref(A, R.astype(f32))
    want = ((f32(a) * AS0).astype(f32) + (f32(b) * R.astype(f32)).astype(f32)).astype(f32)
    assert np.array_equal(S1, want)


@pytest.mark.parametrize("n,kl,kr", [(1000, 9, 9), (3000, 16, 16), (5000, 128, 128),
                                     (777, 128, 128), (70000, 128, 128)])
def test_gram_f64f32_exact_on_integers(gpu_lib, n, kl, kr):
    rng = np.random.default_rng(n + kl + 7)
    L = rng.integers(-8, 9, size=(n, kl)).astype(np.float64)
    R1 = rng.integers(-8, 9, size=(n, kr)).astype(f32)
    assert np.array_equal(gpu_lib.k_gram_f64f32(L, R1), L.T @ R1.astype(np.float64))


@pytest.mark.parametrize("n,sm", [(333, 9), (1000, 16), (999, 48), (2000, 128), (1003, 8), (777, 4),
                                  (4099, 2)])
def test_qupdate_s32_exact_on_integers(gpu_lib, n, sm):
    rng = np.random.default_rng(sm + 1)
    S = rng.integers(-8, 9, size=(n, sm)).astype(f32)
    Q = rng.integers(-8, 9, size=(n, sm)).astype(np.float64)
    beta = rng.integers(-8, 9, size=(sm, sm)).astype(np.float64)
    exact = S.astype(np.float64) - Q @ beta
    assert np.array_equal(gpu_lib.k_qupdate_s32(S, Q, beta), exact)


def test_cvt_f32_f64_exact(gpu_lib):
    S = np.random.default_rng(3).standard_normal((1001, 17)).astype(f32)
    assert np.array_equal(gpu_lib.k_cvt_f32_f64(S), S.astype(np.float64))


def test_colnorm2_fast(gpu_lib):
    R = np.random.default_rng(5).integers(-9, 10, size=(100001, 7)).astype(np.float64)
    assert np.array_equal(gpu_lib.k_colnorm2_fast(R), (R ** 2).sum(axis=0))


def _lap_cond(shape):
    """Exact condition number of the Dirichlet Laplacian on a grid."""
    lo = sum(2 - 2 * np.cos(np.pi / (L + 1)) for L in shape)
    hi = sum(2 - 2 * np.cos(np.pi * L / (L + 1)) for L in shape)
    return hi / lo


PREC_CASES = {
    # name: (matrix, m, s, rhs, cond bound)
    "lap3d16_m16_s8": (lambda: P.laplacian_3d(16), 16, 8, "normal", _lap_cond((16, 16, 16))),
    "lap2d32_m4_s4": (lambda: P.laplacian_2d(32), 4, 4, "uniform", _lap_cond((32, 32))),
    "lap2d24_m3_s3": (lambda: P.laplacian_2d(24), 3, 3, "uniform", _lap_cond((24, 24))),
    "lap2d20_m2_s1": (lambda: P.laplacian_2d(20), 2, 1, "uniform", _lap_cond((20, 20))),
    "spd3d12_m16_s8": (lambda: P.random_spd_like(P.laplacian_3d(12), seed=2), 16, 8, "normal", None),
}


@pytest.mark.parametrize("switch", [0.0, 1e-3])
@pytest.mark.parametrize("restart", [0, 1])
@pytest.mark.parametrize("name", sorted(PREC_CASES))
def test_fp32_refined_solve(gpu_lib, name, restart, switch):
    """restart 0: residual replacement (the recurrence continues); 1: classic
    iterative refinement (the recurrence restarts at every true residual).
    switch: fp32_basis_below (0: FP32 basis throughout; 1e-3, the default: FP64
    basis until every column's true relative residual is below 1e-3)."""
    mk, m, s, rhs, cond = PREC_CASES[name]
    A = mk()
    B = P.rhs_normal(A.n, m, 2) if rhs == "normal" else P.rhs_uniform(A.n, m, 1)
    tol = 1e-10
    out = gpu_lib.bcbcg(A, B, np.zeros_like(B), s, tol=tol, maxiter=400, precision="fp32",
                        refine_restart=restart, fp32_basis_below=switch)
    assert out["status"] == "ok", out["status"]
    As = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
    X = out["X"]
    relres = np.linalg.norm(B - As @ X, axis=0) / np.linalg.norm(B, axis=0)
    assert relres.max() < tol
    assert abs(out["true_relres"] - relres.max()) <= 1e-3 * relres.max() + 1e-16
    # against the FP64 path and the exact solution
    ref = gpu_lib.bcbcg(A, B, np.zeros_like(B), s, tol=tol, maxiter=400)
    Xe = spla.spsolve(As.tocsc(), B)
    Xe = Xe.reshape(A.n, m)
    if cond is None:
        cond = np.linalg.cond(As.toarray()) if A.n <= 4096 else 1e4
    err = np.linalg.norm(X - Xe, axis=0) / np.linalg.norm(Xe, axis=0)
    assert err.max() <= cond * tol
    err64 = np.linalg.norm(X - ref["X"], axis=0) / np.linalg.norm(ref["X"], axis=0)
    assert err64.max() <= 2 * cond * tol
    print(f"{name} restart={restart} switch={switch:g}: iters {out['iters']} (fp64 "
          f"{ref['iters']}), refinements "
          f"{out['refinements']}, true relres {relres.max():.2e}, err vs fp64 {err64.max():.2e}")


def test_fp32_refined_reuse_is_deterministic(gpu_lib):
    A = P.laplacian_3d(12)
    B = P.rhs_normal(A.n, 16, 2)
    sv = gpu_lib.Solver(A, 16, 8, tol=1e-10, maxiter=200, precision="fp32")
    dB = gpu_lib.DeviceBuffer.from_array(B)
    outs = []
    for _ in range(2):
        dX = gpu_lib.DeviceBuffer.from_array(np.zeros_like(B))
        rep = sv.solve(dB, dX)
        outs.append((rep["iters"], dX.download(B.shape)))
    sv.close()
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])


def test_fp32_refined_maxiter(gpu_lib):
    A = P.laplacian_3d(16)
    B = P.rhs_normal(A.n, 16, 2)
    out = gpu_lib.bcbcg(A, B, np.zeros_like(B), 8, tol=1e-14, maxiter=3, precision="fp32")
    assert out["status"] == "maxiter" and out["iters"] == 3
    assert out["true_relres"] > 0


def test_basis_switch_direction(gpu_lib):
    """fp32_basis_below at or above the starting relative residual (1 from
    x0 = 0) switches to the FP32 basis before the first iteration: the same bits
    as the FP32 basis throughout.  A threshold no residual reaches keeps the
    FP64 basis: within one iteration of the FP64 path (a 2D Laplacian of order
    48, smooth right-hand sides, s = 8; the counts are printed)."""
    A = P.laplacian_2d(48)
    B = P.rhs_uniform(A.n, 8, 1)
    kw = dict(tol=1e-10, maxiter=400, precision="fp32")
    a = gpu_lib.bcbcg(A, B, np.zeros_like(B), 8, fp32_basis_below=0.0, **kw)
    b = gpu_lib.bcbcg(A, B, np.zeros_like(B), 8, fp32_basis_below=2.0, **kw)
    assert a["iters"] == b["iters"] and np.array_equal(a["X"], b["X"])
    c = gpu_lib.bcbcg(A, B, np.zeros_like(B), 8, fp32_basis_below=1e-300, **kw)
    ref = gpu_lib.bcbcg(A, B, np.zeros_like(B), 8, tol=1e-10, maxiter=400)
    print(f"switch: FP32 basis {a['iters']} iterations, FP64 basis in the refine loop "
          f"{c['iters']}, FP64 path {ref['iters']}")
    assert c["status"] == "ok" and abs(c["iters"] - ref["iters"]) <= 1
