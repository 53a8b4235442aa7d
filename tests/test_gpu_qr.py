"""Tall-skinny QR (CholeskyQR2) on the GPU.  The reference's TSQR is empty,
so the kernel is pinned by LAPACK (numpy.linalg.qr, its R and Q normalised
to a positive diagonal) and by the QR invariants: Q^T Q = I, Q R = A, R upper
triangular with a positive diagonal, a manufactured A = Q0 R0 recovered, and
a flag on a rank-deficient block."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _lapack(A):
    Ql, Rl = np.linalg.qr(A)
    d = np.sign(np.diag(Rl))
    d[d == 0] = 1.0
    return Ql * d, Rl * d[:, None]


@pytest.mark.parametrize("n,k", [(5000, 1), (5000, 3), (3001, 16), (4000, 48), (20000, 64),
                                 (65536, 128), (777, 100)])
def test_tsqr_matches_lapack_and_invariants(gpu_lib, n, k):
    A = np.random.default_rng(n + k).standard_normal((n, k))
    Q, R, flag = gpu_lib.k_tsqr(A)
    assert flag == 0
    assert np.all(np.tril(R, -1) == 0.0) and np.all(np.diag(R) > 0)
    assert np.abs(Q.T @ Q - np.eye(k)).max() <= 50 * k * U
    assert np.linalg.norm(Q @ R - A) <= 1e-13 * np.linalg.norm(A)
    Ql, Rl = _lapack(A)
    assert np.linalg.norm(R - Rl) <= 1e-12 * np.linalg.norm(Rl)
    assert np.abs(Q - Ql).max() <= 1e-10


def test_tsqr_recovers_manufactured_factors(gpu_lib):
    rng = np.random.default_rng(7)
    Q0, _ = np.linalg.qr(rng.standard_normal((10000, 32)))
    R0 = np.triu(rng.standard_normal((32, 32))) + 0.0
    R0[np.diag_indices(32)] = np.abs(R0[np.diag_indices(32)]) + 1.0
    Q, R, flag = gpu_lib.k_tsqr(Q0 @ R0)
    assert flag == 0
    assert np.linalg.norm(R - R0) <= 1e-12 * np.linalg.norm(R0)
    assert np.abs(Q - Q0).max() <= 1e-11


def test_tsqr_ill_conditioned_stays_orthogonal(gpu_lib):
    """cond(A) = 1e6: one Cholesky QR pass loses ~u cond^2 = 1e-4 of
    orthogonality; the second pass restores O(u)."""
    rng = np.random.default_rng(11)
    n, k = 20000, 24
    Ua, _ = np.linalg.qr(rng.standard_normal((n, k)))
    Va, _ = np.linalg.qr(rng.standard_normal((k, k)))
    A = (Ua * np.logspace(0, -6, k)) @ Va.T
    Q, R, flag = gpu_lib.k_tsqr(A)
    assert flag == 0
    assert np.abs(Q.T @ Q - np.eye(k)).max() <= 100 * k * U
    assert np.linalg.norm(Q @ R - A) <= 1e-13 * np.linalg.norm(A)


def test_tsqr_flags_rank_deficiency(gpu_lib):
    A = np.random.default_rng(3).standard_normal((2000, 8))
    A[:, 5] = A[:, 2]
    _, _, flag = gpu_lib.k_tsqr(A)
    assert flag == 1
