"""cacg_lanczos_interval (host code): the extreme eigenvalues of the Lanczos
tridiagonal by Sturm bisection, against closed forms and LAPACK, and the
widening rule the header states."""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cacg():
    import cacg_amd
    cacg_amd.lib()
    return cacg_amd


def _rule(tmin, tmax, g):
    w = tmax - tmin
    lo = tmin - 0.05 * w
    if tmin > 0:
        lo = max(lo, 0.5 * tmin)
    return lo, min(tmax + 0.05 * w, g[1]) if g[1] > tmax else tmax + 0.05 * w


def test_1d_laplacian_tridiagonal_closed_form(cacg):
    # T = tridiag(-1, 2, -1) of order 30: eigenvalues 2 - 2 cos(k pi / 31)
    k = 30
    a, b = np.full(k, 2.0), np.full(k, -1.0)
    lo, hi = 2 - 2 * np.cos(np.pi / 31), 2 - 2 * np.cos(30 * np.pi / 31)
    g = np.array([0.0, 100.0])
    iv = cacg.lanczos_interval(a, b, g)
    want = _rule(lo, hi, g)
    assert np.allclose(iv, want, rtol=1e-12, atol=0)


def test_random_tridiagonal_against_lapack(cacg):
    rng = np.random.default_rng(5)
    for _ in range(20):
        k = int(rng.integers(2, 60))
        a = rng.uniform(1, 5, k)
        b = rng.uniform(0.1, 1.0, k)
        T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
        ev = np.linalg.eigvalsh(T)
        g = np.array([0.0, 1e9])
        iv = cacg.lanczos_interval(a, b, g)
        assert np.allclose(iv, _rule(ev[0], ev[-1], g), rtol=1e-11, atol=1e-14)


def test_gerschgorin_caps_the_upper_end(cacg):
    a, b = np.array([1.0, 3.0]), np.array([0.5, 0.7])
    ev = np.linalg.eigvalsh(np.array([[1.0, 0.5], [0.5, 3.0]]))
    iv = cacg.lanczos_interval(a, b, np.array([0.0, ev[-1] + 1e-3]))
    assert iv[1] == ev[-1] + 1e-3


def test_invariant_subspace_cuts_the_tridiagonal(cacg):
    # beta_1 = 0: only the leading 2 x 2 block counts; the garbage after it is ignored
    a = np.array([2.0, 2.0, 100.0, -50.0])
    b = np.array([1.0, 0.0, 3.0, 1.0])
    iv = cacg.lanczos_interval(a, b, np.array([0.0, 1e9]))
    assert np.allclose(iv, _rule(1.0, 3.0, [0.0, 1e9]), rtol=1e-12)


def test_rejects_non_finite(cacg):
    with pytest.raises(ValueError):
        cacg.lanczos_interval(np.array([1.0, np.nan]), np.array([1.0, 1.0]), np.array([0, 2.0]))


@pytest.mark.parametrize("shift", [-3.0, -1.0, 0.0])
def test_nonpositive_spectrum_keeps_the_smallest_ritz_value_inside(cacg, shift):
    # an indefinite or singular tridiagonal: the lower end stays at
    # t_min - 0.05 w, below t_min (no clip towards t_min / 2)
    k = 12
    a = np.full(k, 2.0) + shift
    b = np.full(k, -1.0)
    T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
    ev = np.linalg.eigvalsh(T)
    iv = cacg.lanczos_interval(a, b, np.array([-10.0, 1e9]))
    assert iv[0] < ev[0] and iv[1] > ev[-1]
    assert np.allclose(iv, _rule(ev[0], ev[-1], [-10.0, 1e9]), rtol=1e-11, atol=1e-13)
