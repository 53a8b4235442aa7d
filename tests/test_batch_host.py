"""Host-side parts of the batching engine, no GPU: the footprint formula
and refusal to create a batch without a device."""

import pytest

from cacg_amd import batch as BT


def _al(x):
    return (x + 255) // 256 * 256


@pytest.mark.parametrize("n,nnz,m,s,prec", [(1024, 4992, 1, 4, "fp32"), (65536, 326656, 32, 4, "fp32"),
                                             (4096, 20000, 8, 2, "fp64")])
def test_footprint_formula(n, nnz, m, s, prec):
    sm = s * m
    f32 = prec == "fp32"
    ntile, nch = -(-n // 64), -(-n // 2048)
    want = sum(_al(b) for b in [(n + 1) * 4, nnz * 4, nnz * 8, nnz * 4 if f32 else 0,
                                n * m * 8, n * m * 8, n * m * 8, n * sm * (4 if f32 else 8),
                                n * sm * 8, n * sm * 8, sm * sm * 8, sm * sm * 8, sm * sm * 8,
                                sm * sm * 8, sm * m * 8, sm * m * 8, nch * sm * (sm + m) * 8,
                                ntile * m * 8])
    assert BT.footprint(n, nnz, m, s, prec) == want


def test_create_without_device_fails_cleanly():
    import cacg_amd
    if cacg_amd.lib().cacg_device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(RuntimeError):
        BT.Batch()
