"""Host-side parts of the batching engine, no GPU: the footprint formula, the
Python mirrors of the header's structs, refusal of bad options and of a batch
without a device."""

import ctypes

import numpy as np
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


BATCH_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "cacg_batch.h"
int main(void)
{
  printf("opts %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(cacg_batch_opts),
         offsetof(cacg_batch_opts, hbm_bytes), offsetof(cacg_batch_opts, slots),
         offsetof(cacg_batch_opts, s), offsetof(cacg_batch_opts, precision),
         offsetof(cacg_batch_opts, maxiter), offsetof(cacg_batch_opts, max_large),
         offsetof(cacg_batch_opts, graph), offsetof(cacg_batch_opts, tune),
         offsetof(cacg_batch_opts, policy), offsetof(cacg_batch_opts, max_bypass),
         offsetof(cacg_batch_opts, solo_rows));
  printf("stats %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(cacg_batch_stats),
         offsetof(cacg_batch_stats, active), offsetof(cacg_batch_stats, queued),
         offsetof(cacg_batch_stats, finished), offsetof(cacg_batch_stats, admitted),
         offsetof(cacg_batch_stats, retired), offsetof(cacg_batch_stats, rounds),
         offsetof(cacg_batch_stats, arena_used), offsetof(cacg_batch_stats, arena_bytes),
         offsetof(cacg_batch_stats, graph_rounds));
  printf("result %zu %zu %zu %zu %zu\n", sizeof(cacg_batch_result),
         offsetof(cacg_batch_result, status), offsetof(cacg_batch_result, iters),
         offsetof(cacg_batch_result, rounds_queued), offsetof(cacg_batch_result, true_relres));
  return 0;
}
"""


def test_python_batch_structs_match_header(tmp_path):
    import os
    import subprocess
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src, exe = tmp_path / "b.c", tmp_path / "b"
    src.write_text(BATCH_C)
    subprocess.run(["gcc", "-std=gnu11", "-D__HIP_PLATFORM_AMD__", "-I", inc, "-I",
                    "/opt/rocm/include", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line, S in zip(out, (BT.BatchOpts, BT.BatchStats, BT.BatchResult)):
        want = [ctypes.sizeof(S)] + [getattr(S, f).offset for f, _ in S._fields_]
        assert line.split()[1:] == [str(v) for v in want], line


@pytest.mark.parametrize("field,value", [("graph", 2), ("graph", -1), ("tune", 2), ("max_large", -1),
                                         ("s", 0), ("slots", 5000), ("maxiter", 0),
                                         ("policy", 2), ("policy", -1), ("max_bypass", -1),
                                         ("solo_rows", -1)])
def test_batch_options_rejected_without_device(field, value):
    """Invalid batch options are refused (CACG_EINVAL) before any device is touched."""
    o = BT.BatchOpts()
    BT.lib().cacg_batch_default_opts(ctypes.byref(o))
    assert o.graph == 0 and o.tune == 1 and o.max_large == 0
    setattr(o, field, value)
    h = ctypes.c_void_p()
    assert BT.lib().cacg_batch_create(ctypes.byref(h), ctypes.byref(o)) == 3
    assert not h.value


class _FakeBatch:
    """A stand-in for Batch with its call pattern: a request retires two
    rounds after it is submitted.  submit may be called from several threads."""

    def __init__(self, fail_at=None):
        import threading
        self.lock = threading.Lock()
        self.sub, self.ret, self.done = [], [], {}
        self.rounds, self.fail_at = 0, fail_at

    def submit(self, A, B, tol=1e-10):
        with self.lock:
            if self.fail_at is not None and len(self.sub) == self.fail_at:
                raise ValueError("refused")
            rid = len(self.sub)
            self.sub.append((rid, self.rounds, B))
            return rid

    def step(self):
        with self.lock:
            self.rounds += 1
            for rid, r0, B in self.sub:
                if rid not in self.done and self.rounds - r0 >= 2:
                    self.done[rid] = B
                    self.ret.append(rid)
        return {}

    def retired(self):
        with self.lock:
            out, self.ret = self.ret, []
            return out

    def collect(self, rid):
        with self.lock:
            return {"status": "ok", "X": self.done[rid], "iters": 1}


@pytest.mark.parametrize("feeders", [1, 3])
@pytest.mark.parametrize("gap", [0.0, 0.002])
def test_serve_stream_feeders_return_every_request_in_order(feeders, gap):
    stream = [(gap * k, None, np.full((2, 1), float(k))) for k in range(20)]
    res, wall = BT.serve_stream(stream, _FakeBatch(), feeders=feeders)
    assert [float(r["X"][0, 0]) for r in res] == [float(k) for k in range(20)]
    assert all(r["latency"] >= 0 for r in res) and wall >= 0


def test_serve_stream_stops_its_feeders_on_an_error():
    import threading
    before = threading.active_count()
    stream = [(0.001 * k, None, np.zeros((2, 1))) for k in range(30)]
    with pytest.raises(ValueError):
        BT.serve_stream(stream, _FakeBatch(fail_at=5), feeders=3)
    assert threading.active_count() == before
