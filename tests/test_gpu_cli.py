"""The command line program (cacg_amd/cacg, the reference's src/main.c on
libcacg.so) against the reference program's own output: same arguments, same
Matrix Market and CSV inputs, and the same printed lines, Loop residuals to
30 decimals included (the reference order is the default), for the three
solver indices.  Only the timing line differs."""

import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
IO = os.path.join(GOLDEN, "io")
CLI = json.load(open(os.path.join(IO, "manifest.json")))["cli"]
EXE = os.path.join(ROOT, "cacg_amd", "cacg")


def _run(run, env_extra=None, eps=None):
    spec = CLI[run]
    args = [EXE, IO + os.sep, "cli_lap2d16.mtx", os.path.join(IO, f"cli_lap2d16_{spec['m']}.csv"),
            str(spec["m"]), str(spec["solver"]), str(spec["s"])]
    if eps is not None:
        args.append(repr(eps))
    env = dict(os.environ)
    env.pop("CACG_QUIET", None)
    env.pop("CACG_ORDER", None)
    env.update(env_extra or {})
    print(f"cli {run}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout.replace(IO + os.sep, "<io>/")


@pytest.mark.parametrize("run", sorted(CLI))
def test_cli_output_matches_reference_program(run):
    want = open(os.path.join(IO, f"cli_{run}.txt")).read().splitlines()
    out = _run(run)
    got = [l for l in out.splitlines() if not l.startswith("Total time for")]
    assert got == want
    assert sum(l.startswith("Total time for") for l in out.splitlines()) == 1


def test_cli_writes_x_and_takes_epsilon(tmp_path):
    import cacg_amd
    xo = tmp_path / "x.csv"
    # ||B(:,0)|| is about 1e9: epsilon 1 asks for a relative residual near 1e-9
    _run("bcbcg_m4_s3", {"CACG_X_OUT": str(xo), "CACG_QUIET": "1"}, eps=1.0)
    X = cacg_amd.read_csv(str(xo), 256, 4)
    A = cacg_amd.read_mtx(os.path.join(IO, "cli_lap2d16.mtx"))["csr"]
    B = cacg_amd.read_csv(os.path.join(IO, "cli_lap2d16_4.csv"), 256, 4)
    r = np.linalg.norm(B[:, 0] - A.to_dense() @ X[:, 0])
    assert r < 10.0
