"""The input formats (include/cacg_io.h), host code, no GPU.

Bit for bit against what the reference's own reader code builds
(tests/golden/io: readMtx_info_and_coo + Converter_Coo2Csr + parseCSV,
compiled from the reference sources), on files with the quirks that decide the
CSR: scrambled entry order, duplicates, symmetric expansion order, pattern and
integer files, skew-symmetric read as general, CSV spacing, CRLF and short
files.  Round trips through the writers, and the errors where the reference
would abort.  Where the reference build exists (this container), random files
are also checked against the live reference reader."""

import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

IO = os.path.join(GOLDEN, "io")
MANIFEST = json.load(open(os.path.join(IO, "manifest.json")))
READERS = sorted(MANIFEST["readers"])


@pytest.fixture(scope="module")
def cacg():
    import cacg_amd
    cacg_amd.lib()
    return cacg_amd


@pytest.mark.parametrize("name", READERS)
def test_reader_matches_reference_bitwise(cacg, name):
    spec = MANIFEST["readers"][name]
    with np.load(os.path.join(IO, f"{name}.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    out = cacg.read_mtx(os.path.join(IO, spec["mtx"]))
    assert out["num_rows"] == int(g["num_rows"]) and out["num_cols"] == int(g["num_cols"])
    assert np.array_equal(out["rowptr"], g["row_start"])
    assert np.array_equal(out["col"], g["col_idx"])
    assert np.array_equal(out["val"].view(np.int64), g["val"].view(np.int64))
    if spec["csv"]:
        f, rows, cols = spec["csv"]
        B = cacg.read_csv(os.path.join(IO, f), rows, cols)
        assert np.array_equal(B.view(np.int64), g["rhs"].view(np.int64))


def test_symmetric_expansion_order(cacg):
    # (i, j) is followed by (j, i); rows keep file order, no column sort
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "s.mtx")
        with open(p, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n"
                    "3 1 0.5\n1 1 2\n2 1 0.25\n3 3 4\n")
        out = cacg.read_mtx(p)
    assert out["rowptr"].tolist() == [0, 3, 4, 6]
    assert out["col"].tolist() == [2, 0, 1, 0, 0, 2]
    assert out["val"].tolist() == [0.5, 2.0, 0.25, 0.25, 0.5, 4.0]


def test_mtx_round_trip_is_exact(cacg):
    from cacg_amd import problems as P
    A = P.random_spd_like(P.laplacian_2d(10), seed=4)
    rng = np.random.default_rng(1)
    val = A.val * rng.uniform(0.5, 2.0, A.nnz)          # values without short decimals
    val[:7] = [5e-324, -1.7976931348623157e308, 1e-300, -0.0, 1 / 3, 3 * 2.0**-1074, 7.0]
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "a.mtx")
        cacg.write_mtx(p, A.n, A.n, A.rowptr, A.col, val)
        out = cacg.read_mtx(p)
    assert np.array_equal(out["rowptr"], A.rowptr) and np.array_equal(out["col"], A.col)
    assert np.array_equal(out["val"].view(np.int64), val.view(np.int64))


def test_csv_round_trip_is_exact(cacg):
    rng = np.random.default_rng(2)
    B = rng.standard_normal((50, 7)) * 10.0 ** rng.integers(-300, 300, (50, 7))
    B[0, :] = [5e-324, -5e-324, 1.7976931348623157e308, 0.0, -0.0, 1 / 3, 0.1]
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "b.csv")
        cacg.write_csv(p, B)
        text = open(p).read().splitlines()
        R = cacg.read_csv(p, 50, 7)
    # the RHS tools' format: numpy.savetxt(fmt='%.18e', delimiter=',')
    assert text[1] == ",".join("%.18e" % v for v in B[1])
    assert np.array_equal(R.view(np.int64), B.view(np.int64))


def test_csv_reads_numpy_savetxt_files(cacg):
    rng = np.random.default_rng(3)
    B = rng.random((20, 3))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "b.csv")
        np.savetxt(p, B, fmt="%.18e", delimiter=",", newline="\n")
        assert np.array_equal(cacg.read_csv(p, 20, 3), B)


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",          # dense
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",    # complex
    "%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n",       # invalid pair
    "%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n",         # not a matrix
    "%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n",          # bad banner
    "%%MatrixMarket matrix coordinate real\n1 1 1\n1 1 1\n",                 # 4 words
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",         # short
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",         # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n",         # col 0
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 x\n",         # bad value
    "%%MatrixMarket matrix coordinate real general\n2 2\n",                  # size line
])
def test_reader_rejects_what_the_reference_cannot_read(cacg, text):
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "bad.mtx")
        open(p, "w").write(text)
        with pytest.raises(ValueError):
            cacg.read_mtx(p)


def test_missing_files_and_bad_csv(cacg):
    with pytest.raises(OSError):
        cacg.read_mtx("/nonexistent/a.mtx")
    with pytest.raises(OSError):
        cacg.read_csv("/nonexistent/b.csv", 2, 2)
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "b.csv")
        open(p, "w").write("1,2\n3\n")                  # second line: one value of two
        with pytest.raises(ValueError):
            cacg.read_csv(p, 2, 2)


def test_cli_program_is_built_and_checks_arguments():
    exe = os.path.join(ROOT, "cacg_amd", "cacg")
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe, "a", "b"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0 and r.stdout == "Argument setting is wrong.\n"  # src/main.c:37-39


# ---- random files against the live reference reader (this container only) -----------

REF_IO = os.path.join(ROOT, "oracle", "_ref", "io_harness")


def _random_mtx(rng, path):
    n = int(rng.integers(1, 40))
    sym = rng.random() < 0.5
    kind = rng.choice(["real", "integer", "pattern"])
    ents = []
    for _ in range(int(rng.integers(0, 4 * n))):
        i, j = int(rng.integers(0, n)), int(rng.integers(0, n))
        if sym and j > i:
            i, j = j, i
        ents.append((i, j, float(rng.standard_normal() * 10.0 ** rng.integers(-5, 5))))
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {kind} {'symmetric' if sym else 'general'}\n")
        if rng.random() < 0.5:
            f.write("% random comment\n")
        f.write(f"{n} {n} {len(ents)}\n")
        for i, j, v in ents:
            if kind == "pattern":
                f.write(f"{i + 1} {j + 1}\n")
            elif kind == "integer":
                f.write(f"{i + 1} {j + 1} {int(v)}\n")
            else:
                f.write(f"{i + 1} {j + 1} {v!r}\n")


@pytest.mark.skipif(not os.path.exists(REF_IO), reason="reference reader not built here")
def test_random_files_against_live_reference_reader(cacg):
    sys.path.insert(0, GOLDEN)
    from make_io_golden import run_io_harness  # noqa: F401  (reads the harness dump format)
    import make_io_golden as mg
    rng = np.random.default_rng(11)
    with tempfile.TemporaryDirectory() as td:
        for t in range(30):
            p = os.path.join(td, f"r{t}.mtx")
            _random_mtx(rng, p)
            out = cacg.read_mtx(p)
            old = mg.IO
            try:
                mg.IO = td
                g = mg.run_io_harness(os.path.basename(p), None)
            finally:
                mg.IO = old
            assert np.array_equal(out["rowptr"], g["row_start"]), t
            assert np.array_equal(out["col"], g["col_idx"]), t
            assert np.array_equal(out["val"].view(np.int64), g["val"].view(np.int64)), t
