"""Write the golden fixtures of the input formats and the command line program.

Inputs are small text files written by this script into tests/golden/io/
(committed: they are the fixtures' inputs).  Expected outputs come from the
reference's own code, compiled from /root/reference/src by oracle/ref/Makefile:

  <case>.npz   the CSR that readMtx_info_and_coo + Converter_Coo2Csr build
               (oracle/_ref/io_harness), and for CSV cases the block parseCSV
               reads: num_rows, num_cols, row_start, col_idx, val[, rhs]
  cli_<run>.txt  the lines the reference program (oracle/_ref/KSMs, src/main.c)
               prints for one solver run, minus its timing line

    python tests/golden/make_io_golden.py      # needs oracle/_ref (this container)
"""

from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
IO = os.path.join(HERE, "io")
sys.path.insert(0, ROOT)

from cacg_amd import problems as P  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def _fmt(v):
    return f"{v:.17g}"


def write_inputs():
    """name -> (mtx file, csv file or None, rows, cols) ; files written into io/."""
    os.makedirs(IO, exist_ok=True)
    rng = np.random.default_rng(20240)
    files = {}

    # general, entries in a scrambled order with a duplicate, comments, upper-case banner
    A = P.laplacian_2d(5)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    ent = list(zip(rows, A.col, A.val))
    ent.append((3, 3, 0.25))                    # duplicate of a diagonal entry
    order = rng.permutation(len(ent))
    with open(os.path.join(IO, "general_scrambled.mtx"), "w") as f:
        f.write("%%MatrixMarket MATRIX Coordinate Real General\n% a comment\n%\n")
        f.write(f"{A.n} {A.n} {len(ent)}\n")
        for k in order:
            r, c, v = ent[k]
            f.write(f"{r + 1} {c + 1} {_fmt(v)}\n")
    files["general_scrambled"] = ("general_scrambled.mtx", None)

    # symmetric, lower triangle listed column by column (SuiteSparse order), values
    # with full precision
    n = 12
    M = np.zeros((n, n))
    for i in range(n):
        M[i, i] = 4.0 + rng.random()
        for j in range(max(0, i - 3), i):
            if rng.random() < 0.6:
                M[i, j] = -rng.random()
    with open(os.path.join(IO, "symmetric_lower.mtx"), "w") as f:
        lo = [(i, j, M[i, j]) for j in range(n) for i in range(j, n) if M[i, j] != 0.0]
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{n} {n} {len(lo)}\n")
        for i, j, v in lo:
            f.write(f"{i + 1} {j + 1} {_fmt(v)}\n")
    # RHS in the reference tools' format, with a blank-free %.18e layout
    B = rng.random((n, 3))
    np.savetxt(os.path.join(IO, "symmetric_lower_3.csv"), B, fmt="%.18e", delimiter=",",
               newline="\n")
    files["symmetric_lower"] = ("symmetric_lower.mtx", ("symmetric_lower_3.csv", n, 3))

    # pattern symmetric, whitespace variations, blank line before the size line
    with open(os.path.join(IO, "pattern_symmetric.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n%c\n\n")
        f.write("6   6   9\n")
        for r, c in [(1, 1), (2, 1), (3, 3), (4, 2), (5, 5), (6, 4), (6, 6), (2, 2), (4, 4)]:
            f.write(f"  {r}\t{c}  \n")
    files["pattern_symmetric"] = ("pattern_symmetric.mtx", None)

    # integer general, rectangular, an empty row, entries several to a line
    with open(os.path.join(IO, "integer_rect.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n")
        f.write("4 5 6\n1 5 7 3 1 -2\n3 2 4 1 1 9\n4 4 1\n3 5 -3\n")
    files["integer_rect"] = ("integer_rect.mtx", None)

    # skew-symmetric: the reference reads it as general (no expansion)
    with open(os.path.join(IO, "skew.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate real skew-symmetric\n")
        f.write("3 3 2\n2 1 1.5\n3 2 -2.5e-3\n")
    files["skew"] = ("skew.mtx", None)

    # CSV quirks: single column with a trailing blank-free last line without newline,
    # CRLF line ends, spaces after commas, an extra line beyond `rows`
    with open(os.path.join(IO, "quirks_2.csv"), "w", newline="") as f:
        f.write("1.5, -2.25\r\n3e-3,4\r\n  -7.125e+02,  0.1\r\n5,6\r\n9,9\r\n")
    files["csv_quirks"] = ("skew.mtx", ("quirks_2.csv", 4, 2))
    with open(os.path.join(IO, "short_1.csv"), "w") as f:
        f.write("0.5\n0.25")                     # two lines for four rows: rest zero
    files["csv_short"] = ("skew.mtx", ("short_1.csv", 4, 1))
    return files


CLI_RUNS = {
    # run name: (solver, m, s); matrix cli_lap2d16.mtx (symmetric, SuiteSparse order),
    # RHS 1e8 * U[0,1) so the reference's absolute epsilon 20000 takes a few dozen iterations
    "bcg_m4": (0, 4, 1),
    "cbcg_m1_s4": (1, 1, 4),
    "bcbcg_m4_s3": (2, 4, 3),
}


def write_cli_inputs():
    A = P.laplacian_2d(16, shift=0.01)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    lo = sorted(((c, r, v) for r, c, v in zip(rows, A.col, A.val) if r >= c))
    with open(os.path.join(IO, "cli_lap2d16.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.n} {A.n} {len(lo)}\n")
        for c, r, v in lo:                      # column by column, rows ascending
            f.write(f"{r + 1} {c + 1} {_fmt(v)}\n")
    for m in sorted({v[1] for v in CLI_RUNS.values()}):
        B = 1e8 * P.rhs_uniform(A.n, m, 5)
        np.savetxt(os.path.join(IO, f"cli_lap2d16_{m}.csv"), B, fmt="%.18e", delimiter=",",
                   newline="\n")


def run_io_harness(mtx, csv):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "out.bin")
        cmd = [os.path.join(REF, "io_harness"), os.path.join(IO, mtx), out]
        if csv:
            cmd += [os.path.join(IO, csv[0]), str(csv[1]), str(csv[2])]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, timeout=60)
        raw = np.fromfile(out, dtype=np.uint8)
    assert raw[:8].tobytes() == b"CACGIO01"
    nr, nc, nnz = (int(v) for v in raw[8:32].view("<i8"))
    body = raw[32:]
    pos = 0
    rp = body[pos:pos + 8 * (nr + 1)].view("<i8").copy()
    pos += 8 * (nr + 1)
    ci = body[pos:pos + 8 * nnz].view("<i8").copy()
    pos += 8 * nnz
    va = body[pos:pos + 8 * nnz].view("<f8").copy()
    pos += 8 * nnz
    d = {"num_rows": np.array(nr), "num_cols": np.array(nc), "row_start": rp, "col_idx": ci,
         "val": va}
    if csv:
        r, c = (int(v) for v in body[pos:pos + 16].view("<i8"))
        d["rhs"] = body[pos + 16:pos + 16 + 8 * r * c].view("<f8").reshape(r, c).copy()
    return d


def cli_lines(text):
    """The reference program's output minus its timing line."""
    return [l for l in text.splitlines() if not l.startswith("Total time for")]


def run_reference_cli(solver, m, s):
    io = IO + os.sep
    r = subprocess.run([os.path.join(REF, "KSMs"), io, "cli_lap2d16.mtx",
                        os.path.join(IO, f"cli_lap2d16_{m}.csv"), str(m), str(solver), str(s)],
                       check=True, capture_output=True, text=True, timeout=120, cwd=IO)
    return r.stdout


def main():
    if not os.path.exists(os.path.join(REF, "io_harness")):
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    files = write_inputs()
    manifest = {}
    for name, (mtx, csv) in files.items():
        d = run_io_harness(mtx, csv)
        np.savez_compressed(os.path.join(IO, f"{name}.npz"), **d)
        manifest[name] = {"mtx": mtx, "csv": list(csv) if csv else None}
        print(f"{name}: {int(d['num_rows'])} x {int(d['num_cols'])}, nnz {len(d['val'])}"
              + (f", rhs {d['rhs'].shape}" if csv else ""))
    write_cli_inputs()
    cli = {}
    for run, (solver, m, s) in CLI_RUNS.items():
        # the reference prints the path it was given; the fixture stores the
        # directory-independent lines, with the io/ prefix replaced by <io>/
        out = run_reference_cli(solver, m, s).replace(IO + os.sep, "<io>/")
        lines = cli_lines(out)
        with open(os.path.join(IO, f"cli_{run}.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")
        cli[run] = {"solver": solver, "m": m, "s": s,
                    "loops": sum(1 for l in lines if "Loop:" in l)}
        print(f"cli {run}: {len(lines)} lines, {cli[run]['loops']} Loop lines")
    with open(os.path.join(IO, "manifest.json"), "w") as f:
        json.dump({"readers": manifest, "cli": cli}, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
