"""The C-ABI boundary, checked without a GPU: libcacg.so loads, exports every
function include/*.h declares, its types are byte-identical to the
reference's src/DataTypes.h, and the drop-in entry points reject multi-rank
arguments before touching a device."""

import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")
REF_TYPES = "/root/reference/src/DataTypes.h"


def declared_functions():
    names = []
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
        for stmt in text.split(";"):
            stmt = stmt.strip()
            if "(" not in stmt or "typedef" in stmt or "{" in stmt or "}" in stmt and "(" not in stmt.split("}")[-1]:
                stmt = stmt.split("}")[-1].strip()
                if "(" not in stmt or "typedef" in stmt:
                    continue
            m = re.search(r"(\w+)\s*\(", stmt)
            if m:
                names.append(m.group(1))
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_function():
    import cacg_amd
    L = cacg_amd.lib()
    names = declared_functions()
    assert "bcbcg_v1" in names and "cbcg_v1" in names and "cacg_k_gram_fast" in names
    assert "bcg_v1" in names and "cacg_bcg" in names and "cacg_solver_solve_bcg_device" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include HEADER
int main(void) {
  printf("csr %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(csrType_local),
         offsetof(csrType_local, num_rows), offsetof(csrType_local, num_cols),
         offsetof(csrType_local, nnz), offsetof(csrType_local, start),
         offsetof(csrType_local, row_start), offsetof(csrType_local, col_idx),
         offsetof(csrType_local, csrdata));
  printf("dense %zu %zu %zu %zu %zu %zu %zu\n", sizeof(denseType),
         offsetof(denseType, local_num_row), offsetof(denseType, local_num_col),
         offsetof(denseType, global_num_row), offsetof(denseType, global_num_col),
         offsetof(denseType, start_idx), offsetof(denseType, data));
  printf("info %zu %zu %zu %zu\n", sizeof(matInfo), offsetof(matInfo, num_rows),
         offsetof(matInfo, num_cols), offsetof(matInfo, nnz));
  printf("coo %zu %zu %zu %zu\n", sizeof(cooType), offsetof(cooType, rowIdx),
         offsetof(cooType, colIdx), offsetof(cooType, coodata));
  return 0;
}
"""


def _layout(header, tmp_path, tag):
    src = tmp_path / f"layout_{tag}.c"
    exe = tmp_path / f"layout_{tag}"
    src.write_text(LAYOUT_C.replace("HEADER", f'"{header}"'))
    subprocess.run(["gcc", "-std=gnu11", str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_types_byte_identical_to_reference(tmp_path):
    ours = _layout(os.path.join(INCLUDE, "cacg.h"), tmp_path, "ours")
    assert ours.splitlines()[0] == "csr 56 0 8 16 24 32 40 48"
    assert ours.splitlines()[1] == "dense 48 0 8 16 24 32 40"
    assert ours.splitlines()[3] == "coo 24 0 8 16"
    if not os.path.exists(REF_TYPES):
        pytest.skip("reference sources not present (layout checked against the fixed offsets)")
    assert ours == _layout(REF_TYPES, tmp_path, "ref")


OPTS_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "cacg.h"
int main(void) {
  printf("opts %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(cacg_opts),
         offsetof(cacg_opts, stop_mode), offsetof(cacg_opts, eps), offsetof(cacg_opts, tol),
         offsetof(cacg_opts, maxiter), offsetof(cacg_opts, order), offsetof(cacg_opts, precision),
         offsetof(cacg_opts, print_period), offsetof(cacg_opts, inner_tol),
         offsetof(cacg_opts, inner_maxiter), offsetof(cacg_opts, refine_restart),
         offsetof(cacg_opts, bounds), offsetof(cacg_opts, lanczos_steps),
         offsetof(cacg_opts, bound_lo), offsetof(cacg_opts, bound_hi),
         offsetof(cacg_opts, fp32_basis_below));
  printf("report %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(cacg_report),
         offsetof(cacg_report, status), offsetof(cacg_report, iters), offsetof(cacg_report, bounds),
         offsetof(cacg_report, res_hist), offsetof(cacg_report, res_hist_cap),
         offsetof(cacg_report, seconds), offsetof(cacg_report, phase_ms),
         offsetof(cacg_report, phase_calls), offsetof(cacg_report, refinements),
         offsetof(cacg_report, true_relres));
  return 0;
}
"""


def test_python_option_and_report_structs_match_header(tmp_path):
    import cacg_amd
    src, exe = tmp_path / "opts.c", tmp_path / "opts"
    src.write_text(OPTS_C)
    subprocess.run(["gcc", "-std=gnu11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    O, R = cacg_amd.Opts, cacg_amd.Report
    want_o = [ctypes.sizeof(O)] + [getattr(O, f).offset for f, _ in O._fields_]
    want_r = [ctypes.sizeof(R)] + [getattr(R, f).offset for f, _ in R._fields_]
    assert out[0].split()[1:] == [str(v) for v in want_o]
    assert out[1].split()[1:] == [str(v) for v in want_r]


class CsrType(ctypes.Structure):
    _fields_ = [("num_rows", ctypes.c_long), ("num_cols", ctypes.c_long), ("nnz", ctypes.c_long),
                ("start", ctypes.c_long), ("row_start", ctypes.c_void_p),
                ("col_idx", ctypes.c_void_p), ("csrdata", ctypes.c_void_p)]


class DenseType(ctypes.Structure):
    _fields_ = [("local_num_row", ctypes.c_long), ("local_num_col", ctypes.c_long),
                ("global_num_row", ctypes.c_long), ("global_num_col", ctypes.c_long),
                ("start_idx", ctypes.c_long), ("data", ctypes.c_void_p)]


def test_dropin_rejects_multirank_call_without_device():
    import cacg_amd
    L = cacg_amd.lib()
    for fn in (L.bcbcg_v1, L.cbcg_v1):
        fn.argtypes = [CsrType, DenseType, DenseType, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                       ctypes.c_int]
        fn.restype = None
    A = CsrType(4, 4, 4, 0, None, None, None)
    B = DenseType(4, 1, 4, 1, 0, None)
    L.bcbcg_v1(A, B, B, 2, 1.0, 0, 2)     # numprocs = 2
    assert L.cacg_last_status() == 3       # CACG_EINVAL
    L.cbcg_v1(A, B, B, 2, 1.0, 1, 1)       # myid = 1
    assert L.cacg_last_status() == 3
    L.bcg_v1.argtypes = [CsrType, DenseType, DenseType, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.bcg_v1.restype = None
    L.bcg_v1(A, B, B, 1.0, 0, 4)           # numprocs = 4
    assert L.cacg_last_status() == 3
    assert L.cacg_status_string(3) == b"invalid argument"


def test_host_api_rejects_bad_arguments_without_device():
    import numpy as np
    import cacg_amd
    from cacg_amd import problems as P
    A = P.laplacian_2d(4)
    rp = np.ascontiguousarray(A.rowptr)
    rp_bad = rp.copy()
    rp_bad[2] = rp_bad[3] + 1   # decreasing row pointer
    L = cacg_amd.lib()
    lp = ctypes.POINTER(ctypes.c_long)
    dp = ctypes.POINTER(ctypes.c_double)
    B = np.zeros((A.n, 1))
    X = np.zeros((A.n, 1))
    col = np.ascontiguousarray(A.col)
    val = np.ascontiguousarray(A.val)
    rc = L.cacg_bcbcg(A.n, rp_bad.ctypes.data_as(lp), col.ctypes.data_as(lp), val.ctypes.data_as(dp),
                      1, 2, B.ctypes.data_as(dp), X.ctypes.data_as(dp), None, None)
    assert rc == 3
    rc = L.cacg_bcbcg(A.n, rp.ctypes.data_as(lp), col.ctypes.data_as(lp), val.ctypes.data_as(dp),
                      1, 0, B.ctypes.data_as(dp), X.ctypes.data_as(dp), None, None)   # s = 0
    assert rc == 3
    # block CG: bad row pointers, and the FP32 form (FP64 only) are refused
    rc = L.cacg_bcg(A.n, rp_bad.ctypes.data_as(lp), col.ctypes.data_as(lp), val.ctypes.data_as(dp),
                    1, B.ctypes.data_as(dp), X.ctypes.data_as(dp), None, None)
    assert rc == 3
    o = cacg_amd.make_opts(precision="fp32")
    rc = L.cacg_bcg(A.n, rp.ctypes.data_as(lp), col.ctypes.data_as(lp), val.ctypes.data_as(dp),
                    1, B.ctypes.data_as(dp), X.ctypes.data_as(dp), ctypes.byref(o), None)
    assert rc == 3


@pytest.mark.parametrize("bad", [dict(precision="fp32", order="reference"), dict(precision=7),
                                 dict(precision="fp32", inner_tol=-1.0), dict(maxiter=0),
                                 dict(precision="fp32", refine_restart=2), dict(bounds=3),
                                 dict(bounds="given", bound_lo=1.0, bound_hi=1.0),
                                 dict(bounds="lanczos", lanczos_steps=-1),
                                 dict(bounds="lanczos", lanczos_steps=1001)])
def test_options_rejected_without_device(bad):
    """Invalid options are refused (CACG_EINVAL) before any device is touched:
    the FP32-basis form exists in the fast order only."""
    import numpy as np
    import cacg_amd
    from cacg_amd import problems as P
    A = P.laplacian_2d(4)
    L = cacg_amd.lib()
    lp = ctypes.POINTER(ctypes.c_long)
    dp = ctypes.POINTER(ctypes.c_double)
    rp, col, val = (np.ascontiguousarray(a) for a in (A.rowptr, A.col, A.val))
    B, X = np.zeros((A.n, 1)), np.zeros((A.n, 1))
    o = cacg_amd.make_opts(**bad)
    rc = L.cacg_bcbcg(A.n, rp.ctypes.data_as(lp), col.ctypes.data_as(lp), val.ctypes.data_as(dp),
                      1, 2, B.ctypes.data_as(dp), X.ctypes.data_as(dp), ctypes.byref(o), None)
    assert rc == 3


READER_TU_C = r"""
/* a translation unit that includes cacg.h and then declares the reference's
 * reader interface (src/readMTX.h:7-8, src/matrixType.h:5,8 use cooType) */
#include "cacg.h"
void readMtx_info_and_coo(char *path, char *name, matInfo *info, cooType *mat);
void Converter_Coo2Csr(cooType src, csrType_local *dst, matInfo *info);
void delete_cooType(cooType mat);
int main(void) { cooType c = {0, 0, 0}; return c.rowIdx != 0; }
"""


def test_header_declares_every_reference_type(tmp_path):
    src = tmp_path / "reader_tu.c"
    src.write_text(READER_TU_C)
    subprocess.run(["gcc", "-std=gnu11", "-Wall", "-Werror", "-I", INCLUDE, "-c", str(src), "-o",
                    str(tmp_path / "reader_tu.o")], check=True)
