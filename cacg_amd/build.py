"""Build cacg_amd/libcacg.so: HIP kernels for gfx950 (hipcc) plus the C11 host
driver (gcc), linked against the HIP runtime; and cacg_amd/cacg, the command
line program on it.  Run as `python -m cacg_amd.build`
or through __graft_entry__.build().  Always recompiles (a few seconds per
kernel file, in parallel)."""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcacg.so")
CLI = os.path.join(HERE, "cacg")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "kernels"),
]
C_FLAGS = [
    "-O2", "-fPIC", "-std=gnu11", "-pthread", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-Wno-unused-macros", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "kernels"),
]


def _sources():
    kern = sorted(os.path.join(CSRC, "kernels", f) for f in os.listdir(os.path.join(CSRC, "kernels"))
                  if f.endswith(".hip"))
    host = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                  if f.endswith(".c"))
    return kern, host


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "--hip-link" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    kern, host = _sources()
    jobs = []
    for src in kern:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj], obj))
    for src in host:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append((["gcc", *C_FLAGS, "-c", src, "-o", obj], obj))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp, "-lpthread",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    # the command line program of the reference (src/main.c), on libcacg.so
    tmp = CLI + ".tmp"
    _run(["gcc", *C_FLAGS, os.path.join(CSRC, "cli", "cacg_main.c"), "-o", tmp, "-L", HERE,
          "-lcacg", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
