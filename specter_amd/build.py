"""Build recipe for libspx.so (HIP, gfx950) -- compiles in-tree with hipcc.

``python -m specter_amd.build`` or ``specter_amd.build.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspx.so")
# (source, defines): spx_fft.hip is compiled in three parts (see its header)
SOURCES = [("spx_fft.hip", ["-DSPX_PART=1"]), ("spx_fft.hip", ["-DSPX_PART=2"]),
           ("spx_fft.hip", ["-DSPX_PART=3"]), ("spx_ops.hip", []), ("spx_api.hip", []),
           ("spx_noalloc.hip", []), ("spx_wall.hip", [])]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: libspx cannot be built")


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BIN = os.path.join(FORTRAN_DIR, "spx_hd")


def amdflang() -> str | None:
    for cand in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran binding module and the SPX_HD driver against
    libspx.so (skipped when no Fortran compiler is installed)."""
    fc = amdflang()
    if fc is None:
        return None
    bdir = os.path.join(FORTRAN_DIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    cmd = [fc, "-O2", "-J", bdir, os.path.join(FORTRAN_DIR, "spx_bindings.f90"),
           os.path.join(FORTRAN_DIR, "spx_hd.f90"), "-L", HERE, "-lspx",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-o", FORTRAN_BIN + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("amdflang failed building the SPX_HD driver")
    os.replace(FORTRAN_BIN + ".tmp", FORTRAN_BIN)
    return FORTRAN_BIN


def build(verbose: bool = False) -> str:
    """Compile every HIP source into specter_amd/libspx.so (always rebuilds;
    one hipcc process per source, in parallel), then the Fortran driver that
    links it."""
    odir = os.path.join(CSRC, "_build")
    os.makedirs(odir, exist_ok=True)
    objs, procs = [], []
    for src, defs in SOURCES:
        tag = "".join(d.split("=")[-1] for d in defs)
        obj = os.path.join(odir, os.path.splitext(src)[0] + (f"_{tag}" if tag else "") + ".o")
        objs.append(obj)
        cmd = [hipcc(), *HIPCC_FLAGS, *defs, "-c", "-I", CSRC, os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                      text=True))
    failed = False
    for src, pr in zip(SOURCES, procs):
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building libspx.so")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *[f for f in HIPCC_FLAGS if f != "-O3"], *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libspx.so")
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
