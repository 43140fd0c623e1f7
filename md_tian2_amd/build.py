"""In-tree build of the native parts.

* ``md_tian2_amd/libmdt_gpu.so``: HIP kernels + C-ABI, hipcc for gfx950;
* ``md_tian2_amd/fortran/*``: the Fortran host (ISO_C_BINDING module, the
  compute_emt drop-in and the batched driver), when flang is available;
* ``oracle/liboracle_mdt.so``: the CPU oracle (test infrastructure), gcc.

Run ``python -m md_tian2_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "md_tian2_amd")
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB = os.path.join(PKG, "libmdt_gpu.so")

HIP_SOURCES = ["mdt_kernels.hip", "mdt_nene.hip", "mdt_observe.hip", "mdt_sample.hip", "mdt_noalloc.hip", "mdt_api.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay disabled on this pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{res.stdout}")
    return res.stdout


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_hip():
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle_mdt.so")


def build_fortran():
    """Compile the Fortran host against libmdt_gpu.so (skipped without flang)."""
    fc = flang()
    mk = os.path.join(FSRC, "Makefile")
    if fc is None or not os.path.exists(mk):
        return None
    _run(["make", "-s", "-C", FSRC, f"FC={fc}"])
    return FSRC


def build_all(verbose=True):
    out = [build_hip(), build_oracle()]
    f = build_fortran()
    if f:
        out.append(f)
    if verbose:
        for p in out:
            print("built", p)
    return out


if __name__ == "__main__":
    build_all()
    sys.exit(0)
