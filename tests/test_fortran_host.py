"""The Fortran host (md_tian2_amd/fortran): it is built, links
libmdt_gpu.so, and without a GPU stops with the library's message (there
is no CPU fallback)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "md_tian2_amd", "fortran")
EXE = os.path.join(FDIR, "mdt_run")


def test_fortran_objects_built():
    for f in ("mdt_gpu.o", "universe_mod.o", "pes_emt_gpu_mod.o", "mdt_run"):
        assert os.path.exists(os.path.join(FDIR, f)), f


def test_driver_links_the_gpu_library():
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libmdt_gpu.so" in out and "not found" not in out


def test_driver_fails_loudly_without_gpu():
    from md_tian2_amd import mdt
    if mdt.device_count() > 0:
        pytest.skip("a GPU is visible")
    res = subprocess.run([EXE, os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")], cwd=ROOT,
                         capture_output=True, text=True, timeout=60)
    assert res.returncode != 0
    assert "no GPU visible" in res.stdout + res.stderr
