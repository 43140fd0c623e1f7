"""The table-driven FP64 exponential used by the EMT pair kernels
(md_tian2_amd/csrc/mdt_math.h) compiled for the host with gcc and compared
with libm's exp: at most 2 ulp over [-700, 700] (it measures 1 ulp)."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mdt_exp_within_2_ulp():
    src = os.path.join(ROOT, "tests", "native", "exp_check.c")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "exp_check")
        subprocess.run(["gcc", "-O2", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) <= 2.0, out
