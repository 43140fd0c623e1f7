"""GPU run of the Fortran batched driver (mdt_run) on an md_tian.inp-style
deck, checked trajectory by trajectory against the oracle started from the
driver's dumped initial state with the same noise keys (seed, trajectory
id)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, pesfile

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_dump(path):
    raw = open(path, "rb").read()
    ntraj, natoms, nbeads, nsteps = np.frombuffer(raw[:16], dtype=np.int32)
    off = 16
    tid = np.frombuffer(raw[off:off + 4 * ntraj], dtype=np.int32)
    off += 4 * ntraj
    n3 = 3 * natoms * nbeads
    arrs = []
    for count in (ntraj * n3, ntraj * n3, ntraj * n3, ntraj * n3, ntraj * nbeads, ntraj * 9):
        arrs.append(np.frombuffer(raw[off:off + 8 * count], dtype=np.float64))
        off += 8 * count
    r0, v0, r1, v1, ep, cell = arrs
    shp = (ntraj, natoms, nbeads, 3)
    return dict(ntraj=ntraj, natoms=natoms, nbeads=nbeads, nsteps=nsteps, tid=tid, r0=r0.reshape(shp),
                v0=v0.reshape(shp), r1=r1.reshape(shp), v1=v1.reshape(shp), epot=ep.reshape(ntraj, nbeads),
                cell=cell.reshape(ntraj, 3, 3).transpose(0, 2, 1))


def test_fortran_driver_matches_oracle(pes_hau):
    os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")], cwd=ROOT,
                         capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    d = read_dump(os.path.join(ROOT, "gpurun_out", "mdt_run_h_au.bin"))
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, thermal=False)
    assert np.allclose(d["cell"][0], tmpl.simbox, atol=1e-12)
    cols, pesm = pesfile.emt_tables(pes_hau, tmpl.type_names)
    for t in range(d["ntraj"]):
        s = tmpl.copy()
        s.r = d["r0"][t].copy()
        s.v = d["v0"][t].copy()
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=300.0, tproj=0.0)
        st.calc_force()
        st.run(777, int(d["tid"][t]), 1, int(d["nsteps"]))
        assert np.abs(d["r1"][t] - st.r).max() < 1e-9
        assert np.abs(d["v1"][t] - st.v).max() < 1e-9
        assert abs(d["epot"][t, 0] - st.epot[0]) < 1e-8
