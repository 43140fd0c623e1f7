"""Thermal slab library on the GPU (md_tian2_amd.thermal): bare Au(111)
slabs equilibrated under Andersen collisions at Tsurf and written as mxt
lattice files, then drawn by the driver's `conf merge`.

Bars: the library's mean kinetic temperature over the mobile degrees of
freedom within 25 K of Tsurf (equipartition; 64 snapshots of 54 degrees
of freedom, standard error about 7 K); fixed atoms never move; the
driver's merged universes carry the library's lattice unchanged up to
whole cell vectors."""
import os
import subprocess

import numpy as np
import pytest

from md_tian2_amd import conf, thermal

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_thermal_library_and_merge(gpu_ctx, tmp_path):
    latt = tmp_path / "latt_300K"
    temps = thermal.thermal_library(gpu_ctx, str(latt), n=3, nlayers=4, nfixed=2, tsurf=300.0, nconf=64, nrep=16,
                                    equil_steps=3000, every=100, dt=1.0, seed=2)
    print(f"thermal library: {len(temps)} slabs, kinetic temperature {temps.mean():.1f} K "
          f"(min {temps.min():.1f}, max {temps.max():.1f})")
    assert len(temps) == 64
    assert abs(temps.mean() - 300.0) < 25.0
    c1 = conf.read_mxt(str(latt / "mxt_00000001.dat"))
    ideal = thermal.slab_system("Au", 3, 4, 2, 300.0, 2, 1)
    assert c1.natoms == 36 and c1.is_proj == [False] and c1.names == ["Au"]
    fx = c1.fixed[:, 0, 0]
    assert np.array_equal(c1.r[fx], ideal.r[fx]) and not c1.v[fx].any()
    assert np.abs(c1.r[~fx] - ideal.r[~fx]).max() > 1e-3      # the mobile atoms moved
    # the driver draws its lattices from the library
    deck = tmp_path / "merge.inp"
    deck.write_text("\n".join([
        "run md", "start 1", "ntrajs 8", "nsteps 20", "step 0.1", "projectile 1 H 1.00794 ver",
        "lattice 1 Au 196.966569 lan", "Tsurf 300", "Tproj 0", "Einc 2.0", "polar 0", "azimuth r",
        "pip r r 5.0", f"pes '{ROOT}/md_tian2_amd/pes/EMT-HAu.pes'",
        f"conf merge '{ROOT}/tests/golden/conf_hau/proj' 1 '{latt}' 64", f"dump '{tmp_path}/d.bin'"]) + "\n")
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, "--check-input", str(deck)], cwd=str(tmp_path), capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    from md_tian2_amd import sampling
    for tid in range(1, 9):
        m = conf.read_mxt(str(tmp_path / ("d.bin_%08d.mxt" % tid)))
        src = conf.read_mxt(conf.mxt_name(str(latt), sampling.conf_index(2, tid, 1, 64)))
        d = np.einsum("ij,npj->npi", m.isimbox, m.r[1:] - src.r)
        assert np.allclose(d, np.round(d), atol=1e-9)
        assert np.array_equal(m.v[1:], src.v)
    res = subprocess.run([exe, str(deck)], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
