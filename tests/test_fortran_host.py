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


def _check_input(tmp_path, conf_line, extra=""):
    deck = tmp_path / "deck.inp"
    deck.write_text("\n".join([
        "run md", "start 3", "ntrajs 2", "nsteps 10", "step 0.1",
        "projectile 1 H 1.00794 pil", "lattice 1 C 12.0 ver", "Tsurf 300", "Tproj 300",
        conf_line, f"dump '{tmp_path}/init'", extra]) + "\n")
    res = subprocess.run([EXE, "--check-input", str(deck)], cwd=ROOT, capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    from md_tian2_amd import conf
    return [conf.read_mxt(tmp_path / ("init_%08d.mxt" % t)) for t in (3, 4)]


def test_driver_conf_merge_matches_python_reader(tmp_path):
    """conf merge of the reference's H and graphene files with pip, Einc,
    polar, azimuth and force_beads: the driver's initial universes (written
    by its own mxt writer, no GPU) against the Python reader and merge."""
    import numpy as np
    from md_tian2_amd import conf, geometry
    fix = os.path.join(ROOT, "tests", "golden", "conf")
    out = _check_input(tmp_path, f"conf merge '{fix}/proj' 1 '{fix}/latt' 1",
                       "force_beads 4\nEinc 1.92\npolar 50\nazimuth 30\npip r r 6.0")
    ref = conf.merge(conf.read_mxt(os.path.join(fix, "latt", "mxt_00000001.dat")),
                     conf.read_mxt(os.path.join(fix, "proj", "mxt_00000001.dat")))
    m_h = geometry.MASS_AMU["H"] * geometry.AMU2MASS
    vinc = np.sqrt(2 * 1.92 / m_h)
    th, ph = np.radians(50.0), np.radians(30.0)
    vexp = vinc * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), -np.cos(th)])
    for c in out:
        assert (c.natoms, c.nbeads) == (25, 4) and c.names == ["H", "C"] and c.is_proj == [True, False]
        assert list(c.idx) == list(ref.idx)
        # lattice: the file's positions folded as merge_universes does, in every bead
        assert np.abs(c.r[1:] - ref.r[1:, :1]).max() < 1e-10
        assert np.array_equal(c.v[1:], np.repeat(ref.v[1:, :1], 4, axis=1))
        assert np.array_equal(c.fixed[1:], np.repeat(ref.fixed[1:, :1], 4, axis=1))
        # projectile: z at pip, x y inside the cell, incidence velocity
        assert np.allclose(c.r[0, :, 2], 6.0)
        f = np.linalg.solve(c.simbox, c.r[0, 0])
        assert 0 <= f[0] < 1 and 0 <= f[1] < 1
        assert np.allclose(c.v[0], vexp, rtol=1e-12)
    assert not np.array_equal(out[0].r[0], out[1].r[0])  # x, y drawn per trajectory


def test_driver_conf_poscar_matches_python_reader(tmp_path):
    import numpy as np
    from md_tian2_amd import conf
    fix = os.path.join(ROOT, "tests", "golden", "conf")
    out = _check_input(tmp_path, f"conf poscar '{fix}/example.POSCAR'")
    ref = conf.read_poscar(os.path.join(fix, "example.POSCAR"), nprojectiles=1)
    for c in out:
        assert (c.natoms, c.nbeads) == (25, 1) and c.names == ["H", "C"]
        assert np.array_equal(c.r, ref.cartesian()) and np.array_equal(c.fixed, ref.fixed)
        assert np.allclose(c.simbox, ref.simbox, rtol=1e-15)


def test_driver_reads_output_keywords(tmp_path):
    """output / pul / pll / adsorption_distance parse (a format the driver
    does not write is reported and skipped); malformed ones stop the run."""
    fix = os.path.join(ROOT, "tests", "golden", "conf")
    deck = tmp_path / "deck.inp"
    base = ["run md", "start 1", "ntrajs 1", "nsteps 10", "step 0.1", "projectile 1 H 1.00794 ver",
            "lattice 1 C 12.0 ver", f"conf merge '{fix}/proj' 1 '{fix}/latt' 1", f"dump '{tmp_path}/init'"]
    deck.write_text("\n".join(base + ["output scatter 1 energy 10 xyz 5 adsorption_status 2", "pul 6.5",
                                      "pll -3", "adsorption_distance 1.5 2.5"]) + "\n")
    res = subprocess.run([EXE, "--check-input", str(deck)], cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "output format xyz is not written" in res.stdout
    for bad in ("output scatter", "adsorption_distance 1.5"):
        deck.write_text("\n".join(base + [bad]) + "\n")
        res = subprocess.run([EXE, "--check-input", str(deck)], cwd=ROOT, capture_output=True, text=True,
                             timeout=60)
        assert res.returncode != 0, bad


def test_driver_initial_conditions_equal_the_python_sampler(tmp_path):
    """The driver's conf merge universes (file choice, apply_pip,
    force_beads, set_proj_incidence with Einc and azimuth widths) equal the
    Python path through the same library sampler, trajectory by trajectory."""
    import numpy as np
    from md_tian2_amd import conf, geometry, sampling
    fix = os.path.join(ROOT, "tests", "golden", "conf")
    out = _check_input(tmp_path, f"conf merge '{fix}/proj' 1 '{fix}/latt' 1",
                       "force_beads 4\nEinc 1.92 0.2\npolar 50\nazimuth 30 10\npip r r 6.0\nrng_seed 777")
    for tid, c in zip((3, 4), out):
        kp, kl = sampling.conf_index(777, tid, 0, 1), sampling.conf_index(777, tid, 1, 1)
        proj = conf.read_mxt(conf.mxt_name(os.path.join(fix, "proj"), kp))
        latt = conf.read_mxt(conf.mxt_name(os.path.join(fix, "latt"), kl))
        proj.r, proj.is_cart = proj.cartesian(), True
        sampling.apply_pip(proj.r, proj.simbox, sampling.Pip.parse(["r", "r", "6.0"]), 777, tid)
        m = conf.merge(latt, proj)
        r = np.ascontiguousarray(np.repeat(m.r, 4, axis=1))
        v = np.ascontiguousarray(np.repeat(m.v, 4, axis=1))
        mass = np.array([1.00794 if i == 1 else 12.0 for i in m.idx]) * geometry.AMU2MASS
        inc = sampling.Incidence.make(1.92, 50.0, "30", sigma_einc=0.2, sigma_azimuth=10.0)
        sampling.set_proj_incidence(v, m.idx - 1, m.is_proj, mass, inc, 777, tid)
        assert np.abs(c.r - r).max() <= 1e-12
        assert np.abs(c.v - v).max() <= 1e-15
