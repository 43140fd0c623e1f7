"""Trajectory characteristics, energy-file estimators and output formats on
the CPU.

* The oracle's estimators (rpmd.f90 / universe_mod.f90 restatement) are
  pinned by closed forms: a ring polymer of P beads on a circle of radius R
  has r_gyr^2 = R^2 and spring energy m wn^2 (2 R sin(pi/P))^2 / 2; equal
  bead velocities make centroid and bead kinetic energies equal; harmonic
  forces -k (r_b - c) give the virial kinetic energy k R^2 / 2; fixed
  components lower the degrees of freedom of the temperatures.
* The oracle's collect_trajectory_characteristics is driven through a
  hand-built sequence of states whose closest approach (through a periodic
  image), lowest point, turning points, interaction time and adsorption
  toggles are known, including the reference's repeated scaling of
  closest_approach_time by the step.
* The oracle's trajectory loop exits a free projectile at the step a
  straight flight predicts.
* The Fortran edit descriptors of the output writers are checked against
  the Fortran compiler's own output for the same numbers."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, observables, pesfile

KB = 8.61733238496e-5
HBAR = 0.6582119514467406


def _universe(r, v, masses, is_proj_types, idx, cell=None, f=None, fixed=None, tsurf=0.0, tproj=0.0, dt=0.1):
    N, P = r.shape[0], r.shape[1]
    ntypes = len(is_proj_types)
    names = [("H" if p else "Au", bool(p)) for p in is_proj_types]
    fx = np.zeros((N, P, 3), dtype=bool) if fixed is None else fixed
    cell = np.eye(3) * 30.0 if cell is None else cell
    s = geometry.System(names, np.array(idx, dtype=np.int32), np.array(masses, dtype=float), fx,
                        np.array([geometry.PROP_VERLET] * ntypes, dtype=np.int32), cell, r.copy(), v.copy())
    st = oracle.OracleState(s, {k: [0.0] * ntypes for k in pesfile.EMT_KEYS}, [[0] * ntypes] * ntypes,
                            step=dt, tsurf=tsurf, tproj=tproj)
    if f is not None:
        st.f[...] = f
    return st


def test_estimators_ring_polymer_closed_forms():
    P, R, m, k = 8, 0.3, 1.5, 2.0
    c = np.array([5.0, 6.0, 7.0])
    phi = 2 * np.pi * np.arange(P) / P
    ring = c + R * np.stack([np.cos(phi), np.sin(phi), np.zeros(P)], axis=1)
    u = np.array([0.01, -0.02, 0.03])
    r = np.stack([ring, ring + [3.0, 0, 0]])          # atom 0 projectile, atom 1 lattice
    v = np.stack([np.tile(u, (P, 1)), np.zeros((P, 3))])
    f = np.stack([-k * (ring - c), np.zeros((P, 3))])
    tp, ts = 120.0, 300.0
    st = _universe(r, v, [m, 7.0], [True, False], [0, 1], f=f, tsurf=ts, tproj=tp)
    row = st.energy_row()
    ek = 0.5 * m * (u @ u)
    assert row[2] == pytest.approx(R * R, rel=1e-13)            # r_gyr^2 projectile
    assert row[3] == pytest.approx(R * R, rel=1e-13)            # lattice ring, same radius
    assert row[4] == pytest.approx(ek, rel=1e-13)               # centroid ekin
    assert row[6] == pytest.approx(ek, rel=1e-13)               # bead ekin
    assert row[5] == 0.0 and row[7] == 0.0
    assert row[8] == pytest.approx(0.5 * k * R * R, rel=1e-12)  # virial quantum ekin
    wp, wl = KB * tp * P / HBAR, KB * ts * P / HBAR
    chord2 = (2 * R * math.sin(math.pi / P)) ** 2
    spring = 0.5 * (m * wp * wp + 7.0 * wl * wl) * P * chord2 / P
    assert row[11] == pytest.approx(spring, rel=1e-12)
    dof = 3 * 2 * P
    assert row[0] == pytest.approx(m * (P * P * (u @ u)) / KB / dof / P, rel=1e-13)
    assert row[1] == pytest.approx(m * P * (u @ u) / KB / dof, rel=1e-13)
    assert np.allclose(row[14:17], c) and np.allclose(row[17:20], u)
    assert row[13] == pytest.approx(f.sum(), abs=1e-14)
    assert row[12] == pytest.approx(row[6] + row[7] + row[10] + row[11], rel=1e-15)


def test_estimators_single_bead_and_fixed_dof():
    r = np.array([[[1.0, 2.0, 3.0]], [[2.0, 2.0, 3.0]]])
    v = np.array([[[0.1, 0.0, 0.0]], [[0.0, 0.2, 0.0]]])
    fixed = np.zeros((2, 1, 3), dtype=bool)
    fixed[1, 0, :2] = True
    st = _universe(r, v, [1.0, 4.0], [True, False], [0, 1], fixed=fixed, tproj=300.0, tsurf=300.0)
    st.epot[:] = [-1.25]
    row = st.energy_row()
    e = 0.5 * 1.0 * 0.01 + 0.5 * 4.0 * 0.04
    assert row[0] == pytest.approx(2 * e / KB / 4, rel=1e-14)   # dof = 6 - 2
    assert row[0] == row[1]
    assert row[2] == row[3] == row[8] == row[9] == row[11] == 0.0
    assert row[4] == row[6] and row[5] == row[7]
    assert row[10] == -1.25


def test_characteristics_sequence():
    """Projectile over two lattice atoms in a 10 A box; the nearest lattice
    atom sits across the periodic boundary."""
    cell = np.eye(3) * 10.0
    dt = 0.5
    lat = np.array([[9.5, 0.0, 0.0], [5.0, 5.0, 0.0]])
    zs = [3.0, 2.0, 1.5, 1.8, 2.6, 2.2, 3.5]
    vz = [-1.0, -1.0, 1.0, 1.0, -1.0, 1.0, 1.0]   # down, down, up, up, down, up, up
    force_on = [False, True, True, True, False, True, False]
    o = oracle.TrajObs()
    for istep, (z, w, fo) in enumerate(zip(zs, vz, force_on), start=1):
        r = np.zeros((3, 1, 3))
        r[0, 0] = [0.5, 0.0, z]                   # 1.0 A in x from atom 1 through the boundary
        r[1:, 0] = lat
        v = np.zeros((3, 1, 3))
        v[0, 0, 2] = w
        f = np.zeros((3, 1, 3))
        if fo:
            f[0, 0, 1] = 1e-3
        st = _universe(r, v, [1.0, 1.0], [True, False], [0, 1, 1], cell=cell, f=f, dt=dt)
        st.collect(istep, o, adsorption=(2.0, 2.5))
        d = math.hypot(1.0, z)
        if istep == 3:
            assert o.closest_approach == pytest.approx(d, rel=1e-15)
        if istep == 2:
            assert o.is_adsorbed == 0        # sqrt(1 + 4) > 2.0
        if istep == 3:
            assert o.is_adsorbed == 1        # 1.80 < 2.0
        if istep == 4:
            assert o.is_adsorbed == 1        # 2.06 inside 2.5
        if istep == 5:
            assert o.is_adsorbed == 0        # 2.78 > 2.5
    assert o.closest_approach == pytest.approx(math.hypot(1.0, 1.5), rel=1e-15)
    # updated at istep 3 (= 3 * dt), then scaled by dt on each of the 4 later calls
    assert o.closest_approach_time == pytest.approx(3 * dt * dt ** 4, rel=1e-15)
    assert list(o.lowest_z) == [0.5, 0.0, 1.5]
    assert o.interaction_time == pytest.approx(4 * dt)
    assert o.nturning_points == 3                 # down->up, up->down, down->up at steps 3, 5, 6
    assert o.was_moving_down == 0


def test_free_projectile_exits_where_straight_flight_predicts():
    """A projectile 15 A above a lattice atom (beyond the EMT cut-off) moving
    up at vz exits through pul = z0 + 1 at the first step with z > pul."""
    pes = pesfile.load_pes(os.path.join(os.path.dirname(__file__), "..", "md_tian2_amd", "pes", "EMT-HAu.pes"))
    s = geometry.h_on_slab("Au", n=3, nlayers=2, nfixed=2, thermal=False, z0=15.0, einc=1.0,
                           rng=np.random.default_rng(3))
    s.v[0, 0] = [0.0, 0.0, 0.0123]
    cols, pesm = pesfile.emt_tables(pes, s.type_names)
    st = oracle.OracleState(s, cols, pesm, step=0.1)
    st.calc_force()
    assert not st.f[0].any()
    o, iend = st.run_scatter(1, 0, 5000, proj_ul=16.0)
    want = math.floor(1.0 / (0.0123 * 0.1)) + 1
    assert iend == want
    assert st.r[0, 0, 2] > 16.0 and st.r[0, 0, 2] - 0.0123 * 0.1 <= 16.0
    assert o.interaction_time == 0.0 and o.nturning_points == 1   # moving up from the first step
    o2, iend2 = oracle.OracleState(s, cols, pesm, step=0.1).run_scatter(1, 0, 100, proj_ul=16.0)
    assert iend2 == 101                                            # no exit: nsteps + 1
    o3, iend3 = oracle.OracleState(s, cols, pesm, step=0.1).run_scatter(1, 0, 5000, proj_ul=16.0,
                                                                       scatter=False)
    assert iend3 == 5001                                           # the exit needs scatter output


def test_finalize_closest_approach_time_and_final_collection():
    obs = dict(cl_appr=1.2, cl_last=2.7, r_min=[0, 0, 1.0], t_int=3.0, turning_points=2, adsorbed=True,
               adsorbed_prev=True, exit_step=0, cl_appr_step=4, last_step=10)
    fin = observables.finalize(obs, np.array([[0.0, 1e-3, 0.0]]), 0.5, 10, adsorption=(2.0, 2.5))
    assert fin["istep"] == 11 and fin["time"] == 5.0
    assert fin["cl_appr_time"] == 4 * 0.5 * 0.5 ** 7
    assert fin["t_int"] == 3.5                  # the extra collection sees a force
    assert fin["adsorbed"] is False             # 2.7 > 2.5 at the extra collection
    ex = dict(obs, exit_step=7)
    fin = observables.finalize(ex, np.zeros((1, 3)), 0.5, 10, adsorption=(2.0, 2.5))
    assert fin["istep"] == 7 and fin["time"] == 3.0 and fin["t_int"] == 3.0 and fin["adsorbed"] is True
    assert fin["cl_appr_time"] == 4 * 0.5 * 0.5 ** 3


FMT_VALUES = [0.0, 1.0, -1.0, 0.5, 123.456789012345, -392.20940943, 1.0e-5, -2.5e-12, 9.99999999995,
              6.02e23, 1e-100, -1e150, 0.099999999999, 3.0e99, 1234567.8912345, 0.15]


@pytest.mark.skipif(not (shutil.which("flang") or os.path.exists("/opt/rocm/lib/llvm/bin/flang")),
                    reason="no Fortran compiler")
def test_fortran_edit_descriptors(tmp_path):
    fc = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    descs = ["f14.7", "e21.10", "e17.8e2", "e13.6e2", "f11.8", "f16.9"]
    vals = ", ".join("%.17e_8" % v for v in FMT_VALUES)
    src = ["program fmt", "  implicit none", "  real(8) :: x(%d) = [%s]" % (len(FMT_VALUES), vals),
           "  integer :: i", "  do i = 1, size(x)"]
    src += ["    write(*, '(a1,%s,a1)') '|', x(i), '|'" % d for d in descs]
    src += ["  end do", "  write(*, '(a1,i12,a1)') '|', 42, '|'", "end program fmt"]
    f = tmp_path / "fmt.f90"
    f.write_text("\n".join(src) + "\n")
    subprocess.run([fc, "-o", str(tmp_path / "fmt"), str(f)], check=True, capture_output=True)
    out = subprocess.run([str(tmp_path / "fmt")], check=True, capture_output=True, text=True).stdout.splitlines()
    py = []
    for v in FMT_VALUES:
        py += [observables.ffmt(v, 14, 7), observables.efmt(v, 21, 10), observables.efmt(v, 17, 8, 2),
               observables.efmt(v, 13, 6, 2), observables.ffmt(v, 11, 8), observables.ffmt(v, 16, 9)]
    py.append(observables.ifmt(42))
    got = [line.strip()[1:-1] for line in out]
    assert got == py


def test_runner_frames_read_back(tmp_path):
    """input.data frames (output_runner layout) read back by the RuNNer
    reader: cell, symbols, bead-centroid positions in Bohr, centroid forces
    in Ha/Bohr and the bead-averaged energy in Ha, to the printed digits."""
    from md_tian2_amd import nene, surrogate
    rng = np.random.default_rng(5)
    s = geometry.h_on_slab("Au", n=3, nlayers=2, rng=rng, nbeads=4)
    r = s.r + rng.normal(0, 0.05, s.r.shape)
    f = rng.normal(0, 0.3, s.r.shape)
    ep = np.array([-0.31, -0.29, -0.33, -0.30])
    path = tmp_path / "input.data"
    path.write_text(observables.runner_header() + observables.runner_frame_text(
        s.simbox, [s.type_names[k][0] for k in s.idx], r, f, ep))
    (st,) = nene.read_input_data(str(path))
    b2a = observables.BOHR2ANG
    assert np.abs(st["lattice"] - s.simbox.T / b2a).max() <= 5e-9
    assert st["symbols"] == ["H"] + ["Au"] * 18
    assert np.abs(st["xyz"] - r.mean(axis=1) / b2a).max() <= 5e-10
    assert np.abs(st["forces"] - f.mean(axis=1) / observables.HA2EV * b2a).max() <= 5e-9
    assert abs(st["energy"] - ep.mean() / observables.HA2EV) <= 5e-9
    lines = path.read_text().splitlines()
    assert lines[0] == "# This file was generated by MDT2." and lines[2] == "begin" and lines[-1] == "end"
    # the surrogate's frame export is the same format, one frame per row of R
    R = np.stack([r[:, 0, :], r[:, 1, :]])
    F = np.stack([f[:, 0, :], f[:, 1, :]])
    s1 = s.copy()
    surrogate.write_input_data(str(tmp_path / "frames.data"), s1, R, np.array([-1.0, -2.0]), F)
    fr = nene.read_input_data(str(tmp_path / "frames.data"))
    assert len(fr) == 2 and abs(fr[1]["energy"] + 2.0 / observables.HA2EV) <= 5e-9
    assert np.abs(fr[1]["xyz"] - R[1] / b2a).max() <= 5e-10
