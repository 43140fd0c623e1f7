"""Neighbour lists kept across force calls (Verlet skin) in k_replica.

Size classes of more than 256 atoms (and smaller ones where the autotuner
finds it faster) keep their owned half lists in HBM, built with the near
radius + skin (the lists hold the pairs within the near radius r_near,
beyond which a pair weighs less than 2^-56; the pairs beyond it of atoms
whose near sigma is tiny come from a direct scan), while no atom of the
replica moved more than skin / 2 since the build; a kept list gets the
image of every entry recomputed from the current positions.  The reference visits every pair
under the minimum image at every call (src/pes_emt_mod.f90:1204-1206), so
parity is what a rebuild at every call gives:

* skins 0 (rebuild every call), 0.3 A (default) and 0.02 A (a rebuild every
  few steps) give the same trajectories to rounding (only the summation
  order of sigma in registers can differ), over 200 steps in which the
  atoms move beyond half the default skin;
* the forces of the final state equal the oracle's;
* a projectile starting 6.5 A above the surface (tiny sigma: its pairs
  beyond r_near come from the direct scan) moves into the near region on
  kept lists;
* a 4x4 cell (10 A wide, below 2 (r_near + skin), half-cell ties inside the
  cut-off) rebuilds at every call: there the rounded image of a pair need not
  be its nearest, so a list built for one image could miss the pair after it
  rounds to another; the skins agree to rounding there too;
* new positions from set_state drop the kept lists."""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu


def _run(ctx, systems, skin, steps, dt=0.5, seed=4):
    b = mdt.Batch(ctx, systems, dt=dt, tsurf=300.0, tproj=300.0, seed=seed)
    b.set_list_skin(skin)
    b.compute_forces()
    b.step(steps)
    assert not b.status().any()
    return b


def _check_oracle(systems, out, cols, pesm, tol=1e-9):
    """Energies and forces against the oracle.  With beads, the
    reported f of a Langevin atom also holds its ring-polymer spring force
    (the propagator's own use of f), so forces are compared on the other
    atoms (the projectile and the fixed atoms) there."""
    for s, o in zip(systems, out):
        sr = s.copy()
        sr.r = o["r"].copy()
        e, f, d = oracle.emt_energy_forces(sr, cols, pesm)
        assert np.abs(o["epot"] - e).max() <= tol * np.abs(e).max()
        rows = np.ones(s.natoms, dtype=bool)
        if s.nbeads > 1:
            rows = (np.asarray(s.algo)[s.idx] != geometry.PROP_LANGEVIN) | s.fixed[:, 0, :].all(axis=1)
        assert rows.sum() >= 1
        assert np.abs(o["f"][rows] - f[rows]).max() <= tol


@pytest.mark.parametrize("metal,n,nlayers,nbeads", [("Pt", 8, 4, 2), ("Au", 4, 17, 1)])
def test_kept_lists_equal_rebuilt_lists(gpu_ctx, pes_hpt, pes_hau, metal, n, nlayers, nbeads):
    rng = np.random.default_rng(60 + n)
    pes = pes_hpt if metal == "Pt" else pes_hau
    systems = []
    for z in (1.8, 3.0, 6.5):
        s = geometry.h_on_slab(metal, n=n, nlayers=nlayers, nbeads=nbeads, rng=rng, z0=z, einc=0.5, tsurf=300.0)
        s.r[0] += rng.normal(0, 0.05, s.r[0].shape)
        systems.append(s)
    assert systems[0].natoms > 256
    cols, pesm = pesfile.emt_tables(pes, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = {}
    for skin in (0.0, 0.3, 0.02):
        b = _run(gpu_ctx, systems, skin, 200)
        res[skin] = b.get_state()
        if skin == 0.3:
            _check_oracle(systems, res[skin], cols, pesm)
            # new positions: the kept lists are dropped
            b.set_state([s.r for s in systems], [s.v for s in systems])
            b.compute_forces()
            _check_oracle(systems, b.get_state(), cols, pesm)
        b.close()
    moved = max(np.abs(o["r"][1:] - s.r[1:]).max() for s, o in zip(systems, res[0.0]))
    print(f"{metal} {n}x{n}x{nlayers}, {nbeads} beads: atoms moved up to {moved:.3f} A")
    assert moved > 0.15  # the default skin's lists were rebuilt on the way
    for skin in (0.3, 0.02):
        for ref, o in zip(res[0.0], res[skin]):
            assert np.abs(o["r"] - ref["r"]).max() <= 1e-9, skin
            assert np.abs(o["v"] - ref["v"]).max() <= 1e-9 * np.abs(ref["v"]).max(), skin
            assert np.abs(o["epot"] - ref["epot"]).max() <= 1e-9 * np.abs(ref["epot"]).max(), skin


def test_autotuned_keep_equals_rebuilt_lists(gpu_ctx, pes_hau):
    """H/Au(111) 6x6x4 (145 atoms, 15 A cell: kept lists allowed): the
    autotuner times rebuilt and kept lists per size class; whichever it
    picks, the trajectories equal those with lists rebuilt at every call."""
    rng = np.random.default_rng(80)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, rng=rng, z0=z, einc=1.0, tsurf=300.0)
               for z in (2.0, 4.0, 6.0, 7.0)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = {}
    for skin in (0.0, 0.3):
        b = mdt.Batch(gpu_ctx, systems, dt=0.5, tsurf=300.0, tproj=300.0, seed=9)
        b.set_list_skin(skin)
        b.tune(tpa=0)
        shapes = b.class_shapes()
        print(f"skin {skin}: {shapes}")
        assert len(shapes) == 1 and shapes[0]["nmax"] == systems[0].natoms
        if skin == 0.0:
            assert shapes[0]["keep"] == 0
        b.compute_forces()
        b.step(150)
        assert not b.status().any()
        res[skin] = b.get_state()
        if skin == 0.3:
            _check_oracle(systems, res[skin], cols, pesm)
        b.close()
    for ref, o in zip(res[0.0], res[0.3]):
        assert np.abs(o["r"] - ref["r"]).max() <= 1e-9
        assert np.abs(o["v"] - ref["v"]).max() <= 1e-9 * np.abs(ref["v"]).max()
        assert np.abs(o["epot"] - ref["epot"]).max() <= 1e-9 * np.abs(ref["epot"]).max()


@pytest.mark.parametrize("metal,n,nlayers", [("Au", 6, 4), ("Pt", 8, 4)])
def test_list_modes_agree(gpu_ctx, pes_hau, pes_hpt, metal, n, nlayers):
    """mdt_batch_set_keep: lists rebuilt every launch (0), kept and staged in
    LDS during a launch (1), kept and read from HBM (2) give the same
    trajectories to rounding over 150 steps, with the same threads per
    atom; mode 1 runs staged where the lists fit in LDS."""
    rng = np.random.default_rng(85 + n)
    pes = pes_hau if metal == "Au" else pes_hpt
    systems = [geometry.h_on_slab(metal, n=n, nlayers=nlayers, rng=rng, z0=z, einc=1.0, tsurf=300.0)
               for z in (2.0, 6.5, 3.5)]
    cols, pesm = pesfile.emt_tables(pes, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = {}
    for mode in (0, 1, 2):
        b = mdt.Batch(gpu_ctx, systems, dt=0.5, tsurf=300.0, tproj=300.0, seed=9)
        b.set_keep(mode)
        b.tune(tpa=2)
        shapes = b.class_shapes()
        print(f"{metal} mode {mode}: {shapes}")
        assert shapes[0]["keep"] in ((1, 2) if mode == 1 else (mode,))
        b.compute_forces()
        b.step(150)
        assert not b.status().any()
        res[mode] = b.get_state()
        if mode == 1:
            _check_oracle(systems, res[mode], cols, pesm)
        b.close()
    for mode in (1, 2):
        for ref, o in zip(res[0], res[mode]):
            assert np.abs(o["r"] - ref["r"]).max() <= 1e-9, mode
            assert np.abs(o["v"] - ref["v"]).max() <= 1e-9 * np.abs(ref["v"]).max(), mode
            assert np.abs(o["epot"] - ref["epot"]).max() <= 1e-9 * np.abs(ref["epot"]).max(), mode


def test_skin_switch_keeps_batch_usable(gpu_ctx, pes_hpt):
    """Switching the skin off and on moves the lists between LDS and HBM;
    the forces stay the oracle's."""
    rng = np.random.default_rng(70)
    systems = [geometry.h_on_slab("Pt", n=8, nlayers=4, rng=rng, z0=2.0, tsurf=300.0) for _ in range(3)]
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.5, tsurf=300.0, tproj=300.0)
    b.tune(tpa=0)
    for skin in (0.0, 0.3, 0.0, 0.25):
        b.set_list_skin(skin)
        b.compute_forces()
        b.step(20)
        assert not b.status().any()
        _check_oracle(systems, b.get_state(), cols, pesm)
    b.close()
