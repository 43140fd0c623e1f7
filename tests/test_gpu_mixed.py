"""Mixed bead model: the projectile (atom 0) carries the P beads, the slab
atoms are classical (one bead).  The reference has no such model -- its
universe gives every atom the same beads (src/universe_mod.f90:38-67) -- so
parity is anchored on the oracle through what the model is defined by:

* forces: bead b's potential V_b is the full EMT energy of the classical
  slab with projectile bead b, i.e. the oracle's per-bead compute_emt on a
  universe whose slab beads are copies of the classical slab; a classical
  atom's force (and density) is the bead average of the oracle's;
* fixed slab: with every slab atom fixed the mixed and all-beads models are
  the same dynamics, so the projectile follows the reference model exactly;
* distant projectile: with the projectile beyond the cut-off the slab is a
  one-bead system; its Langevin trajectory equals the oracle's one-bead
  trajectory fed with the noise the mixed model draws (slot i*P of atom i);
* Verlet: sum_b (m v_b^2/2 + V_b) + springs + P * classical kinetic energy
  is conserved.
"""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu

STREAM_LANGEVIN = 0  # Philox stream of the Langevin noise


def _spread(s, rng, sig_h, sig_slab):
    """Projectile beads spread around their centroid, slab displaced (the
    same displacement in every bead: the slab is classical)."""
    s.r[0] += rng.normal(0, sig_h, s.r[0].shape)
    d = rng.normal(0, sig_slab, (s.natoms - 1, 1, 3)) * ~s.fixed[1:, :1, :]
    s.r[1:] += d
    return s


def _expected(s, cols, pesm):
    """Mixed-model energies, forces and densities from the oracle's per-bead EMT."""
    e, f, d = oracle.emt_energy_forces(s, cols, pesm)
    f = f.copy()
    d = d.copy()
    f[1:] = f[1:].mean(axis=1, keepdims=True)
    d[1:] = d[1:].mean(axis=1, keepdims=True)
    return e, f, d


def test_mixed_forces_match_bead_average(gpu_ctx, pes_hpt):
    rng = np.random.default_rng(31)
    systems = [
        _spread(geometry.h_on_slab("Pt", n=4, nlayers=4, nbeads=8, rng=rng, z0=1.6, tsurf=100.0,
                                   proj_algo=geometry.PROP_VERLET), rng, 0.15, 0.05),
        _spread(geometry.h_on_slab("Pt", n=4, nlayers=4, nbeads=8, rng=rng, z0=-1.1, tsurf=100.0,
                                   proj_algo=geometry.PROP_VERLET), rng, 0.10, 0.05),   # subsurface
        _spread(geometry.h_on_slab("Pt", n=3, nlayers=5, nbeads=4, rng=rng, z0=2.4, tsurf=100.0,
                                   proj_algo=geometry.PROP_VERLET), rng, 0.20, 0.05),
        _spread(geometry.h_on_slab("Pt", n=3, nlayers=4, nbeads=1, rng=rng, z0=2.0, tsurf=100.0,
                                   proj_algo=geometry.PROP_VERLET), rng, 0.0, 0.05),    # one bead
    ]
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=100.0, tproj=100.0)
    b.set_bead_mode(1)
    b.compute_forces()  # the free ring-polymer step moves the beads first
    out = b.get_state()
    assert not b.status().any()
    for k, (s, o) in enumerate(zip(systems, out)):
        sr = s.copy()
        sr.r = o["r"].copy()
        # the classical atoms are reported in every bead
        assert np.array_equal(o["r"][1:], np.repeat(o["r"][1:, :1], s.nbeads, axis=1))
        e, f, d = _expected(sr, cols, pesm)
        assert np.abs(o["epot"] - e).max() <= 1e-9 * max(1.0, np.abs(e).max()), k
        assert np.abs(o["f"] - f).max() <= 1e-9, (k, np.abs(o["f"] - f).max())
        assert np.abs(o["dens"] - d).max() <= 1e-11, k
        a_ref = np.where(s.fixed, 0.0, o["f"] / s.mass[:, None, None])
        assert np.abs(o["a"] - a_ref).max() <= 1e-12
    b.close()


def test_mixed_fixed_slab_equals_all_beads(gpu_ctx, pes_hpt):
    """Slab fixed: mixed and all-beads models run the same dynamics (PILE
    projectile, COM removal included)."""
    rng = np.random.default_rng(32)
    systems = [geometry.h_on_slab("Pt", n=3, nlayers=4, nfixed=4, nbeads=8, rng=rng, z0=2.0, einc=1.0,
                                  proj_algo=geometry.PROP_PILE, tsurf=100.0) for _ in range(2)]
    for s in systems:
        s.r[0] += rng.normal(0, 0.1, s.r[0].shape)
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = []
    for mode in (0, 1):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=100.0, tproj=100.0, pile_tau=50.0, seed=5)
        b.set_bead_mode(mode)
        b.compute_forces()
        b.step(30)
        res.append(b.get_state())
        b.close()
    for ref, mix in zip(*res):
        assert np.abs(mix["r"] - ref["r"]).max() <= 1e-10
        assert np.abs(mix["v"] - ref["v"]).max() <= 1e-10 * max(1.0, np.abs(ref["v"]).max())
        assert np.abs(mix["epot"] - ref["epot"]).max() <= 1e-9 * np.abs(ref["epot"]).max()
        assert np.abs(mix["f"][0] - ref["f"][0]).max() <= 1e-8


def test_mixed_distant_projectile_slab_matches_one_bead_oracle(gpu_ctx, pes_hau):
    """Projectile 12 A above the surface: the Langevin (+LDFA) slab of the
    mixed model is the oracle's one-bead slab driven by the same noise."""
    rng = np.random.default_rng(33)
    P, nsteps, seed, tid = 8, 20, 77, 3
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nbeads=P, rng=rng, z0=12.0, einc=0.01, tsurf=300.0)
    s.r[0] += rng.normal(0, 0.05, s.r[0].shape)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, [s], dt=0.1, tsurf=300.0, tproj=300.0, seed=seed, traj_ids=[tid])
    b.set_bead_mode(1)
    b.compute_forces()
    b.step(nsteps)
    out = b.get_state()[0]
    b.close()
    s1 = geometry.System(s.type_names, s.idx, s.mass, s.fixed[:, :1], s.algo, s.simbox, s.r[:, :1].copy(),
                         s.v[:, :1].copy())
    st = oracle.OracleState(s1, cols, pesm, step=0.1, tsurf=300.0, tproj=300.0)
    st.calc_force()
    for k in range(nsteps):
        full = oracle.philox_normals(seed, tid, 1 + k, STREAM_LANGEVIN, s.natoms, P)
        st.md_step(randy=np.ascontiguousarray(full[:, :1, :]))  # atom i draws slot i*P + 0
    assert np.abs(out["r"][1:, 0] - st.r[1:, 0]).max() <= 1e-10
    assert np.abs(out["v"][1:, 0] - st.v[1:, 0]).max() <= 1e-10
    assert np.abs(out["epot"] - st.epot[0]).max() <= 1e-9 * abs(st.epot[0])


def _mixed_energy(o, s, tproj):
    P = s.nbeads
    m = s.mass
    kin_h = 0.5 * m[0] * (o["v"][0] ** 2).sum()
    wp = geometry.KB * tproj * P / geometry.HBAR
    rb = o["r"][0]
    spring = 0.5 * m[0] * wp * wp * ((rb - np.roll(rb, -1, axis=0)) ** 2).sum()
    kin_c = 0.5 * (m[1:, None] * (o["v"][1:, 0] ** 2)).sum()
    return kin_h + o["epot"].sum() + spring + P * kin_c


def test_mixed_verlet_energy_conservation(gpu_ctx, pes_hau):
    rng = np.random.default_rng(34)
    P, tproj = 4, 100.0
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=0, nbeads=P, rng=rng, z0=2.2, einc=1.0, tsurf=100.0,
                           proj_algo=geometry.PROP_VERLET, latt_algo=geometry.PROP_VERLET)
    s.r[0] += rng.normal(0, 0.05, s.r[0].shape)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, [s], dt=0.05, tsurf=0.0, tproj=tproj)
    b.set_bead_mode(1)
    b.compute_forces()
    e0 = _mixed_energy(b.get_state()[0], s, tproj)
    dev = 0.0
    kin_c = 0.0
    for _ in range(20):
        b.step(10)
        o = b.get_state()[0]
        dev = max(dev, abs(_mixed_energy(o, s, tproj) - e0))
        kin_c = max(kin_c, 0.5 * (s.mass[1:, None] * (o["v"][1:, 0] ** 2)).sum())
    b.close()
    print("mixed Verlet: max |E - E0| = %.3e eV (classical kinetic energy up to %.3f eV, x P)" % (dev, kin_c))
    # a slab force summed instead of averaged over the beads (or a missing P
    # on the classical kinetic energy) moves E by a sizeable part of P * kin_c
    assert dev <= 2e-3


def test_mixed_layout_checks(gpu_ctx, pes_hau):
    rng = np.random.default_rng(35)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nbeads=4, rng=rng, z0=2.0)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    bad = s.copy()
    bad.idx = bad.idx.copy()
    bad.idx[[0, 1]] = bad.idx[[1, 0]]  # the projectile is not atom 0
    b = mdt.Batch(gpu_ctx, [bad], dt=0.1, tsurf=300.0, tproj=300.0)
    with pytest.raises(mdt.MdtError, match="projectile"):
        b.set_bead_mode(1)
    b.compute_forces()  # the batch stays usable in the all-beads model
    b.close()
    # the mixed model has no network path: refused in either order
    from md_tian2_amd import surrogate
    gpu_ctx.set_nn(surrogate.load_model(surrogate.MODEL_DIR), s.type_names)
    b = mdt.Batch(gpu_ctx, [s], dt=0.1, tsurf=300.0, tproj=300.0)
    b.set_bead_mode(1)
    with pytest.raises(mdt.MdtError, match="EMT potential only"):
        b.set_pes(1)
    b.set_bead_mode(0)
    b.set_pes(1)
    with pytest.raises(mdt.MdtError, match="EMT potential only"):
        b.set_bead_mode(1)
    b.compute_forces()
    b.close()


def test_mixed_list_skin_reuse_equals_rebuild(gpu_ctx, pes_hpt):
    """Classical lists kept across steps (Verlet skin) give the rebuilt
    lists' dynamics: an 8x8 Pt(111) cell (19.2 A wide, above 2 (cut-off +
    skin)) under Langevin at 300 K for 200 steps of 0.5 fs, the slab atoms
    moving beyond half the default skin; skin 0 (rebuild every call), 0.3 A (default) and
    0.02 A (rebuild every few steps) agree to rounding, the forces of the
    final state equal the oracle's per-bead EMT, and new positions from
    set_state drop the kept lists."""
    rng = np.random.default_rng(36)
    systems = [_spread(geometry.h_on_slab("Pt", n=8, nlayers=4, nbeads=8, rng=rng, z0=z, einc=0.5, tsurf=300.0,
                                          proj_algo=geometry.PROP_PILE), rng, 0.1, 0.05) for z in (1.8, 3.0)]
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = {}
    for skin in (0.0, 0.3, 0.02):
        b = mdt.Batch(gpu_ctx, systems, dt=0.5, tsurf=300.0, tproj=300.0, pile_tau=100.0, seed=9)
        b.set_bead_mode(1)
        b.set_list_skin(skin)
        b.compute_forces()
        b.step(200)
        assert not b.status().any()
        res[skin] = b.get_state()
        if skin == 0.3:
            for s, o in zip(systems, res[skin]):
                sr = s.copy()
                sr.r = o["r"].copy()
                e, f, d = _expected(sr, cols, pesm)
                assert np.abs(o["epot"] - e).max() <= 1e-9 * np.abs(e).max()
                assert np.abs(o["f"] - f).max() <= 1e-9
            b.set_state([s.r for s in systems], [s.v for s in systems])
            b.compute_forces()
            for s, o in zip(systems, b.get_state()):
                sr = s.copy()
                sr.r = o["r"].copy()  # the free ring-polymer step moved the beads
                assert np.abs(o["r"][1:] - s.r[1:]).max() == 0.0
                e, f, d = _expected(sr, cols, pesm)
                assert np.abs(o["f"] - f).max() <= 1e-9
                assert np.abs(o["epot"] - e).max() <= 1e-9 * np.abs(e).max()
        b.close()
    moved = max(np.abs(o["r"][1:, 0] - s.r[1:, 0]).max() for s, o in zip(systems, res[0.0]))
    print("slab atoms moved up to %.3f A" % moved)
    assert moved > 0.15  # the default skin's lists were rebuilt on the way
    for skin in (0.3, 0.02):
        for ref, o in zip(res[0.0], res[skin]):
            assert np.abs(o["r"] - ref["r"]).max() <= 1e-9, skin
            assert np.abs(o["v"] - ref["v"]).max() <= 1e-9 * np.abs(ref["v"]).max(), skin
            assert np.abs(o["epot"] - ref["epot"]).max() <= 1e-9 * np.abs(ref["epot"]).max(), skin
    b = mdt.Batch(gpu_ctx, systems, dt=0.5, tsurf=300.0, tproj=300.0)
    with pytest.raises(mdt.MdtError, match="skin"):
        b.set_list_skin(-1.0)
    b.close()
