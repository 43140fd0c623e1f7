"""Oracle checks at the configured slab sizes and bead counts, after steps
of dynamics (so the beads and slab atoms have spread):

* mixed bead model, RPMD H/Pt(111) 8x8x6 with 32 projectile beads (PILE,
  100 K): every one of k_mixed's 32 bead groups is busy and all of them add
  into the shared fixed-point G / density accumulators;
* mixed bead model, Au(111) 12x12x6 (865 atoms) with 16 projectile beads,
  the largest slab and bead count of the uneven workload;
* every atom carrying the beads (the reference's model), 12x12x6 with 4
  beads inside an uneven batch, where that size class keeps its neighbour
  lists in HBM across steps (Verlet skin); there the Langevin slab's beads
  carry calc_force's ring-polymer spring force as well.

After the steps the batch recomputes its forces (calc_force), so that dens
holds the raw EMT density again (a step leaves the LDFA friction of the
Langevin atoms in it, src/md_algo.f90:556-558).  Expected values come from
the oracle's compute_emt (the reference's
compute_emt_2species restated, src/pes_emt_mod.f90:483-1027) evaluated per
bead on one-bead universes.  In the mixed model bead b's universe is the
classical slab with projectile bead b (src/rpmd.f90:309-452 gives every atom
the beads; force_beads, src/md_init.f90:640-669, copies bead 1 into them,
which is this model's slab); a classical atom's force and density are the
bead averages.  Bars as in the other parity tests: energies 1e-9 relative,
forces 1e-9 eV/A, densities 1e-11.
"""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu


def _one_bead(s, r):
    """A one-bead universe of s's atoms at positions r (N, 3)."""
    return geometry.System(s.type_names, s.idx, s.mass, s.fixed[:, :1].copy(), s.algo, s.simbox,
                           np.ascontiguousarray(r[:, None, :]), np.zeros((s.natoms, 1, 3)))


def _oracle_per_bead(s, r, cols, pesm, mixed):
    """epot (P,), f (N, P, 3), dens (N, P) from the oracle bead by bead."""
    N, P = r.shape[0], r.shape[1]
    e, f, d = np.zeros(P), np.zeros((N, P, 3)), np.zeros((N, P))
    for b in range(P):
        rb = r[:, b, :].copy()
        if mixed:
            rb[1:] = r[1:, 0, :]  # the classical slab
        eb, fb, db = oracle.emt_energy_forces(_one_bead(s, rb), cols, pesm)
        e[b], f[:, b], d[:, b] = eb[0], fb[:, 0], db[:, 0]
    if mixed:
        f[1:] = f[1:].mean(axis=1, keepdims=True)
        d[1:] = d[1:].mean(axis=1, keepdims=True)
    return e, f, d


def _langevin_springs(s, r, tproj):
    """The force calc_force gives the beads of Langevin atoms besides the
    potential's: the ring-polymer step is reverted for them and the spring
    force -m(1) (kB Tproj P / hbar)^2 (2 r_b - r_(b-1) - r_(b+1)) is set
    (src/rpmd.f90:418-435, the projectile's mass and temperature for every
    atom); zero for P = 1 and for other propagators."""
    P = r.shape[1]
    out = np.zeros_like(r)
    if P == 1:
        return out
    kt = geometry.KB * tproj * P / geometry.HBAR
    pref = -s.mass[0] * kt * kt
    lan = s.algo[s.idx] == geometry.PROP_LANGEVIN
    out[lan] = pref * (2 * r[lan] - np.roll(r[lan], 1, axis=1) - np.roll(r[lan], -1, axis=1))
    return out


def _check(o, e, f, d, tag):
    de = np.abs(o["epot"] - e).max()
    df = np.abs(o["f"] - f).max()
    dd = np.abs(o["dens"] - d).max()
    print("%s: max dE %.2e eV, max dF %.2e eV/A, max ddens %.2e" % (tag, de, df, dd))
    assert de <= 1e-9 * max(1.0, np.abs(e).max()), tag
    assert df <= 1e-9, tag
    assert dd <= 1e-11, tag


def test_mixed_model_c3_hpt_8x8x6_32_beads(gpu_ctx, pes_hpt):
    g = np.random.default_rng(301)
    systems = [geometry.h_on_slab("Pt", n=8, nlayers=6, nfixed=2, rng=g, nbeads=32, einc=e, polar_deg=20.0,
                                  azimuth_deg=g.uniform(0, 360), z0=z, tsurf=100.0, proj_algo=geometry.PROP_PILE)
               for z, e in ((1.9, 0.8), (3.2, 2.5))]
    for s in systems:
        s.r[0] += g.normal(0, 0.08, s.r[0].shape)
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=100.0, tproj=100.0, pile_tau=200.0, seed=41)
    b.set_bead_mode(1)
    b.compute_forces()
    b.step(120)
    b.compute_forces()
    assert not b.status().any()
    out = b.get_state()
    b.close()
    for k, (s, o) in enumerate(zip(systems, out)):
        spread = np.abs(o["r"][0] - o["r"][0].mean(axis=0)).max()
        assert spread > 0.01  # the 32 beads are distinct
        e, f, d = _oracle_per_bead(s, o["r"], cols, pesm, mixed=True)
        _check(o, e, f, d, "C3 mixed traj %d" % k)


def test_mixed_model_au_12x12x6_16_beads(gpu_ctx, pes_hau):
    g = np.random.default_rng(302)
    systems = [geometry.h_on_slab("Au", n=12, nlayers=6, nfixed=2, rng=g, nbeads=16, einc=3.0, polar_deg=30.0,
                                  azimuth_deg=g.uniform(0, 360), z0=z, tsurf=400.0) for z in (1.5, 2.6)]
    for s in systems:
        s.r[0] += g.normal(0, 0.1, s.r[0].shape)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=400.0, tproj=400.0, seed=42)
    b.set_bead_mode(1)
    b.compute_forces()
    b.step(60)
    b.compute_forces()
    assert not b.status().any()
    out = b.get_state()
    b.close()
    for k, (s, o) in enumerate(zip(systems, out)):
        e, f, d = _oracle_per_bead(s, o["r"], cols, pesm, mixed=True)
        _check(o, e, f, d, "12x12x6 mixed P=16 traj %d" % k)


def test_all_beads_au_12x12x6_in_uneven_batch(gpu_ctx, pes_hau):
    """12x12x6 with 4 beads on every atom (force_beads) between smaller
    trajectories: its size class (865 atoms) keeps HBM lists across the 80
    steps; then every bead of every trajectory against the oracle."""
    g = np.random.default_rng(303)
    spec = [(3, 1, 1.8), (12, 4, 1.4), (6, 1, 2.2), (12, 4, 5.5)]
    systems = []
    for n, p, z in spec:
        s = geometry.h_on_slab("Au", n=n, nlayers=6, nfixed=2, rng=g, nbeads=p, einc=g.uniform(0.5, 4.0),
                               azimuth_deg=g.uniform(0, 360), polar_deg=25.0, z0=z, tsurf=300.0)
        s.r[0] += g.normal(0, 0.05, s.r[0].shape)
        systems.append(s)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, tproj=300.0, seed=43)
    b.tune(tpa=0)
    b.compute_forces()
    b.step(80)
    b.compute_forces()
    assert not b.status().any()
    out = b.get_state()
    b.close()
    for k, (s, o) in enumerate(zip(systems, out)):
        e, f, d = _oracle_per_bead(s, o["r"], cols, pesm, mixed=False)
        _check(o, e, f + _langevin_springs(s, o["r"], 300.0), d, "all beads N=%d P=%d" % (s.natoms, s.nbeads))
