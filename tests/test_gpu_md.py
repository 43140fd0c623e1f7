"""GPU parity of the propagators (Verlet, Langevin+LDFA, Andersen, PILE,
RPMD) against the oracle fed with the same Philox noise.

Single steps agree to ~1e-15 relative; trajectories are compared over tens
of steps (chaotic growth keeps the differences far below the bars used).
The Tsurf = 0 case checks the stated 1e-8 A after 1000 steps.
"""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu

SEED = 4242


def _run_both(ctx, pes, systems, nsteps, tsurf, tproj=0.0, dt=0.1, pile_tau=200.0, andersen_time=30.0):
    cols, pesm = pesfile.emt_tables(pes, systems[0].type_names)
    ctx.set_emt(cols, pesm)
    b = mdt.Batch(ctx, systems, dt=dt, tsurf=tsurf, tproj=tproj, pile_tau=pile_tau,
                  andersen_time=andersen_time, seed=SEED)
    b.compute_forces()
    b.step(nsteps)
    out = b.get_state()
    b.close()
    refs = []
    for t, s in enumerate(systems):
        st = oracle.OracleState(s, cols, pesm, step=dt, tsurf=tsurf, tproj=tproj, pile_tau=pile_tau,
                                andersen_time=andersen_time)
        st.calc_force()
        st.run(SEED, t, 1, nsteps)
        refs.append(st)
    return out, refs


def _assert_close(out, refs, rtol_r, rtol_v):
    for o, st in zip(out, refs):
        assert np.abs(o["r"] - st.r).max() <= rtol_r, np.abs(o["r"] - st.r).max()
        vs = max(1.0, np.abs(st.v).max())
        assert np.abs(o["v"] - st.v).max() <= rtol_v * vs, np.abs(o["v"] - st.v).max()
        assert np.abs(o["epot"] - st.epot).max() <= 1e-8 * max(1.0, np.abs(st.epot).max())


def test_langevin_ldfa_verlet_c1(gpu_ctx, pes_hau):
    """C1-like: H 'ver' on Au(111) 3x3x4, Au 'lan' at 300 K with LDFA friction."""
    g = np.random.default_rng(1)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, einc=1.92, z0=3.0) for _ in range(3)]
    out, refs = _run_both(gpu_ctx, pes_hau, systems, 40, tsurf=300.0)
    _assert_close(out, refs, 1e-9, 1e-9)
    for o, st in zip(out, refs):
        # after propagate_2 dens holds the LDFA friction of Langevin atoms
        assert np.allclose(o["dens"], st.dens, rtol=1e-9, atol=1e-15)


def test_langevin_c2_geometry(gpu_ctx, pes_hau):
    g = np.random.default_rng(2)
    systems = []
    for _ in range(4):
        s = geometry.h_on_slab("Au", n=6, nlayers=4, rng=g, einc=g.uniform(0.5, 5.0),
                               polar_deg=g.uniform(0, 60), azimuth_deg=g.uniform(0, 360), z0=2.5)
        systems.append(s)
    out, refs = _run_both(gpu_ctx, pes_hau, systems, 20, tsurf=300.0)
    _assert_close(out, refs, 1e-9, 1e-9)


def test_tsurf0_1000_steps_within_1e8(gpu_ctx, pes_hau):
    """Tsurf = 0 is deterministic (series branch, zero noise): 1000 FP64 steps
    must stay within 1e-8 A of the oracle."""
    g = np.random.default_rng(3)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, einc=1.92, z0=4.0, thermal=False)]
    out, refs = _run_both(gpu_ctx, pes_hau, systems, 1000, tsurf=0.0)
    dr = np.abs(out[0]["r"] - refs[0].r).max()
    assert dr <= 1e-8, dr


def test_verlet_all_atoms_energy_conservation(gpu_ctx, pes_hau):
    g = np.random.default_rng(4)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=0, rng=g, einc=1.0, z0=2.5,
                           latt_algo=geometry.PROP_VERLET, tsurf=100.0)
    out, refs = _run_both(gpu_ctx, pes_hau, [s], 30, tsurf=0.0)
    _assert_close(out, refs, 1e-9, 1e-9)


def test_andersen_and_pile_p1(gpu_ctx, pes_hau):
    g = np.random.default_rng(5)
    s1 = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, z0=2.5, latt_algo=geometry.PROP_ANDERSEN)
    out, refs = _run_both(gpu_ctx, pes_hau, [s1], 25, tsurf=300.0, andersen_time=1.0)
    _assert_close(out, refs, 1e-9, 1e-9)
    s2 = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, z0=2.5, latt_algo=geometry.PROP_PILE)
    out, refs = _run_both(gpu_ctx, pes_hau, [s2], 25, tsurf=300.0, pile_tau=20.0)
    _assert_close(out, refs, 1e-9, 1e-9)


def test_rpmd_reference_mode_pile_projectile(gpu_ctx, pes_hpt):
    """All atoms P beads (the reference's force_beads): H 'pil', Pt 'lan'
    (spring force with m(1), Tproj for Langevin beads, rpmd.f90:418-435)."""
    g = np.random.default_rng(6)
    systems = [geometry.h_on_slab("Pt", n=3, nlayers=4, nbeads=4, rng=g, z0=2.2,
                                  proj_algo=geometry.PROP_PILE, tsurf=100.0) for _ in range(2)]
    for s in systems:
        s.r[0] += g.normal(0, 0.05, s.r[0].shape)
    out, refs = _run_both(gpu_ctx, pes_hpt, systems, 15, tsurf=100.0, tproj=100.0, pile_tau=200.0)
    _assert_close(out, refs, 1e-8, 1e-8)


def test_rpmd_verlet_slab(gpu_ctx, pes_hau):
    g = np.random.default_rng(7)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nbeads=8, rng=g, z0=2.0,
                           proj_algo=geometry.PROP_VERLET, latt_algo=geometry.PROP_VERLET)
    s.r += g.normal(0, 0.02, s.r.shape)
    out, refs = _run_both(gpu_ctx, pes_hau, [s], 10, tsurf=50.0, tproj=50.0)
    _assert_close(out, refs, 1e-8, 1e-8)


def test_bitwise_deterministic(gpu_ctx, pes_hau):
    g = np.random.default_rng(8)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, rng=g, z0=2.0) for _ in range(8)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    res = []
    for _ in range(2):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=7)
        b.compute_forces()
        b.step(10)
        res.append(b.get_state())
        b.close()
    for a, c in zip(*res):
        assert np.array_equal(a["r"], c["r"]) and np.array_equal(a["v"], c["v"])


def test_step_continuation_matches_single_call(gpu_ctx, pes_hau):
    g = np.random.default_rng(9)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, z0=2.0) for _ in range(2)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b1 = mdt.Batch(gpu_ctx, systems, tsurf=300.0, seed=3)
    b1.compute_forces()
    b1.step(12)
    b2 = mdt.Batch(gpu_ctx, systems, tsurf=300.0, seed=3)
    b2.compute_forces()
    b2.step(5)
    b2.step(7)
    for a, c in zip(b1.get_state(), b2.get_state()):
        assert np.array_equal(a["r"], c["r"])
    b1.close()
    b2.close()
