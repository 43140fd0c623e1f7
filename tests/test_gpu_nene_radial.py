"""Radial-only NENE networks (k_nene_rg / k_nene_rf) against the NENE oracle.

A network whose elements have no angular functions runs the radial kernels:
centres scheduled so that a wavefront's centres share one element, functions
grouped by neighbour element, and, in cells wider than 2 rc in every
direction, neighbour lists from the minimum image alone (one pass, ballot
compaction); narrower cells (3x3 Au(111), 8.65 A) take the image search.
Bars as for the general kernels (tests/test_gpu_nene.py): per frame
|dE|/N <= 2e-4 eV/atom and max |dF| <= 5e-3 eV/A against the FP64 oracle.
The trained H/Au radial surrogate (md_tian2_amd/nn/hau_radial) runs the
mixed-batch workload with the EMT judge.
"""
import numpy as np
import pytest

import nene_oracle as no
import oracle
from md_tian2_amd import geometry, mdt, nene, pesfile, surrogate

pytestmark = pytest.mark.gpu

E_ATOM_TOL = 2e-4
F_TOL = 5e-3
MDT_ST_NN_EXTRAPOLATION = 16


def _frames(rng, spec):
    """spec: (n, layers, H height) per frame; positions jittered."""
    out = []
    for n, nl, z in spec:
        s = geometry.h_on_slab("Au", n=n, nlayers=nl, nfixed=2, rng=rng, einc=rng.uniform(0.5, 5),
                               polar_deg=30, azimuth_deg=rng.uniform(0, 360), tsurf=300.0)
        s.r[:, 0, :] += rng.normal(0, 0.05, s.r[:, 0, :].shape) * (~s.fixed[:, 0, :])
        s.r[0, 0, 2] = z
        out.append(s)
    return out


def _radial_model(frames, rng, margin=0.3, **kw):
    m = nene.hau_radial_design(**kw)
    for z in m.elements:
        m.gmin[z], m.gmax[z], m.gavg[z] = np.zeros(m.nsf(z)), np.ones(m.nsf(z)), np.zeros(m.nsf(z))
    nene.init_weights(m, rng)
    om = no.Model(m)
    Gs = {1: [], 79: []}
    for s in frames:
        el = om.elem_index([1] + [79] * (s.natoms - 1))
        G = om.descriptors(el, s.r[:, 0, :] / nene.BOHR2ANG, s.simbox.T / nene.BOHR2ANG)
        Gs[1].append(G[:1])
        Gs[79].append(G[1:])
    nene.set_scaling(m, {z: np.concatenate(v) for z, v in Gs.items()}, margin=margin)
    return m


def _check_against_oracle(m, frames, out):
    om = no.Model(m)
    worst_e = worst_f = 0.0
    for s, o in zip(frames, out):
        el = om.elem_index([1] + [79] * (s.natoms - 1))
        E, _, F, nex = om.energy_forces(el, s.r[:, 0, :] / nene.BOHR2ANG, s.simbox.T / nene.BOHR2ANG)
        assert nex == 0
        worst_e = max(worst_e, abs(o["epot"][0] - E * nene.HA2EV) / s.natoms)
        worst_f = max(worst_f, np.abs(o["f"][:, 0, :] - F * nene.HABOHR2EVANG).max())
    return worst_e, worst_f


@pytest.mark.parametrize("design", [dict(nh=32, nau=24, nauh=8, rc=6.5), dict()], ids=["r32-6.5A", "r16-5.5A"])
def test_radial_network_batch_forces(gpu_ctx, design):
    """6x6x4 and 9x9x6 cells (minimum-image lists) and 3x3x4 (image search),
    the H from inside the surface to 6 A above it."""
    rng = np.random.default_rng(17)
    spec = [(6, 4, 1.2), (6, 4, 2.5), (6, 4, 6.0), (3, 4, 1.8), (3, 4, 5.0), (9, 6, 0.8), (9, 6, 3.5)]
    frames = _frames(rng, spec)
    m = _radial_model(frames, rng, **design)
    gpu_ctx.set_nn(m, frames[0].type_names)
    b = mdt.Batch(gpu_ctx, frames, dt=0.1, tsurf=300.0)
    b.set_pes(1)
    b.compute_forces()
    out = b.get_state()
    assert not b.status().any()
    for s, o in zip(frames, out):
        a_ref = np.where(s.fixed, 0.0, o["f"] / s.mass[:, None, None])
        assert np.abs(o["a"] - a_ref).max() <= 1e-12
    worst_e, worst_f = _check_against_oracle(m, frames, out)
    print("radial H/Au network %s: max dE/atom %.3e eV, max dF %.3e eV/A" % (design, worst_e, worst_f))
    assert worst_e <= E_ATOM_TOL
    assert worst_f <= F_TOL
    b.close()


def test_radial_drop_in_compute_nene(gpu_ctx):
    """The drop-in entry (mdt_compute_nene) on the radial kernels: one
    universe at a time, 12x12x4 (minimum image) and 3x3x4 (image search)."""
    rng = np.random.default_rng(18)
    frames = _frames(rng, [(12, 4, 2.0), (3, 4, 2.2)])
    m = _radial_model(frames, rng)
    gpu_ctx.set_nn(m, frames[0].type_names)
    outs = []
    for s in frames:
        e_g, f_g, ex = gpu_ctx.compute_nene(s)
        assert not ex
        outs.append({"epot": e_g, "f": f_g})
    worst_e, worst_f = _check_against_oracle(m, frames, outs)
    assert worst_e <= E_ATOM_TOL
    assert worst_f <= F_TOL


def test_radial_extrapolation_falls_back_to_emt(gpu_ctx, pes_hau):
    """A replica whose descriptors leave the scaling range takes EMT forces
    (k_replica on the flagged replicas) equal to the EMT oracle."""
    rng = np.random.default_rng(19)
    m = _radial_model(_frames(rng, [(6, 4, z) for z in (3.0, 4.0, 5.0, 6.0)]), rng, margin=0.0)
    test = _frames(rng, [(6, 4, 0.9), (6, 4, 4.5)])  # the first is far outside the H's training range
    gpu_ctx.set_nn(m, test[0].type_names)
    cols, pesm = pesfile.emt_tables(pes_hau, test[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, test, dt=0.1, tsurf=300.0)
    b.set_pes(1)
    b.compute_forces()
    out = b.get_state()
    fb = b.nn_fallback()
    assert fb[0][0]
    e_r, f_r, _ = oracle.emt_energy_forces(test[0], cols, pesm)
    assert np.abs(out[0]["f"] - f_r).max() <= 1e-9
    assert abs(out[0]["epot"][0] - e_r[0]) <= 1e-9 * max(1.0, abs(e_r[0]))
    b.close()


def _mixed_batch(rng):
    spec = [(3, 1), (6, 4), (9, 1), (12, 16), (3, 4), (6, 1), (9, 16), (12, 1)]
    out, tsurf = [], []
    for n, p in spec:
        ts = float(rng.uniform(100, 500))
        s = geometry.h_on_slab("Au", n=n, nlayers=6, nfixed=2, rng=rng, nbeads=p, einc=rng.uniform(0.2, 6.0),
                               polar_deg=rng.uniform(0, 60), azimuth_deg=rng.uniform(0, 360),
                               z0=rng.uniform(1.5, 4.0), tsurf=ts)
        s.r[0] += rng.normal(0, 0.05, s.r[0].shape)
        out.append(s)
        tsurf.append(ts)
    return out, np.array(tsurf)


def test_uneven_batch_on_the_radial_surrogate(gpu_ctx, pes_hau):
    """The trained radial surrogate on the mixed batch (3x3 .. 12x12 x 6
    layers, 1 / 4 / 16 beads on every atom): every checked replica against
    the NENE oracle, no replica on the fallback, the judge's force RMSE
    within the 5 meV/A class of the surrogate."""
    model = surrogate.load_model(surrogate.MODEL_DIR_RADIAL)
    assert model is not None, "trained surrogate missing (md_tian2_amd/nn/hau_radial)"
    rng = np.random.default_rng(53)
    systems, tsurf = _mixed_batch(rng)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=tsurf, tproj=300.0, seed=5)
    b.set_pes(1, judge_interval=1)
    b.compute_forces()
    b.step(3)
    out = b.get_state()
    rmse, ncomp, _, nfb = b.nn_stats()
    st = b.status(clear=True)
    assert not st.any(), st
    assert nfb == 0
    om = no.Model(model)
    for s, o in zip(systems, out):
        el = om.elem_index([1] + [79] * (s.natoms - 1))
        for bead in sorted({0, s.nbeads - 1}):
            E, _, F, nex = om.energy_forces(el, o["r"][:, bead, :] / nene.BOHR2ANG, s.simbox.T / nene.BOHR2ANG)
            assert nex == 0
            assert abs(o["epot"][bead] - E * nene.HA2EV) / s.natoms <= E_ATOM_TOL
            if s.nbeads == 1:  # no ring-polymer spring forces in f
                assert np.abs(o["f"][:, 0, :] - F * nene.HABOHR2EVANG).max() <= F_TOL
    print(f"radial surrogate, mixed batch: judge RMSE {rmse * 1e3:.2f} meV/A over {ncomp:.0f} components")
    assert 0.0 < rmse < 0.005
    b.close()


def test_radial_surrogate_batch_matches_separate_batches(gpu_ctx, pes_hau):
    """Stepping the uneven radial-surrogate batch equals stepping each
    trajectory alone (same seed and trajectory id), bitwise."""
    model = surrogate.load_model(surrogate.MODEL_DIR_RADIAL)
    rng = np.random.default_rng(54)
    systems, tsurf = _mixed_batch(rng)
    systems, tsurf = systems[:4], tsurf[:4]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=tsurf, tproj=300.0, seed=9, traj_ids=list(range(4)))
    b.set_pes(1)
    b.compute_forces()
    b.step(4)
    together = b.get_state()
    b.close()
    for t, s in enumerate(systems):
        b1 = mdt.Batch(gpu_ctx, [s], dt=0.1, tsurf=tsurf[t], tproj=300.0, seed=9, traj_ids=[t])
        b1.set_pes(1)
        b1.compute_forces()
        b1.step(4)
        alone = b1.get_state()[0]
        b1.close()
        for k in ("r", "v", "f", "epot"):
            assert np.array_equal(together[t][k], alone[k]), (t, k)


def test_kept_lists_equal_rebuilt_lists(gpu_ctx, pes_hau):
    """Neighbour lists kept across steps with the Verlet skin (entries to
    rc + skin, masked beyond rc; k_nene_rf reads k_nene_rg's lists) give the
    energies and forces of lists rebuilt within rc at every launch, to FP32
    summation order, after 40 steps of 6x6x4 trajectories; both agree with
    the NENE oracle."""
    model = surrogate.load_model(surrogate.MODEL_DIR_RADIAL)
    rng = np.random.default_rng(55)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, nfixed=2, rng=rng, einc=rng.uniform(0.5, 5),
                                  polar_deg=30, azimuth_deg=rng.uniform(0, 360), z0=rng.uniform(1.5, 3.0),
                                  tsurf=300.0) for _ in range(6)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=3)
    b.set_pes(1)
    b.compute_forces()
    b.step(40)
    kept = b.get_state()
    assert not b.status().any()
    b.close()
    moved = []
    for s, o in zip(systems, kept):
        s2 = s.copy()
        s2.r = o["r"].copy()
        moved.append(s2)
    b2 = mdt.Batch(gpu_ctx, moved, dt=0.1, tsurf=300.0, seed=3)
    b2.set_list_skin(0.0)
    b2.set_pes(1)
    b2.compute_forces()
    fresh = b2.get_state()
    b2.close()
    for o, q in zip(kept, fresh):
        assert np.abs(o["f"] - q["f"]).max() <= 1e-4
        assert abs(o["epot"][0] - q["epot"][0]) <= 1e-6 * max(1.0, abs(q["epot"][0]))
    worst_e, worst_f = _check_against_oracle(model, moved, kept)
    assert worst_e <= E_ATOM_TOL
    assert worst_f <= F_TOL


def test_radial_surrogate_in_hip_graphs(gpu_ctx, pes_hau):
    """The radial kernels with kept lists replayed from hipGraph chunks (the
    list validity is decided on the device) give bitwise the trajectories of
    eager launches."""
    model = surrogate.load_model(surrogate.MODEL_DIR_RADIAL)
    rng = np.random.default_rng(56)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, nfixed=2, rng=rng, einc=rng.uniform(0.5, 5),
                                  polar_deg=30, azimuth_deg=rng.uniform(0, 360), z0=rng.uniform(1.5, 3.0),
                                  tsurf=300.0) for _ in range(8)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    out = []
    for graph in (0, 10):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=21)
        b.set_pes(1)
        b.set_graph(graph)
        b.compute_forces()
        b.step(30)
        out.append(b.get_state())
        assert not b.status().any()
        b.close()
    for a, g in zip(*out):
        for k in ("r", "v", "f", "epot"):
            assert np.array_equal(a[k], g[k]), k
