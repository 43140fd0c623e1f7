"""Uneven batch on the NN surrogate with the EMT judge (the mixed-batch
workload: slabs of 3x3 .. 12x12 x 6 layers, 1, 4 or 16 beads on every atom,
per-trajectory Tsurf), with the network trained on 4- and 6-layer slabs at
100-500 K (md_tian2_amd/nn/hau_thick).  Each replica is checked on its own: the network's
energy and forces against the NENE oracle on that bead's positions (bars of
tests/test_gpu_nene.py), or, when the replica left the network's training
range, the EMT fallback against the EMT oracle (1e-9).  The EMT judge runs
every step; its force RMSE stays within the surrogate's class."""
import numpy as np
import pytest

import nene_oracle as no
import oracle
from md_tian2_amd import geometry, mdt, nene, pesfile, surrogate

pytestmark = pytest.mark.gpu

E_ATOM_TOL = 2e-4
F_TOL = 5e-3
MDT_ST_NN_EXTRAPOLATION = 16  # status bit of a replica on the EMT fallback


def _batch(rng):
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


def test_uneven_batch_on_the_surrogate(gpu_ctx, pes_hau):
    model = surrogate.load_model(surrogate.MODEL_DIR_THICK)
    assert model is not None, "trained surrogate missing (md_tian2_amd/nn/hau_thick)"
    rng = np.random.default_rng(51)
    systems, tsurf = _batch(rng)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=tsurf, tproj=300.0, seed=5)
    b.set_pes(1, judge_interval=1)
    b.compute_forces()
    b.step(3)
    out = b.get_state()
    rmse, ncomp, _, nfb = b.nn_stats()
    fallback = b.nn_fallback()
    st = b.status(clear=True)
    assert not (st & ~MDT_ST_NN_EXTRAPOLATION).any(), st
    om = no.Model(model)
    n_nn = n_fb = 0
    for s, o, fb in zip(systems, out, fallback):
        el = om.elem_index([1] + [79] * (s.natoms - 1))
        for bead in sorted({0, s.nbeads - 1}):
            sr = s.copy()
            sr.r = o["r"][:, bead:bead + 1, :].copy()
            if not fb[bead]:
                E, _, F, nex = om.energy_forces(el, o["r"][:, bead, :] / nene.BOHR2ANG,
                                                s.simbox.T / nene.BOHR2ANG)
                assert abs(o["epot"][bead] - E * nene.HA2EV) / s.natoms <= E_ATOM_TOL
                if s.nbeads == 1:  # no ring-polymer spring forces in f
                    assert np.abs(o["f"][:, 0, :] - F * nene.HABOHR2EVANG).max() <= F_TOL
                n_nn += 1
            else:
                e_r, f_r, _ = oracle.emt_energy_forces(sr, cols, pesm)
                assert abs(o["epot"][bead] - e_r[0]) <= 1e-9 * max(1.0, abs(e_r[0]))
                if s.nbeads == 1:
                    assert np.abs(o["f"][:, 0, :] - f_r[:, 0, :]).max() <= 1e-9
                n_fb += 1
    print(f"uneven surrogate batch: {n_nn} network replicas, {n_fb} on the EMT fallback checked; "
          f"judge RMSE {rmse * 1e3:.2f} meV/A over {ncomp:.0f} components, fallback replicas {nfb}")
    assert n_nn > 0
    assert 0.0 < rmse < 0.02
    b.close()


def test_uneven_surrogate_batch_matches_separate_batches(gpu_ctx, pes_hau):
    """Stepping the uneven surrogate batch equals stepping each trajectory
    alone (same seed and trajectory id), bitwise."""
    model = surrogate.load_model(surrogate.MODEL_DIR_THICK)
    rng = np.random.default_rng(52)
    systems, tsurf = _batch(rng)
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
