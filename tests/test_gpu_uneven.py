"""Uneven batches: trajectories of different slab sizes and bead counts in
one batch, launched per size class (replicas grouped by atom count, each
class with its own threads per atom, block and LDS shape)."""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu


def _mixed(rng):
    spec = [(3, 1), (6, 4), (9, 1), (3, 4), (6, 1)]  # (n, beads): N = 55, 217, 487
    out = []
    for n, p in spec:
        s = geometry.h_on_slab("Au", n=n, nlayers=6, nfixed=2, rng=rng, nbeads=p, einc=rng.uniform(0.5, 5),
                               tsurf=300.0)
        s.r += rng.normal(0, 0.03, s.r.shape)
        s.r[0, :, 2] = rng.uniform(1.0, 3.0)
        out.append(s)
    return out


def test_uneven_batch_forces(gpu_ctx, pes_hau):
    rng = np.random.default_rng(17)
    systems = _mixed(rng)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    # ring-polymer frequencies scale with kB Tproj P / hbar: P > 1 needs Tproj > 0
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, tproj=300.0)
    b.tune(tpa=0)  # per-class autotune
    b.compute_forces()
    for s, o in zip(systems, b.get_state()):
        # the oracle's calc_force: the free ring-polymer step first when P > 1
        # (src/force.f90:54), spring forces on Langevin beads (src/rpmd.f90:418-435)
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=300.0, tproj=300.0)
        st.calc_force()
        assert np.abs(o["r"] - st.r).max() <= 1e-12
        assert np.abs(o["epot"] - st.epot).max() <= 1e-9 * max(1.0, np.abs(st.epot).max())
        assert np.abs(o["f"] - st.f).max() <= 1e-9
    b.close()


def test_uneven_batch_matches_separate_batches(gpu_ctx, pes_hau):
    """Stepping the mixed batch equals stepping each trajectory alone (same
    seed and trajectory id: the noise is keyed by them, not by the layout)."""
    rng = np.random.default_rng(18)
    systems = _mixed(rng)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, tproj=300.0, seed=21, traj_ids=list(range(5)))
    b.compute_forces()
    b.step(6)
    together = b.get_state()
    b.close()
    for t, s in enumerate(systems):
        b1 = mdt.Batch(gpu_ctx, [s], dt=0.1, tsurf=300.0, tproj=300.0, seed=21, traj_ids=[t])
        b1.compute_forces()
        b1.step(6)
        alone = b1.get_state()[0]
        b1.close()
        for k in ("r", "v", "f", "epot"):
            assert np.array_equal(together[t][k], alone[k]), (t, k)


@pytest.mark.parametrize("bead_mode", [0, 1])
def test_uneven_graph_replay_equals_forked_launches(gpu_ctx, pes_hau, bead_mode):
    """Eager steps fork the size classes (and the mixed-model kernel) onto
    side streams; graph capture keeps them on the batch's stream.  Both give
    bitwise the same state, replay after replay with eager work (state
    downloads) between the replays."""
    rng = np.random.default_rng(23)
    systems = _mixed(rng)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    batches = []
    for chunk in (0, 5):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, tproj=300.0, seed=8)
        b.set_bead_mode(bead_mode)
        b.set_graph(chunk)
        b.compute_forces()
        batches.append(b)
    for rnd in range(3):
        outs = []
        for b in batches:
            b.step(10)  # two replays of the captured chunk per round
            assert not b.status().any()
            outs.append(b.get_state())
        for o0, o1 in zip(*outs):
            for k in ("r", "v", "f", "a", "epot", "dens"):
                assert np.array_equal(o0[k], o1[k]), (rnd, k)
    for b in batches:
        b.close()
