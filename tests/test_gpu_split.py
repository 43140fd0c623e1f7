"""One-bead, one-class EMT batches stepped as two independent halves on two
streams (mdt_batch_set_split) give bitwise the trajectories of the whole
batch: eagerly and from hipGraph chunks (each half its own graph and step
counter), with the neighbour lists kept across steps, for an odd number of
trajectories (unequal halves)."""
import numpy as np
import pytest

from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu


def _run(gpu_ctx, systems, split, graph, keep):
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=17)
    b.set_keep(keep)
    b.set_wide(0)
    b.tune(tpa=2)
    b.set_split(split)
    b.set_graph(graph)
    b.compute_forces()
    b.step(25)
    halves = b.last_split()
    b.step(7)
    st = b.get_state()
    assert not b.status().any()
    b.close()
    return st, halves


@pytest.mark.parametrize("graph,keep", [(0, 1), (10, 1), (10, 0), (10, 2)])
def test_split_equals_whole_batch(gpu_ctx, pes_hau, graph, keep):
    rng = np.random.default_rng(61)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, nfixed=2, rng=rng, einc=rng.uniform(0.5, 5),
                                  polar_deg=rng.uniform(0, 60), azimuth_deg=rng.uniform(0, 360), tsurf=300.0)
               for _ in range(33)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    whole, h0 = _run(gpu_ctx, systems, 0, graph, keep)
    split, h1 = _run(gpu_ctx, systems, 1, graph, keep)
    assert (h0, h1) == (1, 2)
    for a, c in zip(whole, split):
        for k in ("r", "v", "f", "a", "epot", "dens"):
            assert np.array_equal(a[k], c[k]), k


def test_split_refused_mode(gpu_ctx, pes_hau):
    rng = np.random.default_rng(62)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, rng=rng, tsurf=300.0)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, [s, s.copy()], dt=0.1, tsurf=300.0)
    with pytest.raises(mdt.MdtError):
        b.set_split(2)
    b.close()
