"""1024-lane workgroups for the large size classes (mdt_batch_set_wide).

A class whose atoms x threads per atom exceed 512 (k_replica) or whose
classical atoms x threads per atom exceed 512 (k_mixed, the mixed bead
model) may run one 1024-lane workgroup per replica instead of 512 lanes
looping over its atoms.  With the same threads per atom the two widths
accumulate the same terms into the same fixed-point words and sum each
owner's terms in the same lane groups, so trajectories are bitwise equal.
The reference has no launch shape; its forces are the oracle's, which the
config-size tests check for the autotuned shapes."""
import numpy as np
import pytest

from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bead_mode", [0, 1])
def test_wide_workgroups_equal_512_lanes(gpu_ctx, pes_hau, bead_mode):
    rng = np.random.default_rng(91)
    systems = [geometry.h_on_slab("Au", n=n, nlayers=6, nfixed=2, rng=rng, nbeads=p, z0=z, einc=2.0, tsurf=300.0)
               for n, p, z in ((12, 1, 2.0), (12, 4, 3.0), (9, 4, 1.8), (12, 1, 6.0), (9, 1, 2.5))]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    states, blocks = {}, {}
    for wide in (0, 1):
        b = mdt.Batch(gpu_ctx, systems, dt=0.2, tsurf=300.0, tproj=300.0, seed=5)
        b.set_bead_mode(bead_mode)
        b.set_wide(wide)
        b.tune(tpa=1)
        blocks[wide] = [c["block"] for c in b.class_shapes()]
        b.compute_forces()
        b.step(20)
        assert not b.status().any()
        states[wide] = b.get_state()
        b.close()
    print(f"bead mode {bead_mode}: replica-class blocks {blocks}")
    assert max(blocks[0]) <= 512
    assert max(blocks[1]) > 512  # the 865-atom class ran 1024-lane workgroups
    for a, c in zip(states[0], states[1]):
        for k in ("r", "v", "f", "epot"):
            assert np.array_equal(a[k], c[k]), k


def test_set_wide_rejects_bad_mode(gpu_ctx, pes_hau):
    rng = np.random.default_rng(92)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=rng)
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, [s], dt=0.2, tsurf=300.0, tproj=300.0)
    with pytest.raises(mdt.MdtError):
        b.set_wide(2)
    b.close()
