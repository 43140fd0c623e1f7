"""GPU parity of the EMT energy / force / density kernels against the oracle.

FP64 path: both sides are FP64 with different summation orders, so the
bar is 1e-9 eV on energies and 1e-9 eV/A on forces (observed differences are
~1e-13), far inside the 1e-6 eV/atom and 1e-5 eV/A stated for the path.
"""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu

E_TOL = 1e-9
F_TOL = 1e-9


def _cmp(ctx, sysm, pes, e_tol=E_TOL, f_tol=F_TOL):
    cols, pesm = pesfile.emt_tables(pes, sysm.type_names)
    ctx.set_emt(cols, pesm)
    e_g, f_g, d_g = ctx.compute_emt(sysm)
    e_r, f_r, d_r = oracle.emt_energy_forces(sysm, cols, pesm)
    assert np.abs(e_g - e_r).max() <= e_tol * max(1.0, np.abs(e_r).max()), (e_g, e_r)
    assert np.abs(f_g - f_r).max() <= f_tol, np.abs(f_g - f_r).max()
    assert np.allclose(d_g, d_r, rtol=1e-12, atol=1e-15)
    return e_g, f_g


@pytest.mark.parametrize("z", [6.0, 3.0, 1.5, 0.8, -1.2])
def test_hau_3x3x4_projectile_heights(gpu_ctx, pes_hau, z):
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(3))
    s.r[0, 0, 2] = z
    _cmp(gpu_ctx, s, pes_hau)


def test_hau_6x6x4_thermal(gpu_ctx, pes_hau):
    s = geometry.h_on_slab("Au", n=6, nlayers=4, rng=np.random.default_rng(5))
    s.r[:, 0, :] += np.random.default_rng(6).normal(0, 0.05, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 1.9
    _cmp(gpu_ctx, s, pes_hau)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("metal,n", [("Pt", 4), ("Au", 4), ("Pt", 2)])
def test_even_cell_half_cell_ties(gpu_ctx, pes_hau, pes_hpt, metal, n, precision):
    """Perfect fixed layers in a cell of even width put atom pairs exactly
    half a cell apart; within the cut-off (4 x 4 and 2 x 2 cells) the pair's
    image is the one the reference's anint picks (universe_mod.f90:411-428)."""
    pes = pes_hpt if metal == "Pt" else pes_hau
    rng = np.random.default_rng(40 + n)
    s = geometry.h_on_slab(metal, n=n, nlayers=4, rng=rng, z0=1.7)
    mob = ~s.fixed[:, 0, :]
    s.r[:, 0, :] += rng.normal(0, 0.05, s.r[:, 0, :].shape) * mob
    cols, pesm = pesfile.emt_tables(pes, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, [s], dt=0.1)
    b.set_precision(precision)
    b.compute_forces()
    o = b.get_state()[0]
    b.close()
    e_r, f_r, _ = oracle.emt_energy_forces(s, cols, pesm)
    if precision == 0:
        assert np.abs(o["epot"] - e_r).max() <= E_TOL * max(1.0, np.abs(e_r).max())
        assert np.abs(o["f"] - f_r).max() <= F_TOL, np.abs(o["f"] - f_r).max()
    else:  # FP32 pair terms: the H/Pt bounds of the precision tests (a wrong image is ~1e-2)
        assert np.abs(o["epot"] - e_r).max() / s.natoms <= 2e-5
        assert np.abs(o["f"] - f_r).max() <= 2e-3, np.abs(o["f"] - f_r).max()


def test_hpt_8x8x6(gpu_ctx, pes_hpt):
    s = geometry.h_on_slab("Pt", n=8, nlayers=6, rng=np.random.default_rng(8))
    s.r[0, 0, 2] = 1.1
    _cmp(gpu_ctx, s, pes_hpt)


def test_bulk_au_one_species(gpu_ctx, pes_hau):
    s = geometry.bulk_fcc("Au", ncell=3)
    s.r[:, 0, :] += np.random.default_rng(1).normal(0, 0.03, s.r[:, 0, :].shape)
    _cmp(gpu_ctx, s, pes_hau)


def test_beads_are_independent_replicas(gpu_ctx, pes_hau):
    # reference semantics: bead b of atom i interacts with bead b of atom j
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nbeads=4, rng=np.random.default_rng(2))
    g = np.random.default_rng(4)
    s.r += g.normal(0, 0.04, s.r.shape)
    s.r[0, :, 2] = [1.0, 1.3, 1.7, 2.4]
    _cmp(gpu_ctx, s, pes_hau)


def test_compute_emt_accumulates(gpu_ctx, pes_hau):
    # compute_emt adds into epot and f (pes_emt_mod.f90:976, 1001-1011)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(3))
    s.r[0, 0, 2] = 1.5
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    e0, f0, _ = gpu_ctx.compute_emt(s)
    e1 = np.full(1, 2.5)
    f1 = np.ones((s.natoms, 1, 3))
    gpu_ctx.compute_emt(s, epot=e1, f=f1)
    assert np.allclose(e1, e0 + 2.5, rtol=0, atol=1e-12)
    assert np.allclose(f1, f0 + 1.0, rtol=0, atol=1e-12)


def test_batch_forces_many_trajectories(gpu_ctx, pes_hau):
    g = np.random.default_rng(11)
    systems = []
    for t in range(24):
        s = geometry.h_on_slab("Au", n=6, nlayers=4, rng=g, einc=g.uniform(0.5, 5))
        s.r[0, 0, 2] = g.uniform(0.5, 4.0)
        systems.append(s)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0)
    b.compute_forces()
    out = b.get_state()
    for s, o in zip(systems, out):
        e_r, f_r, d_r = oracle.emt_energy_forces(s, cols, pesm)
        assert abs(o["epot"][0] - e_r[0]) <= E_TOL * max(1, abs(e_r[0]))
        assert np.abs(o["f"] - f_r).max() <= F_TOL
        assert np.allclose(o["dens"], d_r, rtol=1e-12, atol=1e-15)
        # set_acceleration: a = f/m where not fixed, 0 where fixed (force.f90:89-98)
        a_ref = np.where(s.fixed, 0.0, f_r / s.mass[:, None, None])
        assert np.abs(o["a"] - a_ref).max() <= 1e-12
    b.close()


@pytest.mark.parametrize("tpa", [1, 2, 4, 8])
def test_launch_shapes_agree_bitwise(gpu_ctx, pes_hau, tpa):
    g = np.random.default_rng(12)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, rng=g) for _ in range(4)]
    for s in systems:
        s.r[0, 0, 2] = 1.4
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    outs = []
    for t in (1, tpa):
        b = mdt.Batch(gpu_ctx, systems)
        b.tune(tpa=t)
        b.compute_forces()
        outs.append(b.get_state())
        b.close()
    for o1, o2 in zip(*outs):
        # same neighbour order; lane partial sums differ only by association
        assert np.abs(o1["f"] - o2["f"]).max() <= 1e-11
        assert abs(o1["epot"][0] - o2["epot"][0]) <= 1e-10


@pytest.mark.parametrize("z", [6.9, 7.4, 7.7])
def test_projectile_at_cutoff_edge(gpu_ctx, pes_hau, z):
    """sigma of the H is ~1e-35..1e-20 here, far below any fixed-point
    quantum of a single int64 word; the GPU's partner sums must still
    reproduce the reference's G = -D/(beta eta2 max(1e-30, sigma)) force."""
    s = geometry.h_on_slab("Au", n=6, nlayers=4, rng=np.random.default_rng(13))
    top = s.r[1, 0].copy()  # an atom of the top layer
    s.r[0, 0, :2] = top[:2]
    s.r[0, 0, 2] = z
    cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
    gpu_ctx.set_emt(cols, pesm)
    e_g, f_g, _ = gpu_ctx.compute_emt(s)
    e_r, f_r, _ = oracle.emt_energy_forces(s, cols, pesm)
    assert abs(e_g[0] - e_r[0]) <= 1e-9 * abs(e_r[0])
    assert np.abs(f_g - f_r).max() <= 1e-9
    # the H force is the s-clamp quirk value, not zero; it agrees to the
    # resolution of the 2^42 fixed-point force sums (2.3e-13 eV/A per term)
    assert abs(f_r[0, 0, 2]) > 1e-12
    assert abs(f_g[0, 0, 2] - f_r[0, 0, 2]) <= 1e-12


def test_largest_supported_trajectory(gpu_ctx, pes_hau):
    """2026 atoms (H over a 15x15x9 Au slab, lists in global memory): the
    largest size class against the oracle."""
    s = geometry.h_on_slab("Au", n=15, nlayers=9, rng=np.random.default_rng(21))
    assert s.natoms == 2026
    s.r[:, 0, :] += np.random.default_rng(22).normal(0, 0.02, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 1.2
    _cmp(gpu_ctx, s, pes_hau)


def test_most_beads(gpu_ctx, pes_hau):
    """64 beads (the largest supported count) on every atom."""
    s = geometry.h_on_slab("Au", n=3, nlayers=4, nbeads=64, rng=np.random.default_rng(23))
    s.r += np.random.default_rng(24).normal(0, 0.03, s.r.shape)
    s.r[0, :, 2] = np.linspace(0.9, 3.0, 64)
    _cmp(gpu_ctx, s, pes_hau)


def test_sizes_beyond_the_limits_are_refused(gpu_ctx, pes_hau):
    """2049 atoms or 65 beads: a clear error before any launch."""
    for kw in (dict(n=16, nlayers=8), dict(n=3, nlayers=4, nbeads=65)):
        s = geometry.h_on_slab("Au", rng=np.random.default_rng(25), **kw)
        cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
        gpu_ctx.set_emt(cols, pesm)
        with pytest.raises(mdt.MdtError, match="supported|natoms|beads"):
            mdt.Batch(gpu_ctx, [s], dt=0.1).compute_forces()


def test_tiny_atoms_far_pairs_at_half_cell_ties(gpu_ctx, pes_hau):
    """Atoms with no pair within the near radius are tiny: their pairs
    between r_near (6.3 A) and the cut-off (7.72 A) come from the direct
    scan, whose minimum image is decided in FP32 and, within 1e-4 of a
    half-cell tie, taken from the unfolded positions by the reference's
    rounding of s = isimbox (r_j - r_i) (src/universe_mod.f90:411-428).  Two
    Au atoms exactly half the 14.5 A x width apart (7.25 A) and two half the
    15 A y width apart (7.5 A) make far pairs of two tiny atoms at ties:
    the image picked sets the force's sign, and sigma ~ 1e-30 puts them in
    the s-clamp regime.  An H atom far from everything is tiny with no
    partner at all."""
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(1))
    pos = np.array([[5.0, 3.0, 22.0],                       # H, alone
                    [1.0, 2.0, 4.0], [1.0 + 7.25, 2.0, 4.0],  # x tie, 7.25 A
                    [3.0, 4.0, 12.0], [3.0, 4.0 + 7.5, 12.0]])  # y tie, 7.5 A
    n = len(pos)
    idx = np.array([0, 1, 1, 1, 1], dtype=np.int32)
    s = geometry.System(list(tmpl.type_names), idx, tmpl.mass[[0, 1, 1, 1, 1]].copy(),
                        np.zeros((n, 1, 3), dtype=bool), tmpl.algo.copy(),
                        np.diag([14.5, 15.0, 30.0]), pos.reshape(n, 1, 3).copy(), np.zeros((n, 1, 3)))
    e_g, f_g = _cmp(gpu_ctx, s, pes_hau)
    e_r, f_r, _ = oracle.emt_energy_forces(s, *pesfile.emt_tables(pes_hau, s.type_names))
    # the tie pairs interact (nonzero forces along the tie axis) and the signs agree
    assert abs(f_r[1, 0, 0]) > 0 and abs(f_r[3, 0, 1]) > 0
    assert np.sign(f_g[1, 0, 0]) == np.sign(f_r[1, 0, 0]) and np.sign(f_g[3, 0, 1]) == np.sign(f_r[3, 0, 1])
    # the same system shifted by whole cell vectors (unfolded positions change
    # the rounded s, so the reference may pick the other image) still agrees
    s2 = s.copy()
    s2.r[2, 0, 0] -= 14.5
    s2.r[4, 0, 1] += 15.0
    _cmp(gpu_ctx, s2, pes_hau)
