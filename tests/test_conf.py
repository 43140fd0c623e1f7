"""Configuration files (mxt, POSCAR, conf merge) on the reference's own
configuration files: conf/proj/mxt_00000001.dat (an H projectile),
conf/latt/mxt_00000001.dat (a 24-atom graphene sheet), conf/example.POSCAR
and conf/cart.POSCAR, copied under tests/golden/conf as fixtures.

The mxt layout is pinned by the files themselves: the two record lengths
(12 and 219 / 1691 bytes) must equal the layout's sizes for the header's
natoms, nbeads, ntypes, and the decoded sheet must be graphene (every atom
with three neighbours at the C-C bond length 1.42 A) -- a wrong array order,
cell orientation or field offset breaks one of these.  Writing a file back
reproduces it byte for byte."""
import os

import numpy as np
import pytest

from md_tian2_amd import conf, geometry

FIX = os.path.join(os.path.dirname(__file__), "golden", "conf")


def _min_image_dists(c):
    r = c.cartesian()[:, 0, :]
    isb = np.linalg.inv(c.simbox)
    d = r[:, None, :] - r[None, :, :]
    s = d @ isb.T
    s -= np.round(s)
    d = s @ c.simbox.T
    return np.sqrt((d * d).sum(-1))


def test_projectile_mxt():
    c = conf.read_mxt(os.path.join(FIX, "proj", "mxt_00000001.dat"))
    assert (c.natoms, c.nbeads, c.ntypes) == (1, 1, 1)
    assert c.names == ["H"] and c.is_proj == [True]
    assert np.allclose(c.simbox @ c.isimbox, np.eye(3), atol=1e-12)
    assert np.isfinite(c.r).all() and np.isfinite(c.v).all()


def test_lattice_mxt_is_graphene():
    c = conf.read_mxt(os.path.join(FIX, "latt", "mxt_00000001.dat"))
    assert (c.natoms, c.nbeads, c.ntypes) == (24, 1, 1)
    assert c.names == ["C"] and c.is_proj == [False]
    assert np.allclose(c.simbox @ c.isimbox, np.eye(3), atol=1e-12)
    d = _min_image_dists(c)
    np.fill_diagonal(d, np.inf)
    nn = np.sort(d, axis=1)[:, :3]
    assert np.abs(nn - 1.42).max() < 0.08, nn.min()
    assert (np.sort(d, axis=1)[:, 3] > 2.2).all()  # the fourth neighbour is the second shell


@pytest.mark.parametrize("name", ["proj", "latt"])
def test_mxt_round_trip_is_byte_exact(tmp_path, name):
    src = os.path.join(FIX, name, "mxt_00000001.dat")
    c = conf.read_mxt(src)
    out = tmp_path / "mxt.dat"
    conf.write_mxt(out, c)
    assert open(src, "rb").read() == open(out, "rb").read()


def test_mxt_rejects_wrong_layout(tmp_path):
    raw = bytearray(open(os.path.join(FIX, "proj", "mxt_00000001.dat"), "rb").read())
    raw[4:8] = (2).to_bytes(4, "little")  # natoms 2: the universe record is too short for it
    p = tmp_path / "bad.dat"
    p.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        conf.read_mxt(p)


def test_example_poscar():
    c = conf.read_poscar(os.path.join(FIX, "example.POSCAR"), nprojectiles=1)
    assert c.names == ["H", "C"] and c.is_proj == [True, False]
    assert (c.natoms, c.nbeads) == (25, 1)
    assert list(c.idx[:2]) == [1, 2] and (c.idx[1:] == 2).all()
    assert c.is_cart and not c.fixed.any()
    assert np.allclose(c.r[0, 0], [5.31845883, 0.46640574, 4.16117769])
    assert np.allclose(np.diag(c.simbox), [8.54759939, 7.40243847, 13.00547707])
    d = _min_image_dists(conf.Conf(c.r[1:], c.v[1:], True, c.fixed[1:], c.idx[1:], c.names, c.is_proj,
                                   c.simbox, c.isimbox))
    np.fill_diagonal(d, np.inf)
    assert np.abs(np.sort(d, axis=1)[:, 0] - 1.42).max() < 0.1  # graphene under the H


def test_cart_poscar():
    c = conf.read_poscar(os.path.join(FIX, "cart.POSCAR"), nprojectiles=1)
    assert c.names == ["H", "Au"] and (c.natoms, c.nbeads) == (3, 1)
    assert np.allclose(c.r[:, 0, :], [[6, 5, 1], [6, 5, 3], [5, 6, 1]])
    assert (c.v == 0).all()


def test_poscar_beads_flags_velocities(tmp_path):
    P = 4
    rng = np.random.default_rng(1)
    r = rng.uniform(0, 1, (2, P, 3))
    v = rng.normal(0, 0.01, (2, P, 3))
    lines = ["beads", "2.0", "5 0 0", "0 6 0", "0 0 7", "H Au", f":{P}: 1 1", "direct"]
    for i in range(2):
        for b in range(P):
            fl = "F T F" if i == 1 else "T T T"
            lines.append("%.12f %.12f %.12f %s" % (*r[i, b], fl))
    lines.append("")
    lines += ["%.15e %.15e %.15e" % tuple(v[i, b]) for i in range(2) for b in range(P)]
    p = tmp_path / "b.POSCAR"
    p.write_text("\n".join(lines) + "\n")
    c = conf.read_poscar(p, nprojectiles=1)
    assert c.nbeads == P and not c.is_cart
    assert np.allclose(c.r, r, atol=1e-12)
    assert np.allclose(c.v, v, rtol=1e-14)
    assert np.allclose(np.diag(c.simbox), [10, 12, 14])          # scaling applies to the cell
    assert np.allclose(c.cartesian(), r * [10, 12, 14], atol=1e-11)  # not to direct coordinates
    assert (c.fixed[1, :, 0] & ~c.fixed[1, :, 1] & c.fixed[1, :, 2]).all() and not c.fixed[0].any()


def test_merge_projectile_and_lattice(tmp_path):
    latt = conf.read_mxt(os.path.join(FIX, "latt", "mxt_00000001.dat"))
    proj = conf.read_mxt(os.path.join(FIX, "proj", "mxt_00000001.dat"))
    m = conf.merge(latt, proj)
    assert m.natoms == 25 and m.names == ["H", "C"] and m.is_proj == [True, False]
    assert list(m.idx) == [1] + [2] * 24
    # positions are the inputs' up to whole cell vectors, centroids folded toward the cell
    d = np.einsum("ij,npj->npi", m.isimbox, m.r - np.concatenate([proj.cartesian(), latt.cartesian()]))
    assert np.allclose(d, np.round(d), atol=1e-9)
    f = np.einsum("ij,npj->npi", m.isimbox, m.r).mean(axis=1)
    assert (f < 1).all() and (f > -1).all()
    out = tmp_path / "m.dat"
    conf.write_mxt(out, m)
    m2 = conf.read_mxt(out)
    assert np.array_equal(m2.r, m.r) and m2.names == m.names and list(m2.idx) == list(m.idx)
    with pytest.raises(ValueError):
        conf.merge(proj, latt)


def test_to_system():
    c = conf.read_poscar(os.path.join(FIX, "cart.POSCAR"), nprojectiles=1)
    s = c.to_system(algos=[geometry.PROP_VERLET, geometry.PROP_LANGEVIN])
    assert s.natoms == 3 and s.type_names == [("H", True), ("Au", False)]
    assert np.allclose(s.mass[1] / geometry.AMU2MASS, geometry.MASS_AMU["Au"])
    assert list(s.algo) == [geometry.PROP_VERLET, geometry.PROP_LANGEVIN]


def test_pes_blocks_of_the_reference_files():
    """The reference's .pes files (tests/golden/pes): nene blocks give the
    network's directory and warning limit, non blocks non-interacting
    pairs; REBO is refused (no GPU kernel)."""
    from md_tian2_amd import pesfile
    d = os.path.join(os.path.dirname(__file__), "golden", "pes")
    p = pesfile.load_pes(os.path.join(d, "NENE-HC.pes"))
    assert p.nene_dir == "RuNNer_input/" and p.nene_max_warnings == 100
    assert (("C", False), ("H", True)) in p.nene_pairs and not p.pairs and not p.params
    q = pesfile.load_pes(os.path.join(d, "NON-H.pes"))
    assert q.non_pairs == {(("H", True), ("H", True))} and not q.pairs
    cols, pesm = pesfile.emt_tables(q, [("H", True)])
    assert pesm == [[0]]
    with pytest.raises(ValueError, match="rebo"):
        pesfile.load_pes(os.path.join(d, "REBO-EMFT_EMT-HCPt.pes"))
