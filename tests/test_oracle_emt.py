"""Pins of the oracle's EMT (pes_emt_mod.f90 restatement) by mathematics
that does not restate its formulas:

* perfect fcc bulk at the EMT lattice constant a0 = sqrt(2) beta s0 has
  sigma = 12 for every atom by construction of gamma1 (pes_emt_mod.f90:579-592),
  so E = 0 and F = 0 up to the theta tail of shells 4-7 (< 1e-9 eV/atom);
* the cohesive energy of the isolated metal atom is -E0 (a textbook EMT
  property, to the accuracy of the s clamp);
* energies are extensive and separable: two atoms farther apart than the
  cut-off add their isolated energies, an H far above the slab adds its
  isolated energy to the slab's, and the 2-species path with the H far away
  equals the 1-species path on the bare slab;
* forces equal -grad E by central finite differences (h = 1e-5 A), sum to
  zero, and transform with rotations, translations and permutations;
* minimum-image vectors on hand-built cells (universe_mod.f90:411-428);
* the H-metal cross terms (chi_12 / chi_21, sigma_12 / sigma_21, V_12 / V_21,
  pes_emt_mod.f90:552-553, 721-759, 767-788, 793-794, 976): two types with
  identical parameters reduce the 2-species path to the 1-species one, and
  H-Au / H-Pt dimers and an H-Au3 cluster (chi_21 = 6.5 and 4.5) match a
  40-digit mpmath evaluation of the published EMT equations, energies,
  densities and forces (its derivative).  Planting a swapped chi, a wrong
  gamma or the metal's eta2 in sigma_21 (energy and force alike) in the
  oracle fails these checks.
"""
import math

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, pesfile


def energy(sysm, pes):
    cols, pesm = pesfile.emt_tables(pes, sysm.type_names)
    e, f, d = oracle.emt_energy_forces(sysm, cols, pesm)
    return e, f, d


def emt_a0(pes, elem):
    beta = (16 * math.pi / 3) ** (1 / 3) / math.sqrt(2)
    return math.sqrt(2) * beta * pes.params[(elem, False)]["s0"]


@pytest.mark.parametrize("elem,fixture", [("Au", "pes_hau"), ("Pt", "pes_hpt")])
def test_bulk_at_emt_lattice_constant_is_zero(elem, fixture, request):
    pes = request.getfixturevalue(fixture)
    s = geometry.bulk_fcc(elem, a=emt_a0(pes, elem), ncell=4)
    e, f, d = energy(s, pes)
    assert abs(e[0]) / s.natoms < 1e-9
    assert np.abs(f).max() < 1e-10
    # dens = n0 at sigma = 12 (s = 0)
    assert np.allclose(d, pes.params[(elem, False)]["n0"], rtol=1e-8)


def test_isolated_atom_cohesive_energy(pes_hau):
    s = geometry.bulk_fcc("Au", a=60.0, ncell=1)  # 4 atoms 30+ A apart
    e, f, _ = energy(s, pes_hau)
    e0 = pes_hau.params[("Au", False)]["e0"]
    assert abs(e[0] / 4 - (-e0)) < 1e-3
    assert np.abs(f).max() == 0.0


def test_two_far_atoms_add(pes_hau):
    s1 = geometry.bulk_fcc("Au", a=60.0, ncell=1)
    s1.r = s1.r[:1].copy()
    s1.idx, s1.mass, s1.fixed, s1.v = s1.idx[:1], s1.mass[:1], s1.fixed[:1], s1.v[:1]
    e1, _, _ = energy(s1, pes_hau)
    s2 = geometry.bulk_fcc("Au", a=60.0, ncell=1)
    e4, _, _ = energy(s2, pes_hau)
    assert abs(e4[0] - 4 * e1[0]) < 1e-12


def test_far_projectile_separates(pes_hau):
    slab = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(0))
    slab.r[0, 0, 2] = 15.0  # beyond 1.5 rcut of every Au
    e_hs, f_hs, _ = energy(slab, pes_hau)
    bare = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(0))
    bare.type_names = [("Au", False)]
    bare.idx = bare.idx[1:] - 1
    bare.r, bare.v, bare.mass, bare.fixed = bare.r[1:], bare.v[1:], bare.mass[1:], bare.fixed[1:]
    bare.algo = bare.algo[1:]
    e_b, f_b, _ = energy(bare, pes_hau)  # 1-species path
    # isolated H: sigma = 0 -> clamped; its energy is the same in any cell
    h = geometry.bulk_fcc("Au", a=60.0, ncell=1)
    h.type_names = [("H", True)]
    h.r, h.v, h.mass, h.fixed, h.idx = h.r[:1], h.v[:1], h.mass[:1], h.fixed[:1], h.idx[:1]
    e_h, _, _ = energy(h, pes_hau)
    assert abs(e_hs[0] - (e_b[0] + e_h[0])) < 1e-10
    assert np.abs(f_hs[1:] - f_b).max() < 1e-12


def _fd_forces(sysm, pes, atoms, h=1e-5):
    out = np.zeros((len(atoms), 3))
    for a_i, a in enumerate(atoms):
        for c in range(3):
            sp = sysm.copy()
            sp.r[a, 0, c] += h
            sm = sysm.copy()
            sm.r[a, 0, c] -= h
            out[a_i, c] = -(energy(sp, pes)[0][0] - energy(sm, pes)[0][0]) / (2 * h)
    return out


@pytest.mark.parametrize("z", [2.5, 1.2, 0.6])
def test_forces_are_minus_gradient_hau(pes_hau, z):
    g = np.random.default_rng(1)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g)
    s.r[:, 0, :] += g.normal(0, 0.06, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = z
    _, f, _ = energy(s, pes_hau)
    atoms = [0, 1, 5, 17, 30]
    fd = _fd_forces(s, pes_hau, atoms)
    assert np.abs(f[atoms, 0, :] - fd).max() < 2e-6


def test_forces_are_minus_gradient_hpt(pes_hpt):
    g = np.random.default_rng(2)
    s = geometry.h_on_slab("Pt", n=3, nlayers=3, nfixed=1, rng=g)
    s.r[:, 0, :] += g.normal(0, 0.05, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 1.0
    _, f, _ = energy(s, pes_hpt)
    atoms = [0, 2, 11, 20]
    fd = _fd_forces(s, pes_hpt, atoms)
    assert np.abs(f[atoms, 0, :] - fd).max() < 2e-5


def test_far_h_force_quirk_is_not_gradient(pes_hau):
    """The s clamp (pes_emt_mod.f90:844) freezes the energy for sigma < 1e-9
    while ds/dr keeps dividing by the 1e-30 clamp (:793,816): the H force
    is then not -dE/dr.  The oracle must reproduce that, not fix it."""
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(0))
    s.r[0, 0, 2] = 6.5
    _, f, _ = energy(s, pes_hau)
    fd = _fd_forces(s, pes_hau, [0])
    assert abs(fd[0, 2]) < 1e-12 < abs(f[0, 0, 2])


def test_sum_of_forces_vanishes(pes_hau):
    g = np.random.default_rng(3)
    s = geometry.h_on_slab("Au", n=6, nlayers=4, rng=g)
    s.r[:, 0, :] += g.normal(0, 0.05, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 1.3
    _, f, _ = energy(s, pes_hau)
    assert np.abs(f.sum(axis=(0, 1))).max() < 1e-11


def _rotz(th):
    c, s_ = math.cos(th), math.sin(th)
    return np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])


def test_rotation_translation_permutation_invariance(pes_hau):
    g = np.random.default_rng(4)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g)
    s.r[:, 0, :] += g.normal(0, 0.05, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 1.4
    e, f, d = energy(s, pes_hau)
    # rotation of positions and cell
    R = _rotz(0.7) @ np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    sr = s.copy()
    sr.r = np.einsum("ij,npj->npi", R, s.r)
    sr.simbox = R @ s.simbox
    er, fr, dr = energy(sr, pes_hau)
    assert abs(er[0] - e[0]) < 1e-11
    assert np.abs(fr - np.einsum("ij,npj->npi", R, f)).max() < 1e-11
    # translation
    st = s.copy()
    st.r = s.r + np.array([3.3, -1.7, 0.9])
    et, ft, _ = energy(st, pes_hau)
    assert abs(et[0] - e[0]) < 1e-11 and np.abs(ft - f).max() < 1e-11
    # permutation of the metal atoms (the projectile stays first)
    perm = np.concatenate([[0], 1 + g.permutation(s.natoms - 1)])
    sp = s.copy()
    sp.r, sp.idx, sp.mass, sp.fixed = s.r[perm], s.idx[perm], s.mass[perm], s.fixed[perm]
    ep, fp, dp = energy(sp, pes_hau)
    assert abs(ep[0] - e[0]) < 1e-11
    assert np.abs(fp - f[perm]).max() < 1e-11
    assert np.allclose(dp, d[perm], rtol=1e-13)


def test_minimum_image_hand_built():
    # cubic 10 A cell: r_j - r_i = (8, -9, 0.5) -> (-2, 1, 0.5)
    s = geometry.bulk_fcc("Au", a=10.0, ncell=1)
    s.r = np.zeros((2, 1, 3))
    s.r[0, 0] = [1.0, 9.5, 2.0]
    s.r[1, 0] = [9.0, 0.5, 2.5]
    s.idx, s.mass, s.fixed, s.v = s.idx[:2], s.mass[:2], s.fixed[:2], s.v[:2]
    st = oracle.OracleState(s, {k: [0.0] for k in pesfile.EMT_KEYS}, [[0]])
    d, vec = st.minimg(0, 1)
    assert np.allclose(vec[0], [-2.0, 1.0, 0.5]) and abs(d[0] - math.sqrt(5.25)) < 1e-14
    # hexagonal cell, 60 degrees: a1=(3,0,0), a2=(1.5, 3 sqrt3/2, 0); r = a1 + a2 - eps
    h = np.array([[3.0, 1.5, 0.0], [0.0, 1.5 * math.sqrt(3), 0.0], [0.0, 0.0, 20.0]])
    s2 = s.copy()
    s2.simbox = h
    s2.r[0, 0] = [0.0, 0.0, 0.0]
    s2.r[1, 0] = h[:, 0] + h[:, 1] - np.array([0.1, 0.0, 0.0])
    st2 = oracle.OracleState(s2, {k: [0.0] for k in pesfile.EMT_KEYS}, [[0]])
    d2, v2 = st2.minimg(0, 1)
    assert np.allclose(v2[0], [-0.1, 0.0, 0.0]) and abs(d2[0] - 0.1) < 1e-13


def test_min_image_truncation_in_small_cell(pes_hau):
    """In the 3x3 Au cell (8.65 A) 1.5 rcut = 7.72 A exceeds half the cell:
    each pair counts once, at its nearest image (pes_emt_mod.f90:661)."""
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=np.random.default_rng(0))
    width = np.linalg.norm(s.simbox[:, 0])
    assert 7.7 > width / 2
    e, f, _ = energy(s, pes_hau)
    assert np.isfinite(e).all() and np.abs(f.sum(axis=(0, 1))).max() < 1e-11


# ---------------------------------------------------------------------------
# H-metal cross terms (chi_12 / chi_21, sigma_12 / sigma_21, V_12 / V_21)
# ---------------------------------------------------------------------------

def _system(types, type_names, pos, box=60.0):
    """A universe of len(pos) atoms in a cubic box, one bead, all mobile."""
    N = len(pos)
    return geometry.System(list(type_names), np.array(types, dtype=np.int32), np.ones(N),
                           np.zeros((N, 1, 3), dtype=bool),
                           np.array([geometry.PROP_VERLET] * len(type_names), dtype=np.int32),
                           np.eye(3) * box, np.array(pos, dtype=np.float64).reshape(N, 1, 3), np.zeros((N, 1, 3)))


def test_species_reduction_equals_one_species(pes_hau):
    """Two EMT types with identical (Au) parameters, type-1 atoms first, must
    give the 1-species result: chi_21 = chi_12 = 1, sigma_1 = sigma_11 +
    sigma_12 and sigma_2 = sigma_22 + sigma_21 are every atom's full sigma,
    and V_11 + V_22 + (V_21 + V_12)/2 is the full pair sum -- so sigma
    placement, the gamma scaling of each block and the half-sum algebra of
    pes_emt_mod.f90:721-759, 767-788, 976 reduce to :230-477."""
    g = np.random.default_rng(11)
    s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g)
    s.r[:, 0, :] += g.normal(0, 0.07, s.r[:, 0, :].shape)
    s.r[0, 0, 2] = 2.1  # the leading atom inside the slab's range
    au = pes_hau.params[("Au", False)]
    one = s.copy()
    one.type_names = [("Au", False)]
    one.idx = np.zeros(s.natoms, dtype=np.int32)
    one.algo = one.algo[:1]
    cols1 = {k: [au[k]] for k in pesfile.EMT_KEYS}
    e1, f1, d1 = oracle.emt_energy_forces(one, cols1, [[pesfile.PES_ID_EMT]])
    for k1 in (1, 4, 10):  # number of leading type-1 atoms
        two = s.copy()
        two.type_names = [("Au", True), ("Au", False)]
        two.idx = np.array([0] * k1 + [1] * (s.natoms - k1), dtype=np.int32)
        cols2 = {k: [au[k], au[k]] for k in pesfile.EMT_KEYS}
        e2, f2, d2 = oracle.emt_energy_forces(two, cols2, [[pesfile.PES_ID_EMT] * 2] * 2)
        assert abs(e2[0] - e1[0]) <= 1e-12 * max(1.0, abs(e1[0])), (k1, e2[0], e1[0])
        assert np.abs(f2 - f1).max() <= 1e-12, k1
        assert np.allclose(d2, d1, rtol=1e-13, atol=0), k1


def _mp_emt(types, pos, par):
    """Energy and embedded densities of an isolated cluster (no periodic
    images) from the published EMT equations as md_tian2 evaluates them,
    in 40-digit arithmetic.  types[i] in {0, 1}: 0 = type 1 (the projectile,
    listed first), 1 = type 2 (the metal); par[t] = dict of eta2, n0, e0,
    lambda, v0, kappa, s0.

      beta = (16 pi / 3)^(1/3) / sqrt(2); r_nn,t = beta s0_t (1, sqrt2, sqrt3)
      r_cut = sqrt3 max_t beta s0_t; r_r = 4 r_cut / (sqrt3 + 2);
      a_cut = ln(10^4) / (r_r - r_cut)
      x_t,k = (12, 6, 24)_k / 12 / (1 + exp(a_cut (r_nn,t,k - r_cut)))
      gamma1_t = sum_k x_t,k exp(-eta2_t (r_nn,t,k - beta s0_t))
      gamma2_t = sum_k x_t,k exp(-kappa_t / beta (r_nn,t,k - beta s0_t))
      chi_21 = n0_1 / n0_2 exp(0.5 / a_Bohr (s0_2 - s0_1)), chi_12 = 1 / chi_21
      pair (r <= 1.5 r_cut): theta = 1 / (1 + exp(a_cut (r - r_cut))),
        fs_t = theta exp(-eta2_t (r - beta s0_t)),
        fv_t = theta exp(-kappa_t / beta (r - beta s0_t));
        a pair of one type adds fs_t to both atoms' own-type sigma and fv_t
        to V_tt; a type-1 / type-2 pair adds fs_1 to the metal's sigma_21,
        fs_2 to the projectile's sigma_12, fv_1 to V_21 and fv_2 to V_12
      sigma_11, sigma_12 / gamma1_1; sigma_22, sigma_21 / gamma1_2;
      V_tt V0_t / gamma2_t; V_21 V0_2 chi_21 / gamma2_2; V_12 V0_1 chi_12 / gamma2_1
      sigma~_1 = sigma_11 + chi_12 sigma_12; sigma~_2 = sigma_22 + chi_21 sigma_21
      s = -ln(max(1e-9, sigma~) / 12) / (beta eta2_t)
      dens = n0_t exp((0.5 / a_Bohr - beta eta2_t) s)
      E = sum_i E0 ((1 + lambda s) e^(-lambda s) - 1)
          - V_11 - V_22 - (V_21 + V_12 - 12 sum_i V0 e^(-kappa s_i)) / 2
    """
    import mpmath as mp

    beta = mp.cbrt(16 * mp.pi / 3) / mp.sqrt(2)
    bohr = mp.mpf("0.529177211")
    P = [{k: mp.mpf(repr(v)) for k, v in par[t].items()} for t in range(2)]
    bs0 = [beta * P[t]["s0"] for t in range(2)]
    rcut = max(bs0) * mp.sqrt(3)
    rr = 4 * rcut / (mp.sqrt(3) + 2)
    acut = mp.log(10000) / (rr - rcut)
    g1, g2 = [], []
    for t in range(2):
        rnn = [bs0[t], bs0[t] * mp.sqrt(2), bs0[t] * mp.sqrt(3)]
        x = [w / mp.mpf(12) / (1 + mp.exp(acut * (rn - rcut))) for w, rn in zip((12, 6, 24), rnn)]
        g1.append(mp.fsum(xk * mp.exp(-P[t]["eta2"] * (rn - bs0[t])) for xk, rn in zip(x, rnn)))
        g2.append(mp.fsum(xk * mp.exp(-P[t]["kappa"] / beta * (rn - bs0[t])) for xk, rn in zip(x, rnn)))
    chi21 = P[0]["n0"] / P[1]["n0"] * mp.exp(mp.mpf("0.5") / bohr * (P[1]["s0"] - P[0]["s0"]))
    chi12 = 1 / chi21
    n = len(types)
    own = [mp.mpf(0)] * n    # sigma_11 (type 1) or sigma_22 (type 2)
    cross = [mp.mpf(0)] * n  # sigma_12 (type 1) or sigma_21 (type 2)
    Vtt = [mp.mpf(0), mp.mpf(0)]
    V21 = V12 = mp.mpf(0)
    for i in range(n):
        for j in range(i + 1, n):
            d = [mp.mpf(pos[j][c]) - mp.mpf(pos[i][c]) for c in range(3)]  # exact for floats
            r = mp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
            if r > mp.mpf("1.5") * rcut:
                continue
            th = 1 / (1 + mp.exp(acut * (r - rcut)))
            fs = [th * mp.exp(-P[t]["eta2"] * (r - bs0[t])) for t in range(2)]
            fv = [th * mp.exp(-P[t]["kappa"] / beta * (r - bs0[t])) for t in range(2)]
            ti, tj = types[i], types[j]
            if ti == tj:
                own[i] += fs[ti]
                own[j] += fs[ti]
                Vtt[ti] += fv[ti]
            else:
                assert ti == 0 and tj == 1, "type-1 atoms come first"
                cross[j] += fs[0]
                cross[i] += fs[1]
                V21 += fv[0]
                V12 += fv[1]
    E = -Vtt[0] * P[0]["v0"] / g2[0] - Vtt[1] * P[1]["v0"] / g2[1]
    E -= (V21 * P[1]["v0"] * chi21 / g2[1] + V12 * P[0]["v0"] * chi12 / g2[0]) / 2
    dens = []
    for i in range(n):
        t = types[i]
        sig = own[i] / g1[t] + (chi12 * cross[i] / g1[0] if t == 0 else chi21 * cross[i] / g1[1])
        s = -mp.log(max(mp.mpf("1e-9"), sig) / 12) / (beta * P[t]["eta2"])
        E += P[t]["e0"] * ((1 + P[t]["lambda"] * s) * mp.exp(-P[t]["lambda"] * s) - 1)
        E += 6 * P[t]["v0"] * mp.exp(-P[t]["kappa"] * s)
        dens.append(P[t]["n0"] * mp.exp((mp.mpf("0.5") / bohr - beta * P[t]["eta2"]) * s))
    return E, dens


def _mp_forces(types, pos, par):
    import mpmath as mp

    F = np.zeros((len(types), 3))
    for a in range(len(types)):
        for c in range(3):
            def e_of(x, a=a, c=c):
                p = [[mp.mpf(float(v)) for v in q] for q in pos]
                p[a][c] = x
                return _mp_emt(types, p, par)[0]
            F[a, c] = -float(mp.diff(e_of, mp.mpf(float(pos[a][c]))))
    return F


_CLUSTERS = {
    "h_au_dimer_1.6": ("hau", [[0, 0, 0], [0.3, -0.4, 1.5]]),
    "h_au_dimer_2.6": ("hau", [[0, 0, 0], [2.0, 1.0, 1.36]]),
    "h_pt_dimer_1.8": ("hpt", [[0, 0, 0], [0.5, 0.2, 1.72]]),
    "h_pt_dimer_3.0": ("hpt", [[0, 0, 0], [1.1, 2.4, 1.4]]),
    "h_au3": ("hau", [[0.1, 0.2, 1.3], [-1.44, -0.83, 0.0], [1.44, -0.83, 0.05], [0.0, 1.66, -0.04]]),
}


@pytest.mark.parametrize("name", sorted(_CLUSTERS))
def test_cross_terms_against_published_equations(name, pes_hau, pes_hpt):
    """H-metal clusters in a 60 A box against an independent 40-digit
    evaluation of the published EMT equations (the docstring of _mp_emt);
    H and metal differ in n0 and s0 (chi_21 = 6.5 for H/Au, 4.5 for H/Pt),
    so a swapped chi_12 / chi_21, a wrong element's eta2 or s0 in sigma_12 /
    sigma_21 or V_12 / V_21, or a wrong gamma fails.  Forces against the
    40-digit derivative of the 40-digit energy (every sigma~ stays above the
    1e-9 s clamp, where the force is -dE/dr)."""
    import mpmath as mp

    mp.mp.dps = 40
    which, pos = _CLUSTERS[name]
    pes = pes_hau if which == "hau" else pes_hpt
    metal = "Au" if which == "hau" else "Pt"
    tn = [("H", True), (metal, False)]
    types = [0] + [1] * (len(pos) - 1)
    cols, pesm = pesfile.emt_tables(pes, tn)
    par = [{k: cols[k][t] for k in pesfile.EMT_KEYS} for t in range(2)]
    chi21 = par[0]["n0"] / par[1]["n0"] * math.exp(0.5 / 0.529177211 * (par[1]["s0"] - par[0]["s0"]))
    assert chi21 > 4.0
    s = _system(types, tn, pos)
    e, f, d = oracle.emt_energy_forces(s, cols, pesm)
    E_mp, dens_mp = _mp_emt(types, pos, par)
    assert abs(e[0] - float(E_mp)) <= 1e-11 * max(1.0, abs(float(E_mp))), (e[0], float(E_mp))
    assert np.allclose(d[:, 0], [float(x) for x in dens_mp], rtol=1e-12, atol=0)
    F_mp = _mp_forces(types, pos, par)
    assert np.abs(f[:, 0, :] - F_mp).max() <= 1e-10 * max(1.0, np.abs(F_mp).max())


def test_cross_term_pin_catches_swapped_chi(pes_hau):
    """The check above is sensitive to the cross terms: the same evaluation
    with chi_12 and chi_21 exchanged (n0 of the two types swapped and s0
    moved so that chi_21 -> 1 / chi_21 while the pair functions stay) moves
    the H-Au dimer's energy by far more than the tolerance."""
    import mpmath as mp

    mp.mp.dps = 40
    which, pos = _CLUSTERS["h_au_dimer_1.6"]
    tn = [("H", True), ("Au", False)]
    cols, pesm = pesfile.emt_tables(pes_hau, tn)
    par = [{k: cols[k][t] for k in pesfile.EMT_KEYS} for t in range(2)]
    E_mp, _ = _mp_emt([0, 1], pos, par)
    swapped = [dict(par[0]), dict(par[1])]
    swapped[0]["n0"], swapped[1]["n0"] = par[1]["n0"] * math.exp(-2 * 0.5 / 0.529177211 * (par[1]["s0"] - par[0]["s0"])), par[0]["n0"]
    swapped_E, _ = _mp_emt([0, 1], pos, swapped)
    s = _system([0, 1], tn, pos)
    e, _, _ = oracle.emt_energy_forces(s, cols, pesm)
    assert abs(e[0] - float(E_mp)) < 1e-10
    assert abs(float(swapped_E) - float(E_mp)) > 1e-3
