"""Pins of the NENE oracle (oracle/nene_oracle.c), the checker of the GPU NN
surrogate.

* Golden: the reference's trained RuNNer carbon network
  (example/md/RuNNer_input, 30 symmetry functions, 30-15-15-1) reproduces
  the DFT energies of its training structures (example/fit/input.data) to
  the fit's accuracy; with the symmetry functions left in file order
  instead of sortsymfunctions order the same check fails.
* Mathematics: forces are -dE/dr (central differences), descriptor
  derivatives match differences of the descriptors, the total force
  vanishes, a 2x supercell doubles the energy with identical atomic
  energies (periodic images complete), energies are invariant under
  translation and rotation, and the analytic training gradient matches
  differences of the loss.
"""
import os

import numpy as np
import pytest

import nene_oracle as no
from md_tian2_amd import nene

FIX = os.path.join(os.path.dirname(__file__), "golden", "runner_c")
HA2MEV = 27211.38602


@pytest.fixture(scope="module")
def cmodel():
    return nene.load(FIX)


@pytest.fixture(scope="module")
def structs():
    return nene.read_input_data(os.path.join(FIX, "input.data"))


def _rmse_e(om, structs):
    d = []
    for s in structs:
        el = om.elem_index(s["symbols"])
        E, _, _, _ = om.energy_forces(el, s["xyz"], s["lattice"])
        d.append((E - s["energy"]) / len(el))
    return np.sqrt(np.mean(np.square(d))) * HA2MEV


def test_runner_carbon_network_energies(cmodel, structs):
    assert len(structs) == 40
    rmse = _rmse_e(no.Model(cmodel), structs)
    assert rmse < 2.0, rmse  # meV/atom (observed 0.77)


def test_unsorted_symmetry_functions_fail(structs):
    m = nene.read_input_nn(os.path.join(FIX, "input.nn"))
    # undo the sortsymfunctions order: file order of the symfunction lines
    order = []
    for line in open(os.path.join(FIX, "input.nn")):
        w = line.split()
        if w and w[0] == "symfunction_short":
            t = int(w[2])
            if t == 2:
                order.append((2, float(w[4]), 0.0, 0.0))
            else:
                order.append((3, float(w[5]), float(w[6]), float(w[7])))
    sf = m.sf[6]
    keyed = {(s.type, s.eta, s.lam, s.zeta): s for s in sf}
    m.sf[6] = [keyed[(t, eta, lam, zeta)] for (t, eta, lam, zeta) in order]
    nene.read_scaling(m, os.path.join(FIX, "scaling.data"))
    nene.read_weights(m, FIX)
    assert _rmse_e(no.Model(m), structs[:10]) > 20.0


def test_forces_vs_dft(cmodel, structs):
    om = no.Model(cmodel)
    pred, ref = [], []
    for s in structs[:10]:
        el = om.elem_index(s["symbols"])
        _, _, F, _ = om.energy_forces(el, s["xyz"], s["lattice"])
        pred.append(F.ravel())
        ref.append(s["forces"].ravel())
    assert np.corrcoef(np.concatenate(pred), np.concatenate(ref))[0, 1] > 0.98


def test_forces_are_minus_gradient(cmodel, structs):
    om = no.Model(cmodel)
    s = structs[3]
    el = om.elem_index(s["symbols"])
    x = s["xyz"].copy()
    E, _, F, _ = om.energy_forces(el, x, s["lattice"])
    assert np.abs(F.sum(0)).max() < 1e-12
    h = 1e-5
    for a in (0, 7, 20):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            fd = -(om.energy_forces(el, xp, s["lattice"])[0] - om.energy_forces(el, xm, s["lattice"])[0]) / (2 * h)
            assert abs(fd - F[a, c]) < 1e-8 * max(1.0, np.abs(F).max()) + 1e-9


def test_descriptor_derivatives(cmodel, structs):
    om = no.Model(cmodel)
    s = structs[5]
    el = om.elem_index(s["symbols"])
    x = s["xyz"].copy()
    G, dG = om.descriptors(el, x, s["lattice"], derivatives=True)
    h = 1e-5
    for a in (2, 11):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += h
            xm[a, c] -= h
            fd = (om.descriptors(el, xp, s["lattice"]) - om.descriptors(el, xm, s["lattice"])) / (2 * h)
            assert np.abs(fd[:, :30] - dG[:, :30, a, c]).max() < 1e-7 * max(1.0, np.abs(G).max())


def test_supercell_and_invariances(cmodel, structs):
    om = no.Model(cmodel)
    s = structs[7]
    el = om.elem_index(s["symbols"])
    lat = s["lattice"]
    E, Ea, F, _ = om.energy_forces(el, s["xyz"], lat)
    # 2x1x1 supercell: every atom keeps its environment
    x2 = np.concatenate([s["xyz"], s["xyz"] + lat[0]])
    l2 = lat.copy()
    l2[0] *= 2
    E2, Ea2, F2, _ = om.energy_forces(np.concatenate([el, el]), x2, l2)
    assert abs(E2 - 2 * E) < 1e-9 * abs(E)
    assert np.abs(Ea2[:len(el)] - Ea).max() < 1e-11
    assert np.abs(F2[:len(el)] - F).max() < 1e-10
    # translation (positions need not be folded into the cell)
    Et, _, Ft, _ = om.energy_forces(el, s["xyz"] + np.array([3.1, -7.2, 40.0]), lat)
    assert abs(Et - E) < 1e-10 and np.abs(Ft - F).max() < 1e-10
    # rotation of the whole periodic system
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    Er, _, Fr, _ = om.energy_forces(el, s["xyz"] @ R.T, lat @ R.T)
    assert abs(Er - E) < 1e-10 and np.abs(Fr - F @ R.T).max() < 1e-10


def test_descriptor_dimer_and_trimer(cmodel):
    """Cluster (no lattice): G2 of a dimer and G4 of an equilateral trimer in
    closed form for the first radial / an angular function."""
    om = no.Model(cmodel)
    sf = cmodel.sf[6]
    d = 2.7
    G = om.descriptors(np.array([0, 0], np.int32), np.array([[0, 0, 0], [d, 0, 0]], float))
    k2 = 1  # eta 0.005
    fc = 0.5 * (np.cos(np.pi * d / sf[k2].rc) + 1)
    assert abs(G[0, k2] - np.exp(-sf[k2].eta * d * d) * fc) < 1e-14
    tri = np.array([[0, 0, 0], [d, 0, 0], [d / 2, d * np.sqrt(3) / 2, 0]])
    G = om.descriptors(np.array([0, 0, 0], np.int32), tri)
    for k, s in enumerate(sf):
        if s.type != 3:
            continue
        fc = 0.5 * (np.cos(np.pi * d / s.rc) + 1)
        pl = 1 + s.lam * 0.5  # cos 60
        want = 2 ** (1 - s.zeta) * pl ** s.zeta * np.exp(-s.eta * 3 * d * d) * fc ** 3 if pl > 0 else 0.0
        assert abs(G[0, k] - want) < 1e-13 * max(1, abs(want)), (k, G[0, k], want)


def test_loss_gradient(cmodel, structs):
    om = no.Model(cmodel)
    frames = []
    for s in structs[:2]:
        el = om.elem_index(s["symbols"])
        frames.append((el, s["xyz"], s["lattice"], s["energy"] * 27.21138602 + 0.3,
                       s["forces"] * 27.21138602 / 0.529177211 * 1.05))
    L, g = om.loss_grad(frames, wE=1.0, wF=2.0)
    w0 = cmodel.weights[6].copy()
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(450, 6, replace=False), [455, 470, 600, 690, 705, 720]])
    h = 1e-6
    for i in idx:
        wp, wm = w0.copy(), w0.copy()
        wp[i] += h
        wm[i] -= h
        Lp, _ = no.Model(cmodel, weights={6: wp}).loss_grad(frames, 1.0, 2.0)
        Lm, _ = no.Model(cmodel, weights={6: wm}).loss_grad(frames, 1.0, 2.0)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(g[i])), (i, fd, g[i])
