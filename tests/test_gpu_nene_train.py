"""GPU training of the NENE surrogate against the oracle.

* The training gradient of the GPU kernels (network forward/backward on BF16
  MFMA with hi/lo split operands, FP32 symmetry functions, the force
  backward through the Jacobian-vector product) equals the oracle's
  analytic FP64 gradient (itself pinned by finite differences) to a relative
  L2 error <= 2e-3; the loss to 1e-4 relative.
* Descriptor statistics (scaling.data values) match the oracle's
  descriptors.
* Training on EMT frames lowers the force and energy errors against EMT.
Frames: H/Au(111) 3x3x4 cells (8.65 A wide, cut-off 6.5 A: several
periodic images of each neighbour).
"""
import numpy as np
import pytest

import nene_oracle as no
import oracle
from md_tian2_amd import geometry, mdt, nene, pesfile

pytestmark = pytest.mark.gpu


def _emt_frames(pes, n, seed):
    g = np.random.default_rng(seed)
    frames, E, F = [], [], []
    cols = pesm = None
    for k in range(n):
        s = geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, einc=g.uniform(0.5, 5))
        s.r[:, 0, :] += g.normal(0, 0.08, s.r[:, 0, :].shape) * (~s.fixed[:, 0, :])
        s.r[0, 0, 2] = g.uniform(1.0, 5.0)
        if cols is None:
            cols, pesm = pesfile.emt_tables(pes, s.type_names)
        e, f, _ = oracle.emt_energy_forces(s, cols, pesm)
        frames.append(s)
        E.append(e[0])
        F.append(f[:, 0, :])
    return frames, np.array(E), F


def _model_for(frames, rng):
    m = nene.hau_design()
    for z in m.elements:
        m.gmin[z], m.gmax[z], m.gavg[z] = np.zeros(m.nsf(z)), np.ones(m.nsf(z)), np.zeros(m.nsf(z))
    nene.init_weights(m, rng)
    return m


def _oracle_G(m, frames):
    om = no.Model(m)
    Gs = {1: [], 79: []}
    for s in frames:
        el = om.elem_index([1] + [79] * (s.natoms - 1))
        G = om.descriptors(el, s.r[:, 0, :] / nene.BOHR2ANG, s.simbox.T / nene.BOHR2ANG)
        Gs[1].append(G[:1, :40])
        Gs[79].append(G[1:, :40])
    return {z: np.concatenate(v) for z, v in Gs.items()}


def _apply_stats(m, mn, mx, av):
    for i, z in enumerate(m.elements):
        n = m.nsf(z)
        m.gmin[z], m.gmax[z], m.gavg[z] = mn[i, :n].copy(), mx[i, :n].copy(), av[i, :n].copy()


def test_descriptor_stats_and_gradient(gpu_ctx, pes_hau):
    frames, E, F = _emt_frames(pes_hau, 6, 101)
    rng = np.random.default_rng(5)
    m = _model_for(frames, rng)
    gpu_ctx.set_nn(m, frames[0].type_names)
    tr = mdt.Trainer(gpu_ctx, frames, E, F, batch_frames=6, val_fraction=0.0, seed=3)
    mn, mx, av = tr.descriptor_stats(2)
    G = _oracle_G(m, frames)
    for i, z in enumerate(m.elements):
        ref_mn, ref_mx = G[z].min(0), G[z].max(0)
        scale = np.maximum(1.0, np.abs(ref_mx))
        assert np.abs(mn[i, :40] - ref_mn).max() <= 1e-5 * scale.max()
        assert np.abs(mx[i, :40] - ref_mx).max() <= 1e-5 * scale.max()
    _apply_stats(m, mn, mx, av)
    gpu_ctx.set_nn(m, frames[0].type_names)
    nw = sum(m.nweights(z) for z in m.elements)
    L_g, g_g = tr.loss_grad(list(range(6)), nw, wE=1.0, wF=1.0)
    om = no.Model(m)
    of = [(om.elem_index([1] + [79] * (s.natoms - 1)), s.r[:, 0, :] / nene.BOHR2ANG, s.simbox.T / nene.BOHR2ANG,
           e, f) for s, e, f in zip(frames, E, F)]
    L_r, g_r = om.loss_grad(of, wE=1.0, wF=1.0)
    print("loss gpu %.8e oracle %.8e; grad rel err %.3e" % (L_g, L_r, np.linalg.norm(g_g - g_r) / np.linalg.norm(g_r)))
    assert abs(L_g - L_r) <= 1e-4 * abs(L_r)
    assert np.linalg.norm(g_g - g_r) <= 2e-3 * np.linalg.norm(g_r)
    tr.close()


def test_training_lowers_error(gpu_ctx, pes_hau):
    frames, E, F = _emt_frames(pes_hau, 96, 202)
    rng = np.random.default_rng(6)
    m = _model_for(frames, rng)
    gpu_ctx.set_nn(m, frames[0].type_names)
    tr = mdt.Trainer(gpu_ctx, frames, E, F, batch_frames=16, val_fraction=0.2, seed=4)
    _apply_stats(m, *tr.descriptor_stats(2))
    gpu_ctx.set_nn(m, frames[0].type_names)
    e0, f0 = tr.eval()
    for _ in range(60):
        tr.epoch(lr=3e-3, wE=1.0, wF=10.0)
    e1, f1 = tr.eval()
    ve, vf = tr.eval(validation=True)
    print("untrained: %.4f eV/atom %.4f eV/A; trained: %.4f eV/atom %.4f eV/A; validation %.4f %.4f"
          % (e0, f0, e1, f1, ve, vf))
    assert f1 < f0 / 3
    assert e1 < e0 / 3
    # the trained weights come back in the weights.XXX.data layout
    w = mdt.nn_weights(gpu_ctx, m)
    assert all(np.isfinite(w[z]).all() and len(w[z]) == m.nweights(z) for z in m.elements)
    tr.close()
