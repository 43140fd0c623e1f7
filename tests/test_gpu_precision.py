"""Precision mode 1 (FP32 pair terms, FP64 accumulation) against the FP64
oracle, and hipGraph replay of the step loop against direct launches.

Error bars of mode 1, per frame, against the oracle (the reference's FP64
compute_emt_2species restated), for H/Au (full FP32 pair terms) and H/Pt
(FP32 cut-off exponential and theta):
  energy  |E_gpu - E_ref| / N <= 1e-6 eV/atom
  forces  max |F_gpu - F_ref| <= 1e-5 eV/A
"""
import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, pesfile

pytestmark = pytest.mark.gpu

E_ATOM_TOL = 1e-6
F_TOL = 1e-5


def _frames_hau(n=6, count=6, seed=21):
    g = np.random.default_rng(seed)
    out = []
    for z in np.linspace(0.7, 6.0, count):
        s = geometry.h_on_slab("Au", n=n, nlayers=4, rng=g, einc=g.uniform(0.5, 5.0))
        s.r[:, 0, :] += g.normal(0, 0.06, s.r[:, 0, :].shape) * (~s.fixed[:, 0, :])
        s.r[0, 0, 2] = z
        out.append(s)
    return out


def _fp32_errors(ctx, systems, pes):
    cols, pesm = pesfile.emt_tables(pes, systems[0].type_names)
    ctx.set_emt(cols, pesm)
    b = mdt.Batch(ctx, systems)
    b.set_precision(1)
    b.compute_forces()
    out = b.get_state()
    b.close()
    res = []
    for s, o in zip(systems, out):
        e_r, f_r, d_r = oracle.emt_energy_forces(s, cols, pesm)
        de = np.abs(o["epot"] - e_r).max() / s.natoms
        df = np.abs(o["f"] - f_r).max()
        dd = np.abs(o["dens"] - d_r).max()
        res.append((de, df, dd))
    return np.array(res)


def test_fp32_pairs_hau_6x6x4_heights(gpu_ctx, pes_hau):
    frames = _frames_hau()
    cols, pesm = pesfile.emt_tables(pes_hau, frames[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, frames[:1])
    b.set_precision(1)
    assert b.precision_arith() == 1  # H 1.25 eV, Au 13.2 eV: full FP32 pair terms
    b.close()
    err = _fp32_errors(gpu_ctx, frames, pes_hau)
    print("FP32 H/Au 6x6x4: max dE/atom %.3e eV, max dF %.3e eV/A" % (err[:, 0].max(), err[:, 1].max()))
    assert err[:, 0].max() <= E_ATOM_TOL
    assert err[:, 1].max() <= F_TOL
    # it is a different arithmetic path: not bitwise the FP64 result
    assert err[:, 1].max() > 1e-12


def test_fp32_pairs_hpt_8x8x6(gpu_ctx, pes_hpt):
    """H/Pt in precision mode 1 at the same bar as H/Au.  Pt's embedding
    stiffness 6 V0 kappa / (beta eta2) = 129 eV (V0 = 21.7 eV) turns the
    ~1e-7 relative rounding of FP32 sigma or pair-potential terms into
    ~6e-6 eV/atom and ~8e-5 eV/A, so mode 1 on this parameter set evaluates
    only the cut-off exponential and theta in FP32 (arithmetic 2)."""
    g = np.random.default_rng(31)
    systems = []
    for z in (0.9, 1.6, 3.0):
        s = geometry.h_on_slab("Pt", n=8, nlayers=6, rng=g)
        s.r[:, 0, :] += g.normal(0, 0.05, s.r[:, 0, :].shape) * (~s.fixed[:, 0, :])
        s.r[0, 0, 2] = z
        systems.append(s)
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems[:1])
    b.set_precision(1)
    assert b.precision_arith() == 2
    b.close()
    err = _fp32_errors(gpu_ctx, systems, pes_hpt)
    print("mode 1 H/Pt 8x8x6: max dE/atom %.3e eV, max dF %.3e eV/A" % (err[:, 0].max(), err[:, 1].max()))
    assert err[:, 0].max() <= E_ATOM_TOL
    assert err[:, 1].max() <= F_TOL
    assert err[:, 1].max() > 1e-13  # not the FP64 path


def test_fp32_pairs_hpt_c3_trajectory_frames(gpu_ctx, pes_hpt):
    """Frames from 100 K RPMD H/Pt(111) 8x8x6 trajectories (32 beads, H PILE,
    slab Langevin+LDFA, Einc U(0.5, 5) eV), every bead a frame, in mode 1
    against the FP64 oracle."""
    g = np.random.default_rng(33)
    systems = [geometry.h_on_slab("Pt", n=8, nlayers=6, nfixed=2, rng=g, nbeads=32, einc=g.uniform(0.5, 5.0),
                                  polar_deg=g.uniform(0, 60), azimuth_deg=g.uniform(0, 360), z0=2.2, tsurf=100.0,
                                  proj_algo=geometry.PROP_PILE) for _ in range(2)]
    cols, pesm = pesfile.emt_tables(pes_hpt, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=100.0, tproj=100.0, pile_tau=200.0, seed=17)
    b.compute_forces()
    b.step(150)
    st = b.get_state()
    b.close()
    frames = []
    for s, o in zip(systems, st):
        for bead in (0, 7, 16, 31):
            f = s.copy()
            f.r = o["r"][:, bead:bead + 1, :].copy()
            f.v, f.fixed = f.v[:, :1, :].copy(), f.fixed[:, :1, :].copy()
            frames.append(f)
    err = _fp32_errors(gpu_ctx, frames, pes_hpt)
    print("mode 1 H/Pt C3 frames: max dE/atom %.3e eV, max dF %.3e eV/A" % (err[:, 0].max(), err[:, 1].max()))
    assert err[:, 0].max() <= E_ATOM_TOL
    assert err[:, 1].max() <= F_TOL


def test_fp32_pairs_along_trajectories(gpu_ctx, pes_hau):
    """Frames taken from 300 K FP64 trajectories (the C2 workload's physics)."""
    g = np.random.default_rng(41)
    systems = [geometry.h_on_slab("Au", n=6, nlayers=4, rng=g, einc=g.uniform(0.5, 5.0),
                                  polar_deg=g.uniform(0, 60), azimuth_deg=g.uniform(0, 360), z0=2.5,
                                  tsurf=300.0) for _ in range(8)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=7)
    b.compute_forces()
    b.step(300)
    st = b.get_state()
    b.close()
    frames = []
    for s, o in zip(systems, st):
        f = s.copy()
        f.r = o["r"].copy()
        frames.append(f)
    err = _fp32_errors(gpu_ctx, frames, pes_hau)
    assert err[:, 0].max() <= E_ATOM_TOL
    assert err[:, 1].max() <= F_TOL


def test_fp32_stepping_stays_close(gpu_ctx, pes_hau):
    """Tsurf = 0, H and slab Verlet/Langevin without noise: FP32-pair and FP64
    trajectories stay within 1e-5 A after 200 steps (forces differ by
    ~1e-7 eV/A)."""
    g = np.random.default_rng(51)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, einc=2.0, z0=3.0) for _ in range(4)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    outs = []
    for mode in (0, 1):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=0.0, seed=3)
        b.set_precision(mode)
        b.compute_forces()
        b.step(200)
        outs.append(b.get_state())
        b.close()
    for o64, o32 in zip(*outs):
        assert np.abs(o64["r"] - o32["r"]).max() <= 1e-5


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_replay_matches_direct_launches(gpu_ctx, pes_hau, mode):
    """mdt_batch_set_graph: chunks replayed from a hipGraph (plus a direct
    remainder) give bitwise the state of direct launches, across calls."""
    g = np.random.default_rng(61)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, einc=g.uniform(0.5, 5), tsurf=300.0)
               for _ in range(16)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    outs = []
    for chunk in (0, 8):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, seed=99)
        b.set_precision(mode)
        b.set_graph(chunk)
        b.compute_forces()
        b.step(21)   # 2 replays + 5 direct steps
        b.step(16)   # 2 replays, step counter continues on the device
        b.step(3)    # direct only
        outs.append(b.get_state())
        assert not b.status().any()
        b.close()
    for o0, o1 in zip(*outs):
        for k in ("r", "v", "f", "a", "epot", "dens"):
            assert np.array_equal(o0[k], o1[k]), k


def test_graph_replay_ring_polymer(gpu_ctx, pes_hau):
    """P = 4 beads (ring-polymer step + PILE noise keyed by the step number)."""
    g = np.random.default_rng(71)
    systems = [geometry.h_on_slab("Au", n=3, nlayers=4, rng=g, nbeads=4, proj_algo=geometry.PROP_PILE, tsurf=300.0)
               for _ in range(4)]
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    outs = []
    for chunk in (0, 5):
        b = mdt.Batch(gpu_ctx, systems, dt=0.1, tsurf=300.0, tproj=300.0, seed=5)
        b.set_graph(chunk)
        b.compute_forces()
        b.step(12)
        outs.append(b.get_state())
        b.close()
    for o0, o1 in zip(*outs):
        assert np.array_equal(o0["r"], o1["r"])
        assert np.array_equal(o0["v"], o1["v"])
