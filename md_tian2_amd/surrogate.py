"""The H/Au NENE surrogate: EMT training frames generated on the GPU, GPU
training, RuNNer-format export.

Training frames are snapshots of batched EMT trajectories of H on 3x3
Au(111) slabs (fcc a = 4.078 A, bottom 2 layers fixed, Langevin+LDFA slab,
dt 0.1 fs), with incidence energy, polar and azimuth sampled per trajectory
so that the H visits the whole approach / bounce / penetration range.

    md_tian2_amd/nn/hau        4-layer slabs at 300 K (the default):
        python -m md_tian2_amd.surrogate train --frames 200000 --out md_tian2_amd/nn/hau
    md_tian2_amd/nn/hau_thick  4- and 6-layer slabs, Tsurf per trajectory in
        100-500 K; the 6-layer frames give the interiors of thicker slabs
        (two layers above and below within the cut-off):
        python -m md_tian2_amd.surrogate train --frames 200000 --layers 4,6 --tsurf 100,500 \
            --out md_tian2_amd/nn/hau_thick
    md_tian2_amd/nn/hau_radial the same frames, radial symmetry functions
        only (nene.hau_radial_design: EMT energies depend on distance sums
        alone):
        python -m md_tian2_amd.surrogate train --design radial --frames 200000 --layers 4,6 \
            --tsurf 100,500 --out md_tian2_amd/nn/hau_radial

Energies and forces are the GPU EMT path's (FP64).  The network is
nene.hau_design (24 radial + 16 angular symmetry functions per element,
6.5 A cut-off, 40-32-32-1) or nene.hau_radial_design (16 radial, 4.5 A,
16-32-32-1), trained by the library's own GPU kernels (BF16
MFMA, Adam) on energy + force loss.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from md_tian2_amd import geometry, mdt, nene, pesfile

HERE = os.path.dirname(os.path.abspath(__file__))
PES_HAU = os.path.join(HERE, "pes", "EMT-HAu.pes")
MODEL_DIR = os.path.join(HERE, "nn", "hau")
MODEL_DIR_THICK = os.path.join(HERE, "nn", "hau_thick")
MODEL_DIR_RADIAL = os.path.join(HERE, "nn", "hau_radial")


def generate_frames(ctx, ntraj=2000, nsteps=1000, every=10, seed=7, n=3, nlayers=4, tsurf=(100.0, 500.0), log=None):
    """EMT frames from batched trajectories (Tsurf per trajectory uniform in
    the range `tsurf`).  Returns (template System, positions (nf, N, 3) [A],
    E (nf,) [eV], F (nf, N, 3) [eV/A]); every frame has the template's atoms
    and cell."""
    g = np.random.default_rng(seed)
    ts = g.uniform(tsurf[0], tsurf[1], ntraj)
    systems = [geometry.h_on_slab("Au", n=n, nlayers=nlayers, nfixed=2, rng=g, einc=g.uniform(0.3, 6.0),
                                  polar_deg=g.uniform(0.0, 60.0), azimuth_deg=g.uniform(0.0, 360.0),
                                  z0=g.uniform(2.0, 6.0), tsurf=ts[k]) for k in range(ntraj)]
    b = mdt.Batch(ctx, systems, dt=0.1, tsurf=ts, seed=seed)
    b.compute_forces()
    R, E, F = [], [], []
    for k in range(nsteps // every):
        b.step(every)
        st = b.get_state()
        R.append(np.stack([o["r"][:, 0, :] for o in st]))
        F.append(np.stack([o["f"][:, 0, :] for o in st]))
        E.append(np.array([o["epot"][0] for o in st]))
        if log and (k + 1) % 10 == 0:
            print(f"[surrogate] {sum(len(e) for e in E)} frames", file=log, flush=True)
    b.close()
    return systems[0], np.concatenate(R), np.concatenate(E), np.concatenate(F)


def write_input_data(path, template, R, E, F):
    """Frames of generate_frames as a RuNNer input.data file (the reference's
    output_runner format, src/output_mod.f90:550-609): Bohr, Hartree."""
    from md_tian2_amd import observables
    symbols = [template.type_names[k][0] for k in template.idx]
    with open(path, "w") as f:
        f.write(observables.runner_header())
        for k in range(len(E)):
            f.write(observables.runner_frame_text(template.simbox, symbols, R[k][:, None, :], F[k][:, None, :],
                                                  np.array([E[k]])))


def train(ctx, sets, epochs=60, batch_frames=64, lr=2e-3, wE=1.0, wF=10.0, seed=11, log=None, design="hau"):
    """Designs, scales and trains the H/Au network on the GPU from frame
    sets [(template System, R, E, F), ...] (generate_frames's output, one
    per slab); returns (model, history) with the trained weights in
    model.weights.  design: a key of nene.DESIGNS."""
    m = nene.DESIGNS[design]()
    for z in m.elements:
        m.gmin[z], m.gmax[z], m.gavg[z] = np.zeros(m.nsf(z)), np.ones(m.nsf(z)), np.zeros(m.nsf(z))
    nene.init_weights(m, np.random.default_rng(seed))
    tmpl = sets[0][0]
    ctx.set_nn(m, tmpl.type_names)
    nat = np.concatenate([np.full(len(E), R.shape[1], np.int32) for _, R, E, _ in sets])
    idx = np.concatenate([np.tile(t.idx, len(E)) for t, _, E, _ in sets])
    cells = np.concatenate([np.tile(np.asarray(t.simbox).T.reshape(-1), len(E)) for t, _, E, _ in sets])
    Rall = np.concatenate([R.reshape(-1, 3) for _, R, _, _ in sets])
    Eall = np.concatenate([E for _, _, E, _ in sets])
    Fall = np.concatenate([F.reshape(-1, 3) for _, _, _, F in sets])
    tr = mdt.Trainer.from_arrays(ctx, nat, idx, cells, Rall, Eall, Fall, batch_frames=batch_frames,
                                 val_fraction=0.05, seed=seed)
    mn, mx, av = tr.descriptor_stats(len(m.elements))
    for i, z in enumerate(m.elements):
        k = m.nsf(z)
        m.gmin[z], m.gmax[z], m.gavg[z] = mn[i, :k].copy(), mx[i, :k].copy(), av[i, :k].copy()
        m.gmax[z] = np.maximum(m.gmax[z], m.gmin[z] + 1e-6)
    eat = Eall / nat
    m.eshort = (float(np.min(eat) / nene.HA2EV), float(np.max(eat) / nene.HA2EV))
    ctx.set_nn(m, tmpl.type_names)
    hist = [("init",) + tr.eval() + tr.eval(validation=True)]
    t0 = time.perf_counter()
    for ep in range(epochs):
        rate = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)) * 0.95 + 0.05)  # cosine decay to 5 %
        loss = tr.epoch(lr=rate, wE=wE, wF=wF)
        if log and (ep % 10 == 9 or ep == epochs - 1):
            e, f = tr.eval(validation=True)
            print(f"[surrogate] epoch {ep + 1}: loss {loss:.4e}  validation {e * 1e3:.2f} meV/atom "
                  f"{f * 1e3:.1f} meV/A  ({time.perf_counter() - t0:.1f} s)", file=log, flush=True)
    hist.append(("trained",) + tr.eval() + tr.eval(validation=True))
    m.weights = mdt.nn_weights(ctx, m)
    tr.close()
    return m, hist


def load_model(directory=MODEL_DIR):
    return nene.load(directory) if os.path.exists(os.path.join(directory, "input.nn")) else None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--frames", type=int, default=200000)
    t.add_argument("--traj", type=int, default=2000)
    t.add_argument("--every", type=int, default=10)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=64)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--layers", default="4", help="slab layer counts of the frame sets (3x3 cells)")
    t.add_argument("--tsurf", default="300", help="Tsurf [K], or a range lo,hi drawn per trajectory")
    t.add_argument("--out", default=MODEL_DIR)
    t.add_argument("--design", default="hau", choices=sorted(nene.DESIGNS),
                   help="symmetry functions: hau (radial + angular) or radial (radial only)")
    t.add_argument("--input-data", default=None,
                   help="also write the EMT frames as RuNNer input.data files (<path>.<layers>)")
    a = ap.parse_args(argv)
    ctx = mdt.Context(0)
    pes = pesfile.load_pes(PES_HAU)
    layers = [int(x) for x in a.layers.split(",")]
    nsteps = a.every * max(1, a.frames // (a.traj * len(layers)))
    s0 = geometry.h_on_slab("Au", n=3, nlayers=4)
    cols, pesm = pesfile.emt_tables(pes, s0.type_names)
    ctx.set_emt(cols, pesm)
    t0 = time.perf_counter()
    tr = [float(x) for x in a.tsurf.split(",")]
    trange = (tr[0], tr[-1])
    sets = [generate_frames(ctx, ntraj=a.traj, nsteps=nsteps, every=a.every, seed=7 + k, nlayers=nl,
                            tsurf=trange, log=sys.stderr) for k, nl in enumerate(layers)]
    print(f"[surrogate] {sum(len(x[2]) for x in sets)} EMT frames ({', '.join(f'{nl} layers' for nl in layers)}) "
          f"in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if a.input_data:
        for nl, st in zip(layers, sets):
            write_input_data(f"{a.input_data}.{nl}", *st)
    m, hist = train(ctx, sets, epochs=a.epochs, batch_frames=a.batch, lr=a.lr, log=sys.stderr, design=a.design)
    nene.write(m, a.out)
    for h in hist:
        print(f"[surrogate] {h[0]}: train {h[1] * 1e3:.2f} meV/atom {h[2] * 1e3:.1f} meV/A; "
              f"validation {h[3] * 1e3:.2f} meV/atom {h[4] * 1e3:.1f} meV/A", file=sys.stderr, flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
