"""Thermal slab libraries on the GPU: the lattice configurations that
`conf merge <proj folder> <n> <lattice folder> <m>` draws from.

md_tian2 users build such a library (the reference's conf/latt_300K
folders, not shipped with it) by running the bare slab thermostatted at
Tsurf and writing mxt snapshots (`output mxt n`, src/output_mod.f90:
612-630).  Here a batch of independent slab replicas is equilibrated at
once and sampled every `every` steps; each snapshot becomes one
lattice-only universe file mxt_%08d.dat (conf.write_mxt, the reference's
mxt layout), numbered from `start` in (snapshot time, replica) order.

Replica r starts from the ideal fcc(111) slab with Maxwell-Boltzmann
velocities at Tsurf drawn by the library's sampler for trajectory id
start + r (bottom `nfixed` layers fixed, as the generated H/metal slabs).
The thermostat is Andersen's (md_algo.f90:402-432) with collision time
`andersen_time` and without the COM velocity removal the reference applies
with Andersen atoms (mdt_batch_set_com_removal(b, 0)): that removal runs
over fixed and mobile atoms and holds the mobile layers about 11 % below
Tsurf.  Langevin gold atoms are no substitute: their LDFA friction is
~1e-5 / fs, so the ideal slab's released surface-relaxation energy would
take tens of ps to leave, and the reference's Langevin step (both halves
reuse one set of normals) settles above Tsurf at such friction.

    python -m md_tian2_amd.thermal --tsurf 300 --n 3 --layers 4 --nconf 200 --out conf/latt_300K
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from md_tian2_amd import conf, geometry, mdt, pesfile, sampling

PES_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pes")


def slab_system(metal, n, nlayers, nfixed, tsurf, seed, traj_id, ny=None, algo=geometry.PROP_ANDERSEN):
    """A lattice-only geometry.System (propagator `algo`), Maxwell-Boltzmann
    velocities from the sampling stream of (seed, traj_id)."""
    a = geometry.LATTICE_A[metal]
    pos, cell = geometry.fcc111_slab(a, n, nlayers, ny=ny)
    N = len(pos)
    per_layer = N // nlayers
    fixed = np.zeros((N, 1, 3), dtype=bool)
    fixed[(nlayers - nfixed) * per_layer:, 0, :] = True
    mass = np.full(N, geometry.MASS_AMU[metal] * geometry.AMU2MASS)
    idx = np.zeros(N, dtype=np.int32)
    v = np.zeros((N, 1, 3))
    sampling.maxwell(v, idx, [0], mass, fixed, tsurf, seed, traj_id)
    return geometry.System([(metal, False)], idx, mass, fixed, np.array([algo], dtype=np.int32), cell,
                           pos[:, None, :].copy(), v)


def thermal_library(ctx, folder, metal="Au", n=3, nlayers=4, nfixed=2, tsurf=300.0, nconf=100, nrep=32,
                    equil_steps=3000, every=100, dt=1.0, seed=2, start=1, pes_path=None, ny=None, andersen_time=50.0,
                    log=None):
    """Writes nconf lattice snapshots into folder; returns their kinetic
    temperatures over the mobile degrees of freedom."""
    pes = pesfile.load_pes(pes_path or os.path.join(PES_DIR, "EMT-H%s.pes" % metal))
    nrep = max(1, min(nrep, nconf))
    systems = [slab_system(metal, n, nlayers, nfixed, tsurf, seed, start + k, ny=ny) for k in range(nrep)]
    cols, pesm = pesfile.emt_tables(pes, systems[0].type_names)
    ctx.set_emt(cols, pesm)
    b = mdt.Batch(ctx, systems, dt=dt, tsurf=tsurf, seed=seed, andersen_time=andersen_time,
                  traj_ids=list(range(start, start + nrep)))
    b.set_com_removal(False)
    b.compute_forces()
    b.step(equil_steps)
    os.makedirs(folder, exist_ok=True)
    s0 = systems[0]
    mobile = ~s0.fixed[:, 0, :]
    temps = []
    k = start
    while k < start + nconf:
        b.step(every)
        for o in b.get_state():
            if k >= start + nconf:
                break
            c = conf.Conf(o["r"].copy(), o["v"].copy(), True, s0.fixed.copy(), np.ones(s0.natoms, dtype=np.int64),
                          [metal], [False], s0.simbox.copy(), np.linalg.inv(s0.simbox))
            conf.write_mxt(conf.mxt_name(folder, k), c)
            ek2 = (s0.mass[:, None] * o["v"][:, 0, :] ** 2)[mobile].sum()
            temps.append(ek2 / (geometry.KB * mobile.sum()))
            k += 1
        if log:
            print(f"[thermal] {k - start} of {nconf} configurations", file=log, flush=True)
    b.close()
    return np.array(temps)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--metal", default="Au")
    ap.add_argument("--n", type=int, default=3)
    ap.add_argument("--ny", type=int, default=None)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--fixed", type=int, default=2)
    ap.add_argument("--tsurf", type=float, default=300.0)
    ap.add_argument("--nconf", type=int, default=100)
    ap.add_argument("--replicas", type=int, default=32)
    ap.add_argument("--equil", type=int, default=3000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--step", type=float, default=1.0)
    ap.add_argument("--andersen-time", type=float, default=50.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--start", type=int, default=1)
    ap.add_argument("--pes", default=None)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    ctx = mdt.Context(0)
    t = thermal_library(ctx, a.out, a.metal, a.n, a.layers, a.fixed, a.tsurf, a.nconf, a.replicas, a.equil,
                        a.every, a.step, a.seed, a.start, a.pes, a.ny, a.andersen_time, log=sys.stderr)
    print(f"[thermal] {len(t)} configurations in {a.out}: kinetic temperature {t.mean():.1f} +- "
          f"{t.std() / np.sqrt(len(t)):.1f} K", file=sys.stderr)
    ctx.close()


if __name__ == "__main__":
    main()
