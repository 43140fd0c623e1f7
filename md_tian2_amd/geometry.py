"""Synthetic H-on-fcc(111) systems for the batched EMT path.

The reference ships no Au/Pt slab geometries (its ``conf/`` holds graphene
only), so slabs are generated here:

* fcc(111), ABC stacking, layer spacing a/sqrt(3), top layer at z = 0 and
  lower layers at negative z; in-plane cell n*a/sqrt(2) * (1, 0, 0) and
  n*a/sqrt(2) * (1/2, sqrt(3)/2, 0); c = 40 A so no z-images are in range;
* the bottom ``nfixed`` layers are frozen in all three components;
* the projectile comes first in the atom order, as md_tian2 requires
  (src/md_init.f90:590-608);
* the projectile starts at absolute z (``pip r r z``, md_init.f90:498-554)
  and its velocity is sqrt(2E/m) (sin t cos p, sin t sin p, -cos t)
  (md_init.f90:861-864).

Arrays use the reference's (3,P,N) Fortran layout, stored in numpy as C-order
(N, P, 3), which is the same memory.  Masses are in program units
(amu * amu2mass, constants.f90:157).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

KB = 8.61733238496e-5  # eV/K, constants.f90:37
HBAR = 0.6582119514467406  # eV fs, constants.f90:38
AMU2MASS = 103.638239276  # constants.f90:157

PROP_VERLET = 1001  # constants.f90:107-112
PROP_LANGEVIN = 1003
PROP_ANDERSEN = 1005
PROP_PILE = 1006

MASS_AMU = {"H": 1.00794, "D": 2.014101778, "Au": 196.966569, "Pt": 195.084}
LATTICE_A = {"Au": 4.078, "Pt": 3.92}


@dataclass
class System:
    """One trajectory's universe (src/universe_mod.f90:38-67)."""

    type_names: list  # [(name, is_proj)] in type-index order
    idx: np.ndarray  # (N,) int32 type index, 0-based
    mass: np.ndarray  # (N,) float64 program units
    fixed: np.ndarray  # (N, P, 3) bool  == is_fixed(3,P,N)
    algo: np.ndarray  # (ntypes,) int32 propagator id per type
    simbox: np.ndarray  # (3, 3) columns are the cell vectors
    r: np.ndarray  # (N, P, 3) float64 == r(3,P,N)
    v: np.ndarray  # (N, P, 3)

    @property
    def natoms(self):
        return int(self.idx.shape[0])

    @property
    def nbeads(self):
        return int(self.r.shape[1])

    @property
    def ntypes(self):
        return len(self.type_names)

    @property
    def is_proj(self):
        return np.array([1 if p else 0 for (_, p) in self.type_names], dtype=np.int32)

    @property
    def isimbox(self):
        return np.linalg.inv(self.simbox)

    def copy(self):
        return System(list(self.type_names), self.idx.copy(), self.mass.copy(),
                      self.fixed.copy(), self.algo.copy(), self.simbox.copy(),
                      self.r.copy(), self.v.copy())

    def with_beads(self, nbeads):
        """force_beads: replicate every atom's bead 1 (md_init.f90:630-671)."""
        rep = lambda a: np.repeat(a[:, :1, :], nbeads, axis=1).copy()
        return System(list(self.type_names), self.idx.copy(), self.mass.copy(),
                      rep(self.fixed), self.algo.copy(), self.simbox.copy(),
                      rep(self.r), rep(self.v))


def fcc111_slab(a: float, n: int, nlayers: int, c: float = 40.0, ny: int | None = None):
    """Positions (M, 3) and cell (3, 3) of an n x ny x nlayers fcc(111) slab.

    Atoms are listed layer by layer from the top (z = 0) downwards.
    """
    ny = n if ny is None else ny
    d = a / math.sqrt(2.0)
    dz = a / math.sqrt(3.0)
    u1 = np.array([d, 0.0, 0.0])
    u2 = np.array([0.5 * d, 0.5 * math.sqrt(3.0) * d, 0.0])
    shift = (u1 + u2) / 3.0
    pos = []
    for layer in range(nlayers):
        off = (layer % 3) * shift
        z = -layer * dz
        for j in range(ny):
            for i in range(n):
                p = i * u1 + j * u2 + off
                pos.append([p[0], p[1], z])
    cell = np.zeros((3, 3))
    cell[:, 0] = n * u1
    cell[:, 1] = ny * u2
    cell[:, 2] = [0.0, 0.0, c]
    pos = np.array(pos, dtype=np.float64)
    # fold into the cell (direct coordinates in [0, 1))
    frac = np.linalg.solve(cell, pos.T).T
    frac[:, :2] -= np.floor(frac[:, :2])
    pos = (cell @ frac.T).T
    return pos, cell


def h_on_slab(metal="Au", n=3, nlayers=4, nfixed=2, a=None, nbeads=1,
              einc=1.92, polar_deg=0.0, azimuth_deg=0.0, z0=6.0,
              tsurf=300.0, proj_algo=PROP_VERLET, latt_algo=PROP_LANGEVIN,
              rng=None, thermal=True, proj="H"):
    """One H/metal(111) trajectory: projectile first, then the slab.

    Projectile x,y uniform over the cell (the intended distribution of
    ``pip r r``; md_init.f90:521-546), Maxwell-Boltzmann slab velocities at
    ``tsurf`` on mobile atoms when ``thermal`` is set.
    """
    rng = np.random.default_rng(0) if rng is None else rng
    a = LATTICE_A[metal] if a is None else a
    slab, cell = fcc111_slab(a, n, nlayers)
    m_slab = len(slab)
    N = m_slab + 1
    type_names = [(proj, True), (metal, False)]
    idx = np.array([0] + [1] * m_slab, dtype=np.int32)
    mass = np.array([MASS_AMU[proj]] + [MASS_AMU[metal]] * m_slab) * AMU2MASS
    fixed = np.zeros((N, 1, 3), dtype=bool)
    per_layer = n * n
    for layer in range(nlayers - nfixed, nlayers):
        fixed[1 + layer * per_layer: 1 + (layer + 1) * per_layer, 0, :] = True
    r = np.zeros((N, 1, 3))
    fx, fy = rng.random(), rng.random()
    r[0, 0, :] = cell[:, 0] * fx + cell[:, 1] * fy
    r[0, 0, 2] = z0
    r[1:, 0, :] = slab
    v = np.zeros((N, 1, 3))
    th = math.radians(polar_deg)
    ph = math.radians(azimuth_deg)
    vinc = math.sqrt(2.0 * einc / mass[0])
    v[0, 0, :] = [vinc * math.sin(th) * math.cos(ph), vinc * math.sin(th) * math.sin(ph),
                  -vinc * math.cos(th)]
    if thermal and tsurf > 0:
        sig = np.sqrt(KB * tsurf / mass[1:])
        v[1:, 0, :] = rng.standard_normal((m_slab, 3)) * sig[:, None]
        v[1:, 0, :][fixed[1:, 0, :]] = 0.0
    algo = np.array([proj_algo, latt_algo], dtype=np.int32)
    sysm = System(type_names, idx, mass, fixed, algo, cell, r, v)
    if nbeads > 1:
        sysm = sysm.with_beads(nbeads)
    return sysm


def bulk_fcc(metal="Au", a=None, ncell=4):
    """Periodic fcc bulk (cubic cell, 4*ncell^3 atoms), all mobile."""
    a = LATTICE_A[metal] if a is None else a
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pos = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                for bvec in basis:
                    pos.append((np.array([i, j, k]) + bvec) * a)
    pos = np.array(pos)
    N = len(pos)
    cell = np.eye(3) * a * ncell
    sysm = System([(metal, False)], np.zeros(N, dtype=np.int32),
                  np.full(N, MASS_AMU[metal] * AMU2MASS), np.zeros((N, 1, 3), dtype=bool),
                  np.array([PROP_VERLET], dtype=np.int32), cell,
                  pos.reshape(N, 1, 3).copy(), np.zeros((N, 1, 3)))
    return sysm
