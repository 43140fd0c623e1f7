"""md_tian2's configuration files: mxt universe dumps, POSCAR, and `conf
merge` of a lattice with a projectile.

* mxt (read_mxt, src/md_init.f90:344-381; written by output_mod.f90:
  620-630): Fortran unformatted sequential, two records framed by 4-byte
  little-endian lengths -- (natoms, nbeads, ntypes) int32, then r(3,P,N)
  f64, v(3,P,N) f64, is_cart logical(4), is_fixed(3,P,N) logical(4),
  idx(N) int32 (1-based types), name(ntypes) character(3), is_proj(ntypes)
  logical(4), simbox(3,3) f64, isimbox(3,3) f64 (column-major, columns are
  the cell vectors).  Positions are cartesian when is_cart, else direct.
  A logical is nonzero when true; the files shipped with md_tian2 hold -1
  (ifort's .true.), gfortran and flang write 1, and a file is written back
  with the value it was read with.
* POSCAR (read_poscar, src/md_init.f90:200-339): comment, scaling, three
  cell vectors, element symbols, counts with an optional ":P:" bead count,
  cartesian|direct, then per atom and bead x y z [and T/F mobility flags,
  F = fixed], optionally a blank line and velocities.  The first
  `nprojectiles` element types are projectiles.
* merge (merge_universes, src/md_init.f90:559-625): projectile atoms first
  (one projectile), lattice types after the projectile's, everything
  cartesian, then each atom's bead centroid folded into the cell by
  subtracting int() of its direct coordinate (truncation toward zero, so
  a negative coordinate is not folded).
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np

from md_tian2_amd import geometry


@dataclass
class Conf:
    r: np.ndarray            # (N, P, 3), cartesian when is_cart, else direct
    v: np.ndarray            # (N, P, 3)
    is_cart: bool
    fixed: np.ndarray        # (N, P, 3) bool
    idx: np.ndarray          # (N,) int, 1-based type
    names: list              # ntypes element symbols
    is_proj: list            # ntypes bools
    simbox: np.ndarray       # (3, 3), columns are the cell vectors
    isimbox: np.ndarray = field(default=None)
    true_code: int = 1       # int32 of .true. in the file: 1 (gfortran, flang) or -1 (ifort)

    @property
    def natoms(self):
        return int(self.idx.shape[0])

    @property
    def nbeads(self):
        return int(self.r.shape[1])

    @property
    def ntypes(self):
        return len(self.names)

    def cartesian(self):
        """Positions in Angstrom (N, P, 3)."""
        if self.is_cart:
            return self.r.copy()
        return np.einsum("ij,npj->npi", self.simbox, self.r)

    def to_system(self, masses_amu=None, algos=None):
        """A geometry.System (positions cartesian, masses in program units).
        masses_amu: per type [amu] (default: geometry.MASS_AMU of the name);
        algos: per type propagator id (default Verlet)."""
        masses_amu = masses_amu or [geometry.MASS_AMU[n] for n in self.names]
        algos = algos or [geometry.PROP_VERLET] * self.ntypes
        m = np.array([masses_amu[i - 1] for i in self.idx]) * geometry.AMU2MASS
        return geometry.System([(n, bool(p)) for n, p in zip(self.names, self.is_proj)],
                               (self.idx - 1).astype(np.int32), m, self.fixed.copy(),
                               np.array(algos, dtype=np.int32), self.simbox.copy(), self.cartesian(),
                               self.v.copy())


def _records(raw):
    out, off = [], 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        if n < 0 or off + 8 + n > len(raw):
            raise ValueError("not a Fortran sequential unformatted file (record length %d at %d)" % (n, off))
        (n2,) = struct.unpack_from("<i", raw, off + 4 + n)
        if n2 != n:
            raise ValueError("record markers differ (%d, %d) at %d" % (n, n2, off))
        out.append(raw[off + 4:off + 4 + n])
        off += 8 + n
    return out


def read_mxt(path):
    recs = _records(open(path, "rb").read())
    if len(recs) != 2 or len(recs[0]) != 12:
        raise ValueError(f"{path}: expected two records (header of 3 int32, universe)")
    N, P, T = struct.unpack("<3i", recs[0])
    n3 = 3 * P * N
    need = 8 * n3 * 2 + 4 + 4 * n3 + 4 * N + 3 * T + 4 * T + 8 * 18
    body = recs[1]
    if len(body) != need:
        raise ValueError(f"{path}: universe record of {len(body)} bytes, {need} expected for "
                         f"natoms={N} nbeads={P} ntypes={T}")
    o = 0

    def take(dtype, count):
        nonlocal o
        a = np.frombuffer(body, dtype=dtype, count=count, offset=o)
        o += a.nbytes
        return a.copy()

    r = take("<f8", n3).reshape(N, P, 3)
    v = take("<f8", n3).reshape(N, P, 3)
    cart_raw = take("<i4", 1)
    fixed_raw = take("<i4", n3)
    idx = take("<i4", N).astype(np.int64)
    names = [body[o + 3 * k:o + 3 * k + 3].decode("ascii").strip() for k in range(T)]
    o += 3 * T
    proj_raw = take("<i4", T)
    simbox = take("<f8", 9).reshape(3, 3).T.copy()   # column-major -> [row][col]
    isimbox = take("<f8", 9).reshape(3, 3).T.copy()
    trues = np.concatenate([cart_raw, fixed_raw, proj_raw])
    trues = trues[trues != 0]
    return Conf(r, v, bool(cart_raw[0]), fixed_raw.reshape(N, P, 3) != 0, idx, names,
                [bool(x) for x in proj_raw], simbox, isimbox, int(trues[0]) if trues.size else 1)


def write_mxt(path, c):
    N, P, T = c.natoms, c.nbeads, c.ntypes
    hdr = struct.pack("<3i", N, P, T)
    isb = c.isimbox if c.isimbox is not None else np.linalg.inv(c.simbox)
    body = b"".join([
        np.ascontiguousarray(c.r, dtype="<f8").tobytes(),
        np.ascontiguousarray(c.v, dtype="<f8").tobytes(),
        struct.pack("<i", c.true_code if c.is_cart else 0),
        (np.ascontiguousarray(c.fixed, dtype="<i4") * c.true_code).astype("<i4").tobytes(),
        np.ascontiguousarray(c.idx, dtype="<i4").tobytes(),
        b"".join(n.ljust(3)[:3].encode("ascii") for n in c.names),
        np.array([c.true_code if p else 0 for p in c.is_proj], dtype="<i4").tobytes(),
        np.ascontiguousarray(np.asarray(c.simbox).T, dtype="<f8").tobytes(),
        np.ascontiguousarray(np.asarray(isb).T, dtype="<f8").tobytes(),
    ])
    with open(path, "wb") as f:
        for rec in (hdr, body):
            f.write(struct.pack("<i", len(rec)))
            f.write(rec)
            f.write(struct.pack("<i", len(rec)))


def read_poscar(path, nprojectiles):
    lines = open(path).read().split("\n")
    k = 1
    scale = float(lines[k].split()[0])
    k += 1
    vec = []
    while len(vec) < 9:
        vec += [float(x) for x in lines[k].split()]
        k += 1
    simbox = np.array(vec[:9]).reshape(3, 3).T * scale  # each line a column
    names = lines[k].split()
    k += 1
    cl = lines[k]
    k += 1
    nbeads = 1
    if ":" in cl:
        a, b = cl.index(":"), cl.rindex(":")
        nbeads = int(cl[a + 1:b])
        if nbeads < 1:
            raise ValueError(f"{path}: bead count {nbeads}")
        cl = cl[b + 1:]
    counts = [int(x) for x in cl.split()]
    if len(counts) != len(names) or min(counts) <= 0:
        raise ValueError(f"{path}: counts {counts} for elements {names}")
    system = lines[k].strip().lower()
    k += 1
    N = sum(counts)
    first = lines[k].split()
    r = np.zeros((N, nbeads, 3))
    fixed = np.zeros((N, nbeads, 3), dtype=bool)
    toks = []
    if len(first) == 3:
        while len(toks) < 3 * N * nbeads:
            toks += lines[k].split()
            k += 1
        r = np.array([float(x) for x in toks[:3 * N * nbeads]]).reshape(N, nbeads, 3)
    elif len(first) == 6:
        while len(toks) < 6 * N * nbeads:
            toks += lines[k].split()
            k += 1
        t = np.array(toks[:6 * N * nbeads]).reshape(N, nbeads, 6)
        r = t[:, :, :3].astype(float)
        fixed = t[:, :, 3:] == "F"
    else:
        raise ValueError(f"{path}: coordinate lines need 3 or 6 entries")
    v = np.zeros((N, nbeads, 3))
    rest = lines[k:]
    if rest and rest[0].strip() == "" and any(x.strip() for x in rest[1:]):
        vt = " ".join(rest[1:]).split()
        v = np.array([float(x) for x in vt[:3 * N * nbeads]]).reshape(N, nbeads, 3)
    elif rest and rest[0].strip() != "":
        raise ValueError(f"{path}: unexpected text after the coordinates")
    if system in ("c", "cart", "cartesian"):
        r = r * scale
        is_cart = True
    elif system in ("d", "direct"):
        is_cart = False
    else:
        raise ValueError(f"{path}: coordinate system '{system}'")
    idx = np.repeat(np.arange(1, len(counts) + 1), counts)
    is_proj = [t < nprojectiles for t in range(len(names))]
    return Conf(r, v, is_cart, fixed, idx, names, is_proj, simbox, np.linalg.inv(simbox))


def merge(lattice, projectile):
    """merge_universes: the projectile (one atom, projectile types only)
    first, the lattice after it, cartesian, centroids folded by int()."""
    if lattice.nbeads != projectile.nbeads:
        raise ValueError("mxt and merge files differ in number of beads")
    if any(lattice.is_proj):
        raise ValueError("mxt file contains projectile atoms")
    if not all(projectile.is_proj):
        raise ValueError("merge file contains lattice atoms")
    if projectile.natoms > 1:
        raise ValueError("multiple projectiles not implemented for use with merge option")
    r = np.concatenate([projectile.cartesian(), lattice.cartesian()])
    v = np.concatenate([projectile.v, lattice.v])
    fixed = np.concatenate([projectile.fixed, lattice.fixed])
    idx = np.concatenate([projectile.idx, lattice.idx + projectile.idx.max()])
    isb = np.linalg.inv(lattice.simbox) if lattice.isimbox is None else lattice.isimbox
    cen = r.mean(axis=1) @ isb.T                       # direct centroids
    shift = np.trunc(cen)                               # int(): toward zero
    d = np.einsum("ij,npj->npi", isb, r) - shift[:, None, :]
    r = np.einsum("ij,npj->npi", lattice.simbox, d)
    return Conf(r, v, True, fixed, idx, projectile.names + lattice.names,
                list(projectile.is_proj) + list(lattice.is_proj), lattice.simbox.copy(), isb.copy(),
                lattice.true_code)


def mxt_name(folder, k):
    """File k (1-based) of an mxt folder: <folder>/mxt_%08d.dat (run_config.f90:472)."""
    return os.path.join(folder, "mxt_%08d.dat" % k)
