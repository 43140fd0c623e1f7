"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the NENE oracle
(nene_oracle.c), the CPU checker of the GPU NN surrogate.  Imported only by
tests/, the smoke check and the benchmark's CPU baseline.

Positions are taken in Bohr and the results are in Hartree and
Hartree/Bohr, as in the network's own units (the reference converts at the
compute_nene boundary, src/pes_nene_mod.f90:7237-7261).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle as _orc

MAXE, MAXSF, MAXL, MAXNODES = 4, 64, 4, 64


class NeneModelC(C.Structure):
    _fields_ = [
        ("nelem", C.c_int), ("z", C.c_int * MAXE), ("nsf", C.c_int * MAXE),
        ("type", (C.c_int * MAXSF) * MAXE), ("e1", (C.c_int * MAXSF) * MAXE), ("e2", (C.c_int * MAXSF) * MAXE),
        ("eta", (C.c_double * MAXSF) * MAXE), ("rs", (C.c_double * MAXSF) * MAXE),
        ("lambda_", (C.c_double * MAXSF) * MAXE), ("zeta", (C.c_double * MAXSF) * MAXE),
        ("rc", (C.c_double * MAXSF) * MAXE),
        ("gmin", (C.c_double * MAXSF) * MAXE), ("gmax", (C.c_double * MAXSF) * MAXE),
        ("gavg", (C.c_double * MAXSF) * MAXE),
        ("scmin", C.c_double), ("scmax", C.c_double), ("scale", C.c_int), ("center", C.c_int),
        ("cutoff_type", C.c_int), ("cutoff_alpha", C.c_double), ("nlayers", C.c_int),
        ("nodes", (C.c_int * (MAXL + 1)) * MAXE), ("act", (C.c_char * MAXL) * MAXE),
        ("woff", C.c_long * MAXE), ("w", C.POINTER(C.c_double)),
    ]


_done = False


def _lib():
    global _done
    L = _orc.lib()
    if not _done:
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        M = C.POINTER(NeneModelC)
        L.orc_nene_descriptors.argtypes = [M, C.c_int, ip, dp, dp, dp, dp]
        L.orc_nene_descriptors.restype = C.c_int
        L.orc_nene_energy_forces.argtypes = [M, C.c_int, ip, dp, dp, dp, dp, dp, ip]
        L.orc_nene_energy_forces.restype = C.c_int
        L.orc_nene_loss_grad.argtypes = [M, C.c_int, ip, C.POINTER(C.c_long), ip, dp, dp, dp, dp, C.c_double,
                                         C.c_double, dp, dp]
        L.orc_nene_loss_grad.restype = C.c_int
        _done = True
    return L


class Model:
    """A nene.NeneModel packed into the oracle's struct (weights kept alive)."""

    def __init__(self, m, weights=None):
        self.m = m
        c = NeneModelC()
        c.nelem = len(m.elements)
        eidx = {z: i for i, z in enumerate(m.elements)}
        ws, off = [], 0
        for ie, z in enumerate(m.elements):
            c.z[ie] = z
            c.nsf[ie] = m.nsf(z)
            if m.nsf(z) > MAXSF or max(m.hidden) > MAXNODES or len(m.hidden) + 1 > MAXL:
                raise ValueError("network larger than the oracle's limits")
            for k, s in enumerate(m.sf[z]):
                c.type[ie][k] = s.type
                c.e1[ie][k] = eidx[s.e1]
                c.e2[ie][k] = eidx[s.e2] if s.type == 3 else 0
                c.eta[ie][k], c.rs[ie][k] = s.eta, s.rs
                c.lambda_[ie][k], c.zeta[ie][k], c.rc[ie][k] = s.lam, s.zeta, s.rc
                c.gmin[ie][k] = m.gmin[z][k]
                c.gmax[ie][k] = m.gmax[z][k]
                c.gavg[ie][k] = m.gavg[z][k]
            for l, n in enumerate(m.nodes(z)):
                c.nodes[ie][l] = n
            for l, a in enumerate(m.activations):
                c.act[ie][l] = a.encode()
            c.woff[ie] = off
            w = (weights[z] if weights is not None else m.weights[z])
            ws.append(np.asarray(w, dtype=np.float64))
            off += m.nweights(z)
        c.scmin, c.scmax = m.scmin, m.scmax
        c.scale, c.center = int(m.scale), int(m.center)
        c.cutoff_type, c.cutoff_alpha = m.cutoff_type, m.cutoff_alpha
        c.nlayers = len(m.hidden) + 1
        self.w = np.ascontiguousarray(np.concatenate(ws))
        c.w = self.w.ctypes.data_as(C.POINTER(C.c_double))
        self.c = c
        self.eidx = eidx

    def elem_index(self, symbols_or_z):
        from md_tian2_amd import nene
        return np.array([self.eidx[s if isinstance(s, (int, np.integer)) else nene.ATOMIC_NUMBER[s]]
                         for s in symbols_or_z], dtype=np.int32)

    def descriptors(self, elem, xyz, lat=None, derivatives=False):
        n = len(elem)
        G = np.zeros((n, MAXSF))
        dG = np.zeros((n, MAXSF, n, 3)) if derivatives else None
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        lt = None if lat is None else np.ascontiguousarray(lat, dtype=np.float64)
        rc = _lib().orc_nene_descriptors(C.byref(self.c), n, _orc._ip(np.ascontiguousarray(elem, np.int32)),
                                         _orc._dp(x), None if lt is None else _orc._dp(lt), _orc._dp(G),
                                         None if dG is None else _orc._dp(dG))
        if rc:
            raise RuntimeError(f"orc_nene_descriptors rc={rc}")
        return (G, dG) if derivatives else G

    def energy_forces(self, elem, xyz, lat=None):
        """(E [Ha], Eatom [Ha], F [Ha/Bohr] (n,3), extrapolation count)."""
        n = len(elem)
        E = np.zeros(1)
        Ea = np.zeros(n)
        F = np.zeros((n, 3))
        nex = C.c_int(0)
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        lt = None if lat is None else np.ascontiguousarray(lat, dtype=np.float64)
        rc = _lib().orc_nene_energy_forces(C.byref(self.c), n, _orc._ip(np.ascontiguousarray(elem, np.int32)),
                                           _orc._dp(x), None if lt is None else _orc._dp(lt), _orc._dp(E),
                                           _orc._dp(Ea), _orc._dp(F), C.byref(nex))
        if rc:
            raise RuntimeError(f"orc_nene_energy_forces rc={rc}")
        return E[0], Ea, F, nex.value

    def loss_grad(self, frames, wE=1.0, wF=1.0):
        """frames: list of (elem, xyz [Bohr], lat [Bohr] or None, Eref [eV], Fref [eV/A]).
        Returns (loss, gradient over the concatenated weights)."""
        nat = np.array([len(f[0]) for f in frames], dtype=np.int32)
        off = np.concatenate([[0], np.cumsum(nat)[:-1]]).astype(np.int64)
        elem = np.ascontiguousarray(np.concatenate([f[0] for f in frames]).astype(np.int32))
        xyz = np.ascontiguousarray(np.concatenate([f[1] for f in frames]), dtype=np.float64)
        periodic = frames[0][2] is not None
        lat = np.ascontiguousarray(np.concatenate([f[2].reshape(9) for f in frames])) if periodic else None
        Eref = np.ascontiguousarray([f[3] for f in frames], dtype=np.float64)
        Fref = np.ascontiguousarray(np.concatenate([f[4] for f in frames]), dtype=np.float64)
        g = np.zeros(len(self.w))
        L = np.zeros(1)
        rc = _lib().orc_nene_loss_grad(C.byref(self.c), len(frames), _orc._ip(nat),
                                       off.ctypes.data_as(C.POINTER(C.c_long)), _orc._ip(elem), _orc._dp(xyz),
                                       None if lat is None else _orc._dp(lat), _orc._dp(Eref), _orc._dp(Fref),
                                       wE, wF, _orc._dp(L), _orc._dp(g))
        if rc:
            raise RuntimeError(f"orc_nene_loss_grad rc={rc}")
        return L[0], g
