"""Initial conditions of a trajectory through the library's sampler
(mdt_sample_* / mdt_apply_pip / mdt_set_proj_incidence, host code): the
reference's apply_pip (src/md_init.f90:498-554), set_proj_incidence
(:829-877) and the configuration-file choice of prepare_next_traj
(:904-926), with every draw a Philox uniform keyed by (seed, trajectory id),
so the Fortran driver and Python build the same universes.

Arrays follow geometry.System / conf.Conf: r, v (N, P, 3) -- the memory
layout of the reference's (3, P, N) --, simbox with the cell vectors as
columns."""
from __future__ import annotations

import ctypes as C

import numpy as np

from md_tian2_amd import mdt


class Pip(C.Structure):
    """mdt_pip: `pip x y z`, each a number or "r"."""
    _fields_ = [("random", C.c_int32 * 3), ("value", C.c_double * 3)]

    @classmethod
    def parse(cls, words):
        p = cls()
        for k, w in enumerate(words):
            if str(w) == "r":
                p.random[k] = 1
            else:
                p.value[k] = float(w)
        return p


class Incidence(C.Structure):
    """mdt_incidence: `Einc e [sigma]`, `polar deg`, `azimuth deg|r [sigma]`."""
    _fields_ = [("einc", C.c_double), ("sigma_einc", C.c_double), ("polar_deg", C.c_double),
                ("azimuth_random", C.c_int32), ("pad_", C.c_int32), ("azimuth_deg", C.c_double),
                ("sigma_azimuth", C.c_double)]

    @classmethod
    def make(cls, einc, polar_deg, azimuth="r", sigma_einc=0.0, sigma_azimuth=0.0):
        x = cls()
        x.einc, x.sigma_einc, x.polar_deg = float(einc), float(sigma_einc), float(polar_deg)
        if str(azimuth) == "r":
            x.azimuth_random = 1
        else:
            x.azimuth_deg = float(azimuth)
        x.sigma_azimuth = float(sigma_azimuth)
        return x


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(rc):
    if rc != 0:
        raise mdt.MdtError(f"sampler returned {rc}")


def uniform(seed, traj_id, slot):
    return mdt.lib().mdt_sample_uniform(seed, traj_id, slot)


def conf_index(seed, traj_id, which, n):
    """File number 1..n of a conf folder (which: 0 projectile, 1 lattice)."""
    return mdt.lib().mdt_sample_conf_index(seed, traj_id, which, n)


def apply_pip(r, simbox, pip, seed, traj_id):
    """apply_pip on a projectile universe's positions r (N, P, 3), in place."""
    assert r.flags.c_contiguous and r.dtype == np.float64
    sb = np.ascontiguousarray(np.asarray(simbox, dtype=np.float64).T.reshape(-1))
    _check(mdt.lib().mdt_apply_pip(C.byref(pip), seed, traj_id, r.shape[0], r.shape[1], _d(sb), _d(r)))


def set_proj_incidence(v, idx, is_proj, mass, inc, seed, traj_id):
    """set_proj_incidence on velocities v (N, P, 3), in place; idx 0-based
    types per atom, is_proj per type, mass per atom (program units)."""
    assert v.flags.c_contiguous and v.dtype == np.float64
    i1 = np.ascontiguousarray(np.asarray(idx) + 1, dtype=np.int32)
    ip = np.ascontiguousarray(np.asarray(is_proj, dtype=np.int32))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    _check(mdt.lib().mdt_set_proj_incidence(C.byref(inc), seed, traj_id, v.shape[0], v.shape[1], _i(i1), _i(ip),
                                            _d(m), _d(v)))


def maxwell(v, idx, is_proj, mass, fixed, temperature, seed, traj_id):
    """Maxwell-Boltzmann velocities for the non-projectile atoms, in place."""
    i1 = np.ascontiguousarray(np.asarray(idx) + 1, dtype=np.int32)
    ip = np.ascontiguousarray(np.asarray(is_proj, dtype=np.int32))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    fx = np.ascontiguousarray(np.asarray(fixed), dtype=np.int32)
    _check(mdt.lib().mdt_sample_maxwell(float(temperature), seed, traj_id, v.shape[0], v.shape[1], _i(i1), _i(ip),
                                        _d(m), _i(fx), _d(v)))
