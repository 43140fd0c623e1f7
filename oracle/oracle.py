"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle_mdt.so`` (built by ``oracle/Makefile``) and exposes
the reference-semantics routines on numpy arrays laid out like the
reference's (3,P,N) Fortran arrays.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAXT = 8
PES_ID_EMT = 2003
STREAM_LANGEVIN, STREAM_PILE, STREAM_ANDERSEN_V, STREAM_ANDERSEN_U = 0, 1, 2, 3


class EmtParams(C.Structure):
    _fields_ = [("ntypes", C.c_int)] + [
        (k, C.c_double * MAXT) for k in ("eta2", "n0", "e0", "lambda_", "v0", "kappa", "s0")]


class Universe(C.Structure):
    _fields_ = [
        ("natoms", C.c_int), ("nbeads", C.c_int), ("ntypes", C.c_int),
        ("r", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("f", C.POINTER(C.c_double)), ("a", C.POINTER(C.c_double)),
        ("m", C.POINTER(C.c_double)), ("is_fixed", C.POINTER(C.c_int)),
        ("idx", C.POINTER(C.c_int)), ("pes", C.POINTER(C.c_int)),
        ("algo", C.POINTER(C.c_int)), ("is_proj", C.POINTER(C.c_int)),
        ("epot", C.POINTER(C.c_double)), ("dens", C.POINTER(C.c_double)),
        ("simbox", C.c_double * 9), ("isimbox", C.c_double * 9),
    ]


class SimParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("step", "Tsurf", "Tproj", "pile_tau", "andersen_time")]


class TrajObs(C.Structure):
    """orc_traj_obs: the module variables of trajectory_info.f90:25-43."""
    _fields_ = [("closest_approach", C.c_double), ("closest_approach_time", C.c_double),
                ("lowest_z", C.c_double * 3), ("interaction_time", C.c_double), ("nturning_points", C.c_int),
                ("was_moving_down", C.c_int), ("is_adsorbed", C.c_int)]


def lib():
    """The oracle library: oracle/liboracle_mdt.so (the checker build,
    -ffp-contract=off), or the build named by MDT_ORACLE_LIB (bench.py's
    CPU-baseline leg times an -O3 -march=native build of the same sources)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("MDT_ORACLE_LIB") or os.path.join(_HERE, "liboracle_mdt.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_calc_force.argtypes = [C.POINTER(Universe), C.POINTER(EmtParams), C.POINTER(SimParams), dp]
        L.orc_calc_force.restype = C.c_int
        L.orc_emt_forces.argtypes = [C.POINTER(Universe), C.POINTER(EmtParams)]
        L.orc_emt_forces.restype = C.c_int
        L.orc_md_step.argtypes = [C.POINTER(Universe), C.POINTER(EmtParams), C.POINTER(SimParams), dp,
                                  dp, dp, dp, dp]
        L.orc_md_step.restype = C.c_int
        L.orc_propagate_1.argtypes = [C.POINTER(Universe), C.POINTER(SimParams), dp, dp, dp, dp, dp]
        L.orc_propagate_1.restype = C.c_int
        L.orc_propagate_2.argtypes = [C.POINTER(Universe), C.POINTER(SimParams), dp]
        L.orc_propagate_2.restype = C.c_int
        L.orc_ldfa.argtypes = [C.POINTER(Universe), C.c_int]
        L.orc_ldfa.restype = C.c_int
        L.orc_remove_com_velocity.argtypes = [C.POINTER(Universe)]
        L.orc_ring_polymer_step.argtypes = [C.POINTER(Universe), C.POINTER(SimParams), dp]
        L.orc_build_cjk.argtypes = [C.c_int, dp]
        L.orc_minimg_beads.argtypes = [C.POINTER(Universe), C.c_int, C.c_int, dp, dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_philox_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_philox_normal.restype = C.c_double
        L.orc_philox_uniform.argtypes = L.orc_philox_normal.argtypes
        L.orc_philox_uniform.restype = C.c_double
        L.orc_philox_fill_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, dp]
        L.orc_run_trajectory.argtypes = [C.POINTER(Universe), C.POINTER(EmtParams), C.POINTER(SimParams),
                                         C.c_uint64, C.c_uint32, C.c_int, C.c_int]
        L.orc_run_trajectory.restype = C.c_int
        L.orc_collect_characteristics.argtypes = [C.POINTER(Universe), C.POINTER(SimParams), C.c_double,
                                                  C.c_double, C.c_int, C.POINTER(TrajObs)]
        L.orc_run_scatter.argtypes = [C.POINTER(Universe), C.POINTER(EmtParams), C.POINTER(SimParams), C.c_uint64,
                                      C.c_uint32, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double,
                                      C.c_double, C.POINTER(TrajObs), C.POINTER(C.c_int)]
        L.orc_run_scatter.restype = C.c_int
        L.orc_energy_row.argtypes = [C.POINTER(Universe), C.POINTER(SimParams), dp]
        L.orc_apply_pip.argtypes = [C.POINTER(Universe), C.POINTER(C.c_int), dp, dp]
        L.orc_set_proj_incidence.argtypes = [C.POINTER(Universe), C.c_double, C.c_double, C.c_double, C.c_int,
                                             C.c_double, C.c_double, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_params(cols, ntypes):
    p = EmtParams()
    p.ntypes = ntypes
    for key, field in (("eta2", "eta2"), ("n0", "n0"), ("e0", "e0"), ("lambda", "lambda_"),
                       ("v0", "v0"), ("kappa", "kappa"), ("s0", "s0")):
        arr = getattr(p, field)
        for t in range(ntypes):
            arr[t] = cols[key][t]
    return p


class OracleState:
    """Owns numpy buffers for one universe and the matching C struct."""

    def __init__(self, sysm, cols, pesm, step=0.1, tsurf=0.0, tproj=0.0, pile_tau=200.0,
                 andersen_time=30.0):
        N, P, T = sysm.natoms, sysm.nbeads, sysm.ntypes
        self.N, self.P, self.T = N, P, T
        self.r = np.ascontiguousarray(sysm.r, dtype=np.float64).copy()
        self.v = np.ascontiguousarray(sysm.v, dtype=np.float64).copy()
        self.f = np.zeros((N, P, 3))
        self.a = np.zeros((N, P, 3))
        self.m = np.ascontiguousarray(sysm.mass, dtype=np.float64).copy()
        self.fixed = np.ascontiguousarray(sysm.fixed.astype(np.int32))
        self.idx = np.ascontiguousarray(sysm.idx, dtype=np.int32).copy()
        self.pes = np.ascontiguousarray(np.array(pesm, dtype=np.int32).T)  # column-major (ti,tj)
        self.algo = np.ascontiguousarray(sysm.algo, dtype=np.int32).copy()
        self.is_proj = np.ascontiguousarray(sysm.is_proj, dtype=np.int32)
        self.epot = np.zeros(P)
        self.dens = np.zeros((N, P))
        self.params = make_params(cols, T)
        self.sp = SimParams(step, tsurf, tproj, pile_tau, andersen_time)
        self.cjk = np.zeros((P, P))  # Fortran column-major storage: cjk[k, j] = cjk(j+1, k+1)
        lib().orc_build_cjk(P, _dp(self.cjk))
        u = Universe()
        u.natoms, u.nbeads, u.ntypes = N, P, T
        u.r, u.v, u.f, u.a = _dp(self.r), _dp(self.v), _dp(self.f), _dp(self.a)
        u.m, u.is_fixed, u.idx, u.pes = _dp(self.m), _ip(self.fixed), _ip(self.idx), _ip(self.pes)
        u.algo, u.is_proj, u.epot, u.dens = _ip(self.algo), _ip(self.is_proj), _dp(self.epot), _dp(self.dens)
        sb = np.asarray(sysm.simbox, dtype=np.float64)
        isb = np.linalg.inv(sb)
        for k in range(9):
            u.simbox[k] = sb.T.reshape(-1)[k]  # column-major
            u.isimbox[k] = isb.T.reshape(-1)[k]
        self.u = u

    def calc_force(self):
        rc = lib().orc_calc_force(C.byref(self.u), C.byref(self.params), C.byref(self.sp), _dp(self.cjk))
        if rc:
            raise RuntimeError(f"oracle calc_force failed rc={rc}")

    def emt_forces(self):
        rc = lib().orc_emt_forces(C.byref(self.u), C.byref(self.params))
        if rc:
            raise RuntimeError(f"oracle compute_emt failed rc={rc}")

    def md_step(self, randy=None, zeta=None, and_z=None, and_u=None):
        z = np.zeros((self.N, self.P, 3))
        args = [np.ascontiguousarray(x if x is not None else z) for x in (randy, zeta, and_z)]
        au = np.ascontiguousarray(and_u if and_u is not None else np.ones((self.N, self.P)))
        rc = lib().orc_md_step(C.byref(self.u), C.byref(self.params), C.byref(self.sp), _dp(self.cjk),
                               _dp(args[0]), _dp(args[1]), _dp(args[2]), _dp(au))
        if rc:
            raise RuntimeError(f"oracle md_step failed rc={rc}")

    def prop1(self, randy=None, zeta=None, and_z=None, and_u=None):
        z = np.zeros((self.N, self.P, 3))
        args = [np.ascontiguousarray(x if x is not None else z) for x in (randy, zeta, and_z)]
        au = np.ascontiguousarray(and_u if and_u is not None else np.ones((self.N, self.P)))
        rc = lib().orc_propagate_1(C.byref(self.u), C.byref(self.sp), _dp(args[0]), _dp(args[1]),
                                   _dp(args[2]), _dp(au), _dp(self.cjk))
        if rc:
            raise RuntimeError(f"oracle propagate_1 failed rc={rc}")

    def prop2(self, randy=None):
        z = np.ascontiguousarray(randy if randy is not None else np.zeros((self.N, self.P, 3)))
        rc = lib().orc_propagate_2(C.byref(self.u), C.byref(self.sp), _dp(z))
        if rc:
            raise RuntimeError(f"oracle propagate_2 failed rc={rc}")

    def ldfa(self, i):
        return lib().orc_ldfa(C.byref(self.u), i)

    def ring_polymer_step(self):
        lib().orc_ring_polymer_step(C.byref(self.u), C.byref(self.sp), _dp(self.cjk))

    def remove_com(self):
        lib().orc_remove_com_velocity(C.byref(self.u))

    def run(self, seed, traj, step0, nsteps):
        rc = lib().orc_run_trajectory(C.byref(self.u), C.byref(self.params), C.byref(self.sp),
                                      seed, traj, step0, nsteps)
        if rc:
            raise RuntimeError(f"oracle run failed rc={rc}")

    def run_scatter(self, seed, traj, nsteps, proj_ul=None, proj_ll=None, scatter=True, adsorption=None,
                    adsorbed=False):
        """The md loop of one trajectory with its exits and characteristics
        (md_tian2.f90:75-101); returns (TrajObs, final istep)."""
        big = np.finfo(np.float64).max
        ul = big if proj_ul is None else proj_ul
        ll = -big if proj_ll is None else proj_ll
        a0, a1 = (big, big) if adsorption is None else adsorption
        o = TrajObs()
        o.is_adsorbed = 1 if adsorbed else 0
        iend = C.c_int(0)
        rc = lib().orc_run_scatter(C.byref(self.u), C.byref(self.params), C.byref(self.sp), seed, traj, nsteps,
                                   ul, ll, 1 if scatter else 0, a0, a1, C.byref(o), C.byref(iend))
        if rc:
            raise RuntimeError(f"oracle run failed rc={rc}")
        return o, iend.value

    def collect(self, istep, obs, adsorption=None):
        big = np.finfo(np.float64).max
        a0, a1 = (big, big) if adsorption is None else adsorption
        lib().orc_collect_characteristics(C.byref(self.u), C.byref(self.sp), a0, a1, istep, C.byref(obs))

    def energy_row(self):
        """Columns 2-21 of output_nrg for the current state."""
        out = np.zeros(20)
        lib().orc_energy_row(C.byref(self.u), C.byref(self.sp), _dp(out))
        return out

    def apply_pip(self, random, value, U):
        rnd = np.ascontiguousarray(random, dtype=np.int32)
        lib().orc_apply_pip(C.byref(self.u), _ip(rnd), _dp(np.ascontiguousarray(value, dtype=np.float64)),
                            _dp(np.ascontiguousarray(U, dtype=np.float64)))

    def set_proj_incidence(self, einc, sigma_einc, polar_deg, az_random, az_deg, sigma_az, rnd):
        lib().orc_set_proj_incidence(C.byref(self.u), einc, sigma_einc, polar_deg, 1 if az_random else 0, az_deg,
                                     sigma_az, _dp(np.ascontiguousarray(rnd, dtype=np.float64)))

    def minimg(self, i, j):
        d = np.zeros(self.P)
        vec = np.zeros((self.P, 3))
        lib().orc_minimg_beads(C.byref(self.u), i, j, _dp(d), _dp(vec))
        return d, vec


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def philox_normals(seed, traj, step, stream, natoms, nbeads):
    out = np.zeros((natoms, nbeads, 3))
    lib().orc_philox_fill_normals(seed, traj, step, stream, natoms, nbeads, _dp(out))
    return out


def emt_energy_forces(sysm, cols, pesm):
    """Reference calc_force on one universe: (epot(P), f(N,P,3), dens(N,P))."""
    st = OracleState(sysm, cols, pesm)
    st.emt_forces()
    return st.epot.copy(), st.f.copy(), st.dens.copy()
