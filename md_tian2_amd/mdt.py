"""Python binding of the C-ABI in include/mdt_gpu.h (libmdt_gpu.so).

The GPU library is the only implementation: there is no CPU fallback, and
every call raises ``MdtError`` if the library is missing or a call fails.
Arrays follow the reference's (3,P,N) Fortran layout, which numpy holds as
C-order (N, P, 3).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import pesfile

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdt_gpu.so")
_LIB = None

PES_ID_EMT = 2003
ENERGY_AND_FORCE = 3001

# exported symbols of include/mdt_gpu.h
SYMBOLS = (
    "mdt_last_error", "mdt_version", "mdt_device_count", "mdt_create", "mdt_destroy",
    "mdt_emt_params_set", "mdt_compute_emt", "mdt_compute_emt_noalloc", "mdt_batch_create", "mdt_batch_destroy",
    "mdt_batch_set_atoms", "mdt_batch_set_cells", "mdt_batch_set_propagation", "mdt_batch_seed",
    "mdt_batch_set_state", "mdt_batch_get_state", "mdt_batch_compute_forces", "mdt_batch_step",
    "mdt_batch_status", "mdt_batch_last_timing", "mdt_batch_tune", "mdt_batch_profile",
    "mdt_batch_force_timing", "mdt_batch_set_precision", "mdt_batch_precision_arith", "mdt_batch_class_shapes", "mdt_batch_set_wide", "mdt_batch_set_keep", "mdt_batch_set_split", "mdt_batch_last_split", "mdt_batch_set_graph", "mdt_batch_set_bead_mode", "mdt_batch_set_com_removal",
    "mdt_batch_set_list_skin",
    "mdt_batch_set_observe", "mdt_batch_get_observables", "mdt_batch_estimators",
    "mdt_sample_uniform", "mdt_sample_conf_index", "mdt_apply_pip", "mdt_set_proj_incidence", "mdt_sample_maxwell",
    "mdt_nn_params_set", "mdt_compute_nene", "mdt_batch_set_pes", "mdt_batch_set_judge_percent", "mdt_batch_nn_stats", "mdt_batch_nn_fallback",
    "mdt_batch_nn_descriptors", "mdt_nn_train_create", "mdt_nn_train_destroy", "mdt_nn_train_descriptor_stats",
    "mdt_nn_train_epoch", "mdt_nn_train_eval", "mdt_nn_train_loss_grad", "mdt_nn_get_weights", "mdt_nn_load",
)


NEST = 20
ESTIMATORS = ("atom_T", "bead_T", "rgyr_sq_p", "rgyr_sq_l", "atom_ekin_p", "atom_ekin_l", "bead_ekin_p",
              "bead_ekin_l", "qntm_ekin_p", "qntm_ekin_l", "atom_epot", "bead_epot", "e_total", "f_total",
              "r_x", "r_y", "r_z", "v_x", "v_y", "v_z")


class TrajObs(C.Structure):
    """mdt_traj_obs of the C interface."""
    _fields_ = [("cl_appr", C.c_double), ("cl_last", C.c_double), ("r_min", C.c_double * 3),
                ("t_int", C.c_double), ("turning_points", C.c_int32), ("moving_down", C.c_int32),
                ("adsorbed", C.c_int32), ("adsorbed_prev", C.c_int32), ("exit_step", C.c_int32),
                ("cl_appr_step", C.c_int32), ("last_step", C.c_int32), ("pad_", C.c_int32)]


class MdtError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MdtError(f"{LIB_PATH} is not built (run python -m md_tian2_amd.build)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, up = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)
        L.mdt_last_error.argtypes = [C.c_char_p, C.c_int]
        L.mdt_device_count.argtypes = [C.POINTER(C.c_int)]
        L.mdt_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.mdt_destroy.argtypes = [vp]
        L.mdt_emt_params_set.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp, dp, dp, ip]
        L.mdt_compute_emt.argtypes = [vp, C.c_int, C.c_int, ip, dp, dp, dp, dp, dp, dp, C.c_int32]
        L.mdt_compute_emt_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int32, vp]
        L.mdt_batch_create.argtypes = [vp, C.c_int, ip, ip, C.POINTER(vp)]
        L.mdt_batch_destroy.argtypes = [vp]
        L.mdt_batch_set_atoms.argtypes = [vp, ip, dp, ip]
        L.mdt_batch_set_cells.argtypes = [vp, dp, dp]
        L.mdt_batch_set_propagation.argtypes = [vp, C.c_int, ip, ip, C.c_double, dp, dp, C.c_double,
                                                C.c_double]
        L.mdt_batch_seed.argtypes = [vp, C.c_uint32, up]
        L.mdt_batch_set_state.argtypes = [vp, dp, dp]
        L.mdt_batch_get_state.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.mdt_batch_compute_forces.argtypes = [vp]
        L.mdt_batch_step.argtypes = [vp, C.c_int]
        L.mdt_batch_status.argtypes = [vp, ip, C.c_int]
        L.mdt_batch_last_timing.argtypes = [vp, dp, C.POINTER(C.c_int)]
        L.mdt_batch_tune.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.mdt_batch_profile.argtypes = [vp, C.c_int]
        L.mdt_batch_force_timing.argtypes = [vp, dp]
        L.mdt_batch_set_precision.argtypes = [vp, C.c_int]
        L.mdt_batch_precision_arith.argtypes = [vp, C.POINTER(C.c_int)]
        L.mdt_batch_class_shapes.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int)]
        L.mdt_batch_set_wide.argtypes = [vp, C.c_int]
        L.mdt_batch_set_keep.argtypes = [vp, C.c_int]
        L.mdt_batch_set_split.argtypes = [vp, C.c_int]
        L.mdt_batch_last_split.argtypes = [vp, C.POINTER(C.c_int)]
        L.mdt_batch_set_bead_mode.argtypes = [vp, C.c_int]
        L.mdt_batch_set_com_removal.argtypes = [vp, C.c_int]
        L.mdt_batch_set_list_skin.argtypes = [vp, C.c_double]
        L.mdt_batch_set_observe.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.mdt_batch_get_observables.argtypes = [vp, C.c_void_p]
        L.mdt_batch_estimators.argtypes = [vp, dp]
        u32 = C.c_uint32
        L.mdt_sample_uniform.argtypes = [u32, u32, u32]
        L.mdt_sample_uniform.restype = C.c_double
        L.mdt_sample_conf_index.argtypes = [u32, u32, C.c_int, C.c_int]
        L.mdt_apply_pip.argtypes = [C.c_void_p, u32, u32, C.c_int, C.c_int, dp, dp]
        L.mdt_set_proj_incidence.argtypes = [C.c_void_p, u32, u32, C.c_int, C.c_int, ip, ip, dp, dp]
        L.mdt_sample_maxwell.argtypes = [C.c_double, u32, u32, C.c_int, C.c_int, ip, ip, dp, ip, dp]
        L.mdt_batch_set_graph.argtypes = [vp, C.c_int]
        L.mdt_nn_params_set.argtypes = [vp, C.c_int, ip, C.c_int, ip, ip, ip, ip, dp, dp, dp, dp, dp, dp, dp, dp,
                                        C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_char_p, dp]
        L.mdt_compute_nene.argtypes = [vp, C.c_int, C.c_int, ip, dp, dp, dp, dp, dp, ip, C.c_int32]
        L.mdt_batch_set_pes.argtypes = [vp, C.c_int, C.c_int]
        L.mdt_batch_set_judge_percent.argtypes = [vp, C.c_int]
        L.mdt_batch_nn_stats.argtypes = [vp, dp, dp, dp, ip]
        L.mdt_batch_nn_fallback.argtypes = [vp, ip]
        L.mdt_batch_nn_descriptors.argtypes = [vp, C.POINTER(C.c_float)]
        L.mdt_nn_train_create.argtypes = [vp, C.c_int, ip, ip, dp, dp, dp, dp, C.c_int, C.c_double, C.c_uint32,
                                          C.POINTER(vp)]
        L.mdt_nn_train_destroy.argtypes = [vp]
        L.mdt_nn_train_descriptor_stats.argtypes = [vp, dp, dp, dp]
        L.mdt_nn_train_epoch.argtypes = [vp, C.c_double, C.c_double, C.c_double, dp]
        L.mdt_nn_train_eval.argtypes = [vp, C.c_int, dp, dp]
        L.mdt_nn_train_loss_grad.argtypes = [vp, C.c_int, ip, C.c_double, C.c_double, dp, dp]
        L.mdt_nn_get_weights.argtypes = [vp, dp]
        L.mdt_nn_load.argtypes = [vp, C.c_char_p, C.c_int, C.c_char_p]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        buf = C.create_string_buffer(2048)
        lib().mdt_last_error(buf, 2048)
        raise MdtError(f"rc={rc}: {buf.value.decode(errors='replace')}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def device_count():
    n = C.c_int(0)
    lib().mdt_device_count(C.byref(n))
    return n.value


class Context:
    """GPU context: HIP stream + EMT parameter table (read_emt equivalent)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().mdt_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().mdt_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_emt(self, cols, pesm):
        """cols: dict of per-type lists (pesfile.emt_tables); pesm: ntypes x ntypes."""
        nt = len(pesm)
        arr = [np.ascontiguousarray(cols[k], dtype=np.float64) for k in pesfile.EMT_KEYS]
        pesf = np.ascontiguousarray(np.array(pesm, dtype=np.int32).T)  # column-major
        _check(lib().mdt_emt_params_set(self.h, nt, *[_dp(a) for a in arr], _ip(pesf)))
        self._keep = (arr, pesf)

    def set_nn(self, model, type_names):
        """Upload a nene.NeneModel; type_names as in geometry.System (symbol, is_proj)."""
        from md_tian2_amd import nene
        el = {z: i for i, z in enumerate(model.elements)}
        eot = np.ascontiguousarray([el.get(nene.ATOMIC_NUMBER.get(t[0], -1), -1) for t in type_names], np.int32)
        nsf = np.ascontiguousarray([model.nsf(z) for z in model.elements], np.int32)
        sfs = [(z, f) for z in model.elements for f in model.sf[z]]
        i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
        f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)
        typ = i32([f.type for _, f in sfs])
        e1 = i32([el[f.e1] for _, f in sfs])
        e2 = i32([el[f.e2] if f.type == 3 else 0 for _, f in sfs])
        eta, rs = f64([f.eta for _, f in sfs]), f64([f.rs for _, f in sfs])
        lam, zeta, rc = f64([f.lam for _, f in sfs]), f64([f.zeta for _, f in sfs]), f64([f.rc for _, f in sfs])
        gmin = f64(np.concatenate([model.gmin[z] for z in model.elements]))
        gmax = f64(np.concatenate([model.gmax[z] for z in model.elements]))
        gavg = f64(np.concatenate([model.gavg[z] for z in model.elements]))
        w = f64(np.concatenate([model.weights[z] for z in model.elements]))
        if len(model.hidden) != 2:
            raise MdtError("the GPU network has exactly two hidden layers")
        _check(lib().mdt_nn_params_set(self.h, len(model.elements), _ip(eot), len(type_names), _ip(nsf), _ip(typ),
                                       _ip(e1), _ip(e2), _dp(eta), _dp(rs), _dp(lam), _dp(zeta), _dp(rc),
                                       _dp(gmin), _dp(gmax), _dp(gavg), model.scmin, model.scmax,
                                       int(model.scale), int(model.center), model.hidden[0], model.hidden[1],
                                       model.activations.encode(), _dp(w)))
        self._keep_nn = (eot, nsf, typ, e1, e2, eta, rs, lam, zeta, rc, gmin, gmax, gavg, w)

    def load_nn(self, directory, type_names):
        """The library's own reader of input.nn / scaling.data / weights.XXX.data."""
        syms = "".join(f"{t[0]:<2s}"[:2] for t in type_names).encode()
        _check(lib().mdt_nn_load(self.h, str(directory).encode(), len(type_names), syms))

    def compute_nene(self, sysm, epot=None, f=None):
        """Drop-in compute_nene on one universe; returns (epot [eV], f [eV/A], extrapolated)."""
        N, P = sysm.natoms, sysm.nbeads
        epot = np.zeros(P) if epot is None else epot
        f = np.zeros((N, P, 3)) if f is None else f
        idx = np.ascontiguousarray(sysm.idx + 1, dtype=np.int32)
        sb = np.ascontiguousarray(np.asarray(sysm.simbox, dtype=np.float64).T)
        isb = np.ascontiguousarray(np.linalg.inv(sysm.simbox).T)
        r = np.ascontiguousarray(sysm.r, dtype=np.float64)
        nex = C.c_int32(0)
        _check(lib().mdt_compute_nene(self.h, N, P, _ip(idx), _dp(sb), _dp(isb), _dp(r), _dp(epot), _dp(f),
                                      C.byref(nex), ENERGY_AND_FORCE))
        return epot, f, bool(nex.value)

    def compute_emt(self, sysm, epot=None, f=None, dens=None):
        """Drop-in compute_emt on one universe; returns (epot, f, dens) accumulated."""
        N, P = sysm.natoms, sysm.nbeads
        epot = np.zeros(P) if epot is None else epot
        f = np.zeros((N, P, 3)) if f is None else f
        dens = np.zeros((N, P)) if dens is None else dens
        idx = np.ascontiguousarray(sysm.idx + 1, dtype=np.int32)
        sb = np.ascontiguousarray(np.asarray(sysm.simbox, dtype=np.float64).T)
        isb = np.ascontiguousarray(np.linalg.inv(sysm.simbox).T)
        r = np.ascontiguousarray(sysm.r, dtype=np.float64)
        _check(lib().mdt_compute_emt(self.h, N, P, _ip(idx), _dp(sb), _dp(isb), _dp(r), _dp(epot), _dp(f),
                                     _dp(dens), ENERGY_AND_FORCE))
        return epot, f, dens


class Batch:
    """A batch of trajectories resident on the GPU (list of geometry.System)."""

    def __init__(self, ctx, systems, dt=0.1, tsurf=0.0, tproj=0.0, pile_tau=200.0, andersen_time=30.0,
                 seed=12345, traj_ids=None):
        self.ctx = ctx
        self.systems = systems
        self.dt = dt
        B = len(systems)
        self.N = np.array([s.natoms for s in systems], dtype=np.int32)
        self.P = np.array([s.nbeads for s in systems], dtype=np.int32)
        self.h = C.c_void_p()
        _check(lib().mdt_batch_create(ctx.h, B, _ip(self.N), _ip(self.P), C.byref(self.h)))
        idx = np.ascontiguousarray(np.concatenate([s.idx + 1 for s in systems]).astype(np.int32))
        mass = np.ascontiguousarray(np.concatenate([s.mass for s in systems]).astype(np.float64))
        fixed = np.ascontiguousarray(np.concatenate([s.fixed.reshape(-1) for s in systems]).astype(np.int32))
        _check(lib().mdt_batch_set_atoms(self.h, _ip(idx), _dp(mass), _ip(fixed)))
        sb = np.ascontiguousarray(np.concatenate([np.asarray(s.simbox).T.reshape(-1) for s in systems]))
        isb = np.ascontiguousarray(np.concatenate([np.linalg.inv(s.simbox).T.reshape(-1) for s in systems]))
        _check(lib().mdt_batch_set_cells(self.h, _dp(sb), _dp(isb)))
        s0 = systems[0]
        algo = np.ascontiguousarray(s0.algo, dtype=np.int32)
        isp = np.ascontiguousarray(s0.is_proj, dtype=np.int32)
        ts = np.ascontiguousarray(np.broadcast_to(np.asarray(tsurf, dtype=np.float64), (B,)))
        tp = np.ascontiguousarray(np.broadcast_to(np.asarray(tproj, dtype=np.float64), (B,)))
        _check(lib().mdt_batch_set_propagation(self.h, s0.ntypes, _ip(algo), _ip(isp), dt, _dp(ts), _dp(tp),
                                               pile_tau, andersen_time))
        tid = None
        if traj_ids is not None:
            tid = np.ascontiguousarray(traj_ids, dtype=np.uint32)
        _check(lib().mdt_batch_seed(self.h, seed, None if tid is None else tid.ctypes.data_as(
            C.POINTER(C.c_uint32))))
        self.n3 = int(sum(3 * s.natoms * s.nbeads for s in systems))
        self.npn = int(sum(s.natoms * s.nbeads for s in systems))
        self.np_ = int(sum(s.nbeads for s in systems))
        r = np.ascontiguousarray(np.concatenate([s.r.reshape(-1) for s in systems]))
        v = np.ascontiguousarray(np.concatenate([s.v.reshape(-1) for s in systems]))
        _check(lib().mdt_batch_set_state(self.h, _dp(r), _dp(v)))

    def set_state(self, rs, vs=None):
        """New positions (and velocities) per trajectory, (natoms, nbeads, 3)
        arrays in the batch's order; forces are stale until compute_forces."""
        r = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in rs]))
        if r.size != self.n3:
            raise ValueError("set_state: %d coordinates for a batch of %d" % (r.size, self.n3))
        v = None
        if vs is not None:
            v = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in vs]))
            if v.size != self.n3:
                raise ValueError("set_state: %d velocities for a batch of %d" % (v.size, self.n3))
        _check(lib().mdt_batch_set_state(self.h, _dp(r), None if v is None else _dp(v)))

    def close(self):
        if self.h:
            lib().mdt_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, tpa=0, nbcap=0):
        out = C.c_int(0)
        _check(lib().mdt_batch_tune(self.h, tpa, nbcap, C.byref(out)))
        return out.value

    def compute_forces(self):
        _check(lib().mdt_batch_compute_forces(self.h))

    def set_precision(self, mode):
        """0: FP64 pair terms (reference arithmetic); 1: FP32 pair terms where
        the EMT types' embedding stiffness keeps them within 1e-6 eV/atom and
        1e-5 eV/A (else only the cut-off exponential), FP64 accumulation."""
        _check(lib().mdt_batch_set_precision(self.h, int(mode)))

    def precision_arith(self):
        """Pair-term arithmetic of the next force launch: 0 FP64, 1 FP32 pair
        terms, 2 FP32 cut-off exponential and theta only."""
        out = C.c_int(0)
        _check(lib().mdt_batch_precision_arith(self.h, C.byref(out)))
        return out.value

    def set_wide(self, mode):
        """Workgroup width of the large size classes: -1 autotuned (default),
        0 always 512 lanes, 1 always 1024 where a class is large enough."""
        _check(lib().mdt_batch_set_wide(self.h, int(mode)))

    def set_keep(self, mode):
        """k_replica's lists: -1 autotuned (default), 0 rebuilt every launch,
        1 kept (Verlet skin) and staged in LDS, 2 kept and read from HBM."""
        _check(lib().mdt_batch_set_keep(self.h, int(mode)))

    def class_shapes(self):
        """Launch shape per size class: dicts with nmax, nrep, tpa, keep
        (lists kept across launches), block, nbcap."""
        n = C.c_int(0)
        _check(lib().mdt_batch_class_shapes(self.h, 0, None, C.byref(n)))
        buf = (C.c_int32 * (6 * max(n.value, 1)))()
        _check(lib().mdt_batch_class_shapes(self.h, n.value, buf, C.byref(n)))
        keys = ("nmax", "nrep", "tpa", "keep", "block", "nbcap")
        return [dict(zip(keys, buf[6 * k:6 * k + 6])) for k in range(n.value)]

    def set_bead_mode(self, mode):
        """0: every atom carries the beads; 1 (mixed): the projectile (atom 0)
        carries them and the other atoms are classical, kept in bead 0."""
        _check(lib().mdt_batch_set_bead_mode(self.h, int(mode)))

    def set_com_removal(self, on):
        """remove_com_velocity after propagate_1 with Andersen / PILE atoms:
        on (the reference) or off (thermostatted bare slabs)."""
        _check(lib().mdt_batch_set_com_removal(self.h, 1 if on else 0))

    def set_list_skin(self, skin):
        """Verlet skin (A) of the neighbour lists kept across force calls
        (mixed-model classical atoms; size classes above 256 atoms) in cells
        wider than 2 (cut-off + skin); 0 rebuilds every list at every call."""
        _check(lib().mdt_batch_set_list_skin(self.h, float(skin)))

    def set_observe(self, proj_ul=None, proj_ll=None, adsorption=None, exit_at_limits=True):
        """Trajectory observation and the exit at the projectile limits
        (`pul`, `pll`, `adsorption_distance start end`); None = the
        reference's default (no limit; adsorption thresholds at +inf)."""
        big = np.finfo(np.float64).max
        ul = big if proj_ul is None else float(proj_ul)
        ll = -big if proj_ll is None else float(proj_ll)
        a0, a1 = (big, big) if adsorption is None else (float(adsorption[0]), float(adsorption[1]))
        _check(lib().mdt_batch_set_observe(self.h, ul, ll, a0, a1, 1 if exit_at_limits else 0))

    def observables(self):
        """Per trajectory dict of mdt_traj_obs: cl_appr [A], cl_last [A],
        r_min (3,) [A], t_int [fs], turning_points, moving_down, adsorbed,
        exit_step (0 while running), cl_appr_step, last_step."""
        out = (TrajObs * len(self.systems))()
        _check(lib().mdt_batch_get_observables(self.h, out))
        return [dict(cl_appr=o.cl_appr, cl_last=o.cl_last, r_min=np.array(o.r_min[:]), t_int=o.t_int,
                     turning_points=o.turning_points, moving_down=bool(o.moving_down), adsorbed=bool(o.adsorbed),
                     adsorbed_prev=bool(o.adsorbed_prev), exit_step=o.exit_step, cl_appr_step=o.cl_appr_step,
                     last_step=o.last_step) for o in out]

    def estimators(self):
        """(ntraj, 20) estimators of the current state: the columns 2-21 of
        the reference's energy file (ESTIMATORS names them)."""
        out = np.zeros((len(self.systems), NEST))
        _check(lib().mdt_batch_estimators(self.h, _dp(out)))
        return out

    def set_pes(self, pes, judge_interval=0):
        """0: EMT; 1: the network (EMT fallback on extrapolation, EMT judge every judge_interval steps)."""
        _check(lib().mdt_batch_set_pes(self.h, int(pes), int(judge_interval)))

    def set_split(self, mode):
        """Two streams for one-bead, one-class EMT batches: -1 auto (>= 1024 trajectories), 0 off, 1 on."""
        _check(lib().mdt_batch_set_split(self.h, int(mode)))

    def last_split(self):
        """2 when the last step call ran as two halves on two streams, else 1."""
        n = C.c_int(0)
        _check(lib().mdt_batch_last_split(self.h, C.byref(n)))
        return n.value

    def set_judge_percent(self, percent):
        """Share of the replicas judged at each judge step (a divisor of 100; default 1)."""
        _check(lib().mdt_batch_set_judge_percent(self.h, int(percent)))

    def nn_stats(self):
        """(judge force RMSE [eV/A] or None, components judged, max |dF|, replicas on the fallback)."""
        sq, n, mx = C.c_double(0), C.c_double(0), C.c_double(0)
        nf = C.c_int32(0)
        _check(lib().mdt_batch_nn_stats(self.h, C.byref(sq), C.byref(n), C.byref(mx), C.byref(nf)))
        return (np.sqrt(sq.value / n.value) if n.value else None), n.value, mx.value, nf.value

    def nn_fallback(self):
        """Per trajectory (P,) bool: replicas that took the EMT fallback in the last force evaluation."""
        f = np.zeros(int(self.P.sum()), dtype=np.int32)
        _check(lib().mdt_batch_nn_fallback(self.h, _ip(f)))
        out, o = [], 0
        for p in self.P:
            out.append(f[o:o + p].astype(bool))
            o += p
        return out

    def nn_descriptors(self):
        """Raw symmetry functions of the last network evaluation: per trajectory (P, N, 64) float32."""
        maxn = int(self.N.max())
        nrep = int(self.P.sum())
        g = np.zeros((nrep, maxn, 64), dtype=np.float32)
        _check(lib().mdt_batch_nn_descriptors(self.h, g.ctypes.data_as(C.POINTER(C.c_float))))
        out, r = [], 0
        for s in self.systems:
            out.append(g[r:r + s.nbeads, :s.natoms].copy())
            r += s.nbeads
        return out

    def set_graph(self, chunk):
        """Replay the step loop as hipGraph chunks of `chunk` steps (0: direct launches)."""
        _check(lib().mdt_batch_set_graph(self.h, int(chunk)))

    def step(self, nsteps=1):
        _check(lib().mdt_batch_step(self.h, nsteps))

    def timing(self):
        ms = C.c_double(0)
        n = C.c_int(0)
        _check(lib().mdt_batch_last_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def profile(self, on=True):
        _check(lib().mdt_batch_profile(self.h, 1 if on else 0))

    def force_timing(self):
        ms = C.c_double(0)
        _check(lib().mdt_batch_force_timing(self.h, C.byref(ms)))
        return ms.value

    def status(self, clear=False):
        st = np.zeros(len(self.systems), dtype=np.int32)
        _check(lib().mdt_batch_status(self.h, _ip(st), 1 if clear else 0))
        return st

    def get_state(self):
        """Per-trajectory lists of dicts with r, v, f, a (N,P,3), epot (P,), dens (N,P)."""
        r, v, f, a = (np.zeros(self.n3) for _ in range(4))
        ep = np.zeros(self.np_)
        de = np.zeros(self.npn)
        _check(lib().mdt_batch_get_state(self.h, _dp(r), _dp(v), _dp(f), _dp(a), _dp(ep), _dp(de)))
        out = []
        o3 = op = opn = 0
        for s in self.systems:
            N, P = s.natoms, s.nbeads
            n3 = 3 * N * P
            out.append(dict(r=r[o3:o3 + n3].reshape(N, P, 3), v=v[o3:o3 + n3].reshape(N, P, 3),
                            f=f[o3:o3 + n3].reshape(N, P, 3), a=a[o3:o3 + n3].reshape(N, P, 3),
                            epot=ep[op:op + P].copy(), dens=de[opn:opn + N * P].reshape(N, P)))
            o3 += n3
            op += P
            opn += N * P
        return out


class Trainer:
    """GPU training set and optimiser for the surrogate held by a Context.

    frames: list of geometry.System (one bead) with reference energies
    E [eV] and forces F [eV/A] (N, 3)."""

    def __init__(self, ctx, frames, E, F, batch_frames=64, val_fraction=0.05, seed=1):
        nat = np.array([s.natoms for s in frames], dtype=np.int32)
        idx = np.concatenate([s.idx + 1 for s in frames])
        sb = np.concatenate([np.asarray(s.simbox).T.reshape(-1) for s in frames])
        r = np.concatenate([s.r[:, 0, :].reshape(-1) for s in frames])
        f = np.concatenate([np.asarray(x).reshape(-1) for x in F])
        self._create(ctx, nat, idx, sb, r, E, f, batch_frames, val_fraction, seed)

    @classmethod
    def from_arrays(cls, ctx, natoms, idx, simbox_colmajor, r, E, F, batch_frames=64, val_fraction=0.05, seed=1):
        """natoms (nf,), idx (sum N, 0-based types), simbox (nf, 9) column-major [A],
        r and F (sum N, 3) [A, eV/A], E (nf,) [eV]."""
        self = cls.__new__(cls)
        self._create(ctx, natoms, np.asarray(idx) + 1, simbox_colmajor, r, E, F, batch_frames, val_fraction, seed)
        return self

    def _create(self, ctx, nat, idx, sb, r, E, f, batch_frames, val_fraction, seed):
        self.ctx = ctx
        nat = np.ascontiguousarray(nat, dtype=np.int32)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        sb = np.ascontiguousarray(sb, dtype=np.float64).reshape(-1)
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        e = np.ascontiguousarray(E, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        self.h = C.c_void_p()
        _check(lib().mdt_nn_train_create(ctx.h, len(nat), _ip(nat), _ip(idx), _dp(sb), _dp(r), _dp(e), _dp(f),
                                         batch_frames, val_fraction, seed, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().mdt_nn_train_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def descriptor_stats(self, nelem):
        mn, mx, av = (np.zeros(nelem * 64) for _ in range(3))
        _check(lib().mdt_nn_train_descriptor_stats(self.h, _dp(mn), _dp(mx), _dp(av)))
        return mn.reshape(nelem, 64), mx.reshape(nelem, 64), av.reshape(nelem, 64)

    def epoch(self, lr=1e-3, wE=1.0, wF=1.0):
        loss = C.c_double(0)
        _check(lib().mdt_nn_train_epoch(self.h, lr, wE, wF, C.byref(loss)))
        return loss.value

    def eval(self, validation=False):
        """(energy RMSE [eV/atom], force RMSE [eV/A])."""
        e, f = C.c_double(0), C.c_double(0)
        _check(lib().mdt_nn_train_eval(self.h, 1 if validation else 0, C.byref(e), C.byref(f)))
        return e.value, f.value

    def loss_grad(self, frames, nweights, wE=1.0, wF=1.0):
        fr = np.ascontiguousarray(frames, dtype=np.int32)
        g = np.zeros(nweights)
        L = C.c_double(0)
        _check(lib().mdt_nn_train_loss_grad(self.h, len(fr), _ip(fr), wE, wF, C.byref(L), _dp(g)))
        return L.value, g


def nn_weights(ctx, model):
    """Current network weights split per element (weights.XXX.data layout)."""
    n = sum(model.nweights(z) for z in model.elements)
    w = np.zeros(n)
    _check(lib().mdt_nn_get_weights(ctx.h, _dp(w)))
    out, o = {}, 0
    for z in model.elements:
        out[z] = w[o:o + model.nweights(z)].copy()
        o += model.nweights(z)
    return out
