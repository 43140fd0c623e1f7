"""Per-trajectory observables and md_tian2's output files for a batch.

The device observes every trajectory after each step (the characteristics of
src/trajectory_info.f90:47-263: closest approach, lowest projectile
position, interaction time, turning points, adsorption) and freezes a
trajectory whose projectile leaves through `pul` / `pll`
(src/md_tian2.f90:88-92); it also evaluates the energy-file estimators
(k_estimators).  This module turns those into the reference's files:

* traj/mxt_fin%08d.dat  -- output_scatter (src/output_mod.f90:635-730),
  initial block after the first calc_force, final block after the loop;
* energy/traj_%08d.dat  -- output_nrg (src/output_mod.f90:170-229) every
  `interval` steps;
* conf/ads_%08d.dat     -- output_adsorption_status (:792-818);
* runner/input.data     -- output_runner (:550-609): RuNNer training
  frames (bead centroids in Bohr, centroid forces in Ha/Bohr, the
  bead-averaged energy in Ha), one file for the run in trajectory order.

Numbers are written with the reference's Fortran edit descriptors (Fw.d,
Ew.d, Ew.dEe, I, L with ifort's default widths for I and L: 12 and 2).

Reference behaviours kept on purpose (they change the printed numbers):
* `time` of the final block is (istep - 1) * step with istep the exit step,
  or nsteps + 1 after a full run;
* after the loop the characteristics are collected once more with that
  istep (md_tian2.f90:100): after a full run this adds one more step to the
  interaction time and applies the adsorption thresholds once more;
* `cl_appr_time` is multiplied by `step` on every collection, not only when
  the closest approach is updated (trajectory_info.f90:111), so the printed
  value is cl_appr_step * step ** (1 + collections since the update);
* the energy file header is written only into the run's first trajectory
  file (overwrite_nrg, output_mod.f90:36, 189-197); every later file is
  appended to, as are the adsorption files after the first.
One difference: the adsorption status starts False for every trajectory
(the reference's is_adsorbed, trajectory_info.f90:39, is never reset and
carries over from the previous trajectory of the same process).
"""
from __future__ import annotations

import math
import os

import numpy as np

RAD2DEG = 180.0 / math.acos(-1.0)  # constants.f90:36, 159
BOHR2ANG = 0.529177211               # constants.f90:160-168
ANG2BOHR = 1.0 / BOHR2ANG
HA2EV = 27.21138602
EV2HA = 1.0 / HA2EV
EVANG2HABOHR = 1.0 / (HA2EV * ANG2BOHR)
BIG = np.finfo(np.float64).max      # huge(0_dp)

ENERGY_LABELS = ("#time/fs", "atom T/K", "bead T/K", "r_gyr_sq_p/A²", "r_gyr_sq_l/A²", "atom ekin proj",
                 "atom ekin latt", "bead ekin proj", "bead ekin latt", "qntm ekin proj", "qntm ekin latt",
                 "atom epot", "bead epot", "e_total", "f_total/(eV/A)", "r_x", "r_y", "r_z", "v_x", "v_y", "v_z")


# ---------------------------------------------------------------- formats
def _nonfinite(x, w):
    s = "NaN" if math.isnan(x) else ("Infinity" if x > 0 else "-Infinity")
    return s.rjust(w) if len(s) <= w else "*" * w


def ffmt(x, w, d):
    """Fortran Fw.d."""
    x = float(x)
    if not math.isfinite(x):
        return _nonfinite(x, w)
    s = "%.*f" % (d, x)
    return s.rjust(w) if len(s) <= w else "*" * w


def efmt(x, w, d, e=None):
    """Fortran Ew.d / Ew.dEe: [-]0.ddddE+xx, mantissa in [0.1, 1)."""
    x = float(x)
    if not math.isfinite(x):
        return _nonfinite(x, w)
    if x == 0.0:
        digits, ex = "0" * d, 0
    else:
        m = "%.*e" % (d - 1, abs(x))
        mant, exp10 = m.split("e")
        digits, ex = mant.replace(".", ""), int(exp10) + 1
    ne = 2 if e is None else e
    if e is None and abs(ex) > 99:
        if abs(ex) > 999:
            return "*" * w
        es = "%+04d" % ex                    # the E is dropped for three digits
    else:
        if abs(ex) >= 10 ** ne:
            return "*" * w
        es = "E" + ("+" if ex >= 0 else "-") + str(abs(ex)).zfill(ne)
    s = ("-" if x < 0 else "") + "0." + digits + es
    return s.rjust(w) if len(s) <= w else "*" * w


def ifmt(n, w=12):
    s = str(int(n))
    return s.rjust(w) if len(s) <= w else "*" * w


def lfmt(b, w=2):
    return ("T" if b else "F").rjust(w)


# --------------------------------------------------- final collection
def finalize(obs, f_proj, dt, nsteps, adsorption=None):
    """The characteristics as the reference prints them after the loop.

    obs: one dict of Batch.observables(); f_proj: (P, 3) projectile forces
    of the final state; nsteps: steps run since observation started.
    Returns istep (exit step or nsteps + 1), time, turn_pnts, cl_appr,
    cl_appr_time (the reference's figure), r_min_p, t_int, adsorbed."""
    a0, a1 = (BIG, BIG) if adsorption is None else adsorption
    t_int = obs["t_int"]
    ads = obs["adsorbed"]
    if obs["exit_step"]:
        istep = obs["exit_step"]                      # the exit step was collected once
    else:
        istep = nsteps + 1                            # collect_trajectory_characteristics(.., nsteps+1)
        if np.any(np.asarray(f_proj) != 0.0):
            t_int = t_int + dt
        here = obs["cl_last"]
        if not ads and here < a0:
            ads = True
        if ads and here > a1:
            ads = False
    cat = 0.0
    if obs["cl_appr_step"] > 0:
        cat = float(obs["cl_appr_step"]) * dt
        for _ in range(istep - obs["cl_appr_step"]):
            cat = cat * dt
            if cat == 0.0 or not math.isfinite(cat):
                break
    return dict(istep=istep, time=(istep - 1) * dt, turn_pnts=obs["turning_points"], cl_appr=obs["cl_appr"],
                cl_appr_time=cat, r_min_p=np.asarray(obs["r_min"]), t_int=t_int, adsorbed=ads)


# ------------------------------------------------------------- writers
def _polar_azi(v):
    vn = math.sqrt(float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
    pol = math.acos(v[2] / vn) * RAD2DEG if vn > 0 else float("nan")
    return pol, math.atan2(v[1], v[0]) * RAD2DEG


def scatter_initial_text(est):
    """The initial block of traj/mxt_fin%08d.dat from one estimator row."""
    pol, azi = _polar_azi(est[17:20])
    lines = ["ekin_p_i = " + ffmt(est[4], 14, 7),
             "ekin_l_i = " + ffmt(est[5], 14, 7),
             "epot_i   = " + efmt(est[10], 21, 10),
             "etotal_i = " + efmt(est[12], 21, 10),
             "r_i      = " + "".join(ffmt(x, 14, 7) for x in est[14:17]),
             "v_i      = " + "".join(ffmt(x, 14, 7) for x in est[17:20]),
             "polar_i  = " + ffmt(180.0 - pol, 14, 7),
             "azi_i    = " + ffmt(azi, 14, 7),
             ""]
    return "\n".join(lines) + "\n"


def scatter_final_text(est, fin):
    pol, azi = _polar_azi(est[17:20])
    lines = ["ekin_p_f     = " + ffmt(est[4], 14, 7),
             "ekin_l_f     = " + ffmt(est[5], 14, 7),
             "epot_f       = " + efmt(est[10], 21, 10),
             "etotal_f     = " + efmt(est[12], 21, 10),
             "r_f          = " + "".join(ffmt(x, 14, 7) for x in est[14:17]),
             "v_f          = " + "".join(ffmt(x, 14, 7) for x in est[17:20]),
             "polar_f      = " + ffmt(pol, 14, 7),
             "azi_f        = " + ffmt(azi, 14, 7),
             "",
             "time         = " + ffmt(fin["time"], 14, 7),
             "turn_pnts    = " + ifmt(fin["turn_pnts"]),
             "cl_appr      = " + ffmt(fin["cl_appr"], 14, 7),
             "cl_appr_time = " + ffmt(fin["cl_appr_time"], 14, 7),
             "r_min_p      = " + "".join(ffmt(x, 14, 7) for x in fin["r_min_p"])]
    return "\n".join(lines) + "\n"


def energy_header():
    """The three header lines of energy/traj_%08d.dat ('(21a17)' pads bytes)."""
    def a17(s):
        b = s.encode("utf-8")
        return (b" " * max(0, 17 - len(b)) + b[:17]).decode("utf-8", errors="replace")
    return ("# All energies are given in eV.\n" + "#" + "%8d" % 1 + "".join("%17d" % i for i in range(2, 22))
            + "\n" + "".join(a17(s) for s in ENERGY_LABELS) + "\n")


def energy_line(istep, dt, est):
    return "".join(efmt(x, 17, 8, 2) for x in [istep * dt] + list(est)) + "\n"


def adsorption_line(istep, dt, adsorbed):
    return efmt((istep - 1) * dt, 13, 6, 2) + lfmt(adsorbed) + "\n"


def runner_header(now=None):
    """The two comment lines output_runner starts input.data with."""
    import datetime
    t = now or datetime.datetime.now()
    return "# This file was generated by MDT2.\n" + "# %4d-%02d-%02d - %02d.%02d\n" % (t.year, t.month, t.day,
                                                                                    t.hour, t.minute)


def runner_frame_text(simbox, symbols, r, f, epot):
    """One RuNNer structure: simbox (3, 3) with the cell vectors as columns
    [A], symbols per atom, r and f (N, P, 3) [A, eV/A], epot (P,) [eV]."""
    lines = ["begin"]
    for k in range(3):
        lines.append("lattice" + "".join(" " + ffmt(x * ANG2BOHR, 11, 8) for x in np.asarray(simbox)[:, k]))
    P = r.shape[1]
    cr = r.sum(axis=1) / P if P > 1 else r[:, 0, :]
    cf = f.sum(axis=1) / P if P > 1 else f[:, 0, :]
    for j in range(r.shape[0]):
        lines.append("atom" + "".join(" " + ffmt(x * ANG2BOHR, 16, 9) for x in cr[j]) + " " + symbols[j][:3].ljust(3)
                     + " " + ffmt(0.0, 11, 8) + " " + ffmt(0.0, 11, 8)
                     + "".join(" " + ffmt(x * EVANG2HABOHR, 11, 8) for x in cf[j]))
    lines.append("charge " + ffmt(0.0, 11, 8))
    lines.append("energy " + ffmt(float(np.sum(epot)) / len(epot) * EV2HA, 11, 8))
    lines.append("end")
    return "\n".join(lines) + "\n"


class Output:
    """The reference's output folders for trajectories traj_ids (md_tian.inp
    `output scatter 1 energy 10 adsorption_status 1`; intervals 0 = off)."""

    def __init__(self, folder, traj_ids, scatter=True, energy=0, adsorption=0, runner=0):
        self.folder = folder
        self.traj_ids = list(traj_ids)
        self.scatter, self.energy, self.adsorption = bool(scatter), int(energy), int(adsorption)
        self.runner = int(runner)
        self.first = min(self.traj_ids) if self.traj_ids else 0
        self.frames = {t: [] for t in self.traj_ids}   # runner frames, written in trajectory order
        for sub, on in (("traj", self.scatter), ("energy", self.energy), ("conf", self.adsorption),
                        ("runner", self.runner)):
            if on:
                os.makedirs(os.path.join(folder, sub), exist_ok=True)
        self._opened = set()

    def path(self, kind, tid):
        if kind == "runner":
            return os.path.join(self.folder, "runner", "input.data")
        name = {"scatter": "traj/mxt_fin%08d.dat", "energy": "energy/traj_%08d.dat",
                "ads": "conf/ads_%08d.dat"}[kind] % tid
        return os.path.join(self.folder, name)

    def _append(self, kind, tid, text, header=""):
        p = self.path(kind, tid)
        key = (kind, tid)
        if key not in self._opened:
            self._opened.add(key)
            if tid == self.first:   # overwrite_nrg / overwrite_ads: only the first trajectory's file is new
                with open(p, "w", encoding="utf-8") as f:
                    f.write(header + text)
                return
        with open(p, "a", encoding="utf-8") as f:
            f.write(text)


def run(batch, nsteps, out, proj_ul=None, proj_ll=None, adsorption=None):
    """md_tian2's trajectory loop (src/md_tian2.f90:68-118) for every
    trajectory of `batch` at once, writing `out`'s files.  The batch holds
    positions and velocities; forces are computed here.  Returns the final
    characteristics (one finalize() dict per trajectory)."""
    dt = batch.dt
    batch.compute_forces()
    B = len(batch.systems)
    est = batch.estimators()
    if out.scatter:
        for t in range(B):
            with open(out.path("scatter", out.traj_ids[t]), "w") as f:
                f.write(scatter_initial_text(est[t]))
    proj_first = all(s.type_names[s.idx[0]][1] for s in batch.systems)
    batch.set_observe(proj_ul, proj_ll, adsorption, exit_at_limits=out.scatter and proj_first)
    intervals = [i for i in (out.energy, out.adsorption, out.runner) if i > 0]
    cur = 0
    while cur < nsteps:
        nxt = nsteps
        for iv in intervals:
            nxt = min(nxt, (cur // iv + 1) * iv)
        batch.step(nxt - cur)
        cur = nxt
        due_e = out.energy and cur % out.energy == 0
        due_a = out.adsorption and cur % out.adsorption == 0
        due_r = out.runner and cur % out.runner == 0
        obs = batch.observables()
        if due_e or due_a or due_r:
            est = batch.estimators() if due_e else None
            state = batch.get_state() if due_r else None
            for t in range(B):
                ex = obs[t]["exit_step"]
                if ex and ex < cur:
                    continue
                tid = out.traj_ids[t]
                if due_e:
                    out._append("energy", tid, energy_line(cur, dt, est[t]), energy_header())
                if due_a:
                    out._append("ads", tid, adsorption_line(cur, dt, obs[t]["adsorbed_prev"]))
                if due_r:
                    s = batch.systems[t]
                    out.frames[tid].append(runner_frame_text(s.simbox, [s.type_names[k][0] for k in s.idx],
                                                             state[t]["r"], state[t]["f"], state[t]["epot"]))
        if all(o["exit_step"] for o in obs):
            break
    obs = batch.observables()
    state = batch.get_state()
    est = batch.estimators()
    fins = []
    for t in range(B):
        fin = finalize(obs[t], state[t]["f"][0], dt, nsteps, adsorption)
        fins.append(fin)
        if out.scatter:
            with open(out.path("scatter", out.traj_ids[t]), "a") as f:
                f.write(scatter_final_text(est[t], fin))
    if out.runner:
        with open(out.path("runner", 0), "w") as f:
            f.write(runner_header())
            for tid in sorted(out.traj_ids):
                f.writelines(out.frames[tid])
    return fins
