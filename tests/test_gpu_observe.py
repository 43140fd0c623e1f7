"""Device-side trajectory observation, the exit at the projectile limits and
the energy-file estimators against the oracle.

The oracle runs md_tian2's loop for each trajectory (propagate, exit test,
collect_trajectory_characteristics, the final collection) with the same
Philox noise; the GPU batch runs all trajectories at once with observation
on.  Scattering H on Au(111) at 2-6 eV: most projectiles come back through
`pul` and are frozen at their exit step, the others run to the end.
Bars: exit steps, turning points, adsorption and the step of the closest
approach exactly; positions, closest approach and lowest point within
1e-7 A after several hundred steps (chaotic growth from ~1e-15); the
printed closest-approach time and interaction time exactly (the same
arithmetic on the same step numbers).  Estimators against the oracle's
restatement on the same state: 1e-10 relative."""

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, mdt, observables, pesfile

pytestmark = pytest.mark.gpu

SEED = 777
DT = 0.1
Z0 = 5.0
PUL = Z0 + 0.3
ADS = (1.9, 2.6)


def _systems(rng, n, nbeads=1):
    out = []
    for _ in range(n):
        out.append(geometry.h_on_slab("Au", n=3, nlayers=4, rng=rng, nbeads=nbeads, einc=rng.uniform(2.0, 6.0),
                                      polar_deg=rng.uniform(0, 45), azimuth_deg=rng.uniform(0, 360), z0=Z0))
    return out


def _oracle_runs(systems, cols, pesm, nsteps, tsurf, tproj):
    refs = []
    for t, s in enumerate(systems):
        st = oracle.OracleState(s, cols, pesm, step=DT, tsurf=tsurf, tproj=tproj)
        st.calc_force()
        o, iend = st.run_scatter(SEED, t, nsteps, proj_ul=PUL, adsorption=ADS)
        refs.append((st, o, iend))
    return refs


def _gpu_run(ctx, systems, nsteps, tsurf, tproj, graph=0):
    b = mdt.Batch(ctx, systems, dt=DT, tsurf=tsurf, tproj=tproj, seed=SEED)
    if graph:
        b.set_graph(graph)
    b.compute_forces()
    b.set_observe(proj_ul=PUL, adsorption=ADS)
    b.step(nsteps)
    return b


def _compare(b, refs, nsteps):
    obs = b.observables()
    state = b.get_state()
    n_exit = 0
    for t, (st, o, iend) in enumerate(refs):
        fin = observables.finalize(obs[t], state[t]["f"][0], DT, nsteps, ADS)
        assert fin["istep"] == iend, (t, fin["istep"], iend)
        n_exit += obs[t]["exit_step"] > 0
        assert np.abs(state[t]["r"] - st.r).max() <= 1e-7, t
        assert np.abs(state[t]["v"] - st.v).max() <= 1e-7, t
        assert abs(fin["cl_appr"] - o.closest_approach) <= 1e-7
        assert fin["cl_appr_time"] == o.closest_approach_time, (t, fin["cl_appr_time"], o.closest_approach_time)
        assert np.abs(fin["r_min_p"] - np.array(o.lowest_z[:])).max() <= 1e-7
        assert fin["t_int"] == o.interaction_time, (t, fin["t_int"], o.interaction_time)
        assert fin["turn_pnts"] == o.nturning_points
        assert fin["adsorbed"] == bool(o.is_adsorbed)
    return n_exit


def test_scatter_observation_and_exit_match_the_oracle(gpu_ctx, pes_hau):
    rng = np.random.default_rng(11)
    systems = _systems(rng, 8)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    nsteps = 700
    refs = _oracle_runs(systems, cols, pesm, nsteps, 300.0, 0.0)
    b = _gpu_run(gpu_ctx, systems, nsteps, 300.0, 0.0)
    n_exit = _compare(b, refs, nsteps)
    print(f"scatter observation: {n_exit} of {len(systems)} trajectories exited through pul")
    assert n_exit > 0
    # frozen trajectories stay as they left; the others keep moving
    before = b.get_state()
    ex = [o["exit_step"] for o in b.observables()]
    b.step(30)
    after = b.get_state()
    for t in range(len(systems)):
        same = all(np.array_equal(before[t][k], after[t][k]) for k in ("r", "v", "f", "epot"))
        assert same == (ex[t] > 0), t
    for t, o in enumerate(b.observables()):   # exits stay; running ones may leave in the extra steps
        assert o["exit_step"] == ex[t] if ex[t] else o["exit_step"] in (0,) + tuple(range(nsteps + 1, nsteps + 31))
    b.close()
    # graph replays observe and freeze identically
    bg = _gpu_run(gpu_ctx, systems, nsteps, 300.0, 0.0, graph=16)
    og, sg = bg.observables(), bg.get_state()
    b2 = _gpu_run(gpu_ctx, systems, nsteps, 300.0, 0.0)
    o2, s2 = b2.observables(), b2.get_state()
    for t in range(len(systems)):
        assert all(np.array_equal(og[t][k], o2[t][k]) for k in og[t]), t
        assert np.array_equal(sg[t]["r"], s2[t]["r"]), t
    bg.close()
    b2.close()


def test_rpmd_observation_matches_the_oracle(gpu_ctx, pes_hau):
    rng = np.random.default_rng(12)
    systems = _systems(rng, 3, nbeads=4)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    nsteps = 500
    refs = _oracle_runs(systems, cols, pesm, nsteps, 300.0, 300.0)
    b = _gpu_run(gpu_ctx, systems, nsteps, 300.0, 300.0)
    _compare(b, refs, nsteps)
    b.close()


@pytest.mark.parametrize("nbeads,mixed", [(1, False), (4, False), (4, True)])
def test_estimators_match_the_oracle(gpu_ctx, pes_hau, nbeads, mixed):
    rng = np.random.default_rng(13 + nbeads)
    systems = _systems(rng, 3, nbeads=nbeads)
    for s in systems:
        s.r[0] += rng.normal(0, 0.05, s.r[0].shape)
        s.v[0] += rng.normal(0, 0.01, s.v[0].shape)
        s.r[0, :, 2] = 2.0 + rng.normal(0, 0.05, nbeads)   # inside the interaction range
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=DT, tsurf=250.0, tproj=120.0, seed=SEED)
    if mixed:
        b.set_bead_mode(1)
    b.compute_forces()
    b.step(5)
    est = b.estimators()
    state = b.get_state()
    for t, s in enumerate(systems):
        st = oracle.OracleState(s, cols, pesm, step=DT, tsurf=250.0, tproj=120.0)
        st.r[...] = state[t]["r"]
        st.v[...] = state[t]["v"]
        st.f[...] = state[t]["f"]
        st.epot[...] = state[t]["epot"]
        ref = st.energy_row()
        scale = np.maximum(np.abs(ref), 1e-12)
        assert (np.abs(est[t] - ref) <= 1e-10 * scale + 1e-13).all(), (t, est[t] - ref)
    b.close()


def test_output_files(gpu_ctx, pes_hau, tmp_path):
    """observables.run writes the reference's scatter, energy and adsorption
    files; their numbers equal the oracle's run within the bars above."""
    rng = np.random.default_rng(14)
    systems = _systems(rng, 4)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    nsteps = 600
    refs = _oracle_runs(systems, cols, pesm, nsteps, 300.0, 0.0)
    b = mdt.Batch(gpu_ctx, systems, dt=DT, tsurf=300.0, tproj=0.0, seed=SEED, traj_ids=list(range(4)))
    out = observables.Output(str(tmp_path), [101, 102, 103, 104], scatter=True, energy=50, adsorption=25,
                             runner=100)
    fins = observables.run(b, nsteps, out, proj_ul=PUL, adsorption=ADS)
    b.close()
    for t, (st, o, iend) in enumerate(refs):
        assert fins[t]["istep"] == iend
        lines = open(out.path("scatter", 101 + t)).read().splitlines()
        assert len(lines) == 23
        kv = {ln.split("=")[0].strip(): ln.split("=")[1].split() for ln in lines if "=" in ln}
        assert int(kv["turn_pnts"][0]) == o.nturning_points
        assert abs(float(kv["cl_appr"][0]) - o.closest_approach) <= 1e-6
        assert float(kv["time"][0]) == pytest.approx((iend - 1) * DT, abs=1e-7)
        rf = st.r[0].mean(axis=0)
        assert np.abs(np.array([float(x) for x in kv["r_f"]]) - rf).max() <= 1e-6
        en = open(out.path("energy", 101 + t)).read().splitlines()
        en = [ln for ln in en if not ln.split()[0].startswith("#")]
        assert len(en) == min(nsteps, iend) // 50
        ads = open(out.path("ads", 101 + t)).read().splitlines()
        assert len(ads) == min(nsteps, iend) // 25
    from md_tian2_amd import nene
    frames = nene.read_input_data(out.path("runner", 0))
    counts = [min(nsteps, iend) // 100 for (_, _, iend) in refs]
    assert len(frames) == sum(counts)
    k = 0
    for t, (st, o, iend) in enumerate(refs):
        k += counts[t]
        if iend > nsteps:   # the trajectory's last frame is its final state
            fr = frames[k - 1]
            assert np.abs(fr["xyz"] * observables.BOHR2ANG - st.r[:, 0, :]).max() <= 2e-7
            assert abs(fr["energy"] * observables.HA2EV - st.epot[0]) <= 1e-6
    head = open(out.path("energy", 101)).read().splitlines()[:3]
    assert head[0] == "# All energies are given in eV." and head[2].split()[0] == "#time/fs"
    assert not open(out.path("energy", 102)).read().startswith("#")


def test_exit_on_the_network_potential(gpu_ctx, pes_hau):
    """With the NENE surrogate as the potential, a projectile flying up out
    of the interaction range leaves through pul at the step straight flight
    predicts and stays frozen; a trajectory at the surface keeps running."""
    from md_tian2_amd import surrogate
    model = surrogate.load_model(surrogate.MODEL_DIR)
    assert model is not None
    rng = np.random.default_rng(15)
    systems = _systems(rng, 2)
    fly = systems[0]
    fly.r[0, 0, 2] = 12.0                 # beyond the network's 6.5 A cut-off
    fly.v[0, 0] = [0.0, 0.0, 0.0123]
    stay = systems[1]
    stay.r[0, 0, 2] = 1.0                 # bound at the surface, at rest
    stay.v[0, 0] = 0.0
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    gpu_ctx.set_nn(model, systems[0].type_names)
    b = mdt.Batch(gpu_ctx, systems, dt=DT, tsurf=300.0, tproj=0.0, seed=SEED)
    b.set_pes(1)
    b.compute_forces()
    b.set_observe(proj_ul=13.0)
    b.step(900)
    obs = b.observables()
    want = int(np.floor(1.0 / (0.0123 * DT))) + 1
    assert obs[0]["exit_step"] == want, obs[0]
    assert obs[1]["exit_step"] == 0
    before = b.get_state()
    b.step(20)
    after = b.get_state()
    assert np.array_equal(before[0]["r"], after[0]["r"]) and np.array_equal(before[0]["f"], after[0]["f"])
    assert not np.array_equal(before[1]["r"], after[1]["r"])
    assert not (b.status() & ~16).any()
    b.close()


def test_observation_in_the_mixed_bead_model(gpu_ctx, pes_hau):
    """Mixed bead model (beads on the projectile only): the device's
    characteristics equal a host recomputation from the state after every
    step (centroids of the projectile's beads, classical slab atoms)."""
    rng = np.random.default_rng(16)
    systems = _systems(rng, 3, nbeads=8)
    cols, pesm = pesfile.emt_tables(pes_hau, systems[0].type_names)
    gpu_ctx.set_emt(cols, pesm)
    b = mdt.Batch(gpu_ctx, systems, dt=DT, tsurf=300.0, tproj=300.0, seed=SEED)
    b.set_bead_mode(1)
    b.compute_forces()
    b.set_observe(proj_ul=PUL, adsorption=ADS)
    big = np.finfo(np.float64).max
    ref = [dict(cl=big, cls=0, rmin=np.full(3, big), tint=0.0, tp=0, down=True, ads=False, exit=0)
           for _ in systems]
    for istep in range(1, 121):
        b.step(1)
        st = b.get_state()
        for t, s in enumerate(systems):
            o = ref[t]
            if o["exit"]:
                continue
            c = st[t]["r"].mean(axis=1)                      # centroids (classical beads are copies)
            ic = np.linalg.inv(s.simbox)
            d = (c[0] - c[1:]) @ ic.T
            d -= np.round(d)
            dist = np.sqrt(((d @ s.simbox.T) ** 2).sum(-1)).min()
            if dist < o["cl"]:
                o["cl"], o["cls"] = dist, istep
            if not o["ads"] and dist < ADS[0]:
                o["ads"] = True
            if o["ads"] and dist > ADS[1]:
                o["ads"] = False
            if c[0, 2] < o["rmin"][2]:
                o["rmin"] = c[0].copy()
            if np.any(st[t]["f"][0] != 0.0):
                o["tint"] += DT
            down = st[t]["v"][0, :, 2].mean() < 0
            if down != o["down"]:
                o["tp"] += 1
                o["down"] = down
            if (st[t]["r"][0, :, 2] > PUL).all():
                o["exit"] = istep
    obs = b.observables()
    for t, o in enumerate(ref):
        g = obs[t]
        assert g["exit_step"] == o["exit"] and g["cl_appr_step"] == o["cls"] and g["turning_points"] == o["tp"]
        assert abs(g["cl_appr"] - o["cl"]) <= 1e-12 and np.abs(g["r_min"] - o["rmin"]).max() <= 1e-12
        assert abs(g["t_int"] - o["tint"]) <= 1e-12 and g["adsorbed"] == o["ads"]
    b.close()
