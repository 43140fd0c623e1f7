"""GPU run of the Fortran batched driver (mdt_run) on an md_tian.inp-style
deck, checked trajectory by trajectory against the oracle started from the
driver's dumped initial state with the same noise keys (seed, trajectory
id)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from md_tian2_amd import geometry, pesfile

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_dump(path):
    raw = open(path, "rb").read()
    ntraj, natoms, nbeads, nsteps = np.frombuffer(raw[:16], dtype=np.int32)
    off = 16
    tid = np.frombuffer(raw[off:off + 4 * ntraj], dtype=np.int32)
    off += 4 * ntraj
    n3 = 3 * natoms * nbeads
    arrs = []
    for count in (ntraj * n3, ntraj * n3, ntraj * n3, ntraj * n3, ntraj * nbeads, ntraj * 9):
        arrs.append(np.frombuffer(raw[off:off + 8 * count], dtype=np.float64))
        off += 8 * count
    r0, v0, r1, v1, ep, cell = arrs
    shp = (ntraj, natoms, nbeads, 3)
    return dict(ntraj=ntraj, natoms=natoms, nbeads=nbeads, nsteps=nsteps, tid=tid, r0=r0.reshape(shp),
                v0=v0.reshape(shp), r1=r1.reshape(shp), v1=v1.reshape(shp), epot=ep.reshape(ntraj, nbeads),
                cell=cell.reshape(ntraj, 3, 3).transpose(0, 2, 1))


def test_fortran_driver_matches_oracle(pes_hau):
    os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")], cwd=ROOT,
                         capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    d = read_dump(os.path.join(ROOT, "gpurun_out", "mdt_run_h_au.bin"))
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, thermal=False)
    assert np.allclose(d["cell"][0], tmpl.simbox, atol=1e-12)
    cols, pesm = pesfile.emt_tables(pes_hau, tmpl.type_names)
    for t in range(d["ntraj"]):
        s = tmpl.copy()
        s.r = d["r0"][t].copy()
        s.v = d["v0"][t].copy()
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=300.0, tproj=0.0)
        st.calc_force()
        st.run(777, int(d["tid"][t]), 1, int(d["nsteps"]))
        assert np.abs(d["r1"][t] - st.r).max() < 1e-9
        assert np.abs(d["v1"][t] - st.v).max() < 1e-9
        assert abs(d["epot"][t, 0] - st.epot[0]) < 1e-8


def test_fortran_nene_dropin_matches_oracle():
    """read_nene_gpu / compute_nene_gpu (Fortran host, library's RuNNer
    reader) on the reference's carbon network and its structures."""
    import nene_oracle as no
    from md_tian2_amd import nene

    fix = os.path.join(ROOT, "tests", "golden", "runner_c")
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "nene_predict")
    res = subprocess.run([exe, fix, os.path.join(fix, "input.data"), "3"], cwd=ROOT, capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    energies, forces = [], []
    for line in res.stdout.splitlines():
        w = line.split()
        if w and w[0] == "energy":
            energies.append(float(w[1]))
            forces.append([])
        elif w and w[0] == "force":
            forces[-1].append([float(x) for x in w[2:5]])
    om = no.Model(nene.load(fix))
    structs = nene.read_input_data(os.path.join(fix, "input.data"), limit=3)
    assert len(energies) == 3
    for e_g, f_g, s in zip(energies, forces, structs):
        el = om.elem_index(s["symbols"])
        E, _, F, _ = om.energy_forces(el, s["xyz"], s["lattice"])
        assert abs(e_g - E * nene.HA2EV) / len(el) <= 2e-4
        assert np.abs(np.array(f_g) - F * nene.HABOHR2EVANG).max() <= 5e-3


@pytest.mark.parametrize("net", ["hau", "hau_radial"])
def test_fortran_driver_network_with_judge(tmp_path, net):
    """The batched driver on a trained H/Au network (nene, judge keywords; the
    radial one on the radial kernels, whose 3x3 cells take the image search):
    the EMT judge's force RMSE stays within the surrogate's 5 meV/A class."""
    deck = open(os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")).read()
    deck = deck.replace("dump 'gpurun_out/mdt_run_h_au.bin'", "")
    deck += f"\nnene 'md_tian2_amd/nn/{net}'\njudge 1\nntrajs 200\n"
    p = tmp_path / "nn.inp"
    p.write_text(deck)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, str(p)], cwd=ROOT, capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    rmse = [float(line.split()[4]) for line in res.stdout.splitlines() if line.startswith("network judge")]
    assert rmse and rmse[0] < (0.02 if net == "hau" else 0.005)


def test_fortran_driver_mixed_beads_matches_library(tmp_path):
    """`beads mixed` (list_skin 0) with force_beads 8 and a PILE projectile: the driver's
    final state equals the Python binding's mixed-model run from the driver's
    dumped initial state with the same seed and trajectory ids."""
    from md_tian2_amd import mdt

    deck = open(os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")).read()
    dump = tmp_path / "mixed.bin"
    deck = deck.replace("dump 'gpurun_out/mdt_run_h_au.bin'", f"dump '{dump}'")
    deck = deck.replace("projectile 1 H 1.00794 ver", "projectile 1 H 1.00794 pil")
    deck = deck.replace("Tproj 0", "Tproj 100\npile_tau 200")
    deck += "\nforce_beads 8\nbeads mixed\nlist_skin 0.0\n"
    p = tmp_path / "mixed.inp"
    p.write_text(deck)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, str(p)], cwd=ROOT, capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    d = read_dump(str(dump))
    assert d["nbeads"] == 8
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, thermal=False, proj_algo=geometry.PROP_PILE,
                              nbeads=8)
    systems = []
    for t in range(d["ntraj"]):
        s = tmpl.copy()
        s.r = d["r0"][t].copy()
        s.v = d["v0"][t].copy()
        systems.append(s)
    pes = pesfile.load_pes(os.path.join(ROOT, "md_tian2_amd", "pes", "EMT-HAu.pes"))
    cols, pesm = pesfile.emt_tables(pes, tmpl.type_names)
    ctx = mdt.Context(0)
    ctx.set_emt(cols, pesm)
    b = mdt.Batch(ctx, systems, dt=0.1, tsurf=300.0, tproj=100.0, pile_tau=200.0, seed=777,
                  traj_ids=[int(x) for x in d["tid"]])
    b.set_bead_mode(1)
    b.set_list_skin(0.0)
    b.compute_forces()
    b.step(int(d["nsteps"]))
    out = b.get_state()
    b.close()
    ctx.close()
    for t, o in enumerate(out):
        assert np.abs(d["r1"][t] - o["r"]).max() < 1e-9
        assert np.abs(d["v1"][t] - o["v"]).max() < 1e-9
        assert np.abs(d["epot"][t] - o["epot"]).max() < 1e-8
        # the slab is classical: every bead reports the same positions
        assert np.array_equal(d["r1"][t][1:], np.repeat(d["r1"][t][1:, :1], 8, axis=1))


def test_fortran_driver_conf_merge_matches_oracle(tmp_path, pes_hau):
    """The driver started from `conf merge` of mxt files (H over a 3x3x4 Au
    lattice with its own velocities, tests/golden/conf_hau): the final
    states against the oracle run from the dumped initial state."""
    deck = open(os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")).read()
    dump = tmp_path / "conf.bin"
    deck = deck.replace("dump 'gpurun_out/mdt_run_h_au.bin'", f"dump '{dump}'")
    deck = deck.replace("slab 4.078 3 3 4 2", "conf merge 'tests/golden/conf_hau/proj' 1 'tests/golden/conf_hau/latt' 1")
    p = tmp_path / "conf.inp"
    p.write_text(deck)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, str(p)], cwd=ROOT, capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    d = read_dump(str(dump))
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, thermal=False)
    cols, pesm = pesfile.emt_tables(pes_hau, tmpl.type_names)
    from md_tian2_amd import conf
    latt = conf.read_mxt(os.path.join(ROOT, "tests", "golden", "conf_hau", "latt", "mxt_00000001.dat"))
    for t in range(d["ntraj"]):
        assert np.array_equal(d["v0"][t][1:], latt.v)  # the lattice keeps the file's velocities
        s = tmpl.copy()
        s.r = d["r0"][t].copy()
        s.v = d["v0"][t].copy()
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=300.0, tproj=0.0)
        st.calc_force()
        st.run(777, int(d["tid"][t]), 1, int(d["nsteps"]))
        assert np.abs(d["r1"][t] - st.r).max() < 1e-9
        assert np.abs(d["v1"][t] - st.v).max() < 1e-9


def test_fortran_driver_scatter_output(tmp_path, pes_hau):
    """`output scatter 1 energy 50 adsorption_status 25` with pul and
    adsorption_distance: the driver's traj/mxt_fin files carry the oracle's
    characteristics (md_tian2's loop with exits, started from the dumped
    initial states) and the initial / final energies of the oracle's state;
    the energy and adsorption files hold one record per interval up to the
    exit step, the header only in the first trajectory's energy file."""
    deck = open(os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")).read()
    dump = tmp_path / "scatter.bin"
    deck = deck.replace("dump 'gpurun_out/mdt_run_h_au.bin'", f"dump '{dump}'")
    deck = deck.replace("'md_tian2_amd/pes/EMT-HAu.pes'", "'%s'" % os.path.join(ROOT, "md_tian2_amd", "pes",
                                                                                   "EMT-HAu.pes"))
    deck = deck.replace("nsteps 30", "nsteps 500").replace("Einc 1.92", "Einc 4.0").replace("pip r r 3.0",
                                                                                          "pip r r 4.0")
    deck += "\npul 4.3\nadsorption_distance 1.9 2.6\noutput scatter 1 energy 50 adsorption_status 25 runner 100\n"
    p = tmp_path / "scatter.inp"
    p.write_text(deck)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
    res = subprocess.run([exe, str(p)], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    d = read_dump(str(dump))
    tmpl = geometry.h_on_slab("Au", n=3, nlayers=4, nfixed=2, thermal=False)
    cols, pesm = pesfile.emt_tables(pes_hau, tmpl.type_names)
    n_exit = nframes = 0
    for t in range(d["ntraj"]):
        tid = int(d["tid"][t])
        s = tmpl.copy()
        s.r = d["r0"][t].copy()
        s.v = d["v0"][t].copy()
        st = oracle.OracleState(s, cols, pesm, step=0.1, tsurf=300.0, tproj=0.0)
        st.calc_force()
        row0 = st.energy_row()
        o, iend = st.run_scatter(777, tid, 500, proj_ul=4.3, adsorption=(1.9, 2.6))
        row1 = st.energy_row()
        n_exit += iend <= 500
        lines = open(tmp_path / "traj" / ("mxt_fin%08d.dat" % tid)).read().splitlines()
        assert len(lines) == 23
        kv = {ln.split("=")[0].strip(): [float(x) for x in ln.split("=")[1].split()] for ln in lines if "=" in ln}
        assert abs(kv["ekin_p_i"][0] - row0[4]) <= 1e-7 and abs(kv["epot_i"][0] - row0[10]) <= 1e-9 * 400
        assert abs(kv["etotal_f"][0] - row1[12]) <= 1e-6 and np.abs(np.array(kv["r_f"]) - row1[14:17]).max() <= 1e-6
        assert kv["time"][0] == pytest.approx((iend - 1) * 0.1, abs=1e-7)
        assert int(kv["turn_pnts"][0]) == o.nturning_points
        assert abs(kv["cl_appr"][0] - o.closest_approach) <= 1e-6
        assert abs(kv["cl_appr_time"][0] - o.closest_approach_time) <= 1e-7
        assert np.abs(np.array(kv["r_min_p"]) - np.array(o.lowest_z[:])).max() <= 1e-6
        en = open(tmp_path / "energy" / ("traj_%08d.dat" % tid)).read().splitlines()
        assert en[0].startswith("#") == (tid == 1)
        en = [ln for ln in en if not ln.split()[0].startswith("#")]
        assert len(en) == min(500, iend) // 50
        ads = open(tmp_path / "conf" / ("ads_%08d.dat" % tid)).read().splitlines()
        assert len(ads) == min(500, iend) // 25
        nframes += min(500, iend) // 100
    print(f"driver scatter output: {n_exit} of {d['ntraj']} trajectories left through pul")
    from md_tian2_amd import nene
    frames = nene.read_input_data(str(tmp_path / "runner" / "input.data"))
    assert len(frames) == nframes and not list((tmp_path / "runner").glob("*.tmp"))


def test_fortran_driver_network_from_the_pes_file(tmp_path):
    """A pes file of `pes nene` blocks (inp_dir naming the trained H/Au
    network, as the reference's NENE decks do) runs the network; with no
    replica extrapolating, the dump equals the run with the `nene` keyword
    and the EMT fallback loaded."""
    pes = tmp_path / "NENE-HAu.pes"
    pes.write_text("pes nene\nH\tAu\tproj\tlatt\n\npes nene\nAu\tAu\tlatt\tlatt\n\npes nene\nH\tH\tproj\tproj\n"
                   f" inp_dir '{ROOT}/md_tian2_amd/nn/hau'\n maxnum_extrapolation_warnings 100\n")
    base = open(os.path.join(ROOT, "tests", "decks", "h_au_3x3x4.inp")).read()
    dumps = []
    for name, edit in (("keyword", lambda d: d + "\nnene 'md_tian2_amd/nn/hau'\n"),
                       ("pesfile", lambda d: d.replace("pes 'md_tian2_amd/pes/EMT-HAu.pes'", f"pes '{pes}'"))):
        dump = tmp_path / f"{name}.bin"
        deck = edit(base.replace("dump 'gpurun_out/mdt_run_h_au.bin'", f"dump '{dump}'"))
        p = tmp_path / f"{name}.inp"
        p.write_text(deck)
        exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "mdt_run")
        res = subprocess.run([exe, str(p)], cwd=ROOT, capture_output=True, text=True, timeout=120)
        print(res.stdout)
        assert res.returncode == 0, res.stdout + res.stderr
        assert "network potential from" in res.stdout
        if name == "keyword":
            assert "trajectories on the EMT fallback at the last step: 0" in res.stdout
        dumps.append(read_dump(str(dump)))
    for k in ("r1", "v1", "epot"):
        assert np.array_equal(dumps[0][k], dumps[1][k]), k


def test_fortran_emt_dropin_read_emt_compute_emt(tmp_path, pes_hau):
    """The Fortran drop-in driven as md_tian2 drives its EMT plugin
    (emt_dropin: read_pes's loop calling read_emt(atoms, unit) per block,
    calc_force's zeroing, compute_emt(atoms, 3001) twice) on universes of
    config 1's and config 2's geometry (H/Au(111) 3x3x4 and 6x6x4, slab
    displaced, H in range) and a 4-bead 3x3x4 one, in one run: energies,
    forces and densities against the oracle's compute_emt
    (src/pes_emt_mod.f90:483-1027), the second call adding exactly the same
    again (epot and f accumulate, src/pes_emt_mod.f90:976, 1001-1011), dens
    set, not accumulated, and reallocated for each universe's shape."""
    from md_tian2_amd import conf

    g = np.random.default_rng(404)
    systems = []
    for n, p, z in ((3, 1, 1.6), (6, 1, 2.3), (3, 4, 1.1)):
        s = geometry.h_on_slab("Au", n=n, nlayers=4, nfixed=2, rng=g, nbeads=p, z0=z)
        s.r += g.normal(0, 0.05, s.r.shape)
        systems.append(s)
    paths = []
    for k, s in enumerate(systems):
        c = conf.Conf(s.r.copy(), s.v.copy(), True, s.fixed.copy(), (s.idx + 1).astype(np.int64),
                      [n for n, _ in s.type_names], [bool(p) for _, p in s.type_names], s.simbox.copy(),
                      np.linalg.inv(s.simbox), 1)
        path = str(tmp_path / ("u%d.dat" % k))
        conf.write_mxt(path, c)
        paths.append(path)
    exe = os.path.join(ROOT, "md_tian2_amd", "fortran", "emt_dropin")
    pes_path = os.path.join(ROOT, "md_tian2_amd", "pes", "EMT-HAu.pes")
    res = subprocess.run([exe, pes_path] + paths, cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    out = {}
    cur = None
    for line in res.stdout.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "universe":
            cur = int(w[1])
            N, P = int(w[2]), int(w[3])
            out[cur] = {c: {"e": np.zeros(P), "f": np.zeros((N, P, 3)), "d": np.zeros((N, P))} for c in (1, 2)}
        elif w[0] == "epot":
            out[cur][int(w[1])]["e"][int(w[2]) - 1] = float(w[3])
        elif w[0] == "force":
            out[cur][int(w[1])]["f"][int(w[2]) - 1, int(w[3]) - 1] = [float(x) for x in w[4:7]]
        elif w[0] == "dens":
            out[cur][int(w[1])]["d"][int(w[2]) - 1, int(w[3]) - 1] = float(w[4])
    assert sorted(out) == [0, 1, 2]
    for k, s in enumerate(systems):
        cols, pesm = pesfile.emt_tables(pes_hau, s.type_names)
        e_r, f_r, d_r = oracle.emt_energy_forces(s, cols, pesm)
        o1, o2 = out[k][1], out[k][2]
        assert o1["f"].shape == f_r.shape
        assert np.abs(o1["e"] - e_r).max() <= 1e-9 * max(1.0, np.abs(e_r).max()), k
        assert np.abs(o1["f"] - f_r).max() <= 1e-9, k
        assert np.abs(o1["d"] - d_r).max() <= 1e-11, k
        assert np.array_equal(o2["e"], 2 * o1["e"]), k
        assert np.array_equal(o2["f"], 2 * o1["f"]), k
        assert np.array_equal(o2["d"], o1["d"]), k
