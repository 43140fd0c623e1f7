! mdt_run.f90 -- batched MD driver on the GPU path (Fortran host).
!
! Replaces the trajectory loop of md_tian2 (src/md_tian2.f90:64-115) for
! H-on-metal scattering: it reads an md_tian.inp-style deck, builds the
! trajectories' initial conditions on the host, and runs all of them at
! once through libmdt_gpu.so (calc_force at t = 0, then nsteps of
! propagate_1 / calc_force / propagate_2).
!
! Keywords (md_tian.inp syntax, src/run_config.f90:224-670, where they
! exist): ntrajs, start, nsteps, step, Einc, polar, azimuth (number or r),
! pip (x y z; r = uniform over the cell), Tsurf, Tproj, pile_tau,
! andersen_time, force_beads, projectile <n> <name> <mass> <algo>,
! lattice <n> <name> <mass> <algo>, pes <file>.  Added for the batch:
! slab <a> <nx> <ny> <nlayers> <nfixed> (fcc(111) slab generated here, as
! the reference ships no metal slabs), rng_seed <n> (key of the noise
! stream), dump <file> (binary initial and final states), precision
! <fp64|fp32_pair> (EMT pair-term arithmetic), graph <chunk> (hipGraph
! replay of the step loop), nene <dir> (the NENE network of dir's input.nn,
! scaling.data, weights.XXX.data as the potential, EMT of `pes` as its
! fallback on extrapolation) and judge <n> (EMT judge of the network on a
! rotating 1 % of the trajectories every n steps), beads <all|mixed> (with
! force_beads > 1: every atom carries the beads, as the reference, or only
! the projectile, the slab staying classical).  conf mxt <file> | mxt
! <folder> <n> | poscar <file> | merge <proj folder> <n> <lattice folder>
! <m> reads the trajectories' universes instead of `slab` (conf_io).
! `mdt_run --check-input <deck>` writes every trajectory's initial universe
! as <dump>_NNNNNNNN.mxt and stops before any GPU call.
! output scatter <n> energy <n> adsorption_status <n> (other formats are
! ignored with a warning), pul <z>, pll <z> and adsorption_distance <start>
! <end> turn on the device-side trajectory observation: trajectories leave
! through pul / pll when scatter output is on, and traj/mxt_fin%08d.dat,
! energy/traj_%08d.dat and conf/ads_%08d.dat are written (scatter_out).
program mdt_run

    use, intrinsic :: iso_c_binding
    use mdt_gpu
    use universe_mod
    use pes_emt_gpu_mod, only : emt_gpu_init, emt_gpu_context, emt_params_from_file
    use conf_io, only : read_mxt, write_mxt, read_poscar, merge_conf, mxt_file, invert_cell
    use scatter_out

    implicit none

    integer, parameter :: dp = c_double
    real(dp), parameter :: pi = acos(-1.0_dp), kB = 8.61733238496e-5_dp, amu2mass = 103.638239276_dp

    ! deck
    integer :: ntrajs = 1, start = 1, nsteps = 0, nbeads = 1, nx = 3, ny = 3, nlayers = 4, nfixed = 2
    integer :: rng_seed = 2   ! key of the noise and sampling streams; 2 = traj_id, the reference's default
    real(dp) :: dt = 0.1_dp, einc = 1.0_dp, polar = 0.0_dp, tsurf = 0.0_dp, tproj = 0.0_dp
    real(dp) :: pile_tau = 200.0_dp, andersen_time = 30.0_dp, alat = 4.078_dp
    real(dp) :: sigma_einc = 0.0_dp, sigma_azimuth = 0.0_dp
    real(dp) :: list_skin = 0.3_dp   ! Verlet skin of the mixed model's classical lists (A)
    real(dp) :: mproj = 1.00794_dp, mlatt = 196.966569_dp
    character(len=8) :: algo_proj = "ver", algo_latt = "lan", azimuth = "0"
    character(len=32) :: pip_w(3) = ["r  ", "r  ", "6.0"]
    character(len=3) :: name_proj = "H", name_latt = "Au"
    character(len=512) :: pes_file = "", dump_file = "", deck = "", nene_dir = ""
    integer :: prec = 0, graph_chunk = 0, judge_interval = 0, bead_mode = 0
    ! conf <mxt|poscar|merge> (src/run_config.f90:440-500)
    character(len=8) :: conf_kind = ""
    character(len=512) :: conf_path = "", proj_path = ""
    integer :: conf_n = 0, proj_n = 0
    logical :: einc_set = .false., pip_set = .false., check_input = .false.
    logical :: pes_has_emt = .false.
    ! output and trajectory limits (src/run_config.f90:508-556, 595-650)
    logical :: out_scatter = .false.
    integer :: out_energy = 0, out_ads = 0, out_runner = 0
    real(dp) :: proj_ul = huge(0.0_dp), proj_ll = -huge(0.0_dp)
    real(dp) :: ads_start = huge(0.0_dp), ads_end = huge(0.0_dp)
    character(len=512) :: arg
    real(dp) :: jsq, jn, jmax
    integer(c_int32_t) :: nfb

    type(universe) :: tmpl
    type(c_ptr) :: ctx, batch
    integer :: natoms, ntot, t, i, nl
    integer(c_int32_t), allocatable :: nat(:), nbd(:), idx(:), fixed(:), tid(:), algo(:), isproj(:)
    real(dp), allocatable :: mass(:), cell(:), icell(:), r(:), v(:), f(:), a(:), epot(:), dens(:)
    real(dp), allocatable :: ts(:), tp(:), r0(:), v0(:)
    real(dp), dimension(2) :: eta2, n0, e0, lambda, v0p, kappa, s0
    real(dp) :: ms, ms_step
    integer(c_int) :: tpa
    integer :: u
    type(mdt_traj_obs), allocatable :: obs(:)
    real(dp), allocatable :: est(:)

    if (command_argument_count() < 1) then
        print '(a)', "usage: mdt_run [--check-input] <deck>"
        stop 2
    end if
    do i = 1, command_argument_count()
        call get_command_argument(i, arg)
        if (trim(arg) == "--check-input") then
            check_input = .true.
        else
            deck = arg
        end if
    end do
    call read_deck(trim(deck))
    if (.not. check_input) then
        call scan_pes(trim(pes_file))
        if (.not. pes_has_emt .and. len_trim(nene_dir) == 0) then
            print '(a)', "Error: the pes file has neither emt nor nene blocks"
            stop 1
        end if
        if (len_trim(nene_dir) > 0) print '(a,a)', "mdt_run: network potential from ", trim(nene_dir)
    end if

    if (len_trim(conf_kind) > 0) then
        call conf_universe(tmpl, start)
        natoms = tmpl%natoms
        nbeads = tmpl%nbeads
    else
        natoms = 1 + nx * ny * nlayers
        tmpl = universe_alloc(nbeads, natoms, 2)
        call build_template(tmpl)
    end if
    if (check_input) then
        ! the initial universe of every trajectory as <dump>_NNNNNNNN.mxt; no GPU
        do t = 1, ntrajs
            call initial_conditions(tmpl, start + t - 1)
            write(arg, '(a, a, i8.8, a)') trim(dump_file), "_", start + t - 1, ".mxt"
            call write_mxt(tmpl, trim(arg), i)
            if (i /= 0) then
                print '(a,a)', "Error: cannot write ", trim(arg)
                stop 1
            end if
        end do
        print '(a,i0,a)', "mdt_run: initial universes of ", ntrajs, " trajectories written"
        stop
    end if
    call emt_gpu_init(0)
    ctx = emt_gpu_context()
    if (pes_has_emt) then
        call emt_params_from_file(tmpl, trim(pes_file), eta2, n0, e0, lambda, v0p, kappa, s0)
        call mdt_check(mdt_emt_params_set(ctx, 2_c_int, eta2, n0, e0, lambda, v0p, kappa, s0, &
                                          int(tmpl%pes, c_int32_t)), "mdt_emt_params_set")
    end if

    ntot = 3 * nbeads * natoms
    allocate(nat(ntrajs), nbd(ntrajs), tid(ntrajs), ts(ntrajs), tp(ntrajs))
    allocate(idx(natoms * ntrajs), mass(natoms * ntrajs), fixed(ntot * ntrajs))
    allocate(cell(9 * ntrajs), icell(9 * ntrajs), r(ntot * ntrajs), v(ntot * ntrajs))
    allocate(f(ntot * ntrajs), a(ntot * ntrajs), epot(nbeads * ntrajs), dens(nbeads * natoms * ntrajs))
    nat = natoms
    nbd = nbeads
    ts = tsurf
    tp = tproj
    do t = 1, ntrajs
        tid(t) = int(start + t - 1, c_int32_t)
        call initial_conditions(tmpl, start + t - 1)
        idx((t - 1) * natoms + 1 : t * natoms) = int(tmpl%idx, c_int32_t)
        mass((t - 1) * natoms + 1 : t * natoms) = tmpl%m
        fixed((t - 1) * ntot + 1 : t * ntot) = reshape(merge(1_c_int32_t, 0_c_int32_t, tmpl%is_fixed), [ntot])
        cell((t - 1) * 9 + 1 : t * 9) = reshape(tmpl%simbox, [9])
        icell((t - 1) * 9 + 1 : t * 9) = reshape(tmpl%isimbox, [9])
        r((t - 1) * ntot + 1 : t * ntot) = reshape(tmpl%r, [ntot])
        v((t - 1) * ntot + 1 : t * ntot) = reshape(tmpl%v, [ntot])
    end do
    r0 = r
    v0 = v
    allocate(algo(2), isproj(2))
    algo = [algo_id(algo_proj), algo_id(algo_latt)]
    isproj = [1_c_int32_t, 0_c_int32_t]

    call mdt_check(mdt_batch_create(ctx, int(ntrajs, c_int), nat, nbd, batch), "mdt_batch_create")
    call mdt_check(mdt_batch_set_atoms(batch, idx, mass, fixed), "mdt_batch_set_atoms")
    call mdt_check(mdt_batch_set_cells(batch, cell, icell), "mdt_batch_set_cells")
    call mdt_check(mdt_batch_set_propagation(batch, 2_c_int, algo, isproj, dt, ts, tp, pile_tau, andersen_time), &
                   "mdt_batch_set_propagation")
    call mdt_check(mdt_batch_seed(batch, int(rng_seed, c_int32_t), tid), "mdt_batch_seed")
    call mdt_check(mdt_batch_set_state(batch, r, v), "mdt_batch_set_state")
    call mdt_check(mdt_batch_set_precision(batch, int(prec, c_int)), "mdt_batch_set_precision")
    call mdt_check(mdt_batch_set_graph(batch, int(graph_chunk, c_int)), "mdt_batch_set_graph")
    call mdt_check(mdt_batch_set_bead_mode(batch, int(bead_mode, c_int)), "mdt_batch_set_bead_mode")
    call mdt_check(mdt_batch_set_list_skin(batch, real(list_skin, c_double)), "mdt_batch_set_list_skin")
    if (len_trim(nene_dir) > 0) then
        call mdt_check(mdt_nn_load(ctx, cstr(nene_dir), 2_c_int, cstr(name2(tmpl%name(1)) // name2(tmpl%name(2)))), &
                       "mdt_nn_load")
        call mdt_check(mdt_batch_set_pes(batch, 1_c_int, int(judge_interval, c_int)), "mdt_batch_set_pes")
    end if
    call mdt_check(mdt_batch_tune(batch, 0_c_int, 0_c_int, tpa), "mdt_batch_tune")
    call mdt_check(mdt_batch_compute_forces(batch), "mdt_batch_compute_forces")
    if (out_scatter .or. out_energy > 0 .or. out_ads > 0 .or. out_runner > 0) then
        call run_observed()
    else
        call mdt_check(mdt_batch_step(batch, int(nsteps, c_int)), "mdt_batch_step")
        call mdt_check(mdt_batch_last_timing(batch, ms, nl), "mdt_batch_last_timing")
    end if
    call mdt_check(mdt_batch_get_state(batch, r, v, f, a, epot, dens), "mdt_batch_get_state")
    if (out_scatter) call scatter_final()
    if (len_trim(nene_dir) > 0) then
        call mdt_check(mdt_batch_nn_stats(batch, jsq, jn, jmax, nfb), "mdt_batch_nn_stats")
        if (jn > 0) print '(a,es12.4,a,i0,a)', "network judge: force RMSE ", sqrt(jsq / jn), " eV/A over ", &
            int(jn), " components"
        print '(a,i0)', "trajectories on the EMT fallback at the last step: ", nfb
    end if

    print '(a,i0,a,i0,a,i0,a,f10.3,a,es12.4,a)', "mdt_run: ", ntrajs, " trajectories x ", nsteps, &
        " steps (tpa ", tpa, ") in ", ms, " ms: ", real(ntrajs, dp) * nsteps / max(ms, 1e-9_dp) * 1e3_dp, &
        " trajectory-steps/s"
    do t = 1, min(ntrajs, 4)
        print '(a,i0,a,3f12.6,a,f14.8)', "  traj ", tid(t), "  r_proj ", r((t - 1) * ntot + 1 : (t - 1) * ntot + 3), &
            "  epot ", epot((t - 1) * nbeads + 1)
    end do

    if (len_trim(dump_file) > 0) then
        open(newunit=u, file=trim(dump_file), access="stream", form="unformatted", status="replace")
        write(u) int(ntrajs, c_int32_t), int(natoms, c_int32_t), int(nbeads, c_int32_t), int(nsteps, c_int32_t)
        write(u) tid, r0, v0, r, v, epot, cell
        close(u)
    end if
    call mdt_check(mdt_batch_destroy(batch), "mdt_batch_destroy")

contains

    ! md_tian2's trajectory loop with its outputs (src/md_tian2.f90:68-118)
    ! for all trajectories at once: the scatter start block, then steps in
    ! chunks that end where an energy or adsorption record is due; a
    ! trajectory frozen at its exit writes no records after its exit step.
    subroutine run_observed()
        integer :: cur, nxt, tt, j
        logical :: due_e, due_a, due_r, e_started, a_started
        character(len=64) :: path
        character(len=3), allocatable :: names(:)
        allocate(est(MDT_NEST * ntrajs), obs(ntrajs))
        if (out_scatter) call make_dir("traj")
        if (out_energy > 0) call make_dir("energy")
        if (out_ads > 0) call make_dir("conf")
        if (out_runner > 0) then
            call make_dir("runner")
            allocate(names(natoms))
            names = [(tmpl%name(tmpl%idx(j)), j = 1, natoms)]
        end if
        call mdt_check(mdt_batch_estimators(batch, est), "mdt_batch_estimators")
        if (out_scatter) then
            do tt = 1, ntrajs
                write(path, '(a12, i8.8, a4)') 'traj/mxt_fin', tid(tt), '.dat'
                call write_scatter_initial(trim(path), est((tt - 1) * MDT_NEST + 1 : tt * MDT_NEST))
            end do
        end if
        call mdt_check(mdt_batch_set_observe(batch, proj_ul, proj_ll, ads_start, ads_end, &
                       merge(1_c_int, 0_c_int, out_scatter .and. tmpl%is_proj(tmpl%idx(1)))), &
                       "mdt_batch_set_observe")
        cur = 0
        ms = 0.0_dp
        e_started = .false.
        a_started = .false.
        do while (cur < nsteps)
            nxt = nsteps
            if (out_energy > 0) nxt = min(nxt, (cur / out_energy + 1) * out_energy)
            if (out_ads > 0) nxt = min(nxt, (cur / out_ads + 1) * out_ads)
            if (out_runner > 0) nxt = min(nxt, (cur / out_runner + 1) * out_runner)
            call mdt_check(mdt_batch_step(batch, int(nxt - cur, c_int)), "mdt_batch_step")
            call mdt_check(mdt_batch_last_timing(batch, ms_step, nl), "mdt_batch_last_timing")
            ms = ms + ms_step
            cur = nxt
            call mdt_check(mdt_batch_get_observables(batch, obs), "mdt_batch_get_observables")
            due_e = out_energy > 0
            if (due_e) due_e = mod(cur, out_energy) == 0
            due_a = out_ads > 0
            if (due_a) due_a = mod(cur, out_ads) == 0
            due_r = out_runner > 0
            if (due_r) due_r = mod(cur, out_runner) == 0
            if (due_e) call mdt_check(mdt_batch_estimators(batch, est), "mdt_batch_estimators")
            if (due_r) call mdt_check(mdt_batch_get_state(batch, r, v, f, a, epot, dens), "mdt_batch_get_state")
            do tt = 1, ntrajs
                if (obs(tt)%exit_step /= 0 .and. obs(tt)%exit_step < cur) cycle
                if (due_e) then
                    write(path, '(a12, i8.8, a4)') 'energy/traj_', tid(tt), '.dat'
                    call energy_record(trim(path), tid(tt) == start .and. .not. e_started, cur, dt, &
                                       est((tt - 1) * MDT_NEST + 1 : tt * MDT_NEST))
                    if (tid(tt) == start) e_started = .true.
                end if
                if (due_a) then
                    write(path, '(a9, i8.8, a4)') 'conf/ads_', tid(tt), '.dat'
                    call adsorption_record(trim(path), tid(tt) == start .and. .not. a_started, cur, dt, &
                                           obs(tt)%adsorbed_prev /= 0)
                    if (tid(tt) == start) a_started = .true.
                end if
                if (due_r) then
                    write(path, '(a14, i8.8, a4)') 'runner/frames_', tid(tt), '.tmp'
                    call runner_frame(trim(path), tmpl%simbox, names, natoms, nbeads, &
                                      r((tt - 1) * ntot + 1 : tt * ntot), f((tt - 1) * ntot + 1 : tt * ntot), &
                                      epot((tt - 1) * nbeads + 1 : tt * nbeads))
                end if
            end do
            if (all(obs%exit_step /= 0)) exit
        end do
        if (out_runner > 0) call runner_join(ntrajs, tid)
    end subroutine run_observed

    ! the final scatter block of every trajectory (the state is in r, v, f)
    subroutine scatter_final()
        integer :: tt, istep
        real(dp) :: t_int, cat
        logical :: ads, fnz
        character(len=64) :: path
        call mdt_check(mdt_batch_get_observables(batch, obs), "mdt_batch_get_observables")
        call mdt_check(mdt_batch_estimators(batch, est), "mdt_batch_estimators")
        do tt = 1, ntrajs
            fnz = any(f((tt - 1) * ntot + 1 : (tt - 1) * ntot + 3 * nbeads) /= 0.0_dp)
            call finalize_characteristics(obs(tt), fnz, dt, nsteps, ads_start, ads_end, istep, t_int, ads, cat)
            write(path, '(a12, i8.8, a4)') 'traj/mxt_fin', tid(tt), '.dat'
            call write_scatter_final(trim(path), est((tt - 1) * MDT_NEST + 1 : tt * MDT_NEST), &
                                     (istep - 1) * dt, int(obs(tt)%turning_points), obs(tt)%cl_appr, cat, &
                                     obs(tt)%r_min)
        end do
    end subroutine scatter_final

    integer(c_int32_t) function algo_id(s)
        character(len=*), intent(in) :: s
        select case (trim(s))
        case ("ver"); algo_id = MDT_PROP_VERLET
        case ("lan"); algo_id = MDT_PROP_LANGEVIN
        case ("and"); algo_id = MDT_PROP_ANDERSEN
        case ("pil"); algo_id = MDT_PROP_PILE
        case default
            print '(a,a)', "Error: unknown propagation algorithm ", trim(s)
            stop 1
        end select
    end function algo_id

    subroutine read_deck(path)
        character(len=*), intent(in) :: path
        character(len=512) :: line, key
        character(len=512) :: w(16)
        integer :: unit, ios, nw, n
        open(newunit=unit, file=path, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            print '(a,a)', "Error: cannot open deck ", path
            stop 1
        end if
        do
            read(unit, '(a)', iostat=ios) line
            if (ios /= 0) exit
            n = index(line, "!")
            if (n > 0) line = line(1:n - 1)
            call words(line, w, nw)
            if (nw < 2) cycle
            key = w(1)
            select case (trim(key))
            case ("ntrajs"); read(w(2), *) ntrajs
            case ("start"); read(w(2), *) start
            case ("nsteps"); read(w(2), *) nsteps
            case ("step"); read(w(2), *) dt
            case ("Einc"); read(w(2), *) einc; einc_set = .true.
                if (nw >= 3) read(w(3), *) sigma_einc
            case ("polar"); read(w(2), *) polar
            case ("azimuth"); azimuth = w(2)
                if (nw >= 3) read(w(3), *) sigma_azimuth
            case ("pip")
                if (nw /= 4) then
                    print '(a)', "Error: pip key needs 3 arguments"
                    stop 1
                end if
                pip_w = w(2:4)
                pip_set = .true.
            case ("Tsurf"); read(w(2), *) tsurf
            case ("Tproj"); read(w(2), *) tproj
            case ("pile_tau"); read(w(2), *) pile_tau
            case ("andersen_time"); read(w(2), *) andersen_time
            case ("force_beads"); read(w(2), *) nbeads
            case ("projectile"); name_proj = w(3); read(w(4), *) mproj; algo_proj = w(5)
            case ("lattice"); name_latt = w(3); read(w(4), *) mlatt; algo_latt = w(5)
            case ("pes"); pes_file = unquote(w(2))
            case ("slab"); read(w(2), *) alat; read(w(3), *) nx; read(w(4), *) ny
                read(w(5), *) nlayers; read(w(6), *) nfixed
            case ("rng_seed")  ! global / traj_id (src/run_config.f90:190-203) or a key
                select case (trim(w(2)))
                case ("global"); rng_seed = 1
                case ("traj_id"); rng_seed = 2
                case default; read(w(2), *) rng_seed
                end select
            case ("dump"); dump_file = unquote(w(2))
            case ("nene"); nene_dir = unquote(w(2))
            case ("judge"); read(w(2), *) judge_interval
            case ("graph"); read(w(2), *) graph_chunk
            case ("list_skin"); read(w(2), *) list_skin
            case ("beads")
                if (trim(w(2)) == "mixed") then
                    bead_mode = 1
                else if (trim(w(2)) == "all") then
                    bead_mode = 0
                else
                    print '(a,a)', "Error: beads must be all or mixed, got ", trim(w(2))
                    stop 1
                end if
            case ("conf")
                conf_kind = w(2)
                select case (trim(conf_kind))
                case ("mxt", "poscar")
                    conf_path = unquote(w(3))
                    if (nw >= 4) read(w(4), *) conf_n   ! a folder of mxt_NNNNNNNN.dat files
                case ("merge")
                    if (nw /= 6) then
                        print '(a)', "Error: conf merge needs projectile folder, count, lattice folder, count"
                        stop 1
                    end if
                    proj_path = unquote(w(3))
                    read(w(4), *) proj_n
                    conf_path = unquote(w(5))
                    read(w(6), *) conf_n
                case default
                    print '(a,a)', "Error: conf must be mxt, poscar or merge, got ", trim(conf_kind)
                    stop 1
                end select
            case ("output")
                if (mod(nw, 2) /= 1) then
                    print '(a)', "Error: number of output arguments must be even"
                    stop 1
                end if
                do n = 2, nw - 1, 2
                    select case (trim(w(n)))
                    case ("scatter"); out_scatter = .true.
                    case ("energy"); read(w(n + 1), *) out_energy
                    case ("adsorption_status"); read(w(n + 1), *) out_ads
                    case ("runner"); read(w(n + 1), *) out_runner
                    case default
                        print '(a,a,a)', "Warning: output format ", trim(w(n)), " is not written by mdt_run"
                    end select
                end do
            case ("pul"); read(w(2), *) proj_ul
            case ("pll"); read(w(2), *) proj_ll
            case ("adsorption_distance")
                if (nw /= 3) then
                    print '(a)', "Error: adsorption_distance key needs 2 arguments"
                    stop 1
                end if
                read(w(2), *) ads_start
                read(w(3), *) ads_end
            case ("precision")
                if (trim(w(2)) == "fp32_pair") then
                    prec = 1
                else if (trim(w(2)) == "fp64") then
                    prec = 0
                else
                    print '(a,a)', "Error: precision must be fp64 or fp32_pair, got ", trim(w(2))
                    stop 1
                end if
            case default
                continue  ! unknown keywords are skipped, as src/run_config.f90:667-669
            end select
        end do
        close(unit)
    end subroutine read_deck

    ! The blocks of the pes file (src/md_init.f90:99-144): `pes emt` blocks
    ! give the EMT parameters; `pes nene` blocks name the network's RuNNer
    ! directory (inp_dir, relative to the working directory as in
    ! read_nene), used unless the deck's `nene` keyword names one; `pes non`
    ! pairs do not interact; any other PES stops the run.
    subroutine scan_pes(path)
        character(len=*), intent(in) :: path
        character(len=1000) :: line
        character(len=512) :: w(8)
        integer :: unit, ios, nw
        character(len=16) :: kind
        open(newunit=unit, file=path, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            print '(a,a)', "Error: cannot open pes file ", path
            stop 1
        end if
        kind = ""
        do
            read(unit, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call words(line, w, nw)
            if (nw == 0) then
                kind = ""
                cycle
            end if
            if (trim(w(1)) == "pes" .and. nw >= 2) then
                kind = w(2)
                select case (trim(kind))
                case ("emt"); pes_has_emt = .true.
                case ("nene", "non")
                case default
                    print '(a,a,a)', "Error: pes ", trim(kind), " is not supported on the GPU path (emt, nene, non)"
                    stop 1
                end select
            else if (trim(kind) == "nene" .and. nw == 2 .and. trim(w(1)) == "inp_dir") then
                if (len_trim(nene_dir) == 0) nene_dir = unquote(w(2))
            end if
        end do
        close(unit)
    end subroutine scan_pes

    ! NUL-terminated copy of a Fortran string for the C-ABI
    function cstr(s) result(c)
        character(len=*), intent(in) :: s
        character(kind=c_char) :: c(len_trim(s) + 1)
        integer :: i
        do i = 1, len_trim(s)
            c(i) = s(i:i)
        end do
        c(len_trim(s) + 1) = c_null_char
    end function cstr

    ! element symbol padded to the 2 characters per type of mdt_nn_load
    function name2(s) result(t)
        character(len=*), intent(in) :: s
        character(len=2) :: t
        t = adjustl(s)
    end function name2

    function unquote(s) result(t)
        character(len=*), intent(in) :: s
        character(len=len(s)) :: t
        t = s
        if (len_trim(t) >= 2) then
            if (t(1:1) == "'" .or. t(1:1) == '"') t = t(2:len_trim(t) - 1)
        end if
    end function unquote

    subroutine words(line, w, nw)
        character(len=*), intent(in) :: line
        character(len=512), intent(out) :: w(:)
        integer, intent(out) :: nw
        integer :: i, st
        logical :: inw
        nw = 0
        inw = .false.
        st = 1
        do i = 1, len_trim(line) + 1
            if (i <= len_trim(line)) then
                if (line(i:i) /= " " .and. iachar(line(i:i)) /= 9) then
                    if (.not. inw) then
                        inw = .true.
                        st = i
                    end if
                    cycle
                end if
            end if
            if (inw) then
                inw = .false.
                if (nw < size(w)) then
                    nw = nw + 1
                    w(nw) = line(st:i - 1)
                end if
            end if
        end do
    end subroutine words

    ! fcc(111) slab: ABC stacking, layers a/sqrt(3) apart from z = 0 down,
    ! in-plane cell nx a/sqrt2 (1,0,0) and ny a/sqrt2 (1/2, sqrt3/2, 0),
    ! c = 40 A; projectile first, bottom nfixed layers frozen.
    subroutine build_template(u)
        type(universe), intent(inout) :: u
        real(dp) :: d, dz, u1(3), u2(3), sh(3), p(3), fr(3)
        integer :: l, i, j, k
        d = alat / sqrt(2.0_dp)
        dz = alat / sqrt(3.0_dp)
        u1 = [d, 0.0_dp, 0.0_dp]
        u2 = [0.5_dp * d, 0.5_dp * sqrt(3.0_dp) * d, 0.0_dp]
        sh = (u1 + u2) / 3.0_dp
        u%simbox = 0.0_dp
        u%simbox(:, 1) = nx * u1
        u%simbox(:, 2) = ny * u2
        u%simbox(3, 3) = 40.0_dp
        u%isimbox = invert3(u%simbox)
        u%name = [name_proj, name_latt]
        u%is_proj = [.true., .false.]
        u%pes = 0
        u%idx(1) = 1
        u%m(1) = mproj * amu2mass
        k = 1
        do l = 0, nlayers - 1
            do j = 0, ny - 1
                do i = 0, nx - 1
                    k = k + 1
                    p = i * u1 + j * u2 + mod(l, 3) * sh
                    p(3) = -l * dz
                    fr = matmul(u%isimbox, p)
                    fr(1:2) = fr(1:2) - floor(fr(1:2))
                    u%r(:, 1, k) = matmul(u%simbox, fr)
                    u%idx(k) = 2
                    u%m(k) = mlatt * amu2mass
                    u%is_fixed(:, :, k) = (l >= nlayers - nfixed)
                end do
            end do
        end do
    end subroutine build_template

    function invert3(m) result(inv)
        real(dp), intent(in) :: m(3,3)
        real(dp) :: inv(3,3), det
        det = m(1,1) * (m(2,2) * m(3,3) - m(2,3) * m(3,2)) - m(1,2) * (m(2,1) * m(3,3) - m(2,3) * m(3,1)) &
            + m(1,3) * (m(2,1) * m(3,2) - m(2,2) * m(3,1))
        inv(1,1) = (m(2,2) * m(3,3) - m(2,3) * m(3,2)) / det
        inv(1,2) = (m(1,3) * m(3,2) - m(1,2) * m(3,3)) / det
        inv(1,3) = (m(1,2) * m(2,3) - m(1,3) * m(2,2)) / det
        inv(2,1) = (m(2,3) * m(3,1) - m(2,1) * m(3,3)) / det
        inv(2,2) = (m(1,1) * m(3,3) - m(1,3) * m(3,1)) / det
        inv(2,3) = (m(1,3) * m(2,1) - m(1,1) * m(2,3)) / det
        inv(3,1) = (m(2,1) * m(3,2) - m(2,2) * m(3,1)) / det
        inv(3,2) = (m(1,2) * m(3,1) - m(1,1) * m(3,2)) / det
        inv(3,3) = (m(1,1) * m(2,2) - m(1,2) * m(2,1)) / det
    end function invert3

    ! Trajectory initial conditions on the generated slab: the projectile
    ! placed by apply_pip (pip, with the slab's cell) and given its incidence
    ! velocity by set_proj_incidence (src/md_init.f90:498-554, 829-877),
    ! mobile slab atoms Maxwell-Boltzmann at Tsurf; every draw from the
    ! library's sampling stream of (rng_seed, trajectory id).
    subroutine initial_conditions(u, traj)
        type(universe), intent(inout) :: u
        integer, intent(in) :: traj
        type(universe) :: p1
        integer(c_int32_t), allocatable :: fx(:)
        integer :: k_
        real(dp) :: cen(3)
        if (len_trim(conf_kind) > 0) then
            call conf_universe(u, traj)
            return
        end if
        call build_template(u)
        p1 = universe_alloc(nbeads, 1, 1)
        p1%r(:, :, 1) = u%r(:, :, 1)
        call mdt_check(mdt_apply_pip(pip_struct(), int(rng_seed, c_int32_t), int(traj, c_int32_t), 1_c_int, &
                                     int(nbeads, c_int), u%simbox, p1%r), "mdt_apply_pip")
        ! the centroid folded into the cell by int() of its direct coordinates,
        ! as merge_universes does after apply_pip (src/md_init.f90:612-621)
        cen = matmul(u%isimbox, sum(p1%r(:, :, 1), dim=2) / nbeads)
        do k_ = 1, nbeads
            u%r(:, k_, 1) = p1%r(:, k_, 1) - matmul(u%simbox, real(int(cen), dp))
        end do
        u%v = 0.0_dp
        call incidence(u, traj)
        allocate(fx(3 * u%nbeads * u%natoms))
        fx = reshape(merge(1_c_int32_t, 0_c_int32_t, u%is_fixed), [size(fx)])
        call mdt_check(mdt_sample_maxwell(tsurf, int(rng_seed, c_int32_t), int(traj, c_int32_t), &
                                          int(u%natoms, c_int), int(u%nbeads, c_int), int(u%idx, c_int32_t), &
                                          int(merge(1, 0, u%is_proj), c_int32_t), u%m, fx, u%v), "mdt_sample_maxwell")
        do k_ = 2, u%natoms
            u%r(:, :, k_) = spread(u%r(:, 1, k_), 2, nbeads)
        end do
    end subroutine initial_conditions

    function pip_struct() result(p)
        type(mdt_pip) :: p
        integer :: k
        p%random = 0
        p%value = 0.0_dp
        do k = 1, 3
            if (trim(pip_w(k)) == "r") then
                p%random(k) = 1
            else
                read(pip_w(k), *) p%value(k)
            end if
        end do
    end function pip_struct

    ! set_proj_incidence through the library's sampler (Einc given)
    subroutine incidence(u, traj)
        type(universe), intent(inout) :: u
        integer, intent(in) :: traj
        type(mdt_incidence) :: inc
        if (.not. einc_set) return
        inc%einc = einc
        inc%sigma_einc = sigma_einc
        inc%polar_deg = polar
        inc%azimuth_random = merge(1, 0, trim(azimuth) == "r")
        inc%pad_ = 0
        inc%azimuth_deg = 0.0_dp
        if (inc%azimuth_random == 0) read(azimuth, *) inc%azimuth_deg
        inc%sigma_azimuth = sigma_azimuth
        call mdt_check(mdt_set_proj_incidence(inc, int(rng_seed, c_int32_t), int(traj, c_int32_t), &
                                              int(u%natoms, c_int), int(u%nbeads, c_int), int(u%idx, c_int32_t), &
                                              int(merge(1, 0, u%is_proj), c_int32_t), u%m, u%v), &
                       "mdt_set_proj_incidence")
    end subroutine incidence

    ! The universe of trajectory `traj` from the conf files (read_geometry,
    ! src/md_init.f90:160-192, and post_process :630-671): the file, or file
    ! k of the folder with k uniform in 1..n (src/run_config.f90:466-496;
    ! drawn from this driver's per-trajectory host stream, not the
    ! reference's), the projectile merged first for `conf merge` and put at
    ! `pip` (x, y: "r" uniform over the cell, else cartesian; z cartesian),
    ! its bead-centroid velocity replaced by the incidence velocity when
    ! Einc is given (set_proj_incidence, :829-877), `force_beads` copies of
    ! bead 1, masses and propagators from the projectile / lattice lines.
    ! Velocities of the lattice are the file's.
    subroutine conf_universe(u, traj)
        type(universe), intent(inout) :: u
        integer, intent(in) :: traj
        type(universe) :: latt, proj, one
        integer :: ios, b, i
        integer(c_int32_t) :: key, tr
        key = int(rng_seed, c_int32_t)
        tr = int(traj, c_int32_t)
        select case (trim(conf_kind))
        case ("poscar")
            call read_poscar(one, trim(conf_path), 1, ios)
            call conf_check(ios, conf_path)
        case ("mxt")
            if (conf_n > 0) then
                call read_mxt(one, mxt_file(trim(conf_path), int(mdt_sample_conf_index(key, tr, 1_c_int, &
                              int(conf_n, c_int)))), ios)
            else
                call read_mxt(one, trim(conf_path), ios)
            end if
            call conf_check(ios, conf_path)
        case ("merge")  ! prepare_next_traj (src/md_init.f90:904-941)
            call read_mxt(proj, mxt_file(trim(proj_path), int(mdt_sample_conf_index(key, tr, 0_c_int, &
                          int(proj_n, c_int)))), ios)
            call conf_check(ios, proj_path)
            call read_mxt(latt, mxt_file(trim(conf_path), int(mdt_sample_conf_index(key, tr, 1_c_int, &
                          int(conf_n, c_int)))), ios)
            call conf_check(ios, conf_path)
            if (pip_set) call mdt_check(mdt_apply_pip(pip_struct(), key, tr, int(proj%natoms, c_int), &
                                                      int(proj%nbeads, c_int), proj%simbox, proj%r), "mdt_apply_pip")
            call merge_conf(latt, proj, one, ios)
            if (ios /= 0) then
                print '(a,i0)', "Error: conf merge: incompatible files, code ", ios
                stop 1
            end if
        end select
        if (one%ntypes /= 2 .or. .not. one%is_proj(1) .or. one%is_proj(2) .or. count(one%idx == 1) /= 1) then
            print '(a)', "Error: the driver runs one projectile atom (type 1) over one lattice type (type 2)"
            stop 1
        end if
        if (nbeads > 1 .and. one%nbeads == 1) then  ! force_beads
            if (trim(conf_kind) == "mxt") then
                print '(a)', "Error: cannot force beads on mxt files"   ! as src/md_init.f90:428
                stop 1
            end if
            u = universe_alloc(nbeads, one%natoms, one%ntypes)
            do b = 1, nbeads
                u%r(:, b, :) = one%r(:, 1, :)
                u%v(:, b, :) = one%v(:, 1, :)
                u%is_fixed(:, b, :) = one%is_fixed(:, 1, :)
            end do
            u%idx = one%idx
            u%name = one%name
            u%is_proj = one%is_proj
            u%simbox = one%simbox
            u%isimbox = one%isimbox
            u%is_cart = .true.
        else
            u = one
        end if
        u%pes = 0
        do i = 1, u%natoms
            u%m(i) = merge(mproj, mlatt, u%idx(i) == 1) * amu2mass
        end do
        call incidence(u, traj)
    end subroutine conf_universe

    subroutine conf_check(ios, path)
        integer, intent(in) :: ios
        character(len=*), intent(in) :: path
        if (ios /= 0) then
            print '(a,a,a,i0)', "Error: cannot read configuration ", trim(path), ", code ", ios
            stop 1
        end if
    end subroutine conf_check

end program mdt_run
