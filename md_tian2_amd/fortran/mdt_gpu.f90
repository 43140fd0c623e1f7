! mdt_gpu.f90 -- ISO_C_BINDING interface to libmdt_gpu.so (include/mdt_gpu.h).
!
! Thin layer only: every binding has the C function's name and arguments;
! arrays are passed as assumed-size dummies, so the caller's Fortran arrays
! (column-major r(3,P,N), dens(P,N), simbox(3,3), pes(ntypes,ntypes)) go
! through unchanged.  Status codes are returned, never handled here.
module mdt_gpu

    use, intrinsic :: iso_c_binding

    implicit none

    integer(c_int), parameter :: MDT_OK = 0
    integer(c_int), parameter :: MDT_EINVAL = -1, MDT_EHIP = -2, MDT_ESTATE = -3
    integer(c_int), parameter :: MDT_ETRAJ = -4, MDT_EUNSUP = -5
    integer(c_int), parameter :: MDT_PES_EMT = 2003
    integer(c_int), parameter :: MDT_PROP_VERLET = 1001, MDT_PROP_LANGEVIN = 1003
    integer(c_int), parameter :: MDT_PROP_ANDERSEN = 1005, MDT_PROP_PILE = 1006

    ! mdt_traj_obs: the characteristics of one trajectory (trajectory_info.f90:25-43)
    type, bind(C) :: mdt_traj_obs
        real(c_double) :: cl_appr, cl_last, r_min(3), t_int
        integer(c_int32_t) :: turning_points, moving_down, adsorbed, adsorbed_prev
        integer(c_int32_t) :: exit_step, cl_appr_step, last_step, pad_
    end type mdt_traj_obs

    integer(c_int), parameter :: MDT_NEST = 20

    ! initial conditions (mdt_sample.hip)
    type, bind(C) :: mdt_pip
        integer(c_int32_t) :: random(3)
        real(c_double) :: value(3)
    end type mdt_pip

    type, bind(C) :: mdt_incidence
        real(c_double) :: einc, sigma_einc, polar_deg
        integer(c_int32_t) :: azimuth_random, pad_
        real(c_double) :: azimuth_deg, sigma_azimuth
    end type mdt_incidence

    interface

        function mdt_last_error(buf, buflen) bind(C, name="mdt_last_error") result(n)
            import :: c_char, c_int
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int), value :: buflen
            integer(c_int) :: n
        end function

        function mdt_version() bind(C, name="mdt_version") result(v)
            import :: c_int
            integer(c_int) :: v
        end function

        function mdt_device_count(count) bind(C, name="mdt_device_count") result(rc)
            import :: c_int
            integer(c_int), intent(out) :: count
            integer(c_int) :: rc
        end function

        function mdt_create(device, ctx) bind(C, name="mdt_create") result(rc)
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: ctx
            integer(c_int) :: rc
        end function

        function mdt_destroy(ctx) bind(C, name="mdt_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: rc
        end function

        function mdt_emt_params_set(ctx, ntypes, eta2, n0, e0, lambda, v0, kappa, s0, pes) &
                bind(C, name="mdt_emt_params_set") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: ntypes
            real(c_double), intent(in) :: eta2(*), n0(*), e0(*), lambda(*), v0(*), kappa(*), s0(*)
            integer(c_int32_t), intent(in) :: pes(*)
            integer(c_int) :: rc
        end function

        function mdt_compute_emt(ctx, natoms, nbeads, idx, simbox, isimbox, r, epot, f, dens, flag) &
                bind(C, name="mdt_compute_emt") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: natoms, nbeads
            integer(c_int32_t), intent(in) :: idx(*)
            real(c_double), intent(in) :: simbox(*), isimbox(*), r(*)
            real(c_double), intent(inout) :: epot(*), f(*), dens(*)
            integer(c_int32_t), value :: flag
            integer(c_int) :: rc
        end function

        function mdt_batch_create(ctx, ntraj, natoms, nbeads, batch) bind(C, name="mdt_batch_create") result(rc)
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: ntraj
            integer(c_int32_t), intent(in) :: natoms(*), nbeads(*)
            type(c_ptr), intent(out) :: batch
            integer(c_int) :: rc
        end function

        function mdt_batch_destroy(batch) bind(C, name="mdt_batch_destroy") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int) :: rc
        end function

        function mdt_batch_set_atoms(batch, idx, mass, is_fixed) bind(C, name="mdt_batch_set_atoms") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), intent(in) :: idx(*), is_fixed(*)
            real(c_double), intent(in) :: mass(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_set_cells(batch, simbox, isimbox) bind(C, name="mdt_batch_set_cells") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), intent(in) :: simbox(*), isimbox(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_set_propagation(batch, ntypes, algo, is_proj, dt, tsurf, tproj, pile_tau, &
                andersen_time) bind(C, name="mdt_batch_set_propagation") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int), value :: ntypes
            integer(c_int32_t), intent(in) :: algo(*), is_proj(*)
            real(c_double), value :: dt, pile_tau, andersen_time
            real(c_double), intent(in) :: tsurf(*), tproj(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_seed(batch, seed, traj_id) bind(C, name="mdt_batch_seed") result(rc)
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), value :: seed
            integer(c_int32_t), intent(in) :: traj_id(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_set_state(batch, r, v) bind(C, name="mdt_batch_set_state") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), intent(in) :: r(*), v(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_get_state(batch, r, v, f, a, epot, dens) bind(C, name="mdt_batch_get_state") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), intent(out) :: r(*), v(*), f(*), a(*), epot(*), dens(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_compute_forces(batch) bind(C, name="mdt_batch_compute_forces") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int) :: rc
        end function

        function mdt_batch_step(batch, nsteps) bind(C, name="mdt_batch_step") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: nsteps
            integer(c_int) :: rc
        end function

        function mdt_batch_status(batch, status, clear) bind(C, name="mdt_batch_status") result(rc)
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: batch
            integer(c_int32_t), intent(out) :: status(*)
            integer(c_int), value :: clear
            integer(c_int) :: rc
        end function

        function mdt_batch_last_timing(batch, ms_total, nlaunch) bind(C, name="mdt_batch_last_timing") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), intent(out) :: ms_total
            integer(c_int), intent(out) :: nlaunch
            integer(c_int) :: rc
        end function

        function mdt_batch_tune(batch, tpa, nbcap, tpa_chosen) bind(C, name="mdt_batch_tune") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: tpa, nbcap
            integer(c_int), intent(out) :: tpa_chosen
            integer(c_int) :: rc
        end function

        function mdt_batch_set_precision(batch, mode) bind(C, name="mdt_batch_set_precision") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function

        function mdt_batch_set_bead_mode(batch, mode) bind(C, name="mdt_batch_set_bead_mode") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function

        function mdt_batch_set_com_removal(batch, on) bind(C, name="mdt_batch_set_com_removal") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: on
            integer(c_int) :: rc
        end function

        function mdt_batch_set_list_skin(batch, skin) bind(C, name="mdt_batch_set_list_skin") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), value :: skin
            integer(c_int) :: rc
        end function

        function mdt_batch_set_keep(batch, mode) bind(C, name="mdt_batch_set_keep") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function

        function mdt_batch_set_split(batch, mode) bind(C, name="mdt_batch_set_split") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function

        function mdt_batch_last_split(batch, halves) bind(C, name="mdt_batch_last_split") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), intent(out) :: halves
            integer(c_int) :: rc
        end function

        function mdt_batch_set_wide(batch, mode) bind(C, name="mdt_batch_set_wide") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function

        function mdt_batch_class_shapes(batch, nmax, shapes, nclass) bind(C, name="mdt_batch_class_shapes") &
                result(rc)
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: nmax
            integer(c_int32_t), intent(out) :: shapes(6, *)
            integer(c_int), intent(out) :: nclass
            integer(c_int) :: rc
        end function

        function mdt_batch_precision_arith(batch, arith) bind(C, name="mdt_batch_precision_arith") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), intent(out) :: arith
            integer(c_int) :: rc
        end function

        function mdt_batch_set_observe(batch, proj_ul, proj_ll, ads_start, ads_end, exit_at_limits) &
                bind(C, name="mdt_batch_set_observe") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), value :: proj_ul, proj_ll, ads_start, ads_end
            integer(c_int), value :: exit_at_limits
            integer(c_int) :: rc
        end function

        function mdt_batch_get_observables(batch, obs) bind(C, name="mdt_batch_get_observables") result(rc)
            import :: c_int, c_ptr, mdt_traj_obs
            type(c_ptr), value :: batch
            type(mdt_traj_obs), intent(out) :: obs(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_estimators(batch, est) bind(C, name="mdt_batch_estimators") result(rc)
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double), intent(out) :: est(*)
            integer(c_int) :: rc
        end function

        function mdt_sample_uniform(seed, traj_id, slot) bind(C, name="mdt_sample_uniform") result(x)
            import :: c_int32_t, c_double
            integer(c_int32_t), value :: seed, traj_id, slot
            real(c_double) :: x
        end function

        function mdt_sample_conf_index(seed, traj_id, which, n) bind(C, name="mdt_sample_conf_index") result(k)
            import :: c_int32_t, c_int
            integer(c_int32_t), value :: seed, traj_id
            integer(c_int), value :: which, n
            integer(c_int) :: k
        end function

        function mdt_apply_pip(pip, seed, traj_id, natoms, nbeads, simbox, r) bind(C, name="mdt_apply_pip") result(rc)
            import :: c_int32_t, c_int, c_double, mdt_pip
            type(mdt_pip), intent(in) :: pip
            integer(c_int32_t), value :: seed, traj_id
            integer(c_int), value :: natoms, nbeads
            real(c_double), intent(in) :: simbox(*)
            real(c_double), intent(inout) :: r(*)
            integer(c_int) :: rc
        end function

        function mdt_set_proj_incidence(inc, seed, traj_id, natoms, nbeads, idx, is_proj, mass, v) &
                bind(C, name="mdt_set_proj_incidence") result(rc)
            import :: c_int32_t, c_int, c_double, mdt_incidence
            type(mdt_incidence), intent(in) :: inc
            integer(c_int32_t), value :: seed, traj_id
            integer(c_int), value :: natoms, nbeads
            integer(c_int32_t), intent(in) :: idx(*), is_proj(*)
            real(c_double), intent(in) :: mass(*)
            real(c_double), intent(inout) :: v(*)
            integer(c_int) :: rc
        end function

        function mdt_sample_maxwell(temperature, seed, traj_id, natoms, nbeads, idx, is_proj, mass, is_fixed, v) &
                bind(C, name="mdt_sample_maxwell") result(rc)
            import :: c_int32_t, c_int, c_double
            real(c_double), value :: temperature
            integer(c_int32_t), value :: seed, traj_id
            integer(c_int), value :: natoms, nbeads
            integer(c_int32_t), intent(in) :: idx(*), is_proj(*), is_fixed(*)
            real(c_double), intent(in) :: mass(*)
            real(c_double), intent(inout) :: v(*)
            integer(c_int) :: rc
        end function

        function mdt_batch_set_graph(batch, chunk) bind(C, name="mdt_batch_set_graph") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: chunk
            integer(c_int) :: rc
        end function

        function mdt_nn_load(ctx, directory, ntypes, type_symbols) bind(C, name="mdt_nn_load") result(rc)
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: directory(*)
            integer(c_int), value :: ntypes
            character(kind=c_char), intent(in) :: type_symbols(*)
            integer(c_int) :: rc
        end function

        function mdt_compute_nene(ctx, natoms, nbeads, idx, simbox, isimbox, r, epot, f, nextrap, flag) &
                bind(C, name="mdt_compute_nene") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: natoms, nbeads
            integer(c_int32_t), intent(in) :: idx(*)
            real(c_double), intent(in) :: simbox(*), isimbox(*), r(*)
            real(c_double), intent(inout) :: epot(*), f(*)
            integer(c_int32_t), intent(out) :: nextrap
            integer(c_int32_t), value :: flag
            integer(c_int) :: rc
        end function

        function mdt_batch_set_pes(batch, pes, judge_interval) bind(C, name="mdt_batch_set_pes") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: pes, judge_interval
            integer(c_int) :: rc
        end function

        function mdt_batch_set_judge_percent(batch, percent) bind(C, name="mdt_batch_set_judge_percent") result(rc)
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
            integer(c_int), value :: percent
            integer(c_int) :: rc
        end function

        function mdt_batch_nn_stats(batch, judge_sq, judge_n, judge_max, nfallback) &
                bind(C, name="mdt_batch_nn_stats") result(rc)
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: batch
            real(c_double), intent(out) :: judge_sq, judge_n, judge_max
            integer(c_int32_t), intent(out) :: nfallback
            integer(c_int) :: rc
        end function

    end interface

contains

    ! Message of the calling thread's last failed call, as a Fortran string.
    function mdt_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char) :: buf(1024)
        integer(c_int) :: n
        integer :: i, m
        n = mdt_last_error(buf, 1024_c_int)
        m = min(int(n), 1023)
        allocate(character(len=m) :: msg)
        do i = 1, m
            msg(i:i) = buf(i)
        end do
    end function mdt_error_message

    ! Stop the program with the library's message when rc is not MDT_OK,
    ! as the reference stops on errors (e.g. src/pes_emt_mod.f90:219-220).
    subroutine mdt_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= MDT_OK) then
            print '(a,a,a,i0,a,a)', "Error in ", where, ": rc=", rc, " ", mdt_error_message()
            stop 1
        end if
    end subroutine mdt_check

end module mdt_gpu
