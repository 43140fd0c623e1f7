! scatter_out.f90 -- md_tian2's per-trajectory output files from the
! batch's device observables and estimators.
!
!   traj/mxt_fin%08d.dat  output_scatter (src/output_mod.f90:635-730): the
!                         initial block after the first calc_force, the
!                         final block after the loop;
!   energy/traj_%08d.dat  output_nrg (:170-229), one record every
!                         `output energy <n>` steps;
!   conf/ads_%08d.dat     output_adsorption_status (:792-818);
!   runner/input.data     output_runner (:550-609): RuNNer frames of bead
!                         centroids and centroid forces, gathered per
!                         trajectory in runner/frames_%08d.tmp and joined in
!                         trajectory order at the end (runner_join).
! The records use the reference's edit descriptors; the reference's bare
! `i` (ifort: width 12) and `L` (width 2) are written as i12 and L2.
!
! finalize_characteristics reproduces the collection md_tian2.f90:100 makes
! after the loop: with istep = exit step (already collected on the device)
! or nsteps + 1 (one more collection of the final state: interaction time
! and adsorption thresholds), and the reference's closest_approach_time,
! scaled by the step at every collection (trajectory_info.f90:111).
! Folders are created with libc mkdir (no child process).
module scatter_out

    use, intrinsic :: iso_c_binding
    use mdt_gpu, only : mdt_traj_obs

    implicit none

    private
    public :: make_dir, write_scatter_initial, write_scatter_final, energy_record, adsorption_record, &
              finalize_characteristics, runner_frame, runner_join

    integer, parameter :: dp = c_double
    real(dp), parameter :: rad2deg = 180.0_dp / acos(-1.0_dp)
    ! constants.f90:160-168
    real(dp), parameter :: bohr2ang = 0.529177211_dp, ang2bohr = 1.0_dp / bohr2ang
    real(dp), parameter :: ha2ev = 27.21138602_dp, ev2ha = 1.0_dp / ha2ev
    real(dp), parameter :: evang2habohr = 1.0_dp / (ha2ev * ang2bohr)

    interface
        function c_mkdir(path, mode) bind(C, name="mkdir") result(rc)
            import :: c_char, c_int
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value :: mode
            integer(c_int) :: rc
        end function
    end interface

contains

    subroutine make_dir(path)
        character(len=*), intent(in) :: path
        integer(c_int) :: rc
        rc = c_mkdir(trim(path) // c_null_char, int(o'755', c_int))  ! an existing folder is fine
    end subroutine make_dir

    subroutine polar_azi(v, polar, azi)
        real(dp), intent(in) :: v(3)
        real(dp), intent(out) :: polar, azi
        polar = acos(v(3) / sqrt(sum(v * v))) * rad2deg
        azi = atan2(v(2), v(1)) * rad2deg
    end subroutine polar_azi

    ! est: the MDT_NEST estimators of the trajectory (mdt_batch_estimators)
    subroutine write_scatter_initial(path, est)
        character(len=*), intent(in) :: path
        real(dp), intent(in) :: est(20)
        real(dp) :: polar, azi
        integer :: u
        call polar_azi(est(18:20), polar, azi)
        open(newunit=u, file=path, status="replace", action="write")
        write(u, '(a11, f14.7)') "ekin_p_i = ", est(5)
        write(u, '(a11, f14.7)') "ekin_l_i = ", est(6)
        write(u, '(a11, e21.10)') "epot_i   = ", est(11)
        write(u, '(a11, e21.10)') "etotal_i = ", est(13)
        write(u, '(a11, 3f14.7)') "r_i      = ", est(15:17)
        write(u, '(a11, 3f14.7)') "v_i      = ", est(18:20)
        write(u, '(a11, f14.7)') "polar_i  = ", 180 - polar
        write(u, '(a11, f14.7)') "azi_i    = ", azi
        write(u, '(a)') ""
        close(u)
    end subroutine write_scatter_initial

    subroutine write_scatter_final(path, est, time, turn_pnts, cl_appr, cl_appr_time, r_min)
        character(len=*), intent(in) :: path
        real(dp), intent(in) :: est(20), time, cl_appr, cl_appr_time, r_min(3)
        integer, intent(in) :: turn_pnts
        real(dp) :: polar, azi
        integer :: u
        call polar_azi(est(18:20), polar, azi)
        open(newunit=u, file=path, status="old", position="append", action="write")
        write(u, '(a15, f14.7)') "ekin_p_f     = ", est(5)
        write(u, '(a15, f14.7)') "ekin_l_f     = ", est(6)
        write(u, '(a15, e21.10)') "epot_f       = ", est(11)
        write(u, '(a15, e21.10)') "etotal_f     = ", est(13)
        write(u, '(a15, 3f14.7)') "r_f          = ", est(15:17)
        write(u, '(a15, 3f14.7)') "v_f          = ", est(18:20)
        write(u, '(a15, f14.7)') "polar_f      = ", polar
        write(u, '(a15, f14.7)') "azi_f        = ", azi
        write(u, '(a)') ""
        write(u, '(a15, f14.7)') "time         = ", time
        write(u, '(a15, i12)') "turn_pnts    = ", turn_pnts
        write(u, '(a15, f14.7)') "cl_appr      = ", cl_appr
        write(u, '(a15, f14.7)') "cl_appr_time = ", cl_appr_time
        write(u, '(a15, 3f14.7)') "r_min_p      = ", r_min
        close(u)
    end subroutine write_scatter_final

    ! new_file: the run's first trajectory (overwrite_nrg, output_mod.f90:36,
    ! 189-197) -- the file is replaced and gets the header; otherwise appended
    subroutine energy_record(path, new_file, istep, step, est)
        character(len=*), intent(in) :: path
        logical, intent(in) :: new_file
        integer, intent(in) :: istep
        real(dp), intent(in) :: step, est(20)
        integer :: u, i
        if (new_file) then
            open(newunit=u, file=path, status="replace", action="write")
            write(u, '(a)') '# All energies are given in eV.'
            write(u, '(a1, i8, 20i17)') "#", (i, i = 1, 21)
            write(u, '(21a17)') '#time/fs', 'atom T/K', 'bead T/K', 'r_gyr_sq_p/A²', &
                'r_gyr_sq_l/A²', 'atom ekin proj', 'atom ekin latt', 'bead ekin proj', &
                'bead ekin latt', 'qntm ekin proj', 'qntm ekin latt', 'atom epot', &
                'bead epot', 'e_total', 'f_total/(eV/A)', 'r_x', 'r_y', 'r_z', 'v_x', 'v_y', 'v_z'
        else
            open(newunit=u, file=path, status="unknown", position="append", action="write")
        end if
        write(u, '(21e17.8e2)') istep * step, est
        close(u)
    end subroutine energy_record

    subroutine adsorption_record(path, new_file, istep, step, adsorbed)
        character(len=*), intent(in) :: path
        logical, intent(in) :: new_file, adsorbed
        integer, intent(in) :: istep
        real(dp), intent(in) :: step
        integer :: u
        if (new_file) then
            open(newunit=u, file=path, status="replace", action="write")
        else
            open(newunit=u, file=path, status="unknown", position="append", action="write")
        end if
        write(u, '(e13.6e2, L2)') (istep - 1) * step, adsorbed
        close(u)
    end subroutine adsorption_record

    ! one RuNNer structure appended to path: simbox(3,3) columns = cell
    ! vectors [A], names(natoms), r and f (3, nbeads, natoms) [A, eV/A],
    ! epot(nbeads) [eV]
    subroutine runner_frame(path, simbox, names, natoms, nbeads, r, f, epot)
        character(len=*), intent(in) :: path
        integer, intent(in) :: natoms, nbeads
        real(dp), intent(in) :: simbox(3, 3), r(3, nbeads, natoms), f(3, nbeads, natoms), epot(nbeads)
        character(len=3), intent(in) :: names(natoms)
        character(len=*), parameter :: lattice_format = '(A7, X, F11.8, X, F11.8, X, F11.8)'
        character(len=*), parameter :: atom_format = &
            '(A4, X, F16.9, X, F16.9, X, F16.9, X, A3, X, F11.8, X, F11.8, X, F11.8, X, F11.8, X, F11.8)'
        character(len=*), parameter :: ce_format = '(A6 ,X , F11.8)'
        real(dp) :: cr(3), cf(3)
        integer :: u, k, j
        open(newunit=u, file=path, status="unknown", position="append", action="write")
        write(u, '(A5)') 'begin'
        do k = 1, 3
            write(u, lattice_format) 'lattice', simbox(:, k) * ang2bohr
        end do
        do j = 1, natoms
            cr = sum(r(:, :, j), dim=2) / nbeads
            cf = sum(f(:, :, j), dim=2) / nbeads
            write(u, atom_format) 'atom', cr * ang2bohr, names(j), 0.0_dp, 0.0_dp, cf * evang2habohr
        end do
        write(u, ce_format) 'charge', 0.0_dp
        write(u, ce_format) 'energy', sum(epot) / nbeads * ev2ha
        write(u, '(A3)') 'end'
        close(u)
    end subroutine runner_frame

    ! runner/input.data: output_runner's header, then every trajectory's
    ! frames in trajectory order (the reference runs them one after another)
    subroutine runner_join(ntraj, tids)
        integer, intent(in) :: ntraj
        integer(c_int32_t), intent(in) :: tids(ntraj)
        integer :: u, v, t, ios, tv(8)
        character(len=512) :: line
        character(len=64) :: path
        logical :: ex
        open(newunit=u, file="runner/input.data", status="replace", action="write")
        write(u, '(a)') '# This file was generated by MDT2.'
        call date_and_time(values=tv)
        write(u, '(a,i4, a, i2.2, a, i2.2, a, i2.2, a, i2.2)') "# ", tv(1), "-", tv(2), "-", tv(3), " - ", tv(5), &
            ".", tv(6)
        do t = 1, ntraj
            write(path, '(a14, i8.8, a4)') 'runner/frames_', tids(t), '.tmp'
            inquire(file=trim(path), exist=ex)
            if (.not. ex) cycle
            open(newunit=v, file=trim(path), status="old", action="read")
            do
                read(v, '(a)', iostat=ios) line
                if (ios /= 0) exit
                write(u, '(a)') trim(line)
            end do
            close(v, status="delete")
        end do
        close(u)
    end subroutine runner_join

    ! o: device observables; fnz: any projectile force component /= 0 in
    ! the final state; nsteps: steps run since mdt_batch_set_observe
    subroutine finalize_characteristics(o, fnz, step, nsteps, ads_start, ads_end, istep, t_int, adsorbed, &
                                        cl_appr_time)
        type(mdt_traj_obs), intent(in) :: o
        logical, intent(in) :: fnz
        real(dp), intent(in) :: step, ads_start, ads_end
        integer, intent(in) :: nsteps
        integer, intent(out) :: istep
        real(dp), intent(out) :: t_int, cl_appr_time
        logical, intent(out) :: adsorbed
        integer :: k
        t_int = o%t_int
        adsorbed = o%adsorbed /= 0
        if (o%exit_step /= 0) then
            istep = o%exit_step
        else
            istep = nsteps + 1
            if (fnz) t_int = t_int + step
            if (.not. adsorbed .and. o%cl_last < ads_start) adsorbed = .true.
            if (adsorbed .and. o%cl_last > ads_end) adsorbed = .false.
        end if
        cl_appr_time = 0.0_dp
        if (o%cl_appr_step > 0) then
            cl_appr_time = dble(o%cl_appr_step) * step
            do k = 1, istep - o%cl_appr_step
                cl_appr_time = cl_appr_time * step
                if (cl_appr_time == 0.0_dp) exit
            end do
        end if
    end subroutine finalize_characteristics

end module scatter_out
