! emt_dropin.f90 -- the EMT drop-in driven the way md_tian2 drives its EMT
! plugin: read_pes's loop (src/md_init.f90:99-144) calls read_emt(atoms,
! unit) on every "pes emt" block, then calc_force's preamble zeroes f and
! epot (src/force.f90:49-51) and calls compute_emt(atoms, energy_and_force)
! (src/force.f90:79).
!
! usage: emt_dropin <pes file> <mxt file> [<mxt file> ...]
! For every universe (mxt file, read with conf_io's read_mxt) compute_emt
! runs twice without zeroing in between, so the second call shows the
! accumulation into epot and f; universes of different sizes in one run
! exercise the reallocation of dens.  Output per universe (k from 0):
!   universe <k> <natoms> <nbeads>
!   epot <call> <bead> <E [eV]>
!   force <call> <atom> <bead> <Fx Fy Fz [eV/A]>
!   dens <call> <atom> <bead> <dens>
program emt_dropin

    use, intrinsic :: iso_c_binding
    use universe_mod
    use conf_io, only : read_mxt
    use pes_emt_gpu_mod, only : read_emt, compute_emt, dens, emt_gpu_finalize

    implicit none

    integer, parameter :: dp = c_double
    integer, parameter :: energy_and_force = 3001   ! src/constants.f90:114
    character(len=1024) :: pes_file, mxt, line
    character(len=64) :: w(4)
    type(universe) :: atoms
    integer :: k, u, ios, nw, call, i, b

    if (command_argument_count() < 2) then
        print '(a)', "usage: emt_dropin <pes file> <mxt file> [<mxt file> ...]"
        stop 2
    end if
    call get_command_argument(1, pes_file)
    do k = 2, command_argument_count()
        call get_command_argument(k, mxt)
        call read_mxt(atoms, trim(mxt), ios)
        if (ios /= 0) then
            print '(a,a)', "Error: cannot read ", trim(mxt)
            stop 1
        end if
        ! read_pes: every "pes emt" block goes to read_emt with the open unit
        atoms%pes = 0
        open(newunit=u, file=trim(pes_file), status="old", action="read", iostat=ios)
        if (ios /= 0) then
            print '(a,a)', "Error: cannot open ", trim(pes_file)
            stop 1
        end if
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call split(line, w, nw)
            if (nw == 2) then
                if (w(1) == "pes" .and. w(2) == "emt") call read_emt(atoms, u)
            end if
        end do
        close(u)
        print '(a,1x,i0,1x,i0,1x,i0)', "universe", k - 2, atoms%natoms, atoms%nbeads
        ! calc_force's preamble, then compute_emt twice (accumulating)
        atoms%f = 0.0_dp
        atoms%epot = 0.0_dp
        do call = 1, 2
            call compute_emt(atoms, energy_and_force)
            do b = 1, atoms%nbeads
                print '(a,1x,i0,1x,i0,es25.16e3)', "epot", call, b, atoms%epot(b)
            end do
            do i = 1, atoms%natoms
                do b = 1, atoms%nbeads
                    print '(a,1x,i0,1x,i0,1x,i0,3es25.16e3)', "force", call, i, b, atoms%f(:, b, i)
                    print '(a,1x,i0,1x,i0,1x,i0,es25.16e3)', "dens", call, i, b, dens(b, i)
                end do
            end do
        end do
    end do
    call emt_gpu_finalize()

contains

    subroutine split(line, w, nw)
        character(len=*), intent(in) :: line
        character(len=64), intent(out) :: w(:)
        integer, intent(out) :: nw
        integer :: i, start
        logical :: inword
        nw = 0
        inword = .false.
        start = 1
        do i = 1, len_trim(line) + 1
            if (i <= len_trim(line)) then
                if (line(i:i) /= " " .and. iachar(line(i:i)) /= 9) then
                    if (.not. inword) then
                        inword = .true.
                        start = i
                    end if
                    cycle
                end if
            end if
            if (inword) then
                inword = .false.
                if (nw < size(w)) then
                    nw = nw + 1
                    w(nw) = line(start:i - 1)
                end if
            end if
        end do
    end subroutine split

end program emt_dropin
