! nene_predict.f90 -- evaluate a NENE network on structures with the GPU
! drop-in (read_nene_gpu / compute_nene_gpu), like RuNNer's prediction mode.
!
! usage: nene_predict <network dir> <input.data> [nstruct]
! Reads up to nstruct structures (default all) of a RuNNer input.data
! (lattice / atom lines in Bohr; every element of the file must be in the
! network) and prints, per structure, "energy <E [eV]> <natoms>
! <extrapolation warnings>" followed by one "force <i> <Fx Fy Fz [eV/A]>"
! line per atom.
program nene_predict

    use, intrinsic :: iso_c_binding
    use universe_mod
    use pes_nene_gpu_mod, only : read_nene_gpu, compute_nene_gpu, nene_extrapolation_warnings
    use pes_emt_gpu_mod, only : emt_gpu_finalize

    implicit none

    integer, parameter :: dp = c_double
    real(dp), parameter :: bohr2ang = 0.529177211_dp   ! src/constants.f90:160
    character(len=512) :: netdir, datafile, arg, line
    character(len=64) :: w(12)
    character(len=3) :: syms(8)
    integer :: nmax = huge(1), nstruct = 0, u, ios, nw, nat, nel, i, k, ilat, before
    real(dp) :: lat(3, 3)
    real(dp), allocatable :: xyz(:, :)
    character(len=3), allocatable :: el(:)
    type(universe) :: atoms

    if (command_argument_count() < 2) then
        print '(a)', "usage: nene_predict <network dir> <input.data> [nstruct]"
        stop 2
    end if
    call get_command_argument(1, netdir)
    call get_command_argument(2, datafile)
    if (command_argument_count() >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) nmax
    end if
    open(newunit=u, file=trim(datafile), status="old", action="read", iostat=ios)
    if (ios /= 0) then
        print '(a,a)', "Error: cannot open ", trim(datafile)
        stop 1
    end if
    allocate(xyz(3, 100000), el(100000))
    nat = 0
    ilat = 0
    do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) exit
        call split(line, w, nw)
        if (nw == 0) cycle
        select case (trim(w(1)))
        case ("begin")
            nat = 0
            ilat = 0
        case ("lattice")
            ilat = ilat + 1
            read(w(2), *) lat(1, ilat)
            read(w(3), *) lat(2, ilat)
            read(w(4), *) lat(3, ilat)
        case ("atom")
            nat = nat + 1
            read(w(2), *) xyz(1, nat)
            read(w(3), *) xyz(2, nat)
            read(w(4), *) xyz(3, nat)
            el(nat) = w(5)
        case ("end")
            if (ilat /= 3) then
                print '(a)', "Error: structure without three lattice vectors"
                stop 1
            end if
            ! types: distinct element symbols in order of appearance
            nel = 0
            do i = 1, nat
                if (.not. any(syms(1:nel) == el(i))) then
                    nel = nel + 1
                    syms(nel) = el(i)
                end if
            end do
            atoms = universe_alloc(1, nat, nel)
            atoms%name(1:nel) = syms(1:nel)
            do i = 1, nat
                do k = 1, nel
                    if (syms(k) == el(i)) atoms%idx(i) = k
                end do
                atoms%r(:, 1, i) = xyz(:, i) * bohr2ang
            end do
            atoms%simbox = lat * bohr2ang      ! columns are the lattice vectors
            atoms%isimbox = inv3(atoms%simbox)
            if (nstruct == 0) call read_nene_gpu(atoms, trim(netdir) // "/")
            atoms%epot = 0.0_dp
            atoms%f = 0.0_dp
            before = nene_extrapolation_warnings
            call compute_nene_gpu(atoms, 3001)
            nstruct = nstruct + 1
            print '(a,es24.15,1x,i0,1x,i0)', "energy ", atoms%epot(1), nat, nene_extrapolation_warnings - before
            do i = 1, nat
                print '(a,i0,3es24.15)', "force ", i, atoms%f(:, 1, i)
            end do
            if (nstruct >= nmax) exit
        end select
    end do
    close(u)
    call emt_gpu_finalize()

contains

    function inv3(a) result(b)
        real(dp), intent(in) :: a(3, 3)
        real(dp) :: b(3, 3), det
        det = a(1,1) * (a(2,2) * a(3,3) - a(2,3) * a(3,2)) - a(1,2) * (a(2,1) * a(3,3) - a(2,3) * a(3,1)) &
            + a(1,3) * (a(2,1) * a(3,2) - a(2,2) * a(3,1))
        b(1,1) = (a(2,2) * a(3,3) - a(2,3) * a(3,2)) / det
        b(1,2) = (a(1,3) * a(3,2) - a(1,2) * a(3,3)) / det
        b(1,3) = (a(1,2) * a(2,3) - a(1,3) * a(2,2)) / det
        b(2,1) = (a(2,3) * a(3,1) - a(2,1) * a(3,3)) / det
        b(2,2) = (a(1,1) * a(3,3) - a(1,3) * a(3,1)) / det
        b(2,3) = (a(1,3) * a(2,1) - a(1,1) * a(2,3)) / det
        b(3,1) = (a(2,1) * a(3,2) - a(2,2) * a(3,1)) / det
        b(3,2) = (a(1,2) * a(3,1) - a(1,1) * a(3,2)) / det
        b(3,3) = (a(1,1) * a(2,2) - a(1,2) * a(2,1)) / det
    end function inv3

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

end program nene_predict
