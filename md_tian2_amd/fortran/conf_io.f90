! conf_io.f90 -- md_tian2's configuration files for the batched driver.
!
!   read_mxt / write_mxt: the universe dump of read_mxt
!     (src/md_init.f90:344-381) and its writer (src/output_mod.f90:620-630):
!     Fortran unformatted sequential, record 1 (natoms, nbeads, ntypes),
!     record 2 r(3,P,N), v(3,P,N), is_cart, is_fixed(3,P,N), idx(N),
!     name(ntypes) character(3), is_proj(ntypes), simbox(3,3),
!     isimbox(3,3).  Logicals are read as 4-byte integers, nonzero = true
!     (the shipped files hold ifort's -1); positions are returned cartesian.
!   read_poscar: read_poscar (src/md_init.f90:200-339): comment, scaling,
!     three cell vectors (columns), element symbols, counts with an
!     optional ":P:" bead count, cartesian|direct, x y z [T/F T/F T/F] per
!     atom and bead (F = fixed), optionally a blank line and velocities;
!     the first nproj types are projectiles; positions returned cartesian.
!   merge_conf: merge_universes (src/md_init.f90:559-625): one projectile
!     atom first, lattice types after the projectile's, then every atom's
!     bead centroid folded into the cell by int() of its direct coordinate.
module conf_io

    use, intrinsic :: iso_fortran_env, only : real64, int32
    use universe_mod

    implicit none

    private
    public :: read_mxt, write_mxt, read_poscar, merge_conf, mxt_file, invert_cell

    integer, parameter :: dp = real64

contains

    function mxt_file(folder, k) result(path)
        character(len=*), intent(in) :: folder
        integer, intent(in) :: k
        character(len=len_trim(folder) + 17) :: path
        write(path, '(a, a, i8.8, a)') trim(folder), "/mxt_", k, ".dat"
    end function mxt_file

    function invert_cell(m) result(inv)
        real(dp), intent(in) :: m(3, 3)
        real(dp) :: inv(3, 3), det
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
    end function invert_cell

    subroutine to_cartesian(u)
        type(universe), intent(inout) :: u
        integer :: b, i
        if (u%is_cart) return
        do i = 1, u%natoms
            do b = 1, u%nbeads
                u%r(:, b, i) = matmul(u%simbox, u%r(:, b, i))
            end do
        end do
        u%is_cart = .true.
    end subroutine to_cartesian

    subroutine read_mxt(u, path, ios)
        type(universe), intent(out) :: u
        character(len=*), intent(in) :: path
        integer, intent(out) :: ios
        integer :: unit, natoms, nbeads, ntypes
        integer(int32) :: cart
        integer(int32), allocatable :: fixed(:, :, :), proj(:), idx(:)
        open(newunit=unit, file=path, form="unformatted", access="sequential", status="old", action="read", &
             iostat=ios)
        if (ios /= 0) return
        read(unit, iostat=ios) natoms, nbeads, ntypes
        if (ios /= 0 .or. natoms < 1 .or. nbeads < 1 .or. ntypes < 1) then
            if (ios == 0) ios = -2
            close(unit)
            return
        end if
        u = universe_alloc(nbeads, natoms, ntypes)
        allocate(fixed(3, nbeads, natoms), proj(ntypes), idx(natoms))
        read(unit, iostat=ios) u%r, u%v, cart, fixed, idx, u%name, proj, u%simbox, u%isimbox
        close(unit)
        if (ios /= 0) return
        u%is_cart = cart /= 0
        u%is_fixed = fixed /= 0
        u%idx = idx
        u%is_proj = proj /= 0
        call to_cartesian(u)
    end subroutine read_mxt

    ! writes with gfortran's / flang's .true. = 1
    subroutine write_mxt(u, path, ios)
        type(universe), intent(in) :: u
        character(len=*), intent(in) :: path
        integer, intent(out) :: ios
        integer :: unit
        open(newunit=unit, file=path, form="unformatted", access="sequential", status="replace", iostat=ios)
        if (ios /= 0) return
        write(unit, iostat=ios) int(u%natoms, int32), int(u%nbeads, int32), int(u%ntypes, int32)
        if (ios == 0) write(unit, iostat=ios) u%r, u%v, merge(1_int32, 0_int32, u%is_cart), &
            merge(1_int32, 0_int32, u%is_fixed), int(u%idx, int32), u%name, merge(1_int32, 0_int32, u%is_proj), &
            u%simbox, u%isimbox
        close(unit)
    end subroutine write_mxt

    ! whitespace-separated words of a line
    subroutine split(line, w, nw)
        character(len=*), intent(in) :: line
        character(len=*), intent(out) :: w(:)
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
    end subroutine split

    subroutine read_poscar(u, path, nproj, ios)
        type(universe), intent(out) :: u
        character(len=*), intent(in) :: path
        integer, intent(in) :: nproj
        integer, intent(out) :: ios
        integer :: unit, nw, ntypes, nbeads, natoms, p1, p2, t, i, b, k
        integer :: counts(16)
        real(dp) :: scale, box(3, 3)
        character(len=1024) :: line, csys
        character(len=64) :: w(16)
        character(len=1) :: mv(3)
        nbeads = 1
        open(newunit=unit, file=path, status="old", action="read", iostat=ios)
        if (ios /= 0) return
        read(unit, '(a)', iostat=ios) line  ! comment
        if (ios == 0) read(unit, *, iostat=ios) scale
        if (ios == 0) read(unit, *, iostat=ios) box
        if (ios == 0) read(unit, '(a)', iostat=ios) line
        if (ios /= 0) goto 9
        call split(line, w, ntypes)
        read(unit, '(a)', iostat=ios) line
        if (ios /= 0) goto 9
        p1 = index(line, ":")
        p2 = index(line, ":", back=.true.)
        if (p1 > 0) then
            read(line(p1 + 1:p2 - 1), *, iostat=ios) nbeads
            if (ios /= 0 .or. nbeads < 1) goto 8
            line = line(p2 + 1:)
        end if
        call split(line, w(ntypes + 1:), nw)
        if (nw /= ntypes) goto 8
        do t = 1, ntypes
            read(w(ntypes + t), *, iostat=ios) counts(t)
            if (ios /= 0 .or. counts(t) < 1) goto 8
        end do
        natoms = sum(counts(1:ntypes))
        u = universe_alloc(nbeads, natoms, ntypes)
        do t = 1, ntypes
            u%name(t) = w(t)(1:3)
            u%is_proj(t) = t <= nproj
        end do
        k = 0
        do t = 1, ntypes
            u%idx(k + 1:k + counts(t)) = t
            k = k + counts(t)
        end do
        read(unit, '(a)', iostat=ios) csys
        if (ios /= 0) goto 9
        read(unit, '(a)', iostat=ios) line
        if (ios /= 0) goto 9
        backspace(unit)
        call split(line, w, nw)
        if (nw == 3) then
            read(unit, *, iostat=ios) u%r
        else if (nw == 6) then
            do i = 1, natoms
                do b = 1, nbeads
                    read(unit, *, iostat=ios) u%r(:, b, i), mv
                    if (ios /= 0) goto 9
                    u%is_fixed(:, b, i) = mv == "F"
                end do
            end do
        else
            goto 8
        end if
        if (ios /= 0) goto 9
        u%v = 0.0_dp
        read(unit, '(a)', iostat=ios) line
        if (ios == 0) then
            if (len_trim(line) /= 0) goto 8
            read(unit, *, iostat=ios) u%v
            if (ios /= 0) goto 9
        end if
        ios = 0
        close(unit)
        u%simbox = box * scale
        u%isimbox = invert_cell(u%simbox)
        csys = adjustl(csys)
        if (csys(1:1) == "c" .or. csys(1:1) == "C") then
            u%r = u%r * scale
            u%is_cart = .true.
        else if (csys(1:1) == "d" .or. csys(1:1) == "D") then
            u%is_cart = .false.
            call to_cartesian(u)
        else
            ios = -3
        end if
        return
8       ios = -2
9       close(unit)
        if (ios == 0) ios = -2
    end subroutine read_poscar

    subroutine merge_conf(latt, proj, out, ios)
        type(universe), intent(in) :: latt, proj
        type(universe), intent(out) :: out
        integer, intent(out) :: ios
        integer :: np, b, i
        real(dp) :: cen(3), sh(3)
        ios = 0
        if (latt%nbeads /= proj%nbeads) ios = -4
        if (any(latt%is_proj)) ios = -5
        if (.not. all(proj%is_proj)) ios = -6
        if (proj%natoms > 1) ios = -7
        if (ios /= 0) return
        np = proj%natoms
        out = universe_alloc(latt%nbeads, latt%natoms + np, latt%ntypes + proj%ntypes)
        out%r(:, :, :np) = proj%r
        out%v(:, :, :np) = proj%v
        out%is_fixed(:, :, :np) = proj%is_fixed
        out%idx(:np) = proj%idx
        out%name(:proj%ntypes) = proj%name
        out%is_proj(:proj%ntypes) = proj%is_proj
        out%r(:, :, np + 1:) = latt%r
        out%v(:, :, np + 1:) = latt%v
        out%is_fixed(:, :, np + 1:) = latt%is_fixed
        out%idx(np + 1:) = latt%idx + maxval(proj%idx)
        out%name(proj%ntypes + 1:) = latt%name
        out%is_proj(proj%ntypes + 1:) = latt%is_proj
        out%simbox = latt%simbox
        out%isimbox = latt%isimbox
        out%is_cart = .true.
        do i = 1, out%natoms
            cen = matmul(out%isimbox, sum(out%r(:, :, i), dim=2) / out%nbeads)
            sh = real(int(cen), dp)
            do b = 1, out%nbeads
                out%r(:, b, i) = matmul(out%simbox, matmul(out%isimbox, out%r(:, b, i)) - sh)
            end do
        end do
    end subroutine merge_conf

end module conf_io
