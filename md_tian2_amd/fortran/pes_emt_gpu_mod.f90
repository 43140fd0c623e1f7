! pes_emt_gpu_mod.f90 -- GPU drop-in for md_tian2's EMT plugin.
!
! read_emt(atoms, inp_unit) and compute_emt(atoms, flag) have the names and
! arguments of the reference's read_emt (src/pes_emt_mod.f90:49-190) and
! compute_emt (:194-225), so that read_pes (src/md_init.f90:99-144) and
! calc_force (src/force.f90:79) call them unchanged once they use this
! module instead of pes_emt_mod.  Semantics kept:
!   * read_emt reads one block from the open unit, after read_pes consumed
!     its "pes emt" line: the interaction line "A B proj|latt proj|latt",
!     then key value lines until a blank line or the end of the file; per
!     element 7 parameters (eta2 n0 E0 lambda V0 kappa s0), the first 7 lines
!     of a mixed block to the first element, the next 7 to the second
!     (:133-179); pes(idx1,idx2) = pes(idx2,idx1) = 2003, and a pair defined
!     twice stops the program (:98-101);
!   * compute_emt adds E_b to epot(b) and F to f(:,b,i) and sets the module
!     array dens(b,i) to the embedded density of every EMT atom, which the
!     propagators read as the LDFA input (src/md_algo.f90:154,512-558); the
!     parameters read so far go to the GPU at the first compute_emt after a
!     read_emt (a mixed block may precede the block that marks its metal
!     as EMT, so each block alone need not be a complete set);
!   * dens is (nbeads, natoms) of the universe last computed: allocated at
!     the first call (the reference allocates it in read_emt, :72) and
!     reallocated when a universe of another shape comes;
!   * flag is ignored (forces are always computed), as in the reference.
! read_emt_gpu(atoms, pes_file) runs read_pes's loop over the "pes emt"
! blocks of a file; compute_emt_gpu is compute_emt.  Errors stop the
! program with the library's message, as the reference stops.
module pes_emt_gpu_mod

    use, intrinsic :: iso_c_binding
    use mdt_gpu
    use universe_mod, only : universe

    implicit none

    private
    public :: dens, emt_gpu_init, emt_gpu_finalize, read_emt_gpu, compute_emt_gpu, emt_gpu_context
    public :: emt_params_from_file, read_emt, compute_emt

    integer, parameter :: dp = c_double
    real(dp), allocatable :: dens(:,:)     ! (nbeads, natoms), as pes_emt_mod's dens
    type(c_ptr), save :: ctx = c_null_ptr
    ! parameters per type read by read_emt (pes_emt_mod's pes_emt), not yet
    ! on the GPU while params_dirty
    real(dp), allocatable, save :: p_eta2(:), p_n0(:), p_e0(:), p_lambda(:), p_v0(:), p_kappa(:), p_s0(:)
    logical, save :: params_dirty = .false.

contains

    subroutine emt_gpu_init(device)
        integer, intent(in) :: device
        if (c_associated(ctx)) return
        call mdt_check(mdt_create(int(device, c_int), ctx), "emt_gpu_init")
    end subroutine emt_gpu_init

    subroutine emt_gpu_finalize()
        integer(c_int) :: rc
        if (.not. c_associated(ctx)) return
        rc = mdt_destroy(ctx)
        ctx = c_null_ptr
    end subroutine emt_gpu_finalize

    function emt_gpu_context() result(c)
        type(c_ptr) :: c
        c = ctx
    end function emt_gpu_context

    ! Reads every "pes emt" block of pes_file into per-type parameter arrays
    ! for the types of atoms (name + proj/latt role) and marks the EMT pairs
    ! in atoms%pes.
    subroutine emt_params_from_file(atoms, pes_file, eta2, n0, e0, lambda, v0, kappa, s0)
        type(universe), intent(inout) :: atoms
        character(len=*), intent(in) :: pes_file
        real(dp), intent(out) :: eta2(:), n0(:), e0(:), lambda(:), v0(:), kappa(:), s0(:)
        character(len=1000) :: line
        character(len=64) :: w(8)
        integer :: unit, ios, nw, idx1, idx2, counter, tgt
        real(dp) :: val
        eta2 = 0; n0 = 0; e0 = 0; lambda = 0; v0 = 0; kappa = 0; s0 = 0
        open(newunit=unit, file=pes_file, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            print '(a,a)', "Error in read_emt_gpu: cannot open ", trim(pes_file)
            stop 1
        end if
        do
            read(unit, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call split(line, w, nw)
            if (nw < 2) cycle
            if (lower(w(1)) /= "pes") cycle
            if (lower(w(2)) /= "emt") cycle
            read(unit, '(a)', iostat=ios) line
            call split(line, w, nw)
            if (ios /= 0 .or. nw /= 4) then
                print '(a)', "Error in read_emt_gpu: need four entries in interaction-defining lines"
                stop 1
            end if
            idx1 = type_index(atoms, w(1), w(3))
            idx2 = type_index(atoms, w(2), w(4))
            atoms%pes(idx1, idx2) = MDT_PES_EMT
            atoms%pes(idx2, idx1) = MDT_PES_EMT
            counter = 1
            do
                read(unit, '(a)', iostat=ios) line
                if (ios /= 0) exit
                call split(line, w, nw)
                if (nw == 0) exit
                if (nw /= 2) then
                    print '(a)', "Error in the PES file: PES parameters must consist of key value pairs."
                    stop 1
                end if
                read(w(2), *) val
                tgt = idx1
                if (counter > 7) tgt = idx2
                select case (lower(w(1)))
                case ("eta2"); eta2(tgt) = val
                case ("n0"); n0(tgt) = val
                case ("e0"); e0(tgt) = val
                case ("lambda"); lambda(tgt) = val
                case ("v0"); v0(tgt) = val
                case ("kappa"); kappa(tgt) = val
                case ("s0"); s0(tgt) = val
                case default
                    print '(a,a)', "Error in the PES file: unknown EMT parameter ", trim(w(1))
                    stop 1
                end select
                counter = counter + 1
            end do
        end do
        close(unit)
    end subroutine emt_params_from_file

    ! read_emt (src/pes_emt_mod.f90:49-190): one "pes emt" block from the
    ! open unit inp_unit, positioned after its "pes emt" line
    subroutine read_emt(atoms, inp_unit)
        type(universe), intent(inout) :: atoms
        integer, intent(in) :: inp_unit
        character(len=1000) :: line
        character(len=64) :: w(8)
        integer :: ios, nw, idx1, idx2, counter, tgt
        real(dp) :: val
        call emt_gpu_init(0)
        if (allocated(p_eta2)) then
            if (size(p_eta2) /= atoms%ntypes) deallocate(p_eta2, p_n0, p_e0, p_lambda, p_v0, p_kappa, p_s0)
        end if
        if (.not. allocated(p_eta2)) then
            allocate(p_eta2(atoms%ntypes), p_n0(atoms%ntypes), p_e0(atoms%ntypes), p_lambda(atoms%ntypes), &
                     p_v0(atoms%ntypes), p_kappa(atoms%ntypes), p_s0(atoms%ntypes))
            p_eta2 = 0; p_n0 = 0; p_e0 = 0; p_lambda = 0; p_v0 = 0; p_kappa = 0; p_s0 = 0
        end if
        read(inp_unit, '(a)', iostat=ios) line
        call split(line, w, nw)
        if (ios /= 0 .or. nw /= 4) then
            print '(a)', "Error in read_emt: need four entries in interaction-defining lines"
            stop 1
        end if
        idx1 = type_index(atoms, w(1), w(3))
        idx2 = type_index(atoms, w(2), w(4))
        if (atoms%pes(idx1, idx2) /= 0) then
            print '(a,4(1x,a))', "Error in read_emt: pes already defined for atoms", trim(w(1)), trim(w(3)), &
                trim(w(2)), trim(w(4))
            stop 1
        end if
        atoms%pes(idx1, idx2) = MDT_PES_EMT
        atoms%pes(idx2, idx1) = MDT_PES_EMT
        counter = 1
        do
            read(inp_unit, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call split(line, w, nw)
            if (nw == 0) exit
            if (nw /= 2) then
                print '(a)', "Error in the PES file: PES parameters must consist of key value pairs."
                stop 1
            end if
            read(w(2), *) val
            tgt = idx1
            if (counter > 7) tgt = idx2
            select case (lower(w(1)))
            case ("eta2"); p_eta2(tgt) = val
            case ("n0"); p_n0(tgt) = val
            case ("e0"); p_e0(tgt) = val
            case ("lambda"); p_lambda(tgt) = val
            case ("v0"); p_v0(tgt) = val
            case ("kappa"); p_kappa(tgt) = val
            case ("s0"); p_s0(tgt) = val
            case default
                print '(a,a)', "Error in the PES file: unknown EMT parameter ", trim(w(1))
                stop 1
            end select
            counter = counter + 1
        end do
        params_dirty = .true.
    end subroutine read_emt

    ! read_pes's loop (src/md_init.f90:99-144) over the "pes emt" blocks of
    ! pes_file; other blocks are skipped
    subroutine read_emt_gpu(atoms, pes_file)
        type(universe), intent(inout) :: atoms
        character(len=*), intent(in) :: pes_file
        character(len=1000) :: line
        character(len=64) :: w(8)
        integer :: unit, ios, nw
        open(newunit=unit, file=pes_file, status="old", action="read", iostat=ios)
        if (ios /= 0) then
            print '(a,a)', "Error in read_emt_gpu: cannot open ", trim(pes_file)
            stop 1
        end if
        do
            read(unit, '(a)', iostat=ios) line
            if (ios /= 0) exit
            call split(line, w, nw)
            if (nw == 2) then
                if (lower(w(1)) == "pes" .and. lower(w(2)) == "emt") call read_emt(atoms, unit)
            end if
        end do
        close(unit)
    end subroutine read_emt_gpu

    ! compute_emt (src/pes_emt_mod.f90:194-225) as calc_force calls it
    ! (src/force.f90:79)
    subroutine compute_emt(atoms, flag)
        type(universe), intent(inout) :: atoms
        integer, intent(in) :: flag
        integer(c_int32_t), allocatable :: idx(:)
        integer(c_int32_t), allocatable :: pesm(:,:)
        if (.not. c_associated(ctx)) then
            print '(a)', "Error in compute_emt: no EMT parameters (read_emt was not called)"
            stop 1
        end if
        if (params_dirty) then
            pesm = int(atoms%pes, c_int32_t)
            call mdt_check(mdt_emt_params_set(ctx, int(atoms%ntypes, c_int), p_eta2, p_n0, p_e0, p_lambda, p_v0, &
                                              p_kappa, p_s0, pesm), "compute_emt")
            params_dirty = .false.
        end if
        if (allocated(dens)) then
            if (size(dens, 1) /= atoms%nbeads .or. size(dens, 2) /= atoms%natoms) deallocate(dens)
        end if
        if (.not. allocated(dens)) then
            allocate(dens(atoms%nbeads, atoms%natoms))
            dens = 0.0_dp
        end if
        idx = int(atoms%idx, c_int32_t)
        call mdt_check(mdt_compute_emt(ctx, int(atoms%natoms, c_int), int(atoms%nbeads, c_int), idx, &
                                       atoms%simbox, atoms%isimbox, atoms%r, atoms%epot, atoms%f, dens, &
                                       int(flag, c_int32_t)), "compute_emt")
    end subroutine compute_emt

    subroutine compute_emt_gpu(atoms, flag)
        type(universe), intent(inout) :: atoms
        integer, intent(in) :: flag
        call compute_emt(atoms, flag)
    end subroutine compute_emt_gpu

    integer function type_index(atoms, name, role) result(idx)
        type(universe), intent(in) :: atoms
        character(len=*), intent(in) :: name, role
        integer :: t
        logical :: proj
        if (role /= "proj" .and. role /= "latt") then
            print '(a)', "Error in read_emt_gpu: interaction must be defined via 'proj' and 'latt' keywords"
            stop 1
        end if
        proj = (role == "proj")
        idx = 0
        do t = 1, atoms%ntypes
            if (trim(atoms%name(t)) == trim(name) .and. (atoms%is_proj(t) .eqv. proj)) idx = t
        end do
        if (idx == 0) then
            print '(a,a,a,a)', "Error in read_emt_gpu: no type ", trim(name), " ", trim(role)
            stop 1
        end if
    end function type_index

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

    function lower(s) result(t)
        character(len=*), intent(in) :: s
        character(len=len(s)) :: t
        integer :: i, c
        t = s
        do i = 1, len(s)
            c = iachar(s(i:i))
            if (c >= 65 .and. c <= 90) t(i:i) = achar(c + 32)
        end do
        t = trim(t)
    end function lower

end module pes_emt_gpu_mod
