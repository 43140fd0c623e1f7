! pes_nene_gpu_mod.f90 -- GPU drop-in for md_tian2's NENE plugin.
!
! read_nene_gpu(atoms, inp_path) and compute_nene_gpu(atoms, flag) take the
! place of read_nene (src/pes_nene_mod.f90:123-5759) and compute_nene
! (:7145-7368) behind calc_force's dispatch (src/force.f90:77-82):
!   * the network is read from inp_path's input.nn, scaling.data and
!     weights.XXX.data by the library (same keywords, symmetry functions in
!     the reference's sorted order);
!   * compute_nene_gpu adds E_b [eV] to epot(b) and F [eV/A] to f(:,b,i)
!     for every bead, as compute_nene does after its Hartree/Bohr
!     conversion (:7245-7250);
!   * symmetry functions outside the scaling range are counted, as the
!     reference's extrapolation warnings (:7260-7261), and do not stop.
! The context is the EMT module's, so both potentials share one GPU stream.
module pes_nene_gpu_mod

    use, intrinsic :: iso_c_binding
    use mdt_gpu
    use pes_emt_gpu_mod, only : emt_gpu_init, emt_gpu_context
    use universe_mod, only : universe

    implicit none

    private
    public :: read_nene_gpu, compute_nene_gpu, nene_extrapolation_warnings

    integer, parameter :: dp = c_double
    integer, save :: nene_extrapolation_warnings = 0

contains

    subroutine read_nene_gpu(atoms, inp_path)
        type(universe), intent(in) :: atoms
        character(len=*), intent(in) :: inp_path
        character(kind=c_char), allocatable :: dir(:), syms(:)
        integer :: i, t
        character(len=2) :: s2
        call emt_gpu_init(0)
        allocate(dir(len_trim(inp_path) + 1))
        do i = 1, len_trim(inp_path)
            dir(i) = inp_path(i:i)
        end do
        dir(len_trim(inp_path) + 1) = c_null_char
        allocate(syms(2 * atoms%ntypes))
        do t = 1, atoms%ntypes
            s2 = adjustl(atoms%name(t))
            syms(2 * t - 1) = s2(1:1)
            syms(2 * t) = s2(2:2)
        end do
        call mdt_check(mdt_nn_load(emt_gpu_context(), dir, int(atoms%ntypes, c_int), syms), "read_nene_gpu")
    end subroutine read_nene_gpu

    subroutine compute_nene_gpu(atoms, flag)
        type(universe), intent(inout) :: atoms
        integer, intent(in) :: flag
        integer(c_int32_t), allocatable :: idx(:)
        integer(c_int32_t) :: nex
        allocate(idx(atoms%natoms))
        idx = int(atoms%idx, c_int32_t)
        call mdt_check(mdt_compute_nene(emt_gpu_context(), int(atoms%natoms, c_int), int(atoms%nbeads, c_int), &
                                        idx, atoms%simbox, atoms%isimbox, atoms%r, atoms%epot, atoms%f, nex, &
                                        int(flag, c_int32_t)), "compute_nene_gpu")
        nene_extrapolation_warnings = nene_extrapolation_warnings + int(nex)
    end subroutine compute_nene_gpu

end module pes_nene_gpu_mod
