! universe_mod.f90 -- the state container the drop-in compiles against.
!
! Only the data layout, with the component names, kinds and ranks of
! md_tian2's type universe (src/universe_mod.f90:38-67), so that
! pes_emt_gpu_mod compiles here unchanged and, inside md_tian2, against the
! original module of the same name.  Arrays are (3, nbeads, natoms) column
! major; the mass is per atom; type properties (pes, algo, is_proj, name)
! are indexed by the atom's type idx.
module universe_mod

    use, intrinsic :: iso_fortran_env, only : real64

    implicit none

    integer, parameter, private :: dp = real64

    type universe
        integer :: natoms
        integer :: nbeads
        integer :: ntypes
        real(dp), allocatable :: r(:,:,:)
        real(dp), allocatable :: v(:,:,:)
        real(dp), allocatable :: f(:,:,:)
        real(dp), allocatable :: a(:,:,:)
        real(dp), allocatable :: m(:)
        logical,  allocatable :: is_fixed(:,:,:)
        integer,  allocatable :: idx(:)
        character(len=3), allocatable :: name(:)
        integer,  allocatable :: pes(:,:)
        integer,  allocatable :: algo(:)
        logical,  allocatable :: is_proj(:)
        real(dp), allocatable :: epot(:)
        integer :: dof
        real(dp) :: simbox(3,3)
        real(dp) :: isimbox(3,3)
        logical :: is_cart
    end type universe

contains

    ! allocation of one universe (the role of new_atoms, src/universe_mod.f90:77-115)
    function universe_alloc(nbeads, natoms, ntypes) result(u)
        integer, intent(in) :: nbeads, natoms, ntypes
        type(universe) :: u
        u%natoms = natoms
        u%nbeads = nbeads
        u%ntypes = ntypes
        allocate(u%r(3, nbeads, natoms), u%v(3, nbeads, natoms), u%f(3, nbeads, natoms))
        allocate(u%a(3, nbeads, natoms), u%m(natoms), u%is_fixed(3, nbeads, natoms))
        allocate(u%idx(natoms), u%name(ntypes), u%pes(ntypes, ntypes), u%algo(ntypes))
        allocate(u%is_proj(ntypes), u%epot(nbeads))
        u%r = 0.0_dp
        u%v = 0.0_dp
        u%f = 0.0_dp
        u%a = 0.0_dp
        u%m = 0.0_dp
        u%is_fixed = .false.
        u%idx = 0
        u%name = ""
        u%pes = 0
        u%algo = 0
        u%is_proj = .false.
        u%epot = 0.0_dp
        u%dof = 0
        u%simbox = 0.0_dp
        u%isimbox = 0.0_dp
        u%is_cart = .true.
    end function universe_alloc

end module universe_mod
