! The part of WannierTools' global state that the surface Green's function
! path reads, under the reference's own module and variable names
! (src/module.f90: prec :5-12, wmpi :14-338, para :340-1148).  The drop-in
! routines in wtsurf_dropin.f90 compile unchanged against either this file
! (standalone wt_surf program) or the reference's module.f90 (GPU drop-in in
! a WannierTools build).  Only names, kinds and shapes follow the reference;
! everything here is plain declarations.
module prec
   implicit none
   integer, parameter :: li = 4
   integer, parameter :: Dp = kind(1.0d0)
end module prec

module wmpi
   use prec
   implicit none
   integer :: cpuid = 0    ! rank; the standalone program is one process
   integer :: num_cpu = 1
end module wmpi

module para
   use prec
   implicit none

   integer, parameter :: stdout = 8
   integer, public, save :: outfileindex = 11932

   ! unit constants (src/module.f90:711-728); eps6 and eps9 are real(4) literals there
   real(dp), parameter :: Angstrom2atomic = 1d0/0.529177210903d0
   real(dp), parameter :: eV2Hartree = 1d0/27.211385d0
   real(dp), parameter :: Pi = 3.14159265358979d0
   real(dp), parameter :: eps6 = 1e-6
   real(dp), parameter :: eps9 = 1e-9
   complex(dp), parameter :: zi = (0.0d0, 1.0d0)

   ! &TB_FILE, &CONTROL, &SYSTEM, &PARAMETERS subset
   character(80) :: Hrfile = 'wannier90_hr.dat'
   character(80) :: Particle = 'electron'
   logical :: SlabSS_calc = .false., SlabArc_calc = .false.
   logical :: SlabSpintexture_calc = .false., SlabQPI_kplane_calc = .false.
   logical :: SlabQPI_calc = .false., SlabQPI_kpath_calc = .false.
   logical :: LOTO_correction = .false.
   integer :: Soc = 0
   real(Dp) :: E_fermi = 0d0
   real(Dp) :: Fermi_broadening = 0.005d0
   real(Dp) :: iso_energy = 0d0
   integer :: OmegaNum = 100
   real(dp) :: OmegaMin = -1d0, OmegaMax = 1d0
   integer :: Nk1 = 10, Nk2 = 10
   integer :: Np = 2

   ! model sizes and the surface-cell Hamiltonian (get_hmnr_cell output)
   integer :: Num_wann = 0
   integer :: Ndim = 0
   integer, public, save :: ijmax = 10
   integer :: nrpts_surfacecell = 0
   complex(dp), allocatable :: HmnR_surfacecell(:, :, :)
   integer, allocatable :: irvec_surfacecell(:, :)
   integer, allocatable :: ndegen_surfacecell(:)

   ! surface orbitals for the spectral traces
   integer :: NtopOrbitals = 0, NBottomOrbitals = 0
   integer, allocatable :: TopOrbitals(:)
   integer, allocatable :: BottomOrbitals(:)

   ! KPATH_SLAB and KPLANE_SLAB
   integer :: knv2 = 0
   real(dp), allocatable :: k2len(:)
   real(dp), allocatable :: k2_path(:, :)
   real(dp) :: K2D_start(2) = (/-0.5d0, -0.5d0/)
   real(dp) :: K2D_vec1(2) = (/1d0, 0d0/)
   real(dp) :: K2D_vec2(2) = (/0d0, 1d0/)

   ! surface reciprocal vectors (1/Bohr)
   real(Dp), public, save :: Ka2(2)
   real(Dp), public, save :: Kb2(2)

   ! rows: x along R1', z along R1' x R2', y = z x x (src/module.f90:839)
   real(dp), public, save :: Urot(3, 3)
end module para
