! GPU drop-in for WannierTools' surface Green's function path.
!
! External subroutines with the reference's names and signatures, reading the
! same module-`para` state and writing the same files, computed on an MI355X
! through libwtsurf (C ABI in include/wtsurf.h, bound in wtsurf_cabi.f90):
!
!   ham_qlayer2qlayer(k, H00new, H01new)      src/ham_qlayer2qlayer.f90:1-102
!   surfgreen_1985(omega, GLL, GRR, GB, H00, H01, ones, eta_broadening)
!                                             src/surfgreen.f90:7-137
!   surfstat                                  src/surfstat.f90:1-605
!   surfstat_jdos                             src/surfstat.f90:606-962
!   SurfaceDOSkk                              src/fermiarc.f90:1-994
!
! ham_qlayer2qlayer and surfgreen_1985 are per-call, for callers that loop
! themselves.  surfstat and SurfaceDOSkk submit their whole (k, omega) grid to
! the device in one call.  This file compiles against the reference's
! module.f90 (para, wmpi) or against para_subset.f90.
!
! Not covered here: the phonon LO-TO branch (the drivers stop with a message
! where the reference would need it) and the gnuplot scripts the reference
! writes beside its data files.
module wtsurf_gpu
   use iso_c_binding
   use wtsurf_cabi
   implicit none
   private
   public :: wtsurf_gpu_ctx, wtsurf_gpu_upload_model, wtsurf_gpu_check, wtsurf_gpu_finalize
   public :: wtsurf_gpu_precision

   type(c_ptr), save :: ctx = c_null_ptr
   logical, save :: model_on_device = .false.
   !> FP64 (the reference's arithmetic) unless the caller selects otherwise
   integer(c_int), save :: wtsurf_gpu_precision = WTSURF_FP64

contains

   !> The context on HIP device WTSURF_DEVICE (default 0), created on first use.
   function wtsurf_gpu_ctx() result(c)
      type(c_ptr) :: c
      integer :: dev, stat
      character(len=16) :: val
      if (.not. c_associated(ctx)) then
         dev = 0
         call get_environment_variable('WTSURF_DEVICE', val, status=stat)
         if (stat == 0) read(val, *, iostat=stat) dev
         if (wtsurf_ctx_create(int(dev, c_int), ctx) /= 0) then
            write(*, '(a, i0)') ' >>> ERROR: no usable HIP device for libwtsurf, device ', dev
            stop 1
         end if
      end if
      c = ctx
   end function wtsurf_gpu_ctx

   subroutine wtsurf_gpu_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= 0) then
         write(*, '(4a)') ' >>> ERROR in ', what, ': ', wtsurf_error_message(ctx)
         stop 1
      end if
   end subroutine wtsurf_gpu_check

   !> Push HmnR_surfacecell / irvec_surfacecell (divided by ndegen_surfacecell,
   !> as ham_qlayer2qlayer applies it, src/ham_qlayer2qlayer.f90:58-60) and the
   !> surface orbital lists to the device.  force=.false. keeps a model that is
   !> already there.
   subroutine wtsurf_gpu_upload_model(force)
      use para
      logical, intent(in) :: force
      complex(dp), allocatable :: hr(:, :, :)
      integer :: ir
      type(c_ptr) :: c
      if (model_on_device .and. .not. force) return
      c = wtsurf_gpu_ctx()
      allocate(hr(Num_wann, Num_wann, nrpts_surfacecell))
      do ir = 1, nrpts_surfacecell
         hr(:, :, ir) = HmnR_surfacecell(:, :, ir)/ndegen_surfacecell(ir)
      end do
      call wtsurf_gpu_check(wtsurf_set_model(c, int(Num_wann, c_int), int(Np, c_int), &
         int(nrpts_surfacecell, c_int), int(ijmax, c_int), irvec_surfacecell, hr), 'wtsurf_set_model')
      deallocate(hr)
      if (allocated(TopOrbitals) .and. allocated(BottomOrbitals)) then
         call wtsurf_gpu_check(wtsurf_set_orbitals(c, int(NtopOrbitals, c_int), TopOrbitals, &
            int(NBottomOrbitals, c_int), BottomOrbitals), 'wtsurf_set_orbitals')
      end if
      model_on_device = .true.
   end subroutine wtsurf_gpu_upload_model

   subroutine wtsurf_gpu_finalize()
      if (c_associated(ctx)) call wtsurf_ctx_destroy(ctx)
      ctx = c_null_ptr
      model_on_device = .false.
   end subroutine wtsurf_gpu_finalize

end module wtsurf_gpu


!> H00(k), H01(k) between principal layers (src/ham_qlayer2qlayer.f90:1-102).
subroutine ham_qlayer2qlayer(k, H00new, H01new)
   use para
   use wtsurf_gpu
   use wtsurf_cabi
   implicit none
   real(Dp), intent(in) :: k(2)
   complex(Dp), intent(out) :: H00new(Ndim, Ndim)
   complex(Dp), intent(out) :: H01new(Ndim, Ndim)
   real(Dp) :: kk(2, 1)
   call wtsurf_gpu_upload_model(.false.)
   kk(:, 1) = k
   call wtsurf_gpu_check(wtsurf_ham_qlayer2qlayer(wtsurf_gpu_ctx(), 1_c_int, kk, H00new, H01new), &
      'wtsurf_ham_qlayer2qlayer')
end subroutine ham_qlayer2qlayer


!> Lopez-Sancho decimation for one (omega, eta) (src/surfgreen.f90:7-137).
!> `ones` is accepted for the signature; the identity is implied.
subroutine surfgreen_1985(omega, GLL, GRR, GB, H00, H01, ones, eta_broadening)
   use para
   use wtsurf_gpu
   use wtsurf_cabi
   implicit none
   real(Dp), intent(in) :: omega
   real(Dp), intent(in) :: eta_broadening
   complex(Dp), intent(in) :: H00(Ndim, Ndim)
   complex(Dp), intent(in) :: H01(Ndim, Ndim)
   complex(Dp), intent(in) :: ones(Ndim, Ndim)
   complex(Dp), intent(inout) :: GLL(Ndim, Ndim)
   complex(Dp), intent(inout) :: GRR(Ndim, Ndim)
   complex(Dp), intent(inout) :: GB(Ndim, Ndim)
   integer(c_int) :: iters
   call wtsurf_gpu_check(wtsurf_surfgreen_1985(wtsurf_gpu_ctx(), int(Ndim, c_int), omega, &
      eta_broadening, H00, H01, GLL, GRR, GB, iters), 'wtsurf_surfgreen_1985')
end subroutine surfgreen_1985


!> Surface spectral function along KPATH_SLAB (src/surfstat.f90:1-605):
!> eta = 3 (OmegaMax-OmegaMin)/OmegaNum, omega_j = OmegaMin + (j-1)(OmegaMax-OmegaMin)/OmegaNum,
!> dos_l/dos_r/dos_bulk traces, log and *_only post-processing, the spin-resolved
!> traces (GRR's over the TopOrbitals indices, as the reference sums them), and
!> the files dos.dat_l, dos.dat_r, dos.dat_bulk, spindos.dat_l, spindos.dat_r
!> in the reference's formats.
subroutine surfstat
   use wmpi
   use para
   use wtsurf_gpu
   use wtsurf_cabi
   implicit none
   integer :: i, j, ikp, doslfile, dosrfile, dosbulkfile, spindoslfile, spindosrfile
   real(dp) :: eta_broadening, s0(3), s1(3)
   real(dp), allocatable :: omega(:), kpts(:, :)
   real(dp), allocatable :: dos_l(:, :), dos_r(:, :), dos_bulk(:, :)
   real(dp), allocatable :: dos_l_only(:, :), dos_r_only(:, :)
   real(dp), allocatable :: dos_l_raw(:, :), dos_r_raw(:, :), dos_bulk_raw(:, :)
   real(dp), allocatable :: spin_l(:, :, :), spin_r(:, :, :)
   integer(c_int), allocatable :: iters(:, :), status(:, :)

   if (index(Particle, 'phonon') /= 0 .and. LOTO_correction) then
      write(*, '(a)') ' >>> ERROR: the GPU surfstat has no LO-TO corrected phonon Hamiltonian'
      stop 1
   end if
   if (cpuid /= 0) return  ! one process drives the GPU

   allocate(omega(omeganum), kpts(2, knv2))
   allocate(dos_l(knv2, omeganum), dos_r(knv2, omeganum), dos_bulk(knv2, omeganum))
   allocate(dos_l_only(knv2, omeganum), dos_r_only(knv2, omeganum))
   allocate(dos_l_raw(knv2, omeganum), dos_r_raw(knv2, omeganum), dos_bulk_raw(knv2, omeganum))
   allocate(iters(knv2, omeganum), status(knv2, omeganum))
   allocate(spin_l(knv2, omeganum, 3), spin_r(knv2, omeganum, 3))

   eta_broadening = (omegamax - omegamin)/dble(omeganum)*3.0d0
   do i = 1, omeganum
      omega(i) = omegamin + (i - 1)*(omegamax - omegamin)/dble(omeganum)
   end do
   if (index(Particle, 'phonon') /= 0) omega = omega*omega
   do ikp = 1, knv2
      kpts(:, ikp) = k2_path(ikp, :)
   end do

   call wtsurf_gpu_upload_model(.true.)
   ! spin-resolved traces of src/surfstat.f90:163-192 from the same GLL, GRR
   call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_SURFSTAT), &
      'wtsurf_set_spin_traces')
   call wtsurf_gpu_check(wtsurf_spectral_grid(wtsurf_gpu_ctx(), int(knv2, c_int), kpts, &
      int(omeganum, c_int), omega, eta_broadening, 1.0d0, wtsurf_gpu_precision, &
      dos_l_raw, dos_r_raw, dos_bulk_raw, iters, status), 'wtsurf_spectral_grid')
   call wtsurf_gpu_check(wtsurf_spin_download(wtsurf_gpu_ctx(), spin_l, spin_r), &
      'wtsurf_spin_download')
   call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_OFF), &
      'wtsurf_set_spin_traces')
   if (any(iand(status, WTSURF_STATUS_SINGULAR) /= 0)) print *, 'something wrong with zgesv'

   ! src/surfstat.f90:237-247
   dos_l = log(abs(dos_l_raw))
   dos_r = log(abs(dos_r_raw))
   dos_bulk = log(abs(dos_bulk_raw) + eps9)
   do ikp = 1, knv2
      do j = 1, omeganum
         dos_l_only(ikp, j) = dos_l_raw(ikp, j) - dos_bulk_raw(ikp, j)
         if (dos_l_only(ikp, j) < 0) dos_l_only(ikp, j) = eps9
         dos_r_only(ikp, j) = dos_r_raw(ikp, j) - dos_bulk_raw(ikp, j)
         if (dos_r_only(ikp, j) < 0) dos_r_only(ikp, j) = eps9
      end do
   end do

   if (index(Particle, 'phonon') /= 0) omega = sqrt(omega)

   outfileindex = outfileindex + 1
   doslfile = outfileindex
   outfileindex = outfileindex + 1
   dosrfile = outfileindex
   outfileindex = outfileindex + 1
   dosbulkfile = outfileindex
   outfileindex = outfileindex + 1
   spindoslfile = outfileindex
   outfileindex = outfileindex + 1
   spindosrfile = outfileindex
   open(unit=doslfile, file='dos.dat_l')
   open(unit=dosrfile, file='dos.dat_r')
   open(unit=dosbulkfile, file='dos.dat_bulk')
   open(unit=spindoslfile, file='spindos.dat_l')
   open(unit=spindosrfile, file='spindos.dat_r')
   write(doslfile, '("#", a12, 6a17)') ' k(1/Ang)', ' E(eV)', 'dos_l', 'dos_l_only'
   write(dosrfile, '("#", a12, 6a17)') ' k(1/Ang)', ' E(eV)', 'dos_r', 'dos_r_only'
   write(dosbulkfile, '("#", a12, 6a17)') ' k(1/Ang)', ' E(eV)', 'dos_bulk'
   ! src/surfstat.f90:280-285
   write(spindoslfile, '("#", a)') ' spin dos_l, the axis is rotated as '
   write(spindoslfile, '("#", a)') " x is along R1', z is along R1'xR2', y is along z x y"
   write(spindoslfile, '("#", a12, 6a17)') ' k(1/Ang)', ' E(eV)', 'sx', 'sy', 'sz'
   write(spindosrfile, '("#", a)') ' spin dos_r, the axis is rotated as '
   write(spindosrfile, '("#", a)') " x is along R1', z is along R1'xR2', y is along z x y"
   write(spindosrfile, '("#", a12, 6a17)') ' k(1/Ang)', ' E(eV)', 'sx', 'sy', 'sz'
   do ikp = 1, knv2
      do j = 1, omeganum
         write(doslfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_l(ikp, j), dos_l_only(ikp, j)
         write(dosrfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_r(ikp, j), dos_r_only(ikp, j)
         write(dosbulkfile, 2003) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_bulk(ikp, j)
         s0 = spin_l(ikp, j, :)
         call rotate(s0, s1)
         write(spindoslfile, 2001) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, s1
         s0 = spin_r(ikp, j, :)
         call rotate(s0, s1)
         write(spindosrfile, 2001) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, s1
      end do
      write(doslfile, *)
      write(dosrfile, *)
      write(dosbulkfile, *)
      write(spindoslfile, *)
      write(spindosrfile, *)
   end do
   close(doslfile)
   close(dosrfile)
   close(dosbulkfile)
   close(spindoslfile)
   close(spindosrfile)
   write(stdout, *) 'ndim', ndim
   write(stdout, *) 'knv2,omeganum,eta_broadening', knv2, omeganum, eta_broadening/eV2Hartree
   write(stdout, *) 'calculate density of state successfully'

2001 format(5(1X, E16.8))
2002 format(4(1X, E16.8))
2003 format(3(1X, E16.8))
end subroutine surfstat


!> Surface spectral function and its k-path joint density of states
!> (src/surfstat.f90:606-962): omega_j = OmegaMin + (j-1) dOmega with
!> dOmega = (OmegaMax-OmegaMin)/OmegaNum, eta = 3 dOmega, the surfstat traces
!> and spin traces written raw (no log, no header) to dos.dat_* and
!> spindos.dat_*, and the jdos of dos_l, dos_r, dos_l_only, dos_r_only along the
!> path, computed on the GPU, to dos.jdat_l, dos.jdat_r.
subroutine surfstat_jdos
   use wmpi
   use para
   use wtsurf_gpu
   use wtsurf_cabi
   implicit none
   integer :: i, j, ikp, doslfile, dosrfile, dosbulkfile, spindoslfile, spindosrfile
   integer :: jdoslfile, jdosrfile
   real(dp) :: eta_broadening, s0(3), s1(3)
   real(dp), allocatable :: omega(:), kpts(:, :)
   real(dp), allocatable :: dos_l(:, :), dos_r(:, :), dos_bulk(:, :)
   real(dp), allocatable :: spin_l(:, :, :), spin_r(:, :, :)
   real(dp), allocatable :: fields(:, :, :), jdos(:, :, :)
   integer(c_int), allocatable :: iters(:, :), status(:, :)

   if (cpuid /= 0) return

   allocate(omega(omeganum), kpts(2, knv2))
   allocate(dos_l(knv2, omeganum), dos_r(knv2, omeganum), dos_bulk(knv2, omeganum))
   allocate(spin_l(knv2, omeganum, 3), spin_r(knv2, omeganum, 3))
   allocate(fields(knv2, omeganum, 4), jdos(knv2, omeganum, 4))
   allocate(iters(knv2, omeganum), status(knv2, omeganum))

   ! src/surfstat.f90:691-696 (the step is formed first, unlike surfstat)
   eta_broadening = (omegamax - omegamin)/dble(omeganum)
   do i = 1, omeganum
      omega(i) = omegamin + (i - 1)*eta_broadening
   end do
   eta_broadening = eta_broadening*3.0d0
   do ikp = 1, knv2
      kpts(:, ikp) = k2_path(ikp, :)
   end do

   call wtsurf_gpu_upload_model(.true.)
   call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_SURFSTAT), &
      'wtsurf_set_spin_traces')
   call wtsurf_gpu_check(wtsurf_spectral_grid(wtsurf_gpu_ctx(), int(knv2, c_int), kpts, &
      int(omeganum, c_int), omega, eta_broadening, 1.0d0, wtsurf_gpu_precision, &
      dos_l, dos_r, dos_bulk, iters, status), 'wtsurf_spectral_grid')
   call wtsurf_gpu_check(wtsurf_spin_download(wtsurf_gpu_ctx(), spin_l, spin_r), &
      'wtsurf_spin_download')
   call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_OFF), &
      'wtsurf_set_spin_traces')
   if (any(iand(status, WTSURF_STATUS_SINGULAR) /= 0)) print *, 'something wrong with zgesv'

   ! src/surfstat.f90:838-845
   fields(:, :, 1) = dos_l
   fields(:, :, 2) = dos_r
   do ikp = 1, knv2
      do j = 1, omeganum
         fields(ikp, j, 3) = dos_l(ikp, j) - dos_bulk(ikp, j)
         if (fields(ikp, j, 3) < 0) fields(ikp, j, 3) = eps9
         fields(ikp, j, 4) = dos_r(ikp, j) - dos_bulk(ikp, j)
         if (fields(ikp, j, 4) < 0) fields(ikp, j, 4) = eps9
      end do
   end do

   outfileindex = outfileindex + 1
   doslfile = outfileindex
   outfileindex = outfileindex + 1
   dosrfile = outfileindex
   outfileindex = outfileindex + 1
   dosbulkfile = outfileindex
   outfileindex = outfileindex + 1
   spindoslfile = outfileindex
   outfileindex = outfileindex + 1
   spindosrfile = outfileindex
   open(unit=doslfile, file='dos.dat_l')
   open(unit=dosrfile, file='dos.dat_r')
   open(unit=dosbulkfile, file='dos.dat_bulk')
   open(unit=spindoslfile, file='spindos.dat_l')
   open(unit=spindosrfile, file='spindos.dat_r')
   do ikp = 1, knv2
      do j = 1, omeganum
         write(doslfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_l(ikp, j), fields(ikp, j, 3)
         write(dosrfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_r(ikp, j), fields(ikp, j, 4)
         write(dosbulkfile, 2003) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, dos_bulk(ikp, j)
         s0 = spin_l(ikp, j, :)
         call rotate(s0, s1)
         write(spindoslfile, 2001) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, s1
         s0 = spin_r(ikp, j, :)
         call rotate(s0, s1)
         write(spindosrfile, 2001) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, s1
      end do
      write(doslfile, *)
      write(dosrfile, *)
      write(dosbulkfile, *)
      write(spindoslfile, *)
      write(spindosrfile, *)
   end do
   close(doslfile)
   close(dosrfile)
   close(dosbulkfile)
   close(spindoslfile)
   close(spindosrfile)

   ! jdos of dos_l, dos_r, dos_l_only, dos_r_only along the path (src/surfstat.f90:891-907)
   call wtsurf_gpu_check(wtsurf_kpath_jdos(wtsurf_gpu_ctx(), int(knv2, c_int), &
      int(omeganum, c_int), 4_c_int, fields, jdos), 'wtsurf_kpath_jdos')

   outfileindex = outfileindex + 1
   jdoslfile = outfileindex
   outfileindex = outfileindex + 1
   jdosrfile = outfileindex
   open(unit=jdoslfile, file='dos.jdat_l')
   open(unit=jdosrfile, file='dos.jdat_r')
   ! both headers name jdos_l, as the reference writes them (src/surfstat.f90:935-936)
   write(jdoslfile, '("#", a12, 3a17)') ' k(1/Ang)', ' E(eV)', 'jdos_l', 'jdos_l_only'
   write(jdosrfile, '("#", a12, 3a17)') ' k(1/Ang)', ' E(eV)', 'jdos_l', 'jdos_l_only'
   do ikp = 1, knv2
      do j = 1, omeganum
         write(jdoslfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, jdos(ikp, j, 1), jdos(ikp, j, 3)
         write(jdosrfile, 2002) k2len(ikp)*Angstrom2atomic, omega(j)/eV2Hartree, jdos(ikp, j, 2), jdos(ikp, j, 4)
      end do
      write(jdoslfile, *)
      write(jdosrfile, *)
   end do
   close(jdoslfile)
   close(jdosrfile)
   write(stdout, *)
   write(stdout, *) 'calculate joint density of state successfully'

2001 format(5(1X, E16.8))
2002 format(4(1X, E16.8))
2003 format(3(1X, E16.8))
end subroutine surfstat_jdos


!> Surface spectral function on KPLANE_SLAB at iso_energy (src/fermiarc.f90:1-494)
!> with its spin texture and QPI (src/fermiarc.f90:85-138, 250-586): k-plane with
!> Nk rounded up to odd (and, for QPI, the KPLANE_SLAB vectors rounded outwards),
!> eta = Fermi_broadening, traces divided by pi, the max/10 thresholds (which
!> also reset the spins), arc.dat_l, arc.dat_r, arc.dat_bulk; with SOC > 0 and
!> spin texture or QPI the spin traces (GRR's over the bottom orbitals) and
!> spindos.dat_l, spindos.dat_r; with QPI the jdos / jsdos autocorrelations on
!> the GPU and arc.jdat_l, arc.jdat_r (and with SOC > 0 arc.jsdat_l,
!> arc.jsdat_r), all in the reference's formats.
subroutine SurfaceDOSkk
   use wmpi
   use para
   use wtsurf_gpu
   use wtsurf_cabi
   implicit none
   integer :: i, j, ikp, nkx, nky, ik1, ik2, arclfile, arcrfile, arcbulkfile, spinfile, jfile
   real(dp) :: omega(1), eta_broadening, dos_l_max, dos_r_max, s0(3), s1(3)
   real(dp) :: K2D_vec_a(2), K2D_vec_b(2)
   logical :: spin, qpi
   integer, allocatable :: ik12(:, :)
   real(dp), allocatable :: k12(:, :), k12_shape(:, :)
   real(dp), allocatable :: dl(:), dr(:), db(:), sl(:, :), sr(:, :)
   real(dp), allocatable :: dos_l(:, :), dos_r(:, :), dos_bulk(:, :)
   real(dp), allocatable :: dos_l_only(:, :), dos_r_only(:, :)
   real(dp), allocatable :: spin_l(:, :, :), spin_r(:, :, :)
   real(dp), allocatable :: fields(:, :, :), corr(:, :), jdos(:, :), jsdos(:, :)
   integer(c_int), allocatable :: iters(:), status(:)

   qpi = SlabQPI_kplane_calc
   spin = Soc > 0 .and. (SlabSpintexture_calc .or. SlabQPI_kplane_calc)
   if (spin .and. index(Particle, 'phonon') /= 0) then
      write(*, '(a)') ' >>> ERROR: no spin texture for a phonon system'
      stop 1
   end if
   if (index(Particle, 'phonon') /= 0 .and. LOTO_correction) then
      write(*, '(a)') ' >>> ERROR: the GPU SurfaceDOSkk has no LO-TO corrected phonon Hamiltonian'
      stop 1
   end if
   if (cpuid /= 0) return

   nkx = Nk1
   nky = Nk2
   if (mod(Nk1, 2) == 0) nkx = Nk1 + 1
   if (mod(Nk2, 2) == 0) nky = Nk2 + 1
   ! QPI needs whole reciprocal vectors: each component rounded away from zero
   ! (src/fermiarc.f90:113-138)
   if (qpi) then
      do i = 1, 2
         if (K2D_vec1(i) > 0) then
            K2D_vec_a(i) = ceiling(K2D_vec1(i))
         else
            K2D_vec_a(i) = floor(K2D_vec1(i))
         end if
         if (K2D_vec2(i) > 0) then
            K2D_vec_b(i) = ceiling(K2D_vec2(i))
         else
            K2D_vec_b(i) = floor(K2D_vec2(i))
         end if
      end do
      write(stdout, '(a)') 'WARNING : KPLANE_SLAB vectors rounded to whole reciprocal vectors for QPI'
      write(stdout, '((a, 2f8.4))') 'The first modified vector in QPI: ', K2D_vec_a
      write(stdout, '((a, 2f8.4))') 'The second modified vector in QPI: ', K2D_vec_b
   else
      K2D_vec_a = K2D_vec1
      K2D_vec_b = K2D_vec2
   end if
   allocate(ik12(2, nkx*nky), k12(2, nkx*nky), k12_shape(2, nkx*nky))
   allocate(dl(nkx*nky), dr(nkx*nky), db(nkx*nky), iters(nkx*nky), status(nkx*nky))
   allocate(dos_l(nkx, nky), dos_r(nkx, nky), dos_bulk(nkx, nky))
   allocate(dos_l_only(nkx, nky), dos_r_only(nkx, nky))
   ikp = 0
   do i = 1, nkx
      do j = 1, nky
         ikp = ikp + 1
         ik12(1, ikp) = i
         ik12(2, ikp) = j
         k12(:, ikp) = K2D_start + (i - 1)*K2D_vec_a/dble(nkx - 1) + (j - 1)*K2D_vec_b/dble(nky - 1)
         k12_shape(:, ikp) = k12(1, ikp)*Ka2 + k12(2, ikp)*Kb2
      end do
   end do

   omega(1) = iso_energy
   eta_broadening = Fermi_broadening
   if (index(Particle, 'phonon') /= 0) omega = omega*omega

   call wtsurf_gpu_upload_model(.true.)
   if (spin) then
      allocate(sl(nkx*nky, 3), sr(nkx*nky, 3), spin_l(nkx, nky, 3), spin_r(nkx, nky, 3))
      call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_ARC), &
         'wtsurf_set_spin_traces')
   end if
   call wtsurf_gpu_check(wtsurf_spectral_grid(wtsurf_gpu_ctx(), int(nkx*nky, c_int), k12, &
      1_c_int, omega, eta_broadening, pi, wtsurf_gpu_precision, dl, dr, db, iters, status), &
      'wtsurf_spectral_grid')
   if (spin) then
      call wtsurf_gpu_check(wtsurf_spin_download(wtsurf_gpu_ctx(), sl, sr), 'wtsurf_spin_download')
      call wtsurf_gpu_check(wtsurf_set_spin_traces(wtsurf_gpu_ctx(), WTSURF_SPIN_OFF), &
         'wtsurf_set_spin_traces')
   end if
   if (any(iand(status, WTSURF_STATUS_SINGULAR) /= 0)) print *, 'something wrong with zgesv'
   do ikp = 1, nkx*nky
      ik1 = ik12(1, ikp)
      ik2 = ik12(2, ikp)
      dos_l(ik1, ik2) = dl(ikp)
      dos_r(ik1, ik2) = dr(ikp)
      dos_bulk(ik1, ik2) = db(ikp)
      if (spin) then
         spin_l(ik1, ik2, :) = sl(ikp, :)
         spin_r(ik1, ik2, :) = sr(ikp, :)
      end if
   end do

   ! src/fermiarc.f90:372-397: below a tenth of the maximum, dos_*_only and the
   ! spins are reset to eps9
   dos_l_max = maxval(dos_l)
   dos_r_max = maxval(dos_r)
   do ikp = 1, nkx*nky
      ik1 = ik12(1, ikp)
      ik2 = ik12(2, ikp)
      dos_l_only(ik1, ik2) = dos_l(ik1, ik2)
      if (dos_l_only(ik1, ik2) < dos_l_max/10d0) then
         dos_l_only(ik1, ik2) = eps9
         if (spin) spin_l(ik1, ik2, :) = eps9
      end if
      dos_r_only(ik1, ik2) = dos_r(ik1, ik2)
      if (dos_r_only(ik1, ik2) < dos_r_max/10d0) then
         dos_r_only(ik1, ik2) = eps9
         if (spin) spin_r(ik1, ik2, :) = eps9
      end if
   end do

   outfileindex = outfileindex + 100
   arclfile = outfileindex
   outfileindex = outfileindex + 1
   arcrfile = outfileindex
   outfileindex = outfileindex + 3
   arcbulkfile = outfileindex
   open(unit=arclfile, file='arc.dat_l')
   open(unit=arcrfile, file='arc.dat_r')
   open(unit=arcbulkfile, file='arc.dat_bulk')
   write(arclfile, '(a)') '# Surface density of states of bottom surface'
   write(arclfile, '(a)') '# The coordinate of k is redefined according to the SURFACE card'
   write(arclfile, '(a)') "# x axis is parallel to R1'"
   write(arclfile, '(a)') "# z axis is parallel to R1'xR2'"
   write(arclfile, '(a)') "# y axis is parallel to z x x"
   write(arclfile, '(30a16)') '#kx', 'ky', 'log(dos)'
   write(arcrfile, '(a)') '# Surface density of states of top surface'
   write(arcbulkfile, '(a)') '#Density of states of bulk system after projection onto the surface BZ'
   write(arcrfile, '(a)') '# The coordinate of k is redefined according to the SURFACE card'
   write(arcrfile, '(a)') "# x axis is parallel to R1'"
   write(arcrfile, '(a)') "# z axis is parallel to R1'xR2'"
   write(arcrfile, '(a)') "# y axis is parallel to z x x"
   write(arcrfile, '(30a16)') '#kx', 'ky', 'log(dos)'
   write(arcbulkfile, '(a)') '#Density of states of bulk system after projection onto the surface BZ'
   write(arcbulkfile, '(a)') '# The coordinate of k is redefined according to the SURFACE card'
   write(arcbulkfile, '(a)') "# x axis is parallel to R1'"
   write(arcbulkfile, '(a)') "# z axis is parallel to R1'xR2'"
   write(arcbulkfile, '(a)') "# y axis is parallel to z x x"
   write(arcbulkfile, '(30a16)') '#kx', 'ky', 'log(dos)'
   do ikp = 1, nkx*nky
      write(arclfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, log(dos_l(ik12(1, ikp), ik12(2, ikp)))
      if (mod(ikp, nky) == 0) write(arclfile, *) ' '
      write(arcrfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, log(dos_r(ik12(1, ikp), ik12(2, ikp)))
      if (mod(ikp, nky) == 0) write(arcrfile, *) ' '
      write(arcbulkfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, log(abs(dos_bulk(ik12(1, ikp), ik12(2, ikp))))
      if (mod(ikp, nky) == 0) write(arcbulkfile, *) ' '
   end do
   close(arclfile)
   close(arcrfile)
   close(arcbulkfile)

   ! spin texture: s / dos where dos_*_only > eps6, in the surface frame
   ! (src/fermiarc.f90:455-488)
   if (spin) then
      do i = 1, 2
         outfileindex = outfileindex + 1
         spinfile = outfileindex
         if (i == 1) open(unit=spinfile, file='spindos.dat_l')
         if (i == 2) open(unit=spinfile, file='spindos.dat_r')
         write(spinfile, '(a)') '# The coordinates of sx,sy,sz are redefined according to the SURFACE card'
         write(spinfile, '(a)') "# x axis is parallel to R1'"
         write(spinfile, '(a)') "# z axis is parallel to R1'xR2'"
         write(spinfile, '(a)') "# y axis is parallel to z x x"
         write(spinfile, '(30a16)') '#kx', 'ky', 'sx', 'sy', 'sz'
         do ikp = 1, nkx*nky
            ik1 = ik12(1, ikp)
            ik2 = ik12(2, ikp)
            if (i == 1) then
               if (.not. dos_l_only(ik1, ik2) > eps6) cycle
               s0 = spin_l(ik1, ik2, :)/dos_l(ik1, ik2)
            else
               if (.not. dos_r_only(ik1, ik2) > eps6) cycle
               s0 = spin_r(ik1, ik2, :)/dos_r(ik1, ik2)
            end if
            call rotate(s0, s1)
            write(spinfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, s1
         end do
         close(spinfile)
      end do
   end if
   write(stdout, *) 'Ndim: ', ndim
   write(stdout, *) 'Nk1,Nk2,eta_broadening: ', nkx, nky, eta_broadening
   write(stdout, *) 'Calculated surface density of state successfully'

   if (.not. qpi) return
   ! QPI (src/fermiarc.f90:496-586): jdos over dos_*_only with moduli nkx-1,
   ! nky-1, normalised by nkx nky; with SOC jsdos over (dos_*_only, sx, sy, sz)
   ! with moduli Nk1-1, Nk2-1, not normalised
   allocate(corr(nkx, nky), jdos(nkx*nky, 2))
   if (Soc > 0) allocate(fields(nkx, nky, 4), jsdos(nkx*nky, 2))
   do i = 1, 2
      if (i == 1) then
         call wtsurf_gpu_check(wtsurf_kplane_autocorrelation(wtsurf_gpu_ctx(), int(nkx, c_int), &
            int(nky, c_int), int(nkx - 1, c_int), int(nky - 1, c_int), 1_c_int, dos_l_only, corr), &
            'wtsurf_kplane_autocorrelation')
      else
         call wtsurf_gpu_check(wtsurf_kplane_autocorrelation(wtsurf_gpu_ctx(), int(nkx, c_int), &
            int(nky, c_int), int(nkx - 1, c_int), int(nky - 1, c_int), 1_c_int, dos_r_only, corr), &
            'wtsurf_kplane_autocorrelation')
      end if
      do ikp = 1, nkx*nky
         jdos(ikp, i) = corr(ik12(1, ikp), ik12(2, ikp))/nkx/nky
      end do
      if (Soc > 0) then
         if (i == 1) then
            fields(:, :, 1) = dos_l_only
            fields(:, :, 2:4) = spin_l
         else
            fields(:, :, 1) = dos_r_only
            fields(:, :, 2:4) = spin_r
         end if
         call wtsurf_gpu_check(wtsurf_kplane_autocorrelation(wtsurf_gpu_ctx(), int(nkx, c_int), &
            int(nky, c_int), int(Nk1 - 1, c_int), int(Nk2 - 1, c_int), 4_c_int, fields, corr), &
            'wtsurf_kplane_autocorrelation')
         do ikp = 1, nkx*nky
            jsdos(ikp, i) = corr(ik12(1, ikp), ik12(2, ikp))
         end do
      end if
   end do
   write(stdout, *) 'The calculation of joint density of state was done.'
   do i = 1, 2
      outfileindex = outfileindex + 1
      jfile = outfileindex
      if (i == 1) then
         open(unit=jfile, file='arc.jdat_l')
         write(jfile, '(a)') '# Surface joint density of states of bottom surface'
      else
         open(unit=jfile, file='arc.jdat_r')
         write(jfile, '(a)') '# Surface joint density of states of top surface'
      end if
      write(jfile, '(a)') '# The coordinate of k is redefined according to the SURFACE card'
      write(jfile, '(a)') "# x axis is parallel to R1'"
      write(jfile, '(a)') "# z axis is parallel to R1'xR2'"
      write(jfile, '(a)') "# y axis is parallel to z x x"
      write(jfile, '(30a16)') '#kx', 'ky', 'log(dos)'
      do ikp = 1, nkx*nky
         write(jfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, log(abs(jdos(ikp, i)))
         if (mod(ikp, nky) == 0) write(jfile, *) ' '
      end do
      close(jfile)
   end do
   if (Soc > 0) then
      do i = 1, 2
         outfileindex = outfileindex + 1
         jfile = outfileindex
         if (i == 1) then
            open(unit=jfile, file='arc.jsdat_l')
            write(jfile, '(a)') '# Surface joint density of states(in consideration of spin contribution) of bottom surface'
         else
            open(unit=jfile, file='arc.jsdat_r')
            write(jfile, '(a)') '# Surface joint density of states(in consideration of spin contribution) of top surface'
         end if
         write(jfile, '(a)') '# The coordinate of k is redefined according to the SURFACE card'
         write(jfile, '(a)') "# x axis is parallel to R1'"
         write(jfile, '(a)') "# z axis is parallel to R1'xR2'"
         write(jfile, '(a)') "# y axis is parallel to z x x"
         write(jfile, '(30a16)') '#kx', 'ky', 'log(dos)'
         do ikp = 1, nkx*nky
            write(jfile, '(30f16.8)') k12_shape(:, ikp)*Angstrom2atomic, log(abs(jsdos(ikp, i)))
            if (mod(ikp, nky) == 0) write(jfile, *) ' '
         end do
         close(jfile)
      end do
   end if
   write(stdout, *) 'calculate joint density of state successfully'
end subroutine SurfaceDOSkk
