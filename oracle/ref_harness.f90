! TEST INFRASTRUCTURE ONLY -- never part of the product path.
!
! C-ABI harness around the reference's own Fortran routines, compiled from the
! reference sources by oracle/Makefile into oracle/_ref/libwtref.so.  It lets
! the tests and the golden-vector generator call, unchanged:
!   ham_qlayer2qlayer  (src/ham_qlayer2qlayer.f90:1-102)
!   surfgreen_1985     (src/surfgreen.f90:7-137)
!   inv                (src/inverse.f90:1-58)
!   mat_mul            (src/mat_mul.f90:3-27)
! The model is placed in the module `para` globals exactly where
! get_hmnr_cell leaves it (src/readHmnR.f90:578-766): HmnR_surfacecell already
! divided by ndegen, ndegen_surfacecell = 1.
module wtref_harness
   use iso_c_binding
   use para, only: Num_wann, Np, Ndim, ijmax, nrpts_surfacecell, &
                   HmnR_surfacecell, irvec_surfacecell, ndegen_surfacecell
   implicit none
   external :: ham_qlayer2qlayer, surfgreen_1985, inv, mat_mul

contains

   subroutine wtref_set_model(nw, np_in, nr, ijmax_in, irvec, hr) &
         bind(C, name='wtref_set_model')
      integer(c_int), value, intent(in) :: nw, np_in, nr, ijmax_in
      integer(c_int), intent(in) :: irvec(3, nr)
      complex(c_double_complex), intent(in) :: hr(nw, nw, nr)

      Num_wann = nw
      Np = np_in
      Ndim = nw*np_in
      ijmax = ijmax_in
      nrpts_surfacecell = nr
      if (allocated(HmnR_surfacecell)) deallocate(HmnR_surfacecell)
      if (allocated(irvec_surfacecell)) deallocate(irvec_surfacecell)
      if (allocated(ndegen_surfacecell)) deallocate(ndegen_surfacecell)
      allocate(HmnR_surfacecell(nw, nw, nr))
      allocate(irvec_surfacecell(3, nr))
      allocate(ndegen_surfacecell(nr))
      HmnR_surfacecell = hr
      irvec_surfacecell = irvec
      ndegen_surfacecell = 1
   end subroutine wtref_set_model

   subroutine wtref_ham_qlayer2qlayer(k, h00, h01) bind(C, name='wtref_ham_qlayer2qlayer')
      real(c_double), intent(in) :: k(2)
      complex(c_double_complex), intent(out) :: h00(Ndim, Ndim), h01(Ndim, Ndim)
      call ham_qlayer2qlayer(k, h00, h01)
   end subroutine wtref_ham_qlayer2qlayer

   subroutine wtref_surfgreen_1985(omega, eta, h00, h01, gll, grr, gb) &
         bind(C, name='wtref_surfgreen_1985')
      real(c_double), value, intent(in) :: omega, eta
      complex(c_double_complex), intent(in) :: h00(Ndim, Ndim), h01(Ndim, Ndim)
      complex(c_double_complex), intent(out) :: gll(Ndim, Ndim), grr(Ndim, Ndim), gb(Ndim, Ndim)
      complex(c_double_complex), allocatable :: ones(:, :)
      integer :: i
      allocate(ones(Ndim, Ndim))
      ones = (0d0, 0d0)
      do i = 1, Ndim
         ones(i, i) = (1d0, 0d0)
      end do
      call surfgreen_1985(omega, gll, grr, gb, h00, h01, ones, eta)
      deallocate(ones)
   end subroutine wtref_surfgreen_1985

   subroutine wtref_inv(n, a) bind(C, name='wtref_inv')
      integer(c_int), value, intent(in) :: n
      complex(c_double_complex), intent(inout) :: a(n, n)
      integer :: nn
      nn = n
      call inv(nn, a)
   end subroutine wtref_inv

   subroutine wtref_mat_mul(n, a, b, c) bind(C, name='wtref_mat_mul')
      integer(c_int), value, intent(in) :: n
      complex(c_double_complex), intent(in) :: a(n, n), b(n, n)
      complex(c_double_complex), intent(out) :: c(n, n)
      integer :: nn
      nn = n
      call mat_mul(nn, a, b, c)
   end subroutine wtref_mat_mul

end module wtref_harness
