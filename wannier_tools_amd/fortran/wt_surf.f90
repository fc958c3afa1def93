! wt_surf: standalone driver for the surface Green's function path on the GPU.
!
! Run in a WannierTools input directory: reads wt.in (the part the path
! uses) and the Wannier90 *_hr.dat it names, then runs SlabSS_calc (surfstat,
! dos.dat_*, spindos.dat_*), SlabQPI_kpath_calc (surfstat_jdos, dos.jdat_*),
! SlabArc_calc, SlabQPI_calc / SlabQPI_kplane_calc
! and SlabSpintexture_calc (SurfaceDOSkk: arc.dat_*, arc.jdat_*, arc.jsdat_*,
! spindos.dat_*) exactly as the reference's main program dispatches them
! (src/main.f90:714-765), and reports the time of each in WT.out.
!
! Input subset (names, defaults and units as src/readinput.f90):
!   &TB_FILE    Hrfile, Particle
!   &CONTROL    SlabSS_calc, SlabArc_calc, SlabQPI_calc, SlabQPI_kplane_calc,
!               SlabQPI_kpath_calc, SlabSpintexture_calc (other flags are accepted
!               and ignored)
!   &SYSTEM     SOC, E_FERMI; Add_Zeeman_Field with Zeeman_energy_in_eV or
!               Effective_gfactor, Bmagnitude, Btheta, Bphi, Bx, By, Bz (SOC = 1);
!               settings that change H(R) in ways not built here (Zeeman with
!               SOC = 0, electric fields, sparse or non-orthogonal H(R)) stop
!               the run with a message
!   &PARAMETERS Eta_Arc, E_arc, OmegaNum, OmegaMin, OmegaMax, Nk1, Nk2, NP,
!               topsurface_atom_index
!   cards       LATTICE, ATOM_POSITIONS, PROJECTORS, SURFACE, KPATH_SLAB,
!               KPLANE_SLAB, SURFACE_ATOMS
! The SURFACE card defines the surface cell: R1', R2', R3' = U (a1, a2, a3) with
! U settled as the reference does (third row 0 0 1 unless given, handedness
! fix, search for R3' when the first two rows do not span a cell of the
! crystal's volume), and H(R) is remapped onto that cell (get_hmnr_cell,
! src/readHmnR.f90:578-766) before anything reaches the device.


!> The surface cell of the SURFACE card and H(R) remapped onto it (host-side
!> geometry, once per model): U with the reference's handedness fix and its
!> search for R3' (src/readinput.f90:1575-1668, 4247-4463), the atoms in the new
!> cell shifted so the top-surface atom sits at 0.95 along R3'
!> (src/readinput.f90:1749-1781), and get_hmnr_cell (src/readHmnR.f90:578-766).
module wt_surf_cell
   use para
   implicit none
   private
   public :: surface_umatrix, surface_cell_positions, remap_hr, inv3, cart

   integer, parameter :: max_ir = 8, search_rmax = 6

contains

   !> Inverse of a 3 x 3 matrix (cofactors).
   function inv3(a) result(b)
      real(dp), intent(in) :: a(3, 3)
      real(dp) :: b(3, 3), det
      b(1, 1) = a(2, 2)*a(3, 3) - a(2, 3)*a(3, 2)
      b(1, 2) = a(1, 3)*a(3, 2) - a(1, 2)*a(3, 3)
      b(1, 3) = a(1, 2)*a(2, 3) - a(1, 3)*a(2, 2)
      b(2, 1) = a(2, 3)*a(3, 1) - a(2, 1)*a(3, 3)
      b(2, 2) = a(1, 1)*a(3, 3) - a(1, 3)*a(3, 1)
      b(2, 3) = a(1, 3)*a(2, 1) - a(1, 1)*a(2, 3)
      b(3, 1) = a(2, 1)*a(3, 2) - a(2, 2)*a(3, 1)
      b(3, 2) = a(1, 2)*a(3, 1) - a(1, 1)*a(3, 2)
      b(3, 3) = a(1, 1)*a(2, 2) - a(1, 2)*a(2, 1)
      det = a(1, 1)*b(1, 1) + a(1, 2)*b(2, 1) + a(1, 3)*b(3, 1)
      b = b/det
   end function inv3

   !> R1 . (R2 x R3)
   real(dp) function triple(r1, r2, r3)
      real(dp), intent(in) :: r1(3), r2(3), r3(3)
      triple = r1(1)*(r2(2)*r3(3) - r2(3)*r3(2)) + r1(2)*(r2(3)*r3(1) - r2(1)*r3(3)) &
         + r1(3)*(r2(1)*r3(2) - r2(2)*r3(1))
   end function triple

   !> Cartesian vector of lattice coordinates v; lat(:, i) = a_i.
   function cart(lat, v) result(r)
      real(dp), intent(in) :: lat(3, 3), v(3)
      real(dp) :: r(3)
      r = v(1)*lat(:, 1) + v(2)*lat(:, 2) + v(3)*lat(:, 3)
   end function cart

   !> U from the SURFACE rows (third row 0 0 1 unless given), as
   !> src/readinput.f90:1618-1668 settles it.
   subroutine surface_umatrix(lat, u)
      real(dp), intent(in) :: lat(3, 3)
      real(dp), intent(inout) :: u(3, 3)
      real(dp) :: vol0, vol2
      vol0 = triple(lat(:, 1), lat(:, 2), lat(:, 3))
      vol2 = triple(cart(lat, u(1, :)), cart(lat, u(2, :)), cart(lat, u(3, :)))
      if (vol2 < 0) u(3, :) = -u(3, :)
      if (abs(abs(vol2) - abs(vol0)) > 0.001d0) call find_third_vector(lat, u, vol0)
   end subroutine surface_umatrix

   !> Rows 1 and 2 replaced by the shortest lattice vectors along them, row 3 the
   !> shortest lattice vector giving the smallest non-zero volume
   !> (src/readinput.f90:4247-4463).
   subroutine find_third_vector(lat, u, vol0)
      real(dp), intent(in) :: lat(3, 3), vol0
      real(dp), intent(inout) :: u(3, 3)
      integer :: i1, i2, i3, row
      real(dp) :: v(3), cr(3), best, smallest, shortest, vol, len3, r1(3), r2(3), urow(3)
      logical :: done
      do row = 1, 2
         best = 9999999d0
         urow = u(row, :)
         do i1 = -search_rmax, search_rmax
            do i2 = -search_rmax, search_rmax
               do i3 = -search_rmax, search_rmax
                  if (i1 == 0 .and. i2 == 0 .and. i3 == 0) cycle
                  v = (/dble(i1), dble(i2), dble(i3)/)
                  cr(1) = urow(2)*v(3) - v(2)*urow(3)
                  cr(2) = urow(3)*v(1) - v(3)*urow(1)
                  cr(3) = urow(1)*v(2) - v(1)*urow(2)
                  if (sum(abs(cr)) < eps9 .and. sum(v*urow) > 0) then
                     len3 = sqrt(sum(cart(lat, v)**2))
                     if (len3 < best) then
                        best = len3
                        u(row, :) = v
                     end if
                  end if
               end do
            end do
         end do
      end do
      r1 = cart(lat, u(1, :))
      r2 = cart(lat, u(2, :))
      smallest = 9999999d0
      shortest = 9999999d0
      do i1 = -search_rmax, search_rmax
         do i2 = -search_rmax, search_rmax
            do i3 = -search_rmax, search_rmax
               if (i1 == 0 .and. i2 == 0 .and. i3 == 0) cycle
               vol = abs(triple(r1, r2, cart(lat, (/dble(i1), dble(i2), dble(i3)/))))
               if (vol < eps9) cycle
               if (vol < smallest) smallest = vol
            end do
         end do
      end do
      do i1 = -search_rmax, search_rmax
         do i2 = -search_rmax, search_rmax
            do i3 = -search_rmax, search_rmax
               if (i1 == 0 .and. i2 == 0 .and. i3 == 0) cycle
               v = (/dble(i1), dble(i2), dble(i3)/)
               vol = abs(triple(r1, r2, cart(lat, v)))
               if (abs(vol - smallest) < eps9) shortest = min(shortest, sqrt(sum(cart(lat, v)**2)))
            end do
         end do
      end do
      done = .false.
      do i1 = -search_rmax, search_rmax
         do i2 = -search_rmax, search_rmax
            do i3 = -search_rmax, search_rmax
               if (done) exit
               if (i1 == 0 .and. i2 == 0 .and. i3 == 0) cycle
               v = (/dble(i1), dble(i2), dble(i3)/)
               vol = abs(triple(r1, r2, cart(lat, v)))
               if (abs(vol - smallest) < eps9 .and. abs(sqrt(sum(cart(lat, v)**2)) - shortest) < eps9) then
                  u(3, :) = v
                  done = .true.
               end if
            end do
         end do
      end do
      if (triple(cart(lat, u(1, :)), cart(lat, u(2, :)), cart(lat, u(3, :))) < 0) u(3, :) = -u(3, :)
      if (abs(triple(cart(lat, u(1, :)), cart(lat, u(2, :)), cart(lat, u(3, :))) - vol0) > eps9) then
         write(*, '(a)') ' >>> ERROR: SURFACE card: no cell of the crystal''s volume spans the first two'
         write(*, '(a)') '     vectors; the Miller-index construction is not supported here'
         stop 1
      end if
   end subroutine find_third_vector

   !> Each component into [0, 1) (src/readinput.f90:4495-4522).
   subroutine home(x)
      real(dp), intent(inout) :: x(3)
      integer :: i
      do i = 1, 3
         do while (.not. (x(i) >= 0d0 .and. x(i) < 1d0))
            if (x(i) < 0d0) then
               x(i) = x(i) + 1d0
            else
               x(i) = x(i) - 1d0
            end if
         end do
      end do
   end subroutine home

   !> Fractional positions in the new cell; top is the 1-based top-surface atom
   !> (0: the largest third coordinate, first of equals).
   subroutine surface_cell_positions(u, pos, top, posn)
      real(dp), intent(in) :: u(3, 3), pos(:, :)
      integer, intent(in) :: top
      real(dp), intent(out) :: posn(:, :)
      real(dp) :: ui(3, 3), shift(3)
      integer :: ia, it
      ui = inv3(u)
      do ia = 1, size(pos, 2)
         posn(:, ia) = ui(1, :)*pos(1, ia) + ui(2, :)*pos(2, ia) + ui(3, :)*pos(3, ia)
         call home(posn(:, ia))
      end do
      it = top
      if (it == 0) then
         it = 1
         do ia = 2, size(pos, 2)
            if (posn(3, ia) > posn(3, it)) it = ia
         end do
      end if
      shift = (/0d0, 0d0, 0.95d0 - posn(3, it)/)
      do ia = 1, size(pos, 2)
         posn(:, ia) = posn(:, ia) + shift
         call home(posn(:, ia))
      end do
   end subroutine surface_cell_positions

   !> get_hmnr_cell: <i,0|H|j,R> / ndegen(R) moves to
   !> R' = U^-T (R + tau_j - tau_i) - (tau'_j - tau'_i), rounded; dropped beyond
   !> max_ir; R' listed with the first component fastest; a later hopping on an
   !> occupied (R', i, j) replaces the earlier one.
   subroutine remap_hr(u, atom_of, pos, posn, nr, irvec, hmnr, ndeg)
      real(dp), intent(in) :: u(3, 3), pos(:, :), posn(:, :)
      integer, intent(in) :: atom_of(:), nr, irvec(:, :), ndeg(:)
      complex(dp), intent(in) :: hmnr(:, :, :)
      integer, allocatable :: rp(:, :, :, :), map(:, :, :)
      logical, allocatable :: used(:, :, :)
      real(dp) :: ui(3, 3), rmn(3), rnew(3)
      integer :: ir, i, j, n, a, b, c, nw
      nw = size(hmnr, 1)
      ui = inv3(u)
      allocate(rp(3, nw, nw, nr), map(-max_ir:max_ir, -max_ir:max_ir, -max_ir:max_ir))
      allocate(used(-max_ir:max_ir, -max_ir:max_ir, -max_ir:max_ir))
      used = .false.
      do ir = 1, nr
         do j = 1, nw
            do i = 1, nw
               rmn = irvec(:, ir) + pos(:, atom_of(j)) - pos(:, atom_of(i))
               rnew = ui(1, :)*rmn(1) + ui(2, :)*rmn(2) + ui(3, :)*rmn(3) &
                  - (posn(:, atom_of(j)) - posn(:, atom_of(i)))
               rp(:, i, j, ir) = nint(rnew)
               if (any(abs(rp(:, i, j, ir)) > max_ir)) cycle
               used(rp(1, i, j, ir), rp(2, i, j, ir), rp(3, i, j, ir)) = .true.
            end do
         end do
      end do
      nrpts_surfacecell = count(used)
      if (allocated(irvec_surfacecell)) deallocate(irvec_surfacecell)
      if (allocated(HmnR_surfacecell)) deallocate(HmnR_surfacecell)
      if (allocated(ndegen_surfacecell)) deallocate(ndegen_surfacecell)
      allocate(irvec_surfacecell(3, nrpts_surfacecell))
      allocate(HmnR_surfacecell(nw, nw, nrpts_surfacecell), ndegen_surfacecell(nrpts_surfacecell))
      n = 0
      map = 0
      do c = -max_ir, max_ir
         do b = -max_ir, max_ir
            do a = -max_ir, max_ir
               if (.not. used(a, b, c)) cycle
               n = n + 1
               irvec_surfacecell(:, n) = (/a, b, c/)
               map(a, b, c) = n
            end do
         end do
      end do
      HmnR_surfacecell = (0d0, 0d0)
      ndegen_surfacecell = 1
      do ir = 1, nr
         do j = 1, nw
            do i = 1, nw
               if (any(abs(rp(:, i, j, ir)) > max_ir)) cycle
               n = map(rp(1, i, j, ir), rp(2, i, j, ir), rp(3, i, j, ir))
               HmnR_surfacecell(i, j, n) = hmnr(i, j, ir)/ndeg(ir)
            end do
         end do
      end do
   end subroutine remap_hr

end module wt_surf_cell


module wt_surf_input
   use para
   use wt_surf_cell
   implicit none
   private
   public :: read_wt_in, read_hr, dump_cell

   ! the surface cell for read_hr's remap: U, the atoms' fractional positions
   ! in the crystal's and in the new cell, and the atom of every orbital
   real(dp) :: cell_u(3, 3)
   real(dp), allocatable :: cell_pos(:, :), cell_posn(:, :)
   integer, allocatable :: orb_atom(:)

   ! Zeeman term added to H(R = 0) by read_hr (Add_Zeeman_Field): energy in
   ! Hartree and the field's direction cosines
   logical :: add_zeeman = .false.
   real(dp) :: zeeman_au = 0d0, bdir(3) = 0d0

   integer, parameter :: maxkv = 512
   character(len=64) :: keys(maxkv)
   character(len=256) :: vals(maxkv)
   integer :: nkv = 0

contains

   function lower(s) result(t)
      character(len=*), intent(in) :: s
      character(len=len(s)) :: t
      integer :: i, c
      t = s
      do i = 1, len(s)
         c = iachar(s(i:i))
         if (c >= iachar('A') .and. c <= iachar('Z')) t(i:i) = achar(c + 32)
      end do
   end function lower

   !> Collect name = value pairs of every namelist group in the file.
   subroutine scan_namelists(fname)
      character(len=*), intent(in) :: fname
      character(len=1024) :: line
      character(len=8192) :: buf
      integer :: u, ios, p, q, e
      logical :: ingroup, inquote
      nkv = 0
      open(newunit=u, file=fname, status='old', action='read', iostat=ios)
      if (ios /= 0) then
         write(*, '(3a)') ' >>> ERROR: cannot open ', trim(fname), ' in the current directory'
         stop 1
      end if
      ingroup = .false.
      buf = ''
      do
         read(u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         p = index(line, '!')
         if (p > 0) line = line(1:p - 1)
         line = adjustl(line)
         if (.not. ingroup) then
            if (line(1:1) == '&') then
               ingroup = .true.
               buf = ''
               p = index(line, ' ')
               line = line(p:)
            else
               cycle
            end if
         end if
         ! a '/' outside quotes ends the group
         inquote = .false.
         e = 0
         do q = 1, len_trim(line)
            if (line(q:q) == "'" .or. line(q:q) == '"') inquote = .not. inquote
            if (line(q:q) == '/' .and. .not. inquote) then
               e = q
               exit
            end if
         end do
         if (e > 0) then
            buf = trim(buf)//', '//line(1:e - 1)
            call split_pairs(buf)
            ingroup = .false.
         else
            buf = trim(buf)//', '//trim(line)
         end if
      end do
      close(u)
   end subroutine scan_namelists

   subroutine split_pairs(buf)
      character(len=*), intent(in) :: buf
      integer :: i, start, eq
      logical :: inquote
      character(len=512) :: tok
      start = 1
      inquote = .false.
      do i = 1, len_trim(buf) + 1
         if (i <= len_trim(buf)) then
            if (buf(i:i) == "'" .or. buf(i:i) == '"') inquote = .not. inquote
            if (inquote .or. buf(i:i) /= ',') cycle
         end if
         tok = adjustl(buf(start:i - 1))
         start = i + 1
         eq = index(tok, '=')
         if (eq == 0 .or. nkv >= maxkv) cycle
         nkv = nkv + 1
         keys(nkv) = lower(adjustl(tok(1:eq - 1)))
         vals(nkv) = adjustl(tok(eq + 1:))
      end do
   end subroutine split_pairs

   logical function find(key, val)
      character(len=*), intent(in) :: key
      character(len=*), intent(out) :: val
      integer :: i
      find = .false.
      val = ''
      do i = nkv, 1, -1
         if (trim(keys(i)) == lower(key)) then
            val = vals(i)
            find = .true.
            return
         end if
      end do
   end function find

   subroutine get_real(key, x)
      character(len=*), intent(in) :: key
      real(dp), intent(inout) :: x
      character(len=256) :: v
      integer :: i, ios
      if (.not. find(key, v)) return
      do i = 1, len_trim(v)
         if (v(i:i) == 'd' .or. v(i:i) == 'D') v(i:i) = 'e'
      end do
      read(v, *, iostat=ios) x
      if (ios /= 0) call bad(key, v)
   end subroutine get_real

   subroutine get_int(key, x)
      character(len=*), intent(in) :: key
      integer, intent(inout) :: x
      character(len=256) :: v
      integer :: ios
      if (.not. find(key, v)) return
      read(v, *, iostat=ios) x
      if (ios /= 0) call bad(key, v)
   end subroutine get_int

   subroutine get_logical(key, x)
      character(len=*), intent(in) :: key
      logical, intent(inout) :: x
      character(len=256) :: v
      integer :: ios
      if (.not. find(key, v)) return
      read(v, *, iostat=ios) x
      if (ios /= 0) call bad(key, v)
   end subroutine get_logical

   subroutine get_string(key, x)
      character(len=*), intent(in) :: key
      character(len=*), intent(inout) :: x
      character(len=256) :: v
      integer :: n
      if (.not. find(key, v)) return
      v = adjustl(v)
      n = len_trim(v)
      if (n >= 2 .and. (v(1:1) == "'" .or. v(1:1) == '"')) v = v(2:n - 1)
      x = trim(v)
   end subroutine get_string

   subroutine bad(key, v)
      character(len=*), intent(in) :: key, v
      write(*, '(4a)') ' >>> ERROR: cannot read ', trim(key), ' = ', trim(v)
      stop 1
   end subroutine bad

   !> Position the unit just after the card line `name`; false if absent.
   logical function seek_card(u, name)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      character(len=256) :: line
      integer :: ios, p
      rewind(u)
      seek_card = .false.
      do
         read(u, '(a)', iostat=ios) line
         if (ios /= 0) return
         p = index(line, '!')
         if (p > 0) line = line(1:p - 1)
         if (trim(adjustl(lower(line))) == lower(name)) then
            seek_card = .true.
            return
         end if
      end do
   end function seek_card

   !> Integers of one line, "a b c" or ranges "a-b" (SURFACE_ATOMS).
   subroutine parse_index_list(line, out, n)
      character(len=*), intent(in) :: line
      integer, allocatable, intent(out) :: out(:)
      integer, intent(out) :: n
      integer :: tmp(4096), i, start, dash, a, b, k, ios
      character(len=64) :: tok
      character(len=len(line)) :: s
      s = line
      i = index(s, '!')
      if (i > 0) s = s(1:i - 1)
      do i = 1, len_trim(s)
         if (s(i:i) == ',') s(i:i) = ' '
      end do
      n = 0
      start = 1
      do i = 1, len_trim(s) + 1
         if (i <= len_trim(s)) then
            if (s(i:i) /= ' ') cycle
         end if
         tok = s(start:i - 1)
         start = i + 1
         if (len_trim(tok) == 0) cycle
         dash = index(tok, '-')
         if (dash > 1) then
            read(tok(1:dash - 1), *, iostat=ios) a
            read(tok(dash + 1:), *, iostat=ios) b
         else
            read(tok, *, iostat=ios) a
            b = a
         end if
         do k = a, b
            n = n + 1
            tmp(n) = k
         end do
      end do
      allocate(out(n))
      out = tmp(1:n)
   end subroutine parse_index_list

   subroutine read_wt_in(fname)
      character(len=*), intent(in) :: fname
      real(dp) :: eta_arc, e_arc, lat(3, 3), r1(3), r2(3), s(3, 3), ra2(2), rb2(2)
      real(dp) :: vol, kstart(2), kend(2), kp(2, 64), ke(2, 64), k1(2), k2(2), t1, temp
      real(dp) :: row3(3), linv(3, 3)
      integer :: u, ios, i, j, ia, io, natoms, nk2lines, ntop_atoms, nbot_atoms, nproj_tot
      integer :: top_atom
      integer, allocatable :: nprojs(:), orbitals_start(:), topatoms(:), botatoms(:)
      character(len=256) :: line
      character(len=16) :: name1, name2
      character(len=64) :: word
      logical :: have_surface, angstrom

      call scan_namelists(fname)
      call get_string('Hrfile', Hrfile)
      call get_string('Particle', Particle)
      call get_logical('SlabSS_calc', SlabSS_calc)
      call get_logical('SlabArc_calc', SlabArc_calc)
      call get_logical('SlabSpintexture_calc', SlabSpintexture_calc)
      call get_logical('SlabQPI_kplane_calc', SlabQPI_kplane_calc)
      call get_logical('SlabQPI_calc', SlabQPI_calc)
      call get_logical('SlabQPI_kpath_calc', SlabQPI_kpath_calc)
      SlabQPI_kplane_calc = SlabQPI_kplane_calc .or. SlabQPI_calc  ! src/readinput.f90:171
      call get_logical('LOTO_correction', LOTO_correction)
      call get_int('SOC', Soc)
      call get_real('E_FERMI', E_fermi)
      ! src/readinput.f90:588-599, 665-666, 721-724
      eta_arc = -999d0
      e_arc = -999d0
      call get_real('Eta_Arc', eta_arc)
      call get_real('E_arc', e_arc)
      call get_real('Fermi_broadening', Fermi_broadening)
      call get_real('iso_energy', iso_energy)
      call get_int('OmegaNum', OmegaNum)
      call get_real('OmegaMin', OmegaMin)
      call get_real('OmegaMax', OmegaMax)
      call get_int('Nk1', Nk1)
      call get_int('Nk2', Nk2)
      call get_int('NP', Np)
      top_atom = 0
      call get_int('topsurface_atom_index', top_atom)
      call read_fields()
      if (eta_arc > -998d0) Fermi_broadening = eta_arc
      if (e_arc > -998d0) iso_energy = e_arc
      iso_energy = iso_energy*eV2Hartree
      Fermi_broadening = Fermi_broadening*eV2Hartree
      OmegaMin = OmegaMin*eV2Hartree
      OmegaMax = OmegaMax*eV2Hartree

      open(newunit=u, file=fname, status='old', action='read')
      ! LATTICE: unit line, then a1, a2, a3 as rows
      if (.not. seek_card(u, 'LATTICE')) stop ' >>> ERROR: wt.in has no LATTICE card'
      read(u, *) word
      do i = 1, 3
         read(u, *) lat(:, i)
      end do
      ! the unit line counts as Angstrom when it says so, Bohr otherwise (src/readinput.f90:772-787)
      angstrom = index(lower(word), 'ang') > 0
      if (angstrom) lat = lat*Angstrom2atomic
      ! ATOM_POSITIONS: fractional, or Cartesian in the lattice's unit; the "D"
      ! test is case-sensitive as in src/readinput.f90:871-933
      if (.not. seek_card(u, 'ATOM_POSITIONS')) stop ' >>> ERROR: wt.in has no ATOM_POSITIONS card'
      read(u, *) natoms
      read(u, *) word
      allocate(cell_pos(3, natoms), cell_posn(3, natoms))
      linv = inv3(lat)
      do ia = 1, natoms
         read(u, *) name1, cell_pos(:, ia)
         if (index(word, 'D') == 0) then
            if (angstrom) cell_pos(:, ia) = cell_pos(:, ia)*Angstrom2atomic
            cell_pos(:, ia) = matmul(linv, cell_pos(:, ia))
         end if
      end do
      if (.not. seek_card(u, 'PROJECTORS')) stop ' >>> ERROR: wt.in has no PROJECTORS card'
      allocate(nprojs(natoms), orbitals_start(natoms + 1))
      read(u, *) nprojs
      nproj_tot = sum(nprojs)
      orbitals_start(1) = 1
      do i = 1, natoms
         orbitals_start(i + 1) = orbitals_start(i) + nprojs(i)
      end do
      Num_wann = nproj_tot
      if (Soc > 0) Num_wann = 2*nproj_tot
      Ndim = Num_wann*Np
      ! atom of every orbital, spin down repeating spin up (spinorbital_to_atom_index)
      allocate(orb_atom(Num_wann))
      do ia = 1, natoms
         do j = 1, nprojs(ia)
            orb_atom(orbitals_start(ia) + j - 1) = ia
            if (Soc > 0) orb_atom(orbitals_start(ia) + j - 1 + nproj_tot) = ia
         end do
      end do

      ! SURFACE: R1', R2' (and optionally R3') in units of a1, a2, a3
      s = 0d0
      s(1, 1) = 1d0
      s(2, 2) = 1d0
      s(3, 3) = 1d0
      have_surface = seek_card(u, 'SURFACE')
      if (have_surface) then
         read(u, *) s(1, :)
         read(u, *) s(2, :)
         read(u, *, iostat=ios) row3
         if (ios == 0) s(3, :) = row3
      end if
      call surface_umatrix(lat, s)
      cell_u = s
      call surface_cell_positions(cell_u, cell_pos, top_atom, cell_posn)
      r1 = cart(lat, s(1, :))
      r2 = cart(lat, s(2, :))
      ! surface frame and reciprocal vectors (src/readinput.f90:1812-1855)
      Urot = 0d0
      Urot(1, :) = r1/sqrt(sum(r1**2))
      Urot(3, 1) = r1(2)*r2(3) - r1(3)*r2(2)
      Urot(3, 2) = r1(3)*r2(1) - r1(1)*r2(3)
      Urot(3, 3) = r1(1)*r2(2) - r1(2)*r2(1)
      Urot(3, :) = Urot(3, :)/sqrt(sum(Urot(3, :)**2))
      Urot(2, 1) = Urot(3, 2)*Urot(1, 3) - Urot(3, 3)*Urot(1, 2)
      Urot(2, 2) = Urot(3, 3)*Urot(1, 1) - Urot(3, 1)*Urot(1, 3)
      Urot(2, 3) = Urot(3, 1)*Urot(1, 2) - Urot(3, 2)*Urot(1, 1)
      Urot(2, :) = Urot(2, :)/sqrt(sum(Urot(2, :)**2))
      ra2(1) = sum(Urot(1, :)*r1)
      ra2(2) = sum(Urot(2, :)*r1)
      rb2(1) = sum(Urot(1, :)*r2)
      rb2(2) = sum(Urot(2, :)*r2)
      vol = abs(ra2(1)*rb2(2) - rb2(1)*ra2(2))
      if (vol < 1e-6) stop ' >>> ERROR: cell_volume equal zero'
      Ka2(1) = 2d0*pi/vol*rb2(2)
      Ka2(2) = -2d0*pi/vol*rb2(1)
      Kb2(1) = -2d0*pi/vol*ra2(2)
      Kb2(2) = 2d0*pi/vol*ra2(1)

      ! KPATH_SLAB (src/readinput.f90:2134-2213), Nk = Nk1 points per line
      knv2 = 0
      if (seek_card(u, 'KPATH_SLAB')) then
         read(u, *) nk2lines
         do i = 1, nk2lines
            read(u, *) name1, kp(:, i), name2, ke(:, i)
         end do
         knv2 = Nk1*nk2lines
         allocate(k2_path(knv2, 2), k2len(knv2))
         t1 = 0d0
         do j = 1, nk2lines
            do i = 1, Nk1
               kstart = kp(:, j)
               kend = ke(:, j)
               k1 = kstart(1)*Ka2 + kstart(2)*Kb2
               k2 = kend(1)*Ka2 + kend(2)*Kb2
               k2_path(i + (j - 1)*Nk1, :) = kstart + (kend - kstart)*(i - 1)/dble(Nk1 - 1)
               temp = dsqrt((k2(1) - k1(1))**2 + (k2(2) - k1(2))**2)/dble(Nk1 - 1)
               if (i > 1) t1 = t1 + temp
               k2len(i + (j - 1)*Nk1) = t1
            end do
         end do
      else if (SlabSS_calc .or. SlabQPI_kpath_calc) then
         stop 'ERROR: please set KPATH_SLAB for slab band structure calculation'
      end if

      ! KPLANE_SLAB (src/readinput.f90:2216-2262)
      if (seek_card(u, 'KPLANE_SLAB')) then
         read(u, *) K2D_start
         read(u, *) K2D_vec1
         read(u, *) K2D_vec2
      else if (SlabArc_calc .or. SlabSpintexture_calc) then
         stop 'ERROR: please set KPLANE_SLAB for arc or spintexture calculations'
      end if

      ! SURFACE_ATOMS -> TopOrbitals / BottomOrbitals (src/readinput.f90:2851-2962)
      if (seek_card(u, 'SURFACE_ATOMS')) then
         read(u, '(a)') line
         call parse_index_list(line, topatoms, ntop_atoms)
         read(u, '(a)') line
         call parse_index_list(line, botatoms, nbot_atoms)
      else
         ntop_atoms = natoms
         nbot_atoms = natoms
         allocate(topatoms(natoms), botatoms(natoms))
         topatoms = (/(i, i=1, natoms)/)
         botatoms = topatoms
      end if
      close(u)
      NtopOrbitals = 0
      do i = 1, ntop_atoms
         NtopOrbitals = NtopOrbitals + nprojs(topatoms(i))
      end do
      if (Soc > 0) NtopOrbitals = NtopOrbitals*2
      allocate(TopOrbitals(NtopOrbitals))
      io = 0
      do i = 1, ntop_atoms
         ia = topatoms(i)
         do j = 1, nprojs(ia)
            io = io + 1
            TopOrbitals(io) = orbitals_start(ia) + j - 1
            if (Soc > 0) TopOrbitals(io + NtopOrbitals/2) = orbitals_start(ia) + j - 1 + nproj_tot
         end do
      end do
      NBottomOrbitals = 0
      do i = 1, nbot_atoms
         NBottomOrbitals = NBottomOrbitals + nprojs(botatoms(i))
      end do
      if (Soc > 0) NBottomOrbitals = NBottomOrbitals*2
      allocate(BottomOrbitals(NBottomOrbitals))
      io = 0
      do i = 1, nbot_atoms
         ia = botatoms(i)
         do j = 1, nprojs(ia)
            io = io + 1
            BottomOrbitals(io) = orbitals_start(ia) + j - 1
            if (Soc > 0) BottomOrbitals(io + NBottomOrbitals/2) = orbitals_start(ia) + j - 1 + nproj_tot
         end do
      end do
   end subroutine read_wt_in

   !> The &SYSTEM / &TB_FILE settings that change H(R) before the remap.  The
   !> Zeeman term with SOC = 1 is applied by read_hr; H(R) sources and terms
   !> this host does not build (sparse or non-orthogonal H(R), the SOC = 0 spin
   !> extension, electric fields) stop the run rather than being ignored.
   subroutine read_fields()
      logical :: flag, ortho, sparse, sparse_hr, hrfile
      real(dp) :: ez, gfac, bmag, btheta, bphi, bx, by, bz, ef1, ef2
      ! src/readinput.f90:47-50 and 66: the H(R) source
      hrfile = .true.
      sparse = .false.
      sparse_hr = .false.
      ortho = .true.
      call get_logical('Is_HrFile', hrfile)
      call get_logical('Is_Sparse', sparse)
      call get_logical('Is_Sparse_Hr', sparse_hr)
      call get_logical('Orthogonal_Basis', ortho)
      if (.not. hrfile .or. sparse .or. sparse_hr .or. .not. ortho) then
         write(*, '(a)') ' >>> ERROR: only a dense Wannier90 *_hr.dat in an orthogonal basis is read'
         write(*, '(a)') '     here (Is_HrFile = T, Is_Sparse(_Hr) = F, Orthogonal_Basis = T)'
         stop 1
      end if
      ! src/readinput.f90:355-369: field defaults
      ef1 = 0d0
      ef2 = 0d0
      call get_real('Electric_field_in_eVpA', ef1)
      call get_real('Symmetrical_Electric_field_in_eVpA', ef2)
      if (ef1 /= 0d0 .or. ef2 /= 0d0) then
         write(*, '(a)') ' >>> ERROR: electric fields on H(R) (Electric_field_in_eVpA,'
         write(*, '(a)') '     Symmetrical_Electric_field_in_eVpA) are not supported here'
         stop 1
      end if
      flag = .false.
      call get_logical('Add_Zeeman_Field', flag)
      if (.not. flag) return
      if (Soc == 0) then
         write(*, '(a)') ' >>> ERROR: Add_Zeeman_Field with SOC = 0 (spin-extended H(R)) is not'
         write(*, '(a)') '     supported here; give the spinful H(R) with SOC = 1'
         stop 1
      end if
      ez = 0d0
      gfac = 2d0
      bmag = 0d0
      btheta = -99999d0
      bphi = -99999d0
      bx = 0d0
      by = 0d0
      bz = 0d0
      call get_real('Zeeman_energy_in_eV', ez)
      call get_real('Effective_gfactor', gfac)
      call get_real('Bmagnitude', bmag)
      call get_real('Btheta', btheta)
      call get_real('Bphi', bphi)
      call get_real('Bx', bx)
      call get_real('By', by)
      call get_real('Bz', bz)
      ! the field's direction and magnitude in Tesla: Bmagnitude with Btheta and
      ! Bphi first, else Bx, By, Bz, else Btheta and Bphi alone with magnitude 1
      ! (src/readinput.f90:464-565)
      if (abs(bmag) > eps6) then
         if (abs(btheta + 99999d0) < eps6) btheta = 0d0
         if (abs(bphi + 99999d0) < eps6) bphi = 0d0
         bdir = (/sin(btheta/180d0*pi)*cos(bphi/180d0*pi), sin(btheta/180d0*pi)*sin(bphi/180d0*pi), &
                  cos(btheta/180d0*pi)/)
      else if (abs(bx) + abs(by) + abs(bz) > eps6) then
         bmag = sqrt(bx*bx + by*by + bz*bz)
         bdir = (/bx, by, bz/)/bmag
      else
         if (abs(btheta + 99999d0) < eps6) btheta = 0d0
         if (abs(bphi + 99999d0) < eps6) bphi = 0d0
         bdir = (/sin(btheta/180d0*pi)*cos(bphi/180d0*pi), sin(btheta/180d0*pi)*sin(bphi/180d0*pi), &
                  cos(btheta/180d0*pi)/)
         bmag = 1d0
      end if
      ! Tesla -> atomic units (src/readinput.f90:574-577, constants of
      ! src/module.f90:710-715 as single-precision literals), then the Zeeman
      ! energy g mu_B B unless given in eV (src/readHmnR.f90:1459-1464)
      bmag = bmag*real(1.602176634e-19, dp)*real(5.29177210903e-11, dp)**2/real(1.054571817e-34, dp)
      if (abs(ez) < eps9) then
         zeeman_au = gfac*0.5d0*bmag
      else
         zeeman_au = ez*eV2Hartree
      end if
      add_zeeman = .true.
   end subroutine read_fields

   !> Wannier90 *_hr.dat (src/readHmnR.f90:3-305 for the orthogonal basis):
   !> E_fermi off the R = 0 diagonal, eV -> Hartree, then remapped onto the
   !> surface cell (get_hmnr_cell, with 1/ndegen applied there).
   subroutine read_hr()
      integer :: u, ios, nw, nr, ir, n, i, r(3), mm, nn, nused
      real(dp) :: re, im
      integer, allocatable :: ndeg(:), irv(:, :)
      complex(dp), allocatable :: hmnr(:, :, :)
      character(len=256) :: line
      open(newunit=u, file=trim(Hrfile), status='old', action='read', iostat=ios)
      if (ios /= 0) then
         write(*, '(3a)') ' >>> ERROR: cannot open the Hrfile ', trim(Hrfile)
         stop 1
      end if
      read(u, '(a)') line
      read(u, *) nw
      read(u, *) nr
      if (nw /= Num_wann) then
         write(*, '(a, i0, a, i0)') ' >>> ERROR: Hrfile num_wann ', nw, ' but PROJECTORS give ', Num_wann
         stop 1
      end if
      allocate(ndeg(nr), irv(3, nr), hmnr(nw, nw, nr))
      read(u, *) ndeg
      hmnr = (0d0, 0d0)
      !> every line is placed by its own R and (m, n), R vectors numbered in
      !> order of first appearance (src/readHmnR.f90:101-124), so the order of
      !> the lines does not matter; the search starts at the newest R
      nused = 0
      do i = 1, nr*nw*nw
         read(u, *, iostat=ios) r, mm, nn, re, im
         if (ios /= 0) then
            write(*, '(2a)') ' >>> ERROR: short or malformed H(R) lines in ', trim(Hrfile)
            stop 1
         end if
         ir = 0
         do n = nused, 1, -1
            if (all(irv(:, n) == r)) then
               ir = n
               exit
            end if
         end do
         if (ir == 0) then
            if (nused == nr) then
               write(*, '(a, i0, 2a)') ' >>> ERROR: more than ', nr, ' R vectors in ', trim(Hrfile)
               stop 1
            end if
            nused = nused + 1
            ir = nused
            irv(:, ir) = r
         end if
         if (mm < 1 .or. mm > nw .or. nn < 1 .or. nn > nw) then
            write(*, '(2a)') ' >>> ERROR: orbital index out of range in ', trim(Hrfile)
            stop 1
         end if
         hmnr(mm, nn, ir) = dcmplx(re, im)
      end do
      close(u)
      if (nused /= nr) then
         write(*, '(a, i0, a, i0, 2a)') ' >>> ERROR: ', nused, ' R vectors, header says ', nr, &
            ' in ', trim(Hrfile)
         stop 1
      end if
      do ir = 1, nr
         if (all(irv(:, ir) == 0)) then
            do i = 1, nw
               hmnr(i, i, ir) = hmnr(i, i, ir) - E_fermi
            end do
         end if
      end do
      if (index(Particle, 'phonon') /= 0) then
         hmnr = hmnr*eV2Hartree*eV2Hartree
      else
         hmnr = hmnr*eV2Hartree
      end if
      ! Zeeman term (B . sigma) E_Z / 2 on the R = 0 block, spin-up orbitals
      ! 1..nw/2 paired with nw/2+1..nw (src/readHmnR.f90:1475-1484, SOC = 1)
      if (add_zeeman) then
         ir = 0
         do n = 1, nr
            if (all(irv(:, n) == 0)) ir = n
         end do
         if (ir == 0) then
            write(*, '(a)') ' >>> ERROR: Add_Zeeman_Field needs an R = 0 block in the Hrfile'
            stop 1
         end if
         do n = 1, nw/2
            hmnr(n, n, ir) = hmnr(n, n, ir) + zeeman_au/2d0*bdir(3)
            hmnr(n + nw/2, n + nw/2, ir) = hmnr(n + nw/2, n + nw/2, ir) - zeeman_au/2d0*bdir(3)
            hmnr(n, n + nw/2, ir) = hmnr(n, n + nw/2, ir) + zeeman_au/2d0*bdir(1) &
               - (0d0, 1d0)*zeeman_au/2d0*bdir(2)
            hmnr(n + nw/2, n, ir) = hmnr(n + nw/2, n, ir) + zeeman_au/2d0*bdir(1) &
               + (0d0, 1d0)*zeeman_au/2d0*bdir(2)
         end do
      end if
      call remap_hr(cell_u, orb_atom, cell_pos, cell_posn, nr, irv, hmnr, ndeg)
   end subroutine read_hr

   !> Urot, Ka2, Kb2, TopOrbitals, BottomOrbitals and HmnR_surfacecell (Hartree,
   !> ndegen applied) as plain text, full precision.
   subroutine dump_cell(fname)
      character(len=*), intent(in) :: fname
      integer :: u, ir, i, j
      open(newunit=u, file=fname)
      write(u, '(a)') '# Urot rows, Ka2, Kb2, Num_wann Np nrpts, top, bottom, then R1 R2 R3 i j Re Im'
      do i = 1, 3
         write(u, '(3es25.16)') Urot(i, :)
      end do
      write(u, '(2es25.16)') Ka2
      write(u, '(2es25.16)') Kb2
      write(u, '(3i8)') Num_wann, Np, nrpts_surfacecell
      write(u, '(*(i6))') TopOrbitals
      write(u, '(*(i6))') BottomOrbitals
      do ir = 1, nrpts_surfacecell
         do j = 1, Num_wann
            do i = 1, Num_wann
               write(u, '(5i6, 2es25.16)') irvec_surfacecell(:, ir), i, j, &
                  HmnR_surfacecell(i, j, ir)/ndegen_surfacecell(ir)
            end do
         end do
      end do
      close(u)
   end subroutine dump_cell

end module wt_surf_input


!> A vector in the surface frame: s1 = Urot s0 (src/readinput.f90:3722-3737).
!> In a WannierTools build the reference's own rotate is used instead.
subroutine rotate(R1, R2)
   use para, only: dp, Urot
   implicit none
   real(dp), intent(in) :: R1(3)
   real(dp), intent(inout) :: R2(3)
   integer :: i
   do i = 1, 3
      R2(i) = Urot(i, 1)*R1(1) + Urot(i, 2)*R1(2) + Urot(i, 3)*R1(3)
   end do
end subroutine rotate


program wt_surf
   use para
   use wt_surf_input
   use wtsurf_gpu
   implicit none
   integer :: c0, c1, rate
   character(len=64) :: arg
   open(unit=stdout, file='WT.out')
   write(stdout, '(a)') ' wt_surf: WannierTools surface Green''s function path on the GPU (libwtsurf)'
   call read_wt_in('wt.in')
   call read_hr()
   ! `wt_surf --dump-cell`: write the surface cell and the remapped H(R) to
   ! surface_cell.dat and run no driver (no device work)
   if (command_argument_count() >= 1) then
      call get_command_argument(1, arg)
      if (trim(arg) == '--dump-cell') then
         call dump_cell('surface_cell.dat')
         SlabSS_calc = .false.
         SlabQPI_kpath_calc = .false.
         SlabArc_calc = .false.
         SlabQPI_calc = .false.
         SlabQPI_kplane_calc = .false.
         SlabSpintexture_calc = .false.
      end if
   end if
   write(stdout, '(a, i0, a, i0, a, i0)') ' Num_wann ', Num_wann, '  Np ', Np, '  Ndim ', Ndim
   if (SlabSS_calc) then
      write(stdout, *) ' '
      write(stdout, *) 'Start to calculate surface state'
      call system_clock(c0, rate)
      call surfstat
      call system_clock(c1)
      write(stdout, '(a, f16.3, a)') ' Time cost for SlabSS_calc is about ', dble(c1 - c0)/rate, ' s'
   end if
   if (SlabArc_calc) then
      write(stdout, *) ' '
      write(stdout, *) 'Start to calculate fermi arc'
      call system_clock(c0, rate)
      call SurfaceDOSkk
      call system_clock(c1)
      write(stdout, '(a, f16.3, a)') ' Time cost for SlabArc is about ', dble(c1 - c0)/rate, ' s'
   end if
   if (SlabQPI_kpath_calc) then
      write(stdout, *) ' '
      write(stdout, *) 'Start to calculate the surface QPI in kpath mode'
      call system_clock(c0, rate)
      call surfstat_jdos
      call system_clock(c1)
      write(stdout, '(a, f16.3, a)') ' Time cost for SlabQPI is about ', dble(c1 - c0)/rate, ' s'
   end if
   if (SlabQPI_calc .or. SlabQPI_kplane_calc) then
      write(stdout, *) ' '
      write(stdout, *) 'Start to calculate the surface QPI'
      call system_clock(c0, rate)
      call SurfaceDOSkk
      call system_clock(c1)
      write(stdout, '(a, f16.3, a)') ' Time cost for SlabQPI is about ', dble(c1 - c0)/rate, ' s'
   end if
   if (SlabSpintexture_calc) then
      write(stdout, *) ' '
      write(stdout, *) 'Start to calculate the spin texture for surface'
      call system_clock(c0, rate)
      call SurfaceDOSkk
      call system_clock(c1)
      write(stdout, '(a, f16.3, a)') ' Time cost for SlabSpintexture is about ', dble(c1 - c0)/rate, ' s'
   end if
   call wtsurf_gpu_finalize()
   write(stdout, *) 'Congratulations! you finished the calculation.'
   close(stdout)
end program wt_surf
