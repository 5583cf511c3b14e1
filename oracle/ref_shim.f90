! C entry points onto the reference Fortran routines, for the oracle cross-check only.
!
! TEST INFRASTRUCTURE ONLY: linked with the reference sources into oracle/_ref/libqmlref.so by
! oracle/Makefile. The reference routines take assumed-shape dummies, so each wrapper gives
! them an explicit interface and hands over arrays in the layout f2py would have produced
! (qml/representations.py:109-115, qml/kernels.py:58-61, qml/distance.py:55-57).

subroutine ref_cm(z, r_t, natoms, nmax, sorted, cm) bind(C, name="ref_cm")
  ! r_t is the C-ordered (natoms, 3) coordinate array, i.e. Fortran (3, natoms); the
  ! reference wants coordinates(natoms, 3) (qml/frepresentations.f90:57).
  use iso_c_binding
  implicit none
  integer(c_int), value :: natoms, nmax, sorted
  real(c_double), intent(in) :: z(natoms), r_t(3, natoms)
  real(c_double), intent(out) :: cm((nmax * (nmax + 1)) / 2)
  real(c_double), allocatable :: coords(:, :)
  interface
    subroutine fgenerate_coulomb_matrix(atomic_charges, coordinates, nmax, cm)
      double precision, dimension(:), intent(in) :: atomic_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: nmax
      double precision, dimension(((nmax + 1) * nmax) / 2), intent(out) :: cm
    end subroutine
    subroutine fgenerate_unsorted_coulomb_matrix(atomic_charges, coordinates, nmax, cm)
      double precision, dimension(:), intent(in) :: atomic_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: nmax
      double precision, dimension(((nmax + 1) * nmax) / 2), intent(out) :: cm
    end subroutine
  end interface
  allocate(coords(natoms, 3))
  coords = transpose(r_t)
  if (sorted /= 0) then
    call fgenerate_coulomb_matrix(z, coords, nmax, cm)
  else
    call fgenerate_unsorted_coulomb_matrix(z, coords, nmax, cm)
  end if
  deallocate(coords)
end subroutine

subroutine ref_kernel(kind, a, na, b, nb, d, k, sigma) bind(C, name="ref_kernel")
  ! kind: 0 Laplacian, 1 Gaussian, 2 Manhattan distance, 3 L2 distance.
  ! a, b are C-ordered (na, d) / (nb, d), i.e. Fortran (d, na) / (d, nb); k is (na, nb)
  ! column-major.
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, na, nb, d
  real(c_double), intent(in) :: a(d, na), b(d, nb)
  real(c_double), intent(inout) :: k(na, nb)
  real(c_double), value :: sigma
  interface
    subroutine flaplacian_kernel(a, na, b, nb, k, sigma)
      double precision, dimension(:,:), intent(in) :: a, b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
    end subroutine
    subroutine fgaussian_kernel(a, na, b, nb, k, sigma)
      double precision, dimension(:,:), intent(in) :: a, b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
    end subroutine
    subroutine fmanhattan_distance(a, b, d)
      double precision, dimension(:,:), intent(in) :: a, b
      double precision, dimension(:,:), intent(inout) :: d
    end subroutine
    subroutine fl2_distance(a, b, d)
      double precision, dimension(:,:), intent(in) :: a, b
      double precision, dimension(:,:), intent(inout) :: d
    end subroutine
  end interface
  select case (kind)
  case (0)
    call flaplacian_kernel(a, na, b, nb, k, sigma)
  case (1)
    call fgaussian_kernel(a, na, b, nb, k, sigma)
  case (2)
    call fmanhattan_distance(a, b, k)
  case (3)
    call fl2_distance(a, b, k)
  end select
end subroutine
