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

subroutine ref_atomic_cm(z, r_t, natoms, cent1, ncent, nmax, sorting, cc, cd, ic, id, cm) &
    bind(C, name="ref_atomic_cm")
  ! cent1: 1-based central atom indices (qml/representations.py:205-220); cm is Fortran
  ! (ncent, nmax(nmax+1)/2) as the reference returns it (frepresentations.f90:225/428).
  ! sorting: 0 row-norm (fgenerate_local_coulomb_matrix), 1 distance
  ! (fgenerate_atomic_coulomb_matrix). The cut-offs are intent(inout) there, so copies go in.
  use iso_c_binding
  implicit none
  integer(c_int), value :: natoms, ncent, nmax, sorting
  real(c_double), intent(in) :: z(natoms), r_t(3, natoms)
  integer(c_int), intent(in) :: cent1(ncent)
  real(c_double), value :: cc, cd, ic, id
  real(c_double), intent(out) :: cm(ncent, (nmax * (nmax + 1)) / 2)
  real(c_double), allocatable :: coords(:, :)
  integer, allocatable :: idx(:)
  double precision :: c1, c2, c3, c4
  interface
    subroutine fgenerate_local_coulomb_matrix(central_atom_indices, central_natoms, &
        & atomic_charges, coordinates, natoms, nmax, cent_cutoff, cent_decay, &
        & int_cutoff, int_decay, cm)
      integer, intent(in) :: central_natoms
      integer, dimension(:), intent(in) :: central_atom_indices
      double precision, dimension(:), intent(in) :: atomic_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: natoms
      integer, intent(in) :: nmax
      double precision, intent(inout) :: cent_cutoff, cent_decay, int_cutoff, int_decay
      double precision, dimension(central_natoms, ((nmax + 1) * nmax) / 2), intent(out):: cm
    end subroutine
    subroutine fgenerate_atomic_coulomb_matrix(central_atom_indices, central_natoms, &
        & atomic_charges, coordinates, natoms, nmax, cent_cutoff, cent_decay, &
        & int_cutoff, int_decay, cm)
      integer, dimension(:), intent(in) :: central_atom_indices
      integer, intent(in) :: central_natoms
      double precision, dimension(:), intent(in) :: atomic_charges
      double precision, dimension(:,:), intent(in) :: coordinates
      integer, intent(in) :: natoms
      integer, intent(in) :: nmax
      double precision, intent(inout) :: cent_cutoff, cent_decay, int_cutoff, int_decay
      double precision, dimension(central_natoms, ((nmax + 1) * nmax) / 2), intent(out):: cm
    end subroutine
  end interface
  allocate(coords(natoms, 3))
  allocate(idx(ncent))
  coords = transpose(r_t)
  idx = cent1
  c1 = cc
  c2 = cd
  c3 = ic
  c4 = id
  if (sorting == 0) then
    call fgenerate_local_coulomb_matrix(idx, ncent, z, coords, natoms, nmax, c1, c2, c3, c4, cm)
  else
    call fgenerate_atomic_coulomb_matrix(idx, ncent, z, coords, natoms, nmax, c1, c2, c3, c4, cm)
  end if
  deallocate(coords)
  deallocate(idx)
end subroutine

subroutine ref_local_kernels(kind, q1, na, q2, nb, d, n1, nm1, n2, nm2, sigmas, nsig, k) &
    bind(C, name="ref_local_kernels")
  ! kind: 0 Laplacian (fget_local_kernels_laplacian), 1 Gaussian (fget_local_kernels_gaussian).
  ! q1, q2 are C-ordered (atoms, d), i.e. Fortran (d, atoms) as qml/kernels.py:256/298 pass
  ! A.T; k is Fortran (nsig, nm1, nm2).
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, na, nb, d, nm1, nm2, nsig
  real(c_double), intent(in) :: q1(d, na), q2(d, nb), sigmas(nsig)
  integer(c_int), intent(in) :: n1(nm1), n2(nm2)
  real(c_double), intent(out) :: k(nsig, nm1, nm2)
  integer, allocatable :: m1(:), m2(:)
  interface
    subroutine fget_local_kernels_gaussian(q1, q2, n1, n2, sigmas, nm1, nm2, nsigmas, kernels)
      double precision, dimension(:,:), intent(in) :: q1
      double precision, dimension(:,:), intent(in) :: q2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
    subroutine fget_local_kernels_laplacian(q1, q2, n1, n2, sigmas, nm1, nm2, nsigmas, kernels)
      double precision, dimension(:,:), intent(in) :: q1
      double precision, dimension(:,:), intent(in) :: q2
      integer, dimension(:), intent(in) :: n1
      integer, dimension(:), intent(in) :: n2
      double precision, dimension(:), intent(in) :: sigmas
      integer, intent(in) :: nm1
      integer, intent(in) :: nm2
      integer, intent(in) :: nsigmas
      double precision, dimension(nsigmas,nm1,nm2), intent(out) :: kernels
    end subroutine
  end interface
  allocate(m1(nm1))
  allocate(m2(nm2))
  m1 = n1
  m2 = n2
  if (kind == 1) then
    call fget_local_kernels_gaussian(q1, q2, m1, m2, sigmas, nm1, nm2, nsig, k)
  else
    call fget_local_kernels_laplacian(q1, q2, m1, m2, sigmas, nm1, nm2, nsig, k)
  end if
  deallocate(m1)
  deallocate(m2)
end subroutine

subroutine ref_kernel_ext(kind, a, na, b, nb, d, k, sigma, gammas, ng, order, p, ip) &
    bind(C, name="ref_kernel_ext")
  ! kind: 0 linear (flinear_kernel), 1 Sargan (fsargan_kernel), 2 Matern L2
  ! (fmatern_kernel_l2), 3 integer p-distance (fp_distance_integer), 4 real p-distance
  ! (fp_distance_double). a, b are C-ordered (na, d) / (nb, d); k is (na, nb) column-major.
  use iso_c_binding
  implicit none
  integer(c_int), value :: kind, na, nb, d, ng, order, ip
  real(c_double), intent(in) :: a(d, na), b(d, nb), gammas(*)
  real(c_double), intent(inout) :: k(na, nb)
  real(c_double), value :: sigma, p
  interface
    subroutine flinear_kernel(a, na, b, nb, k)
      double precision, dimension(:,:), intent(in) :: a, b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
    end subroutine
    subroutine fsargan_kernel(a, na, b, nb, k, sigma, gammas, ng)
      double precision, dimension(:,:), intent(in) :: a, b
      double precision, dimension(:), intent(in) :: gammas
      integer, intent(in) :: na, nb, ng
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
    end subroutine
    subroutine fmatern_kernel_l2(a, na, b, nb, k, sigma, order)
      double precision, dimension(:,:), intent(in) :: a, b
      integer, intent(in) :: na, nb
      double precision, dimension(:,:), intent(inout) :: k
      double precision, intent(in) :: sigma
      integer, intent(in) :: order
    end subroutine
    subroutine fp_distance_integer(a, b, d, p)
      double precision, dimension(:,:), intent(in) :: a, b
      double precision, dimension(:,:), intent(inout) :: d
      integer, intent(in) :: p
    end subroutine
    subroutine fp_distance_double(a, b, d, p)
      double precision, dimension(:,:), intent(in) :: a, b
      double precision, dimension(:,:), intent(inout) :: d
      double precision, intent(in) :: p
    end subroutine
  end interface
  integer :: n1, n2, o, pi
  double precision :: s, pp
  n1 = na
  n2 = nb
  s = sigma
  o = order
  pi = ip
  pp = p
  select case (kind)
  case (0)
    call flinear_kernel(a, n1, b, n2, k)
  case (1)
    call fsargan_kernel(a, n1, b, n2, k, s, gammas(1:ng), ng)
  case (2)
    call fmatern_kernel_l2(a, n1, b, n2, k, s, o)
  case (3)
    call fp_distance_integer(a, b, k, pi)
  case (4)
    call fp_distance_double(a, b, k, pp)
  end select
end subroutine
