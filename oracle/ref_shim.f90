! Forwarding shim for the reference's bonded kernels (oracle; test infrastructure only).
!
! The reference binds cbf/caf/cdf (+ _d) to Python through f2py
! (setup.py:26-38, force.py:8-23).  f2py needs meson, absent from this image,
! so this file gives the same subroutines C entry points instead: explicit-shape
! bind(C) wrappers that call the reference subroutines, compiled from their own
! sources (compute_bond_forces*.f90, compute_angle_forces*.f90,
! compute_dihedral_forces*.f90, dipole_reconstruction*.f90).  No arithmetic
! happens here.  Arrays use Fortran (column-major) order exactly as f2py would
! pass np.asfortranarray inputs (main.py:500-506).
module hymd_ref_shim
  use iso_c_binding
  implicit none
  interface
    subroutine cbf(f, r, box, a, b, r0, k, energy, bond_pr)
      real(4), dimension(:,:), intent(in out) :: f
      real(4), dimension(:,:), intent(in)     :: r
      real(8), dimension(:),   intent(in)     :: box
      integer, dimension(:),   intent(in)     :: a, b
      real(8), dimension(:),   intent(in)     :: r0, k
      real(8),                 intent(out)    :: energy
      real(4), dimension(3),   intent(out)    :: bond_pr
    end subroutine
    subroutine cbf_d(f, r, box, a, b, r0, k, energy, bond_pr)
      real(8), dimension(:,:), intent(in out) :: f
      real(8), dimension(:,:), intent(in)     :: r
      real(8), dimension(:),   intent(in)     :: box
      integer, dimension(:),   intent(in)     :: a, b
      real(8), dimension(:),   intent(in)     :: r0, k
      real(8),                 intent(out)    :: energy
      real(8), dimension(3),   intent(out)    :: bond_pr
    end subroutine
    subroutine caf(f, r, box, a, b, c, t0, k, energy, angle_pr)
      real(4), dimension(:,:), intent(inout) :: f
      real(4), dimension(:,:), intent(in)    :: r
      real(8), dimension(:),   intent(in)    :: box
      integer, dimension(:),   intent(in)    :: a, b, c
      real(8), dimension(:),   intent(in)    :: t0, k
      real(8),                 intent(out)   :: energy
      real(4), dimension(3),   intent(out)   :: angle_pr
    end subroutine
    subroutine caf_d(f, r, box, a, b, c, t0, k, energy, angle_pr)
      real(8), dimension(:,:), intent(inout) :: f
      real(8), dimension(:,:), intent(in)    :: r
      real(8), dimension(:),   intent(in)    :: box
      integer, dimension(:),   intent(in)    :: a, b, c
      real(8), dimension(:),   intent(in)    :: t0, k
      real(8),                 intent(out)   :: energy
      real(8), dimension(3),   intent(out)   :: angle_pr
    end subroutine
    subroutine cdf(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, energy)
      real(4), intent(inout) :: force(:,:)
      real(4), intent(in)    :: r(:,:)
      real(4), intent(inout) :: dipoles(:,:,:)
      real(4), intent(inout) :: transfer_matrix(:,:,:,:)
      real(8), intent(in)    :: box(:)
      integer, intent(in)    :: a(:), b(:), c(:), d(:), dtype(:), bb_index(:), dipole_flag
      real(4), intent(in)    :: coeff(:,:,:)
      real(8), intent(out)   :: energy
    end subroutine
    subroutine cdf_d(force, r, dipoles, transfer_matrix, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, energy)
      real(8), intent(inout) :: force(:,:)
      real(8), intent(in)    :: r(:,:)
      real(8), intent(inout) :: dipoles(:,:,:)
      real(8), intent(inout) :: transfer_matrix(:,:,:,:)
      real(8), intent(in)    :: box(:)
      integer, intent(in)    :: a(:), b(:), c(:), d(:), dtype(:), bb_index(:), dipole_flag
      real(4), intent(in)    :: coeff(:,:,:)
      real(8), intent(out)   :: energy
    end subroutine
  end interface

contains

  subroutine ref_cbf(n, m, f, r, box, a, b, r0, k, energy, bond_pr) bind(C, name="ref_cbf")
    integer(c_int64_t), value :: n, m
    real(c_float)  :: f(n, 3), r(n, 3), bond_pr(3)
    real(c_double) :: box(3), r0(m), k(m), energy
    integer(c_int32_t) :: a(m), b(m)
    call cbf(f, r, box, a, b, r0, k, energy, bond_pr)
  end subroutine

  subroutine ref_cbf_d(n, m, f, r, box, a, b, r0, k, energy, bond_pr) bind(C, name="ref_cbf_d")
    integer(c_int64_t), value :: n, m
    real(c_double) :: f(n, 3), r(n, 3), bond_pr(3)
    real(c_double) :: box(3), r0(m), k(m), energy
    integer(c_int32_t) :: a(m), b(m)
    call cbf_d(f, r, box, a, b, r0, k, energy, bond_pr)
  end subroutine

  subroutine ref_caf(n, m, f, r, box, a, b, c, t0, k, energy, angle_pr) bind(C, name="ref_caf")
    integer(c_int64_t), value :: n, m
    real(c_float)  :: f(n, 3), r(n, 3), angle_pr(3)
    real(c_double) :: box(3), t0(m), k(m), energy
    integer(c_int32_t) :: a(m), b(m), c(m)
    call caf(f, r, box, a, b, c, t0, k, energy, angle_pr)
  end subroutine

  subroutine ref_caf_d(n, m, f, r, box, a, b, c, t0, k, energy, angle_pr) bind(C, name="ref_caf_d")
    integer(c_int64_t), value :: n, m
    real(c_double) :: f(n, 3), r(n, 3), angle_pr(3)
    real(c_double) :: box(3), t0(m), k(m), energy
    integer(c_int32_t) :: a(m), b(m), c(m)
    call caf_d(f, r, box, a, b, c, t0, k, energy, angle_pr)
  end subroutine

  subroutine ref_cdf(n, m, force, r, dipoles, tmat, box, a, b, c, d, coeff, dtype, bb_index, &
                     dipole_flag, energy) bind(C, name="ref_cdf")
    integer(c_int64_t), value :: n, m
    integer(c_int32_t), value :: dipole_flag
    real(c_float)  :: force(n, 3), r(n, 3), dipoles(m, 4, 3), tmat(m, 6, 3, 3), coeff(m, 6, 5)
    real(c_double) :: box(3), energy
    integer(c_int32_t) :: a(m), b(m), c(m), d(m), dtype(m), bb_index(m)
    call cdf(force, r, dipoles, tmat, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, energy)
  end subroutine

  subroutine ref_cdf_d(n, m, force, r, dipoles, tmat, box, a, b, c, d, coeff, dtype, bb_index, &
                       dipole_flag, energy) bind(C, name="ref_cdf_d")
    integer(c_int64_t), value :: n, m
    integer(c_int32_t), value :: dipole_flag
    real(c_double) :: force(n, 3), r(n, 3), dipoles(m, 4, 3), tmat(m, 6, 3, 3)
    real(c_float)  :: coeff(m, 6, 5)
    real(c_double) :: box(3), energy
    integer(c_int32_t) :: a(m), b(m), c(m), d(m), dtype(m), bb_index(m)
    call cdf_d(force, r, dipoles, tmat, box, a, b, c, d, coeff, dtype, bb_index, dipole_flag, energy)
  end subroutine

end module hymd_ref_shim
