! Volume quadrature and the field part of the free energy, without GPU dependence (used
! by module wormlike_MDE and by the CPU check program fe_check).
!
! periodic_simpson_grid: periodic Simpson volume integral, weights 2/3 (even index) and 4/3
!   (odd index) per axis times dx dy dz (simposon_3D_1D_mpi, utility_mod.f90:238-298).
! field_energy: (1/V) int [ sum_{a<b} chi_ab phi_a phi_b - sum_a w_a phi_a
!   - xi (1 - sum_a phi_a) ], xi = mean of the block fields at each point; for two blocks
!   this is the reference's chiN RA RB - wA RA - wB RB + (wA + wB)/2 (RA + RB - 1)
!   (SCFT_Density_FE.f90:296-298, :326-331).  Only the upper triangle of chi is read.
module scft_energy
  implicit none
contains

  real(8) function periodic_simpson_grid(f, nx, ny, nz, box) result(s)
    integer, intent(in) :: nx, ny, nz
    real(8), intent(in) :: f(0:), box(3)
    integer :: x, y, z
    real(8) :: cx, cy, cz
    s = 0d0
    do z = 0, nz - 1
       cz = merge(4d0/3d0, 2d0/3d0, mod(z, 2) == 1)
       do y = 0, ny - 1
          cy = merge(4d0/3d0, 2d0/3d0, mod(y, 2) == 1)
          do x = 0, nx - 1
             cx = merge(4d0/3d0, 2d0/3d0, mod(x, 2) == 1)
             s = s + cx * cy * cz * f(x + nx * (y + ny * z))
          end do
       end do
    end do
    s = s * (box(1) / nx) * (box(2) / ny) * (box(3) / nz)
  end function periodic_simpson_grid

  real(8) function field_energy(w, phi, chi, nblk, nx, ny, nz, box) result(fe)
    integer, intent(in) :: nblk, nx, ny, nz
    real(8), intent(in) :: w(0:, :), phi(0:, :), chi(3, 3), box(3)
    real(8), allocatable :: ar(:)
    integer :: r, a, b, n
    real(8) :: xi, sphi, e
    n = nx * ny * nz
    allocate(ar(0:n-1))
    do r = 0, n - 1
       xi = sum(w(r, 1:nblk)) / nblk
       sphi = sum(phi(r, 1:nblk))
       e = 0d0
       do a = 1, nblk
          do b = a + 1, nblk
             e = e + chi(a, b) * phi(r, a) * phi(r, b)
          end do
          e = e - w(r, a) * phi(r, a)
       end do
       ar(r) = e - xi * (1d0 - sphi)
    end do
    fe = periodic_simpson_grid(ar, nx, ny, nz, box) / (box(1) * box(2) * box(3))
    deallocate(ar)
  end function field_energy

end module scft_energy
