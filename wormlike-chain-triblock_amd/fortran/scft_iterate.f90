! SCFT field update for an AB or ABC wormlike chain around the GPU propagator module
! (module wormlike_MDE): simple mixing, then Anderson mixing, with the reference's
! residuals and convergence test (AB_iterate, SCFT_iterate.f90:140-333, :1146-1150),
! generalised from two to n_blk species.
!
! Forces and outputs per species a (xi = mean_b w_b at each point, s_a = sum_{b /= a}
! chi_ab (phi_b - f_b), sbar = mean_a s_a, kappa = half the mean pair chi):
!   t_a     = s_a - sbar + kappa (sum_b phi_b - 1) + xi - w_a           (simple mixing, :145-147)
!   Wout_a  = w_a + t_a,  dW_a = t_a                                    (Anderson, :201-204)
! For two blocks t_a is exactly the reference's force chiN (phi_B - f_B) + (w_A + w_B)/2 - w_A;
! for more, the common mode sum_a t_a = n kappa (sum_b phi_b - 1) drives incompressibility
! as the reference's does for two (the plain extension sum_{b /= a} chi_ab (phi_b - f_b) +
! xi - w_a would instead drive sum_b c_b phi_b, c_b = sum_{a /= b} chi_ab, which is not the
! total density unless every chi is equal).  The reference's two-species Anderson output
! chiN phi_B + (w_A + w_B - chiN)/2 differs from Wout by the constant chiN (f_B - 1/2) per
! species, which shifts the converged fields by a constant and leaves phi and F unchanged.
! Residual per species: max_r |t_a| while mixing simply (:150-151), and
! sqrt(<dW_a, dW_a> / <Wg_a, Wg_a>) in the Anderson phase (:206-215), where <a, b> is the
! periodic-Simpson volume average (dot_product_mpi, :1200-1237).  Converged when every
! species' residual is below the tolerance (:1146).
! The reference divides by <Wout_a, Wout_a>, which depends on the field gauge: adding one
! constant to every field changes nothing else (phi, F, t and the whole mixing trajectory)
! but moves that denominator, and from zero-mean fields it is nearly zero.  Wg_a is Wout_a
! in a fixed gauge: Wout_a - <xi> + xbar, with xbar = (1/n) sum_a sum_{b /= a} chi_ab f_b the
! exchange field of the disordered melt, so the residual is a relative change of a field of
! size O(chi N) whatever constant the fields carry (the reference's own initial fields,
! init_w in init_array_2dcomp_mod.f90:843-1065, have means of that order).  The reference's extra exit after 150
! iterations with residuals below 5e-4 (:1151-1154) is not replicated.
module scft_iterate
  use wormlike_MDE
  implicit none

  integer :: n_simple = 20          ! Num_simple_mixing_AB (global_parameters.f90:132)
  integer :: n_hist = 10            ! Anderson_nim_AB (read from input in the reference)
  real(8) :: lam_simple = 0.05d0    ! lanbtWA / lanbtWB
  real(8) :: lam_anderson = 0.1d0   ! lambda_WAB_anderson (global_parameters.f90:135)
  ! 0: the reference's update w = sum_m c_m Wout_m + lambda sum_m c_m dW_m (:311-324);
  ! 1: w = sum_m c_m w_m + lambda sum_m c_m dW_m (mixing of the inputs; the reference's form
  !    is this one with lambda + 1)
  integer :: anderson_form = 0
  ! adaptive step (no counterpart in the reference, whose step is fixed): the Anderson step
  ! grows 5 % per iteration that lowers the largest residual, up to lam_max; when that
  ! residual exceeds ten times the best since the last restart the history restarts from the
  ! current entry (this iteration's update is a plain step) and the step halves (>= 0.01)
  logical :: adaptive = .false.
  real(8) :: lam_max = 1d0
  real(8) :: resid(3) = 0d0         ! residual per species of the last update
  integer :: restarts = 0
  integer, private :: nvalid = 0    ! history entries usable by the next Anderson step
  real(8), private :: lam_cur = 0.1d0, r_best = 1d300

  real(8), allocatable, private :: w_out(:, :, :), d_w(:, :, :)  ! (0:N-1, n_blk, 0:n_hist)

contains

  subroutine scft_iterate_init()
    integer :: n
    n = SIDEx * SIDEy * SIDEz
    if (allocated(w_out)) deallocate(w_out, d_w)
    allocate(w_out(0:n-1, n_blk, 0:n_hist), d_w(0:n-1, n_blk, 0:n_hist))
    w_out = 0d0
    d_w = 0d0
    nvalid = 0
    restarts = 0
    lam_cur = lam_anderson
    r_best = 1d300
  end subroutine scft_iterate_init

  ! ring slot of history entry i (Anderson_index, SCFT_iterate.f90:1-24)
  integer function hist_slot(i)
    integer, intent(in) :: i
    hist_slot = modulo(i, n_hist + 1)
  end function hist_slot

  ! periodic-Simpson volume average of a*b (dot_product_mpi)
  real(8) function vdot(a, b)
    real(8), intent(in) :: a(0:), b(0:)
    vdot = periodic_simpson(a * b) / (box_x * box_y * box_z)
  end function vdot

  ! LU with partial pivoting and solve of the n x n system a x = b (dgetrf + dgetrs as
  ! called by lapacksolver, matrix_inverse_mod.f90:129-200); b returns x
  subroutine lu_solve(a, b, n)
    integer, intent(in) :: n
    real(8), intent(inout) :: a(n, n), b(n)
    integer :: i, j, k, p
    real(8) :: t, row(n)
    do k = 1, n
      p = k
      do i = k + 1, n
        if (abs(a(i, k)) > abs(a(p, k))) p = i
      end do
      if (p /= k) then
        row = a(k, :); a(k, :) = a(p, :); a(p, :) = row
        t = b(k); b(k) = b(p); b(p) = t
      end if
      if (a(k, k) == 0d0) cycle
      do i = k + 1, n
        t = a(i, k) / a(k, k)
        a(i, k:n) = a(i, k:n) - t * a(k, k:n)
        b(i) = b(i) - t * b(k)
      end do
    end do
    do k = n, 1, -1
      t = b(k)
      do j = k + 1, n
        t = t - a(k, j) * b(j)
      end do
      if (a(k, k) /= 0d0) then
        b(k) = t / a(k, k)
      else
        b(k) = 0d0
      end if
    end do
  end subroutine lu_solve

  ! One field update after density(): iteration number iter (1-based, the reference's
  ! n_iter), interaction matrix chi(a, b), block fractions f.  Sets resid and converged.
  subroutine scft_update(iter, chi, f, tol, converged)
    integer, intent(in) :: iter
    real(8), intent(in) :: chi(3, 3), f(3), tol
    logical, intent(out) :: converged
    real(8), allocatable :: t(:, :), wo(:, :), xi(:), wi(:, :), di(:, :), u(:, :), v(:), wg(:), sbar(:)
    integer :: n, a, b, k, m, q, nr, km, kq
    real(8) :: xbar, rmax, kappa
    n = SIDEx * SIDEy * SIDEz
    allocate(t(0:n-1, n_blk), wo(0:n-1, n_blk), xi(0:n-1), sbar(0:n-1))
    xi = sum(w_blk(:, 1:n_blk), dim=2) / n_blk
    ! kappa: half the mean pair interaction (the reference's AB force has chiN / 2)
    kappa = 0d0
    do a = 1, n_blk
      do b = a + 1, n_blk
        kappa = kappa + chi(a, b)
      end do
    end do
    if (n_blk > 1) kappa = kappa / (n_blk * (n_blk - 1))
    do a = 1, n_blk
      t(:, a) = 0d0
      do b = 1, n_blk
        if (b /= a) t(:, a) = t(:, a) + chi(a, b) * (R_blk(:, b) - f(b))
      end do
    end do
    sbar = sum(t(:, 1:n_blk), dim=2) / n_blk
    do a = 1, n_blk
      t(:, a) = t(:, a) - sbar + kappa * (sum(R_blk(:, 1:n_blk), dim=2) - 1d0) + xi - w_blk(:, a)
      wo(:, a) = w_blk(:, a) + t(:, a)
    end do
    resid = 0d0
    if (iter <= n_simple) then
      if (iter == n_simple) then  ! first Anderson history entry
        w_out(:, :, 0) = wo
        d_w(:, :, 0) = wo - w_blk(:, 1:n_blk)
        nvalid = 1
      end if
      do a = 1, n_blk
        resid(a) = maxval(abs(t(:, a)))
        w_blk(:, a) = w_blk(:, a) + lam_simple * t(:, a)
      end do
    else
      k = hist_slot(iter - n_simple)
      nr = min(nvalid, n_hist)
      w_out(:, :, k) = wo
      d_w(:, :, k) = wo - w_blk(:, 1:n_blk)
      allocate(wg(0:n-1))
      xbar = 0d0
      do a = 1, n_blk
        do b = 1, n_blk
          if (b /= a) xbar = xbar + chi(a, b) * f(b)
        end do
      end do
      xbar = xbar / n_blk
      wg = 1d0
      xbar = xbar - vdot(xi, wg)
      do a = 1, n_blk
        wg = w_out(:, a, k) + xbar
        resid(a) = sqrt(abs(vdot(d_w(:, a, k), d_w(:, a, k)) / vdot(wg, wg)))
      end do
      deallocate(wg)
      allocate(u(nr, nr), v(nr), wi(0:n-1, n_blk), di(0:n-1, n_blk))
      u = 0d0
      v = 0d0
      do m = 1, nr
        km = hist_slot(k - m)
        do q = m, nr
          kq = hist_slot(k - q)
          do a = 1, n_blk
            u(m, q) = u(m, q) + vdot(d_w(:, a, k) - d_w(:, a, km), d_w(:, a, k) - d_w(:, a, kq))
          end do
          u(q, m) = u(m, q)
        end do
        do a = 1, n_blk
          v(m) = v(m) + vdot(d_w(:, a, k) - d_w(:, a, km), d_w(:, a, k))
        end do
      end do
      call lu_solve(u, v, nr)
      rmax = maxval(resid(1:n_blk))
      if (adaptive .and. rmax > 10d0 * r_best) then
        nvalid = 1
        restarts = restarts + 1
        lam_cur = max(0.5d0 * lam_cur, 0.01d0)
        nr = 0
        r_best = rmax
      else
        nvalid = nvalid + 1
        if (adaptive .and. rmax < r_best) lam_cur = min(lam_max, 1.05d0 * lam_cur)
        if (rmax < r_best) r_best = rmax
      end if
      wi = w_out(:, :, k)
      di = d_w(:, :, k)
      do m = 1, nr
        km = hist_slot(k - m)
        wi = wi + v(m) * (w_out(:, :, km) - w_out(:, :, k))
        di = di + v(m) * (d_w(:, :, km) - d_w(:, :, k))
      end do
      if (anderson_form == 1) wi = wi - di  ! sum_m c_m w_m
      w_blk(:, 1:n_blk) = wi + lam_cur * di
      deallocate(u, v, wi, di)
    end if
    converged = all(resid(1:n_blk) < tol)
    deallocate(t, wo, xi, sbar)
  end subroutine scft_update

end module scft_iterate
