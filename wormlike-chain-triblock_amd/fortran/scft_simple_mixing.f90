! Plain SCFT with simple mixing for an AB or ABC wormlike chain, driving the GPU
! propagator solve through module wormlike_MDE exactly like the reference driver loop
! (SCFT_control.f90:237-257: MDE_q -> MDE_qstar -> density -> free_energy -> iterate).
!
! Mixing (reference AB rule SCFT_iterate.f90:145-153 generalised to n blocks):
!   xi  = mean_b w_b
!   t_a = sum_{b /= a} chi_ab (phi_b - f_b) + xi - w_a,   w_a <- w_a + step * t_a
!   residual = max_{a, r} |t_a(r)|
!
! usage: scft_simple_mixing <input.bin> <output.txt> [final_w.bin]
! input (stream, native endian): int32 nx, ny, nz, nblk, Ns, Nblk(3), Lmax(3), niter;
!   real64 box(3), lambda(3), chi(3,3) (row a, column b), step; real64 w(nx*ny*nz, nblk)
! output: one line per iteration: iter, F, lnQ, residual, mean phi per block
program scft_simple_mixing
  use wormlike_MDE
  implicit none
  character(len=1024) :: fin, fout, fw
  integer :: u, hdr(13), niter, it, a, b, n, nargs
  real(8) :: box(3), lam(3), chi9(9), chi(3, 3), step, resid, xi, f(3)
  real(8), allocatable :: t(:, :)

  nargs = command_argument_count()
  if (nargs < 2) then
     write(*, '(a)') 'usage: scft_simple_mixing <input.bin> <output.txt> [final_w.bin]'
     error stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  fw = ''
  if (nargs >= 3) call get_command_argument(3, fw)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  SIDEx = hdr(1); SIDEy = hdr(2); SIDEz = hdr(3)
  n_blk = hdr(4); Nmax = hdr(5)
  Ns_blk = hdr(6:8); L_bar_blk = hdr(9:11); niter = hdr(12)
  read(u) box, lam, chi9, step
  box_x = box(1); box_y = box(2); box_z = box(3)
  lambda_blk = lam
  do a = 1, 3
     do b = 1, 3
        chi(a, b) = chi9(3 * (a - 1) + b)
     end do
  end do
  call MDE_setup()
  n = SIDEx * SIDEy * SIDEz
  read(u) w_blk
  close(u)
  allocate(t(0:n-1, n_blk))
  f = 0d0
  do a = 1, n_blk
     f(a) = dble(Ns_blk(a)) / dble(Nmax)
  end do

  open(newunit=u, file=trim(fout), status='replace', action='write')
  do it = 1, niter
     call MDE_q(.true.)
     call MDE_qstar(.true.)
     call density()
     call free_energy(chi)
     resid = 0d0
     do a = 1, n_blk
        t(:, a) = 0d0
     end do
     do a = 1, n_blk
        do b = 1, n_blk
           if (b /= a) t(:, a) = t(:, a) + chi(a, b) * (R_blk(:, b) - f(b))
        end do
     end do
     do a = 1, n_blk
        t(:, a) = t(:, a) - w_blk(:, a)
     end do
     do b = 0, n - 1
        xi = sum(w_blk(b, 1:n_blk)) / n_blk
        t(b, 1:n_blk) = t(b, 1:n_blk) + xi
     end do
     resid = maxval(abs(t))
     write(u, '(i6, 3es25.16e3, 3es25.16e3)') it, FE_global, pff_global, resid, &
          (sum(R_blk(:, a)) / n, a = 1, n_blk)
     w_blk = w_blk + step * t
  end do
  close(u)
  if (len_trim(fw) > 0) then
     open(newunit=u, file=trim(fw), access='stream', form='unformatted', status='replace')
     write(u) w_blk
     close(u)
  end if
  call MDE_release()
end program scft_simple_mixing
