! Plain SCFT for an AB or ABC wormlike chain, driving the GPU propagator solve through
! module wormlike_MDE the way the reference driver loop does (SCFT_control.f90:237-262:
! MDE_q -> MDE_qstar -> density -> free_energy -> field update, until converged or the
! iteration limit), with simple or simple-then-Anderson mixing (module scft_iterate).
!
! usage: scft_driver <input.bin> <output.txt> [final_w.bin] [--readW] [--dumpW]
!   --readW  start from the field file W0.dat in the working directory instead of the
!            fields of input.bin (the reference's wAB_init_type < 0 -> readW,
!            SCFT_control.f90:79-83)
!   --dumpW  write the final fields to W0.dat (W_dump, SCFT_dump.f90:273-293)
!   --spectrum-fp32 <tol>  iterate with the half spectrum stored as FP32 (all arithmetic
!            FP64) until the largest residual is below tol, then switch to FP64 for the
!            remaining iterations; convergence is only declared in FP64, so the final
!            fields, residual and free energy come from FP64 sweeps
!   --split-operator-slot  set up and solve through the split-operator call slot
!            (split_operator_MDE_init / AB_diblock_sph_driver) instead of MDE_setup /
!            MDE_q / MDE_qstar (same results: the slot runs the same solve)
!   --adaptive <lam_max>  adaptive Anderson step (module scft_iterate: grows to lam_max while
!            the residual falls, history restart and half step when it grows tenfold past the
!            best since the last restart)
!   --logs   write the reference's run logs in the working directory (created empty at
!            start as initialize_mod.f90:296-306 does; SCFT_dump.f90:35-56 and :183-197
!            append after every iteration): error.dat (iteration, residual per species),
!            FreeEG.dat (iteration, F, ln Q, field energy; iteration, box x, "FE-FH",
!            F - sum_{a<b} chi_ab f_a f_b; then "dx,dy" and "box size" lines until
!            converged), and when converged result.txt (appended) and result.dat
!            (replaced): FREE_ENERGY, FREE_ENERGY-FH, box size, converge (records up to
!            1024 characters, so no list-directed line wraps)
! input (stream, native endian):
!   int32  nx, ny, nz, nblk, Ns, Nblk(3), Lmax(3), max_iter, method (0 simple mixing only,
!          1 simple then Anderson), n_simple, n_hist, Anderson form (0 reference, 1 mixing
!          of the inputs; see module scft_iterate)
!   real64 box(3), lambda(3), chi(3,3) (row a, column b), lam_simple, lam_anderson, tol
!   real64 w(nx*ny*nz, nblk)
! output: one line per iteration: iter, F, lnQ, max residual, residual per species (3),
!   mean phi per block (3), max |w| after the update; then
!   "# iterations <n> converged <0|1> seconds <wall>"
!   (wall time of the loop: propagators, densities and field updates)
program scft_driver
  use wormlike_MDE
  use scft_iterate
  use field_io
  implicit none
  character(len=1024) :: fin, fout, fw, arg
  integer :: u, hdr(16), max_iter, method, it, a, b, n, nargs, done, i, npos, stat
  logical :: read_w, dump_w, sph_slot, fp32_on, next_is_tol, logs, next_is_lmax
  real(8) :: fp32_tol, lmax_arg
  integer(8) :: c0, c1, rate
  real(8) :: box(3), lam(3), chi9(9), chi(3, 3), f(3), tol, mix(2)
  logical :: converged

  nargs = command_argument_count()
  fin = ''; fout = ''; fw = ''
  read_w = .false.; dump_w = .false.; sph_slot = .false.; logs = .false.
  fp32_on = .false.; next_is_tol = .false.; fp32_tol = 0d0
  next_is_lmax = .false.; lmax_arg = 1d0
  npos = 0
  do i = 1, nargs
     call get_command_argument(i, arg)
     if (next_is_tol) then
        read(arg, *) fp32_tol
        fp32_on = fp32_tol > 0d0
        next_is_tol = .false.
     else if (next_is_lmax) then
        read(arg, *) lmax_arg
        adaptive = .true.
        next_is_lmax = .false.
     else if (trim(arg) == '--spectrum-fp32') then
        next_is_tol = .true.
     else if (trim(arg) == '--readW') then
        read_w = .true.
     else if (trim(arg) == '--dumpW') then
        dump_w = .true.
     else if (trim(arg) == '--split-operator-slot') then
        sph_slot = .true.
     else if (trim(arg) == '--logs') then
        logs = .true.
     else if (trim(arg) == '--adaptive') then
        next_is_lmax = .true.
     else
        npos = npos + 1
        if (npos == 1) fin = arg
        if (npos == 2) fout = arg
        if (npos == 3) fw = arg
     end if
  end do
  if (npos < 2 .or. npos > 3 .or. next_is_tol .or. next_is_lmax) then
     write(*, '(a)') 'usage: scft_driver <input.bin> <output.txt> [final_w.bin] [--readW] [--dumpW] &
          &[--spectrum-fp32 <tol>] [--split-operator-slot] [--logs] [--adaptive <lam_max>]'
     error stop 2
  end if

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  SIDEx = hdr(1); SIDEy = hdr(2); SIDEz = hdr(3)
  n_blk = hdr(4); Nmax = hdr(5)
  Ns_blk = hdr(6:8); L_bar_blk = hdr(9:11)
  max_iter = hdr(12); method = hdr(13)
  n_simple = hdr(14); n_hist = max(1, hdr(15)); anderson_form = hdr(16)
  if (method == 0) n_simple = max_iter + 1
  read(u) box, lam, chi9, mix, tol
  lam_simple = mix(1); lam_anderson = mix(2)
  lam_max = lmax_arg
  box_x = box(1); box_y = box(2); box_z = box(3)
  lambda_blk = lam
  do a = 1, 3
     do b = 1, 3
        chi(a, b) = chi9(3 * (a - 1) + b)
     end do
  end do
  chi_blk = chi
  if (sph_slot) then
     call split_operator_MDE_init()
  else
     call MDE_setup()
  end if
  n = SIDEx * SIDEy * SIDEz
  read(u) w_blk
  close(u)
  if (read_w) then
     call read_w_file(w_file_name(0), w_blk, n, n_blk, stat)
     if (stat /= 0) then
        write(*, '(a, a, a, i0)') 'scft_driver: cannot read ', w_file_name(0), ', status ', stat
        error stop 1
     end if
  end if
  call scft_iterate_init()
  if (fp32_on) call MDE_set_spectrum_fp32(.true.)
  f = 0d0
  do a = 1, n_blk
     f(a) = dble(Ns_blk(a)) / dble(Nmax)
  end do

  if (logs) call logs_create()
  open(newunit=u, file=trim(fout), status='replace', action='write')
  converged = .false.
  done = 0
  call system_clock(c0, rate)
  do it = 1, max_iter
     if (sph_slot) then
        call AB_diblock_sph_driver()
     else
        call MDE_q(.true.)
        call MDE_qstar(.true.)
     end if
     call density()
     call free_energy()
     call scft_update(it, chi, f, tol, converged)
     if (fp32_on .and. (converged .or. maxval(resid(1:n_blk)) < fp32_tol)) then
        ! the rest, including the convergence test, in FP64
        call MDE_set_spectrum_fp32(.false.)
        fp32_on = .false.
        converged = .false.
     end if
     write(u, '(i6, 3es25.16e3, 3es25.16e3, 3es25.16e3, es25.16e3)') it, FE_global, pff_global, &
          maxval(resid(1:n_blk)), resid, (sum(R_blk(:, a)) / n, a = 1, n_blk), &
          (0d0, a = n_blk + 1, 3), maxval(abs(w_blk))
     flush(u)
     if (logs) call logs_append(it, converged)
     done = it
     if (converged) exit
  end do
  call system_clock(c1)
  write(u, '(a, i0, a, i0, a, es16.8)') '# iterations ', done, ' converged ', merge(1, 0, converged), &
       ' seconds ', dble(c1 - c0) / dble(rate)
  close(u)
  if (dump_w) then
     call write_w_file(w_file_name(0), w_blk, n, n_blk, stat)
     if (stat /= 0) then
        write(*, '(a, a, a, i0)') 'scft_driver: cannot write ', w_file_name(0), ', iostat ', stat
        error stop 1
     end if
  end if
  if (len_trim(fw) > 0) then
     open(newunit=u, file=trim(fw), access='stream', form='unformatted', status='replace')
     write(u) w_blk
     close(u)
  end if
  call MDE_release()

contains

  subroutine logs_create()
    integer :: v
    open(newunit=v, file='FreeEG.dat', status='replace')
    close(v)
    open(newunit=v, file='error.dat', status='replace')
    close(v)
    open(newunit=v, file='result.txt', status='replace')
    close(v)
  end subroutine logs_create

  subroutine logs_append(iter, conv)
    integer, intent(in) :: iter
    logical, intent(in) :: conv
    integer :: v, p, q
    real(8) :: fh, r3(3)
    fh = 0d0
    do p = 1, n_blk
       do q = p + 1, n_blk
          fh = fh + chi(p, q) * f(p) * f(q)
       end do
    end do
    r3 = 0d0
    r3(1:n_blk) = resid(1:n_blk)
    open(newunit=v, file='error.dat', status='old', position='append', recl=1024)
    write(v, *) iter, r3
    close(v)
    open(newunit=v, file='FreeEG.dat', status='old', position='append', recl=1024)
    if (conv) then
       write(v, *) iter, FE_global, pff_global, tempEE_global, FE_global - fh
       write(v, *) iter, box_x, 'FE-FH', FE_global - fh
    else
       write(v, *) iter, FE_global, pff_global, tempEE_global
       write(v, *) iter, box_x, 'FE-FH', FE_global - fh
       write(v, *) 'dx,dy', box_x / SIDEx, box_y / SIDEy
       write(v, *) 'box size', box_x, box_y, box_z
    end if
    close(v)
    if (conv) then
       open(newunit=v, file='result.txt', status='old', position='append', recl=1024)
       call result_lines(v, fh, conv)
       close(v)
       open(newunit=v, file='result.dat', status='replace', recl=1024)
       call result_lines(v, fh, conv)
       close(v)
    end if
  end subroutine logs_append

  subroutine result_lines(v, fh, conv)
    integer, intent(in) :: v
    real(8), intent(in) :: fh
    logical, intent(in) :: conv
    write(v, *) 'FREE_ENERGY', FE_global, pff_global, tempEE_global, FE_global - fh
    write(v, *) 'FREE_ENERGY-FH', FE_global - fh
    write(v, *) 'box size', box_x, box_y, box_z
    write(v, *) 'converge', conv
  end subroutine result_lines
end program scft_driver
