! Field part of the free energy without a GPU: reads from <in.bin> (stream) int32 nx, ny,
! nz, nblk, then real64 box(3), chi(3,3) (row a, column b), w(n, nblk), phi(n, nblk), and
! writes field_energy (module scft_energy) to <out.bin> as one real64.
! usage: fe_check <in.bin> <out.bin>
program fe_check
  use scft_energy
  implicit none
  character(len=1024) :: fin, fout
  integer :: u, hdr(4), n, a, b
  real(8) :: box(3), chi9(9), chi(3, 3), fe
  real(8), allocatable :: w(:, :), phi(:, :)

  if (command_argument_count() /= 2) then
    write(*, '(a)') 'usage: fe_check <in.bin> <out.bin>'
    error stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  n = hdr(1) * hdr(2) * hdr(3)
  read(u) box, chi9
  do a = 1, 3
     do b = 1, 3
        chi(a, b) = chi9(3 * (a - 1) + b)
     end do
  end do
  allocate(w(0:n - 1, hdr(4)), phi(0:n - 1, hdr(4)))
  read(u) w, phi
  close(u)
  fe = field_energy(w, phi, chi, hdr(4), hdr(1), hdr(2), hdr(3), box)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) fe
  close(u)
end program fe_check
