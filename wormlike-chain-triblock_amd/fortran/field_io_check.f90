! Round trip of module field_io without a GPU: reads w(0:n-1, 1:nblk) as raw doubles from
! <in.bin>, writes it as a field file <w.dat>, reads that back and writes the result as raw
! doubles to <out.bin>.  With <in.bin> = '-' it only reads <w.dat> (a field file written
! elsewhere) into <out.bin>.
! usage: field_io_check <n> <nblk> <in.bin|-> <w.dat> <out.bin>
program field_io_check
  use field_io
  implicit none
  character(len=1024) :: a(5)
  integer :: n, nblk, i, u, stat
  real(8), allocatable :: w(:, :), r(:, :)

  if (command_argument_count() /= 5) then
    write(*, '(a)') 'usage: field_io_check <n> <nblk> <in.bin|-> <w.dat> <out.bin>'
    error stop 2
  end if
  do i = 1, 5
    call get_command_argument(i, a(i))
  end do
  read(a(1), *) n
  read(a(2), *) nblk
  allocate(w(0:n - 1, nblk), r(0:n - 1, nblk))
  if (trim(a(3)) /= '-') then
    open(newunit=u, file=trim(a(3)), access='stream', form='unformatted', status='old')
    read(u) w
    close(u)
    call write_w_file(trim(a(4)), w, n, nblk, stat)
    if (stat /= 0) then
      write(*, '(a, i0)') 'write_w_file failed, iostat ', stat
      error stop 1
    end if
  end if
  call read_w_file(trim(a(4)), r, n, nblk, stat)
  if (stat /= 0) then
    write(*, '(a, i0)') 'read_w_file failed, status ', stat
    error stop 1
  end if
  open(newunit=u, file=trim(a(5)), access='stream', form='unformatted', status='replace')
  write(u) r
  close(u)
end program field_io_check
