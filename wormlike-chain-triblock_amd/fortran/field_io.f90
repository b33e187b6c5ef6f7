! Field files of the SCFT driver in the reference's per-rank text format: file
! 'W<rank>.dat', one line per local grid point K = 0 .. N-1 (x fastest), holding
! K and the block fields at K (W_dump, SCFT_dump.f90:273-293, writes "K, WA(K), WB(K)";
! readW, init_array_2dcomp_mod.f90:1070-1091, reads them back list-directed).  Three-block
! chains add the C field as a fourth column.  The values are written with 17 significant
! digits, so a dump read back by the list-directed reader returns every double exactly.
! Standalone: no dependency on the propagator module.
module field_io
  implicit none
  private
  public :: w_file_name, write_w_file, read_w_file

contains

  function w_file_name(rank) result(name)
    integer, intent(in) :: rank
    character(len=:), allocatable :: name
    character(len=16) :: buf
    write(buf, '(i0)') rank
    name = 'W' // trim(buf) // '.dat'
  end function w_file_name

  ! w(0:n-1, 1:nblk) -> path; stat = 0 on success, else the iostat of the failing statement
  subroutine write_w_file(path, w, n, nblk, stat)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n, nblk
    real(8), intent(in) :: w(0:n - 1, nblk)
    integer, intent(out) :: stat
    integer :: u, k
    open(newunit=u, file=path, status='replace', action='write', iostat=stat)
    if (stat /= 0) return
    do k = 0, n - 1
      write(u, '(i0, *(1x, es25.16e3))', iostat=stat) k, w(k, 1:nblk)
      if (stat /= 0) exit
    end do
    close(u)
  end subroutine write_w_file

  ! path -> w(0:n-1, 1:nblk), list-directed as the reference reads it; stat = 0 on success,
  ! -1 if a line's index is not its position, else the iostat of the failing statement
  subroutine read_w_file(path, w, n, nblk, stat)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n, nblk
    real(8), intent(out) :: w(0:n - 1, nblk)
    integer, intent(out) :: stat
    integer :: u, k, idx
    open(newunit=u, file=path, status='old', action='read', iostat=stat)
    if (stat /= 0) return
    do k = 0, n - 1
      read(u, *, iostat=stat) idx, w(k, 1:nblk)
      if (stat /= 0) exit
      if (idx /= k) then
        stat = -1
        exit
      end if
    end do
    close(u)
  end subroutine read_w_file

end module field_io
