! Fortran drop-in for the reference's MDE module on MI355X.
!
! The reference driver calls, once per SCFT iteration (SCFT_control.f90:243-257),
!     call MDE_q(.true.); call MDE_qstar(.true.); call density; call free_energy
! on procedures of `module wormlike_MDE` (wormlike_MDE_2dcomp_mod.f90:227-244,
! 961-973) and `density` / `free_energy` (SCFT_Density_FE.f90:3, 273), all reading and
! writing module globals.  This file keeps those names and that calling style and
! forwards the work through ISO_C_BINDING to libwlc_mde.so (include/wlc_mde.h).
!
! module wlc_mde_c    : bind(C) interfaces of the C ABI
! module wormlike_MDE : the globals a driver sets (grid, box, blocks, fields) and
!                       MDE_setup, MDE_q, MDE_qstar, density, free_energy, MDE_release
module wlc_mde_c
  use iso_c_binding
  implicit none

  type, bind(C) :: wlc_params
     integer(c_int) :: nx, ny, nz
     real(c_double) :: box(3)
     integer(c_int) :: nblk, Ns
     integer(c_int) :: ns_blk(3), lmax(3)   ! C: Nblk, Lmax
     real(c_double) :: lambda(3)
     integer(c_int) :: keep_qstar
     integer(c_int) :: ckpt_every            ! 0: automatic forward store / checkpoints
     integer(c_int) :: spectrum_fp32
     integer(c_int) :: reserved(5)
  end type wlc_params

  interface
     integer(c_int) function wlc_mde_create(p, h) bind(C, name="wlc_mde_create")
       import :: wlc_params, c_int, c_ptr
       type(wlc_params), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function wlc_mde_create

     integer(c_int) function wlc_mde_set_box(h, box) bind(C, name="wlc_mde_set_box")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: box(3)
     end function wlc_mde_set_box

     integer(c_int) function wlc_mde_set_w(h, w) bind(C, name="wlc_mde_set_w")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: w(*)
     end function wlc_mde_set_w

     integer(c_int) function wlc_mde_forward(h) bind(C, name="wlc_mde_forward")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function wlc_mde_forward

     integer(c_int) function wlc_mde_backward(h) bind(C, name="wlc_mde_backward")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function wlc_mde_backward

     integer(c_int) function wlc_mde_density(h, phi, Q, totden) bind(C, name="wlc_mde_density")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double) :: phi(*)
       real(c_double) :: Q, totden
     end function wlc_mde_density

     integer(c_int) function wlc_mde_set_spectrum_fp32(h, on) bind(C, name="wlc_mde_set_spectrum_fp32")
       import :: c_ptr, c_int
       type(c_ptr), value :: h
       integer(c_int), value :: on
     end function wlc_mde_set_spectrum_fp32

     subroutine wlc_mde_destroy(h) bind(C, name="wlc_mde_destroy")
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine wlc_mde_destroy

     type(c_ptr) function wlc_mde_last_error() bind(C, name="wlc_mde_last_error")
       import :: c_ptr
     end function wlc_mde_last_error

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  subroutine wlc_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    type(c_ptr) :: p
    character(kind=c_char), pointer :: msg(:)
    integer :: n, i
    character(len=512) :: text
    if (rc == 0) return
    p = wlc_mde_last_error()
    n = int(c_strlen(p))
    text = ''
    if (n > 0) then
       call c_f_pointer(p, msg, [n])
       do i = 1, min(n, 512)
          text(i:i) = msg(i)
       end do
    end if
    write(*,'(a,a,a,i0,a,a)') 'wormlike_MDE: ', where, ' failed, status ', rc, ': ', trim(text)
    error stop 1
  end subroutine wlc_check

end module wlc_mde_c


module wormlike_MDE
  use iso_c_binding
  use wlc_mde_c
  use scft_energy
  implicit none

  ! ---- globals the driver sets (the reference keeps these in global_para) ----
  integer :: SIDEx = 0, SIDEy = 0, SIDEz = 0          ! grid
  real(8) :: box_x = 0d0, box_y = 0d0, box_z = 0d0    ! box edges (SIDE*d)
  integer :: n_blk = 3                                ! blocks (2: AB, 3: ABC)
  integer :: Nmax = 0                                 ! contour steps, ds = 1/Nmax
  integer :: Ns_blk(3) = 0                            ! NA, NB, NC
  integer :: L_bar_blk(3) = 0                         ! Lmax per block
  real(8) :: lambda_blk(3) = 1d0                      ! Loa/(2 kappa) per block
  real(8), allocatable :: w_blk(:,:)                  ! (0:LOCAL_SIZE-1, 1:n_blk), x fastest
  real(8) :: chi_blk(3,3) = 0d0                       ! chi_ab N (upper triangle read; AB: NXab)
  ! ---- outputs (density / free_energy) ----
  real(8), allocatable :: R_blk(:,:)                  ! phi per block, same layout as w_blk
  real(8) :: pff_global = 0d0                         ! ln Q   (SCFT_Density_FE.f90:91)
  real(8) :: totden = 0d0
  real(8) :: FE_global = 0d0                          ! F per chain (SCFT_Density_FE.f90:333)
  real(8) :: tempEE_global = 0d0                      ! field-energy volume average (SCFT_Density_FE.f90:326-331)

  type(c_ptr), private, save :: handle = c_null_ptr

contains

  ! create the GPU solver from the globals (replaces the derived-parameter part of
  ! initialize, init_arrays_2decomp allocations and init_inverse_sparse_GAB)
  subroutine MDE_setup()
    type(wlc_params) :: p
    integer :: b
    if (c_associated(handle)) call MDE_release()
    p%nx = SIDEx; p%ny = SIDEy; p%nz = SIDEz
    p%box = [box_x, box_y, box_z]
    p%nblk = n_blk
    p%Ns = Nmax
    p%ns_blk = 0; p%lmax = 0; p%lambda = 0d0
    do b = 1, n_blk
       p%ns_blk(b) = Ns_blk(b)
       p%lmax(b) = L_bar_blk(b)
       p%lambda(b) = lambda_blk(b)
    end do
    p%keep_qstar = 0
    p%ckpt_every = 0
    p%spectrum_fp32 = 0
    p%reserved = 0
    call wlc_check(wlc_mde_create(p, handle), 'wlc_mde_create')
    if (.not. allocated(w_blk)) allocate(w_blk(0:SIDEx*SIDEy*SIDEz-1, n_blk))
    if (.not. allocated(R_blk)) allocate(R_blk(0:SIDEx*SIDEy*SIDEz-1, n_blk))
  end subroutine MDE_setup

  ! box change (the reference rebuilds G^-1 tables: init_inverse_sparse_GAB)
  subroutine MDE_set_box(bx, by, bz)
    real(8), intent(in) :: bx, by, bz
    box_x = bx; box_y = by; box_z = bz
    call wlc_check(wlc_mde_set_box(handle, [bx, by, bz]), 'wlc_mde_set_box')
  end subroutine MDE_set_box

  ! forward propagator (wormlike_MDE_2dcomp_mod.f90:231); uploads the current fields
  subroutine MDE_q(R2C_FFT)
    logical, optional :: R2C_FFT
    call wlc_check(wlc_mde_set_w(handle, w_blk), 'wlc_mde_set_w')
    call wlc_check(wlc_mde_forward(handle), 'wlc_mde_forward')
  end subroutine MDE_q

  ! backward propagator (wormlike_MDE_2dcomp_mod.f90:961)
  subroutine MDE_qstar(R2C_FFT)
    logical, optional :: R2C_FFT
    call wlc_check(wlc_mde_backward(handle), 'wlc_mde_backward')
  end subroutine MDE_qstar

  ! FP32 storage of the half spectrum (all arithmetic stays FP64) on or off for the next
  ! sweeps: an SCFT loop can iterate in FP32 far from convergence and finish in FP64
  subroutine MDE_set_spectrum_fp32(on)
    logical, intent(in) :: on
    integer(c_int) :: v
    v = 0
    if (on) v = 1
    call wlc_check(wlc_mde_set_spectrum_fp32(handle, v), 'wlc_mde_set_spectrum_fp32')
  end subroutine MDE_set_spectrum_fp32

  ! Call slot of the reference's split-operator MDE (split_operator_MDE_mod.f90:24-115
  ! split_operator_MDE_init, called whenever the cell changes; :120-238
  ! AB_diblock_sph_driver, both propagators).  A driver built for that variant keeps its
  ! calls; here they run the spherical-harmonic BDF solve of MDE_q / MDE_qstar (the
  ! split-operator numerics on the (theta, phi) grid are not rebuilt), and density
  ! consumes the result as after MDE_q / MDE_qstar.
  subroutine split_operator_MDE_init()
    if (c_associated(handle)) then
       call MDE_set_box(box_x, box_y, box_z)
    else
       call MDE_setup()
    end if
  end subroutine split_operator_MDE_init

  subroutine AB_diblock_sph_driver()
    call MDE_q(.true.)
    call MDE_qstar(.true.)
  end subroutine AB_diblock_sph_driver

  ! Q, totden and phi per block (SCFT_Density_FE.f90:3-159)
  subroutine density()
    real(c_double) :: Q, td
    call wlc_check(wlc_mde_density(handle, R_blk, Q, td), 'wlc_mde_density')
    pff_global = log(Q)
    totden = td
  end subroutine density

  ! F = -ln Q + (1/V) int [ sum_{a<b} chi_ab phi_a phi_b - sum_a w_a phi_a - xi (1 - sum_a phi_a) ],
  ! xi = mean of the block fields; for two blocks this is SCFT_Density_FE.f90:296-298, :333.
  ! Like the reference's free_energy() it reads the module globals (w_blk, R_blk, pff_global
  ! and the interaction matrix chi_blk, the reference's NXab); a driver may pass chi instead.
  subroutine free_energy(chi)
    real(8), intent(in), optional :: chi(3, 3)
    if (present(chi)) then
       tempEE_global = field_energy(w_blk, R_blk, chi, n_blk, SIDEx, SIDEy, SIDEz, &
            [box_x, box_y, box_z])
    else
       tempEE_global = field_energy(w_blk, R_blk, chi_blk, n_blk, SIDEx, SIDEy, SIDEz, &
            [box_x, box_y, box_z])
    end if
    FE_global = -pff_global + tempEE_global
  end subroutine free_energy

  ! periodic Simpson volume integral over the module grid (utility_mod.f90:238-298)
  real(8) function periodic_simpson(f) result(s)
    real(8), intent(in) :: f(0:)
    s = periodic_simpson_grid(f, SIDEx, SIDEy, SIDEz, [box_x, box_y, box_z])
  end function periodic_simpson

  subroutine MDE_release()
    if (c_associated(handle)) call wlc_mde_destroy(handle)
    handle = c_null_ptr
  end subroutine MDE_release

end module wormlike_MDE
