! gsis_capi.f90 -- ISO_C_BINDING interface to libgsis.so, the MI355X GSIS
! kinetic loop (include/gsis_capi.h).
!
! A Fortran host (Hybrid2D's opencfd_ec2d main loop, src/opencfd_ec2d.f90:45-56)
! uses this module to replace Hybrid_Time_advance_LU_SGS (src/Hybrid_Sub.f90:4-160)
! and DETFP_res (src/GKUA_DETFP_RES.f90:4-60) by calls into the library.  Every
! array is passed in the reference's own layout (column-major, ghost lower
! bounds as in src/COMMONDATA.f90): contiguous allocatables such as B%FRD or
! B%S_GKUA are passed unchanged.  Arguments the C side accepts as NULL are
! type(c_ptr), value (pass c_loc(array) or c_null_ptr).
module gsis_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GSIS_OK = 0, GSIS_EARG = 1, GSIS_EHIP = 2, GSIS_ENAN = 3, &
                               GSIS_EUNSUPPORTED = 4, GSIS_ETEMP = 5
  integer(c_int32_t), parameter :: GSIS_BC_WALL = -2, GSIS_BC_FARFIELD = -4, GSIS_BC_PERIODIC = -5
  ! symmetry face; a code >= 1 is an inner face (the neighbour block number)
  integer(c_int32_t), parameter :: GSIS_BC_SYMMETRY = -3
  integer(c_int32_t), parameter :: GSIS_SCHEME_UP1 = 0, GSIS_SCHEME_NND = 1, GSIS_SCHEME_WENO3 = 3
  integer(c_int32_t), parameter :: GSIS_MODE_M0 = 0, GSIS_MODE_M1 = 1
  integer(c_int32_t), parameter :: GSIS_PREC_FP64 = 0, GSIS_PREC_FP32 = 1
  integer(c_int32_t), parameter :: GSIS_KERNEL_AUTO = 0, GSIS_KERNEL_SPLIT = 1, GSIS_KERNEL_TILES = 2, &
                                   GSIS_KERNEL_STRIP16 = 3

  ! BC_MSG_TYPE (src/COMMONDATA.f90:34-41) as filled by read_bcin (src/sub_init.f90:266-278)
  type, bind(C) :: gsis_bc_t
    integer(c_int32_t) :: face, ist, iend, jst, jend, code
    real(c_double) :: uw, vw
  end type gsis_bc_t

  ! everything one step reads from the module globals of COMMONDATA
  type, bind(C) :: gsis_params_t
    integer(c_int32_t) :: nx, ny, nvi, nvj
    type(c_ptr) :: vi, vj, wi, wj
    type(c_ptr) :: x, y
    type(c_ptr) :: vol, si, sj, ni1, ni2, nj1, nj2
    integer(c_int32_t) :: nbc
    type(c_ptr) :: bc
    real(c_double) :: cfc, x1s, cpr, twall, dt
    real(c_double) :: ma, gamma, re, pr, cv, cp, aoa
    integer(c_int32_t) :: scheme, mode, precision, if_viscous
    integer(c_int32_t) :: local_dt
    real(c_double) :: cfl, dtmax, dtmin
    real(c_double) :: ns_tol
    integer(c_int32_t) :: ns_maxit
    integer(c_int32_t) :: kstep_show
    integer(c_int32_t) :: split_kernels
    integer(c_int32_t) :: walk_tiles
  end type gsis_params_t

  interface
    integer(c_int) function gsis_create(p, h) bind(C, name="gsis_create")
      import :: c_int, c_ptr, gsis_params_t
      type(gsis_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function gsis_upload_frd(h, frd) bind(C, name="gsis_upload_frd")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: frd(*)
    end function
    integer(c_int) function gsis_upload_macro(h, s, cf, u, u0) bind(C, name="gsis_upload_macro")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, s, cf, u, u0
    end function
    integer(c_int) function gsis_refresh_moments(h) bind(C, name="gsis_refresh_moments")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function gsis_step(h, nsteps) bind(C, name="gsis_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function gsis_residual(h, res_rms) bind(C, name="gsis_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: res_rms(4)
    end function
    integer(c_int) function gsis_download_macro(h, s, s0, cf, u, taonsi_g, taonsj_g, qi_g, qj_g, swall) &
        bind(C, name="gsis_download_macro")
      import :: c_int, c_ptr
      type(c_ptr), value :: h, s, s0, cf, u, taonsi_g, taonsj_g, qi_g, qj_g, swall
    end function
    integer(c_int) function gsis_download_frd(h, frd) bind(C, name="gsis_download_frd")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: frd(*)
    end function
    integer(c_int) function gsis_download_frd_noalloc(h, frd) bind(C, name="gsis_download_frd_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: frd(*)
    end function
    ! the uniform Maxwellian fill of GKUA_sudufenbu on the device (no FRD array)
    integer(c_int) function gsis_fill_frd_uniform(h, g, hv) bind(C, name="gsis_fill_frd_uniform")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: g(*), hv(*)
    end function
    ! MassCorrect (src/GKUA_DETFP.f90:153-172) after the moments of every later step
    integer(c_int) function gsis_set_mass_correct(h, on) bind(C, name="gsis_set_mass_correct")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: on
    end function
    ! the learned HoT closure in the synthetic solve (net from gsis_mlp_create, or c_null_ptr)
    integer(c_int) function gsis_set_closure(h, net, gate, every) bind(C, name="gsis_set_closure")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h, net
      real(c_double), value :: gate
      integer(c_int32_t), value :: every
    end function
    integer(c_int) function gsis_set_closure_warmup(h, net, nsteps, every) bind(C, name="gsis_set_closure_warmup")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, net
      integer(c_int32_t), value :: nsteps, every
    end function
    integer(c_int) function gsis_mlp_create(net) bind(C, name="gsis_mlp_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: net
    end function
    integer(c_int) function gsis_mlp_set_params(net, params) bind(C, name="gsis_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(in) :: params(*)
    end function
    integer(c_int) function gsis_mlp_set_scaling(net, scaling) bind(C, name="gsis_mlp_set_scaling")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(in) :: scaling(*)
    end function
    integer(c_int) function gsis_mlp_destroy(net) bind(C, name="gsis_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: net
    end function
    ! batches of handles (hs: an array of handles)
    integer(c_int) function gsis_batch_create(hs, n, b) bind(C, name="gsis_batch_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: b
    end function
    integer(c_int) function gsis_batch_step(b, nsteps) bind(C, name="gsis_batch_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function gsis_batch_tune(b, steps_per_trial, nw, order, ms_per_step) bind(C, name="gsis_batch_tune")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: b
      integer(c_int32_t), value :: steps_per_trial
      integer(c_int32_t), intent(out) :: nw, order
      real(c_double), intent(out) :: ms_per_step
    end function
    integer(c_int) function gsis_batch_set_ns_graph(b, mode) bind(C, name="gsis_batch_set_ns_graph")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function gsis_batch_ns_graph(b, count) bind(C, name="gsis_batch_ns_graph")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: b
      integer(c_int32_t), intent(out) :: count
    end function
    integer(c_int) function gsis_batch_destroy(b) bind(C, name="gsis_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function
    integer(c_int) function gsis_counters(h, kstep, tt, ns_iter) bind(C, name="gsis_counters")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: kstep, ns_iter
      real(c_double), intent(out) :: tt
    end function
    integer(c_int) function gsis_set_counters(h, kstep, tt) bind(C, name="gsis_set_counters")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: kstep
      real(c_double), value :: tt
    end function
    ! GKUA_inner_mesco_boundary + GKUA_inner_macro_boundary for one inner
    ! subface (src/GKUA_boundary.f90:400-633); ksub, msub 1-based, orient = Bc%orient
    integer(c_int) function gsis_inner_exchange(dst, ksub, src, msub, orient) &
        bind(C, name="gsis_inner_exchange")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: dst, src
      integer(c_int32_t), value :: ksub, msub, orient
    end function
    integer(c_int) function gsis_destroy(h) bind(C, name="gsis_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function gsis_last_error_c() bind(C, name="gsis_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! The reference stops on fatal errors (src/IO.f90:236-244,
  ! src/sub_init.f90:25-29); so does a failed library call.
  subroutine gsis_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= GSIS_OK) then
      write(*, '(a,a,a,i0,a,a)') 'gsis: ', what, ' failed (code ', rc, '): ', trim(gsis_last_error())
      stop 1
    end if
  end subroutine gsis_check

  function gsis_last_error() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: k
    msg = ''
    p = gsis_last_error_c()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do k = 1, 512
      if (cs(k) == c_null_char) exit
      msg(k:k) = cs(k)
    end do
  end function gsis_last_error

end module gsis_capi
