! ISO_C_BINDING interface to libdiablo_hip.so (include/diablo_hip.h).
!
! The derived types mirror dh_params and dh_host_state field by field; the
! dh_upload_arrays / dh_download_arrays wrappers take the COMMON arrays of
! the reference's header (header:212-258) directly, so a driver needs no
! TARGET attribute on its COMMON blocks.
module diablo_hip
  use iso_c_binding
  implicit none
  integer, parameter :: DH_MAX_TH = 4

  type, bind(C) :: dh_params
    integer(c_int32_t) :: nx, ny, nz, n_th
    real(c_double) :: nu, beta, lx, lz, nu_v_scale
    real(c_double) :: delta_t, cfl
    integer(c_int32_t) :: variable_dt, update_dt
    integer(c_int32_t) :: f_type
    real(c_double) :: px0, omega0, amp_omega0
    real(c_double) :: i_ro, grav_x, grav_y, grav_z
    integer(c_int32_t) :: u_bc_ymin, v_bc_ymin, w_bc_ymin, u_bc_ymax, v_bc_ymax, w_bc_ymax
    real(c_double) :: u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1
    real(c_double) :: u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1
    integer(c_int32_t) :: th_bc_ymin(DH_MAX_TH), th_bc_ymax(DH_MAX_TH)
    real(c_double) :: th_bc_ymin_c1(DH_MAX_TH), th_bc_ymax_c1(DH_MAX_TH)
    real(c_double) :: ri(DH_MAX_TH), pr(DH_MAX_TH)
    real(c_double) :: drhodx(DH_MAX_TH), drhodz(DH_MAX_TH)
    integer(c_int32_t) :: create_new_flow
  end type dh_params

  type, bind(C) :: dh_host_state
    type(c_ptr) :: cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th
  end type dh_host_state

  interface
    integer(c_int) function dh_create(p, ctx) bind(C, name='dh_create')
      import :: c_int, c_ptr, dh_params
      type(dh_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function dh_create
    integer(c_int) function dh_destroy(ctx) bind(C, name='dh_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function dh_destroy
    type(c_ptr) function dh_last_error(ctx) bind(C, name='dh_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function dh_last_error
    integer(c_int) function dh_set_grid(ctx, gy, gyf, dy, dyf) bind(C, name='dh_set_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: gy(*), gyf(*), dy(*), dyf(*)
    end function dh_set_grid
    integer(c_int) function dh_upload(ctx, st) bind(C, name='dh_upload')
      import :: c_int, c_ptr, dh_host_state
      type(c_ptr), value :: ctx
      type(dh_host_state), intent(in) :: st
    end function dh_upload
    integer(c_int) function dh_download(ctx, st) bind(C, name='dh_download')
      import :: c_int, c_ptr, dh_host_state
      type(c_ptr), value :: ctx
      type(dh_host_state), intent(in) :: st
    end function dh_download
    integer(c_int) function dh_set_clock(ctx, time, delta_t, time_step) bind(C, name='dh_set_clock')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: time, delta_t
      integer(c_int32_t), value :: time_step
    end function dh_set_clock
    integer(c_int) function dh_get_clock(ctx, time, delta_t, h_bar3, time_step) bind(C, name='dh_get_clock')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time, delta_t, h_bar3(3)
      integer(c_int32_t), intent(out) :: time_step
    end function dh_get_clock
    integer(c_int) function dh_rk_substep(ctx, rk_step) bind(C, name='dh_rk_substep')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rk_step
    end function dh_rk_substep
    integer(c_int) function dh_time_step(ctx, nsteps) bind(C, name='dh_time_step')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function dh_time_step
    integer(c_int) function dh_synchronize(ctx) bind(C, name='dh_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function dh_synchronize
    integer(c_int) function dh_rem_div(ctx) bind(C, name='dh_rem_div')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function dh_rem_div
    integer(c_int) function dh_poisson_p_chan(ctx) bind(C, name='dh_poisson_p_chan')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function dh_poisson_p_chan
    integer(c_int) function dh_save_stats(ctx, ly, out, ubulk) bind(C, name='dh_save_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: ly
      real(c_double), intent(out) :: out(*)
      real(c_double), intent(out) :: ubulk
    end function dh_save_stats
    integer(c_int) function dh_set_mixed_precision(ctx, on) bind(C, name='dh_set_mixed_precision')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function dh_set_mixed_precision
    ! LES = .TRUE. with LES_MODEL_TYPE = 1 (header:168, les.f:1-336)
    integer(c_int) function dh_set_les(ctx, on) bind(C, name='dh_set_les')
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function dh_set_les
    ! learned closure in place of the Smagorinsky NU_T (no reference counterpart)
    integer(c_int) function dh_set_les_closure(ctx, w, n) bind(C, name='dh_set_les_closure')
      import :: c_int, c_int32_t, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: w(*)
      integer(c_int32_t), value :: n
    end function dh_set_les_closure
  end interface

contains

  ! Text of the last error as a Fortran string.
  function dh_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(dh_last_error(ctx), s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function dh_error_text

  integer(c_int) function dh_upload_arrays(ctx, cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th) result(rc)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), target, intent(in) :: cu1(*), cu2(*), cu3(*), cp(*), cf1(*), cf2(*), cf3(*)
    complex(c_double_complex), target, intent(in) :: cth(*), cfth(*)
    real(c_double), target, intent(in) :: th(*)
    type(dh_host_state) :: st
    st%cu1 = c_loc(cu1); st%cu2 = c_loc(cu2); st%cu3 = c_loc(cu3); st%cp = c_loc(cp)
    st%cf1 = c_loc(cf1); st%cf2 = c_loc(cf2); st%cf3 = c_loc(cf3)
    st%cth = c_loc(cth); st%cfth = c_loc(cfth); st%th = c_loc(th)
    rc = dh_upload(ctx, st)
  end function dh_upload_arrays

  integer(c_int) function dh_download_arrays(ctx, cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th) result(rc)
    type(c_ptr), intent(in) :: ctx
    complex(c_double_complex), target, intent(inout) :: cu1(*), cu2(*), cu3(*), cp(*), cf1(*), cf2(*), cf3(*)
    complex(c_double_complex), target, intent(inout) :: cth(*), cfth(*)
    real(c_double), target, intent(inout) :: th(*)
    type(dh_host_state) :: st
    st%cu1 = c_loc(cu1); st%cu2 = c_loc(cu2); st%cu3 = c_loc(cu3); st%cp = c_loc(cp)
    st%cf1 = c_loc(cf1); st%cf2 = c_loc(cf2); st%cf3 = c_loc(cf3)
    st%cth = c_loc(cth); st%cfth = c_loc(cfth); st%th = c_loc(th)
    rc = dh_download(ctx, st)
  end function dh_download_arrays

end module diablo_hip
