!===============================================================================
! iso_c_binding drop-in of the MI355X MHD step for the reference driver.
!
! Links in place of the reference's external subroutines compute_dt
! (src/subroutines/physics.f90:18), godunov (src/subroutines/scheme.f90:104),
! fargo_update (src/subroutines/fargo.f90:17), boundary
! (src/subroutines/boundary.f90:18) and dissipation
! (src/subroutines/physics.f90:272), which the unchanged driver
! (src/dumses.f90:82,88,97,101,104) and init (src/subroutines/init.f90:95)
! call.
! All state stays in the reference's modules; the shim forwards to libdmx.so
! (include/dmx.h) and keeps uin on the device between host synchronisation
! points.  The host copy is refreshed where the OpenACC build of the
! reference updates it (src/dumses.f90:62,122,133,138,144): the shim wraps
! output, history and special, whose reference versions are compiled under
! the names dmx_ref_output, dmx_ref_history and dmx_ref_special.
!
! Precision: DMX_PRECISION=fp32 in the environment selects the float32-bulk
! engine; the default is float64.  Build: dumses-hybrid_amd/fortran/dropin.mk
!===============================================================================
module dmx_bind
  use iso_c_binding
  implicit none

  type, bind(C) :: dmx_params_t
     integer(c_int32_t) :: nx, ny, nz
     integer(c_int32_t) :: iso, slope_type, riemann, riemann2d
     integer(c_int32_t) :: bdtype(3)
     integer(c_int32_t) :: fargo, precision
     real(c_double) :: xmin, xmax, ymin, ymax, zmin, zmax
     real(c_double) :: gamma, ciso, omega0, courant
     real(c_double) :: nu, eta
  end type dmx_params_t

  interface
     integer(c_int) function dmx_create(p, ctx) bind(C, name='dmx_create')
       import :: c_int, c_ptr, dmx_params_t
       type(dmx_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function dmx_create
     integer(c_int) function dmx_upload(ctx, uin) bind(C, name='dmx_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uin(*)
     end function dmx_upload
     integer(c_int) function dmx_download(ctx, uin) bind(C, name='dmx_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: uin(*)
     end function dmx_download
     integer(c_int) function dmx_compute_dt(ctx, dt) bind(C, name='dmx_compute_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt
     end function dmx_compute_dt
     integer(c_int) function dmx_godunov(ctx, time, dt) bind(C, name='dmx_godunov')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time, dt
     end function dmx_godunov
     integer(c_int) function dmx_fargo(ctx, dt) bind(C, name='dmx_fargo')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
     end function dmx_fargo
     integer(c_int) function dmx_boundary(ctx, time, dt) bind(C, name='dmx_boundary')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time, dt
     end function dmx_boundary
     integer(c_int) function dmx_dissipation(ctx, time, dt) bind(C, name='dmx_dissipation')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time, dt
     end function dmx_dissipation
     integer(c_int) function dmx_tune(ctx, best_ms, shape) bind(C, name='dmx_tune')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: best_ms
       integer(c_int32_t), intent(out) :: shape(4)
     end function dmx_tune
     type(c_ptr) function dmx_last_error() bind(C, name='dmx_last_error')
       import :: c_ptr
     end function dmx_last_error
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: on_device = .false.   ! device copy authoritative
  logical, save :: host_stale = .false.  ! host copy older than the device

contains

  subroutine dmx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: n
    if (rc == 0) return
    call c_f_pointer(dmx_last_error(), msg, [512])
    text = ''
    do n = 1, 512
       if (msg(n) == c_null_char) exit
       text(n:n) = msg(n)
    end do
    print '(A,A,A,I0,A,A)', 'dmx: ', what, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine dmx_check

  ! create the context from the reference's module state
  subroutine dmx_ensure_ctx()
    use input_params
    type(dmx_params_t) :: p
    character(len=16) :: prec
    integer :: d, plen, pstat
    if (c_associated(ctx)) return
    ! source terms (rhs, src/dumses.f90:84-93) run on the host copy between
    ! godunov and the ghost fill, and the gravity predictor inside godunov
    ! (src/subroutines/scheme.f90:141) has no device counterpart: refuse
    ! rather than step with a stale host state
    if (rhs) then
       print '(A)', 'dmx: rhs=.true. (source_term) is not supported by the device step'
       error stop 1
    endif
    p%nx = nx; p%ny = ny; p%nz = nz
#if ISO == 1
    p%iso = 1
#else
    p%iso = 0
#endif
    p%slope_type = slope_type
    p%riemann = iriemann
    p%riemann2d = iriemann2d
    do d = 1, 3
       select case (trim(boundary_type(d)))
       case ('periodic')
          p%bdtype(d) = 0
       case ('zero_gradient')
          p%bdtype(d) = 1
       case ('shearingbox')
          p%bdtype(d) = 2
       case default
          p%bdtype(d) = -1   ! rejected by dmx_create (analytic: host callback)
       end select
    end do
    p%fargo = merge(1, 0, fargo)
    call get_environment_variable('DMX_PRECISION', prec, plen, pstat)
    p%precision = 0
    if (pstat == 0 .and. (prec(1:4) == 'fp32' .or. prec(1:4) == 'FP32')) p%precision = 1
    p%xmin = xmin; p%xmax = xmax; p%ymin = ymin; p%ymax = ymax
    p%zmin = zmin; p%zmax = zmax
    p%gamma = gamma; p%ciso = ciso; p%omega0 = Omega0; p%courant = courant
    p%nu = nu; p%eta = eta
    call dmx_check(dmx_create(p, ctx), 'dmx_create')
  end subroutine dmx_ensure_ctx

  ! hand the host state to the device (start of the time loop, after init
  ! and restart_run may have rewritten uin); on the first upload the launch
  ! shapes are autotuned on that state (it is left unchanged; float64
  ! results do not depend on the shape) unless DMX_TUNE=0
  subroutine dmx_to_device()
    use variables, only: uin
    logical, save :: tuned = .false.
    character(len=8) :: tv
    integer :: tlen, tstat
    real(c_double) :: ms
    integer(c_int32_t) :: shape(4)
    if (on_device) return
    call dmx_ensure_ctx()
    call dmx_check(dmx_upload(ctx, uin), 'dmx_upload')
    on_device = .true.
    host_stale = .false.
    if (.not. tuned) then
       tuned = .true.
       call get_environment_variable('DMX_TUNE', tv, tlen, tstat)
       if (.not. (tstat == 0 .and. tv(1:1) == '0')) call dmx_check(dmx_tune(ctx, ms, shape), 'dmx_tune')
    end if
  end subroutine dmx_to_device

  ! bring the device state back before host code reads uin
  subroutine dmx_to_host()
    use variables, only: uin
    if (.not. (on_device .and. host_stale)) return
    call dmx_check(dmx_download(ctx, uin), 'dmx_download')
    host_stale = .false.
  end subroutine dmx_to_host

end module dmx_bind

!-------------------------------------------------------------------------------
! replacements of the reference's step routines
!-------------------------------------------------------------------------------
subroutine compute_dt(dt)
  use iso_c_binding
  use precision, only: dp
  use dmx_bind
  implicit none
  real(dp), intent(out) :: dt
  real(c_double) :: cdt
  call dmx_to_device()
  call dmx_check(dmx_compute_dt(ctx, cdt), 'dmx_compute_dt')
  dt = cdt
end subroutine compute_dt

subroutine godunov
  use dmx_bind
  use variables, only: time, dt
  implicit none
  call dmx_to_device()
  call dmx_check(dmx_godunov(ctx, time, dt), 'dmx_godunov')
  host_stale = .true.
end subroutine godunov

subroutine fargo_update
  use dmx_bind
  use variables, only: dt
  implicit none
  call dmx_to_device()
  call dmx_check(dmx_fargo(ctx, dt), 'dmx_fargo')
  host_stale = .true.
end subroutine fargo_update

subroutine boundary
  use dmx_bind
  use variables, only: uin, time, dt
  implicit none
  if (.not. on_device) then
     ! init-time fill of the host state (src/subroutines/init.f90:95)
     call dmx_ensure_ctx()
     call dmx_check(dmx_upload(ctx, uin), 'dmx_upload')
     call dmx_check(dmx_boundary(ctx, time, dt), 'dmx_boundary')
     call dmx_check(dmx_download(ctx, uin), 'dmx_download')
  else
     call dmx_check(dmx_boundary(ctx, time, dt), 'dmx_boundary')
     host_stale = .true.
  endif
end subroutine boundary

subroutine dissipation
  use dmx_bind
  use variables, only: time, dt
  implicit none
  call dmx_to_device()
  call dmx_check(dmx_dissipation(ctx, time, dt), 'dmx_dissipation')
  host_stale = .true.
end subroutine dissipation

!-------------------------------------------------------------------------------
! host synchronisation around the reference's I/O
!-------------------------------------------------------------------------------
subroutine output
  use dmx_bind
  implicit none
  call dmx_to_host()
  call dmx_ref_output
end subroutine output

subroutine history
  use dmx_bind
  implicit none
  call dmx_to_host()
  call dmx_ref_history
end subroutine history

subroutine special
  use dmx_bind
  implicit none
  call dmx_to_host()
  call dmx_ref_special
end subroutine special
