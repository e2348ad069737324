!===============================================================================
! Test harness around the compiled reference (checker only).
!
! Linked against the reference's own objects, built from the unmodified
! sources by oracle/ref.mk.  It replaces only the reference's main program
! (src/dumses.f90): it runs the same initialisation (init_param / init, which
! injects the initial state through the reference's binary restart reader,
! src/subroutines/restart.f90:212-223) and then the same step order as the
! reference time loop (src/dumses.f90:75-105):
!     compute_dt -> godunov -> [fargo_update] -> boundary -> [dissipation]
!     -> time += dt
! but for a fixed number of steps instead of up to tlim, and it writes the
! state and the per-step dt as raw little-endian float64 streams.
!
! A second mode ("kernels") feeds random states through the reference's 1D and
! 2D Riemann solvers (src/modules/solvers/solver.f90, solver_magnetic.f90) and
! writes their outputs, to pin the solver restatements function by function.
! A third mode ("history") injects the state and calls the reference's own
! history routine (src/subroutines/history.f90:18-165) once, which writes its
! volume sums to history.txt at the precision of its format.
!===============================================================================
program ref_harness
  use params
  use input_params
  use variables
  use mpi_var
  implicit none

  integer :: nsteps, dumpevery, istep, u
  integer(kind=8) :: c0, c1, crate
  character(len=16) :: mode
  namelist /harness_params/ nsteps, dumpevery, mode

  nsteps = 1
  dumpevery = 0
  mode = 'step'
  open(newunit=u, file='harness', status='old')
  read(u, harness_params)
  close(u)

  if (trim(mode) == 'kernels') then
     call run_kernels
     stop
  endif

  ndump = 0
  nspec = 0
  call init_param
  call init

  if (trim(mode) == 'history') then
     dt = 0.d0
     call history
     stop
  endif

  open(unit=21, file='dt.raw', access='stream', form='unformatted', &
       status='replace')
  call system_clock(count=c0, count_rate=crate)
  do istep = 1, nsteps
     call compute_dt(dt)
     call godunov
     if (fargo) call fargo_update
     call boundary
     if ((nu > 0.d0) .or. (eta > 0.d0)) call dissipation
     time = time + dt
     write(21) dt
     if (dumpevery > 0) then
        if (mod(istep, dumpevery) == 0) call dump_state(istep)
     endif
  end do
  call system_clock(count=c1)
  close(21)
  ! wall time of the step loop alone (no initialisation or dump I/O)
  open(newunit=u, file='loop_seconds.txt', status='replace')
  write(u, '(ES24.16)') real(c1 - c0, kind=8) / real(crate, kind=8)
  close(u)
  call dump_state(nsteps)

contains

  subroutine dump_state(istep)
    integer, intent(in) :: istep
    character(len=32) :: fname
    write(fname, '("uin_",I6.6,".raw")') istep
    open(unit=22, file=trim(fname), access='stream', form='unformatted', &
         status='replace')
    write(22) uin
    close(22)
  end subroutine dump_state

  ! Reads nstate random states from kin.raw and writes solver outputs.
  ! kin.raw layout: int32 n, float64 gamma, then n x 16 (1D) and n x 32 (2D)
  ! float64 values, then float64 dt.
  subroutine run_kernels
    use solver, only: hlld1 => hlld, hll1 => hll, llf1 => llf, upwind1 => upwind
    use solver_magnetic, only: hlld2 => hlld, llf2 => llf, upwind2 => upwind, &
         hlla2 => hlla
    integer :: n, i, iv
    real(dp) :: gam, cis, dtk
    real(dp), allocatable :: s1(:,:), s2(:,:)
    real(dp) :: ql(8), qr(8), f(8), st(7)
    real(dp) :: qLL(8), qRL(8), qLR(8), qRR(8), e
    open(unit=23, file='kin.raw', access='stream', form='unformatted', &
         status='old')
    read(23) n
    read(23) gam, cis, dtk
    allocate(s1(16, n), s2(32, n))
    read(23) s1
    read(23) s2
    close(23)
    open(unit=24, file='kout.raw', access='stream', form='unformatted', &
         status='replace')
    ! 1D solvers: hlld (8 flux + 7 star), hll, llf, upwind (8 flux each)
    do i = 1, n
       ql = s1(1:8, i); qr = s1(9:16, i)
       call hlld1(ql, qr, f, st(1), st(2), st(3), st(4), st(5), st(6), st(7), &
            gam, cis)
       write(24) f, st
       ql = s1(1:8, i); qr = s1(9:16, i)
       call hll1(ql, qr, f, gam, cis)
       write(24) f
       ql = s1(1:8, i); qr = s1(9:16, i)
       call llf1(ql, qr, f, gam, cis)
       write(24) f
       ql = s1(1:8, i); qr = s1(9:16, i)
       call upwind1(ql, qr, f, gam, cis)
       write(24) f
    end do
    ! 2D EMF solvers: hlld, llf, upwind, hlla (1 value each)
    do i = 1, n
       do iv = 1, 8
          qLL(iv) = s2(iv, i); qRL(iv) = s2(8+iv, i)
          qLR(iv) = s2(16+iv, i); qRR(iv) = s2(24+iv, i)
       end do
       call hlld2(qLL, qRL, qLR, qRR, dtk, e, gam, cis)
       write(24) e
       call llf2(qLL, qRL, qLR, qRR, dtk, e, gam, cis)
       write(24) e
       call upwind2(qLL, qRL, qLR, qRR, dtk, e, gam, cis)
       write(24) e
       call hlla2(qLL, qRL, qLR, qRR, dtk, e, gam, cis)
       write(24) e
    end do
    close(24)
  end subroutine run_kernels

end program ref_harness
