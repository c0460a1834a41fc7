!-----------------------------------------------------------------------
! tgvsolver_hip -- the reference's TGV box solver program
! (miniapps/tgv_solver_3d/tgvsolver.F90:1086-1116, program boxsolver) with
! every field resident on the GPU.  Same call sequence, same state.dat
! (nstep, time, kinetic energy, enstrophy per step, list-directed).
!
! usage: tgvsolver_hip [im [jm [km [maxstep [deltat [noise_amp [seed]]]]]]]
! defaults are the reference's: 128^3, 100 steps, deltat 1e-3, no noise.
! The main solver's datin-driven run (astr run) is the program astr_hip
! (astr_hip_run.F90).
!-----------------------------------------------------------------------
program tgvsolver_hip
  use iso_c_binding, only: c_long_long
  use astr_hip_tgv
  implicit none
  integer :: nargs, v(4), i
  character(len=64) :: arg
  real(8) :: tstart, tfinish

  call cpu_time(tstart)

  nargs = command_argument_count()
  v = [128, 128, 128, 100]
  do i = 1, min(nargs, 4)
    call get_command_argument(i, arg)
    read(arg, *) v(i)
  end do
  if (nargs >= 5) then
    call get_command_argument(5, arg)
    read(arg, *) deltat
  end if
  if (nargs >= 6) then
    call get_command_argument(6, arg)
    read(arg, *) noise_amp
  end if
  if (nargs >= 7) then
    call get_command_argument(7, arg)
    read(arg, *) noise_seed
  end if

  call program_init(v(1), v(2), v(3), v(4))
  call solver_init
  call allocommarray
  call flowfield_init
  call steploop
  call solver_end

  call cpu_time(tfinish)
  print *, ' ** steps: ', maxstep, ' wall (host cpu) time: ', tfinish - tstart

end program tgvsolver_hip
