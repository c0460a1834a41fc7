!-----------------------------------------------------------------------
! tgvsolver_hip -- the reference's TGV box solver program
! (miniapps/tgv_solver_3d/tgvsolver.F90:1086-1116, program boxsolver) with
! every field resident on the GPU.  Same call sequence, same state.dat
! (nstep, time, kinetic energy, enstrophy per step, list-directed).
!
! usage: tgvsolver_hip [im [jm [km [maxstep [deltat [noise_amp [seed]]]]]]]
!        tgvsolver_hip --datin INPUTFILE CONTROLLER
! defaults are the reference's: 128^3, 100 steps, deltat 1e-3, no noise.
! With --datin the grid, filter and reference values come from the
! reference's datin input file and the step count and time step from its
! controller file (readinput / readcont, src/readwrite.F90:466-769).
!-----------------------------------------------------------------------
program tgvsolver_hip
  use iso_c_binding, only: c_long_long
  use astr_hip_tgv
  implicit none
  integer :: nargs, v(4), i
  character(len=64) :: arg
  character(len=512) :: path
  real(8) :: tstart, tfinish

  call cpu_time(tstart)

  nargs = command_argument_count()
  if (nargs >= 3) then
    call get_command_argument(1, arg)
    if (trim(arg) == '--datin') then
      call get_command_argument(2, path)
      call readinput(trim(path))
      call get_command_argument(3, path)
      call readcont(trim(path))
      print *, ' ** note: this run uses the miniapp discretisation of'
      print *, '    miniapps/tgv_solver_3d/tgvsolver.F90 (filter after the RK'
      print *, '    update, Sutherland 110.4 K), not the stage order of astr run'
      call solver_init
      call allocommarray
      call flowfield_init
      call steploop
      call solver_end
      call cpu_time(tfinish)
      print *, ' ** steps: ', maxstep, ' wall (host cpu) time: ', tfinish - tstart
      stop
    end if
  end if

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
