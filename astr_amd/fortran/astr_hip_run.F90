!-----------------------------------------------------------------------
! astr_hip_run.F90 -- the main solver's run ("astr run", src/astr.F90:39-93)
! on the device, for the cases this build's main-path semantics cover: a
! periodic box (flowtype tgv, 3-D or 2-D, and 2dvort), the convective
! scheme 643c, an explicit upwind-biased one (odd first digit, suffix 'e',
! with recon_schem, lchardecomp, bfacmpld, shkcrt) or the compact
! upwind-biased 543c (with lchardecomp, bfacmpld, shkcrt), diffusion 643c,
! rk3, the compact filter, no species, no turbulence model.
!
! module astr_hip_main   the reference's module state (commvar, commarray)
!                        and subroutines under their own names: readinput,
!                        readcont, refcal, allocommarray, gridgen, solvrinit,
!                        geomcal, flowinit, updateq, steploop,
!                        time_integration_rk, updatefvar, statout,
!                        writechkpt, readcheckpoint.  Every field lives in
!                        HBM; gridgen's coordinates, the initial primitives
!                        and the checkpoints pass through host arrays.
! program astr_hip       usage: astr_hip run [inputfile]  (default
!                        datin/input.dat; the controller is datin/controller
!                        next to the input file)
!-----------------------------------------------------------------------
module astr_hip_main
  use iso_c_binding
  use ieee_arithmetic, only: ieee_value, ieee_positive_inf
  use astr_hip_capi
  implicit none

  ! commvar (src/commvar.F90:18-98)
  character(len=64) :: flowtype = 'tgv'
  integer :: ia = 0, ja = 0, ka = 0, im = 0, jm = 0, km = 0, hm = 5, numq = 5, ndims = 3
  integer :: nstep = 0, maxstep = 0
  integer :: feqchkpt = 1000000, feqwsequ, feqslice, feqlist, feqavg
  logical :: diffterm = .true., lrestart = .false.
  real(8) :: time = 0.d0, deltat = 1.d-3
  real(8) :: gamma = 1.4d0, prandtl = 0.72d0, ref_tem = 273.15d0, reynolds = 1600.d0
  real(8) :: mach = 0.1d0, alfa_filter = 0.49d0, ref_len = 1.d0
  real(8) :: const1, const2, const5, const6, tempconst
  real(8) :: roinf, uinf, vinf, winf, tinf, pinf
  character(len=512) :: contfile = 'datin/controller'
  ! the convective scheme (readwrite.F90:541-546)
  logical :: lupwind = .false., lcompact = .false., lchardecomp = .false.
  integer :: recon_schem = 0
  real(8) :: bfacmpld = 0.3d0, shkcrt = 0.05d0

  ! commarray (src/commarray.F90:273-285), the computed points 0..n
  real(8), allocatable, target :: x(:,:,:,:), jacob(:,:,:), dxi(:,:,:,:,:)
  real(8), allocatable, target :: q(:,:,:,:), rho(:,:,:), vel(:,:,:,:), prs(:,:,:), tmp(:,:,:)

  type(c_ptr) :: ctx = c_null_ptr
  logical :: fsopen = .false.
  integer :: fsfile
  real(8), parameter :: pi = 4.d0 * atan(1.d0)

contains

  subroutine input_stop(what)
    character(len=*), intent(in) :: what
    print *, ' !! input not supported by the device solver: ', what
    stop 1
  end subroutine input_stop

  ! readinput (src/readwrite.F90:466-716, the build without COMB): the
  ! datin input file, record by record as the reference reads it
  subroutine readinput(inputfile)
    character(len=*), intent(in) :: inputfile
    character(len=64) :: gridfile
    character(len=16) :: conschm, dschm, rkscheme, turbmode, iomode
    character(len=255) :: lineread
    logical :: lihomo, ljhomo, lkhomo, nondimen, lfilter, lreadgrid, lfftk, limmbou, ltimrpt
    integer :: fh, n, kcutoff, num_species, ninit, bctype(6), nconv, ios

    open(newunit=fh, file=trim(inputfile), action='read', status='old')
    read(fh, '(////)')
    read(fh, *) flowtype
    read(fh, '(/)')
    read(fh, *) ia, ja, ka
    read(fh, '(/)')
    read(fh, *) lihomo, ljhomo, lkhomo
    read(fh, '(/)')
    read(fh, *) nondimen, diffterm, lfilter, lreadgrid, lfftk, limmbou, ltimrpt
    read(fh, '(/)')
    read(fh, *) lrestart
    read(fh, '(/)')
    read(fh, *) alfa_filter, kcutoff
    if (.not. nondimen) call input_stop('dimensional input (nondimen = f)')
    read(fh, '(/)')
    read(fh, *) ref_tem, reynolds, mach
    read(fh, '(/)')
    read(fh, *) conschm, dschm, rkscheme
    read(fh, '(/)')
    read(fh, *) recon_schem, lchardecomp, bfacmpld, shkcrt
    read(fh, '(/)')
    read(fh, *) num_species
    read(fh, '(/)')
    read(fh, *) turbmode, iomode
    read(fh, '(/)')
    do n = 1, 6
      read(fh, *) bctype(n)
    end do
    read(fh, '(/)')
    read(fh, *) ninit
    read(fh, '(/)')
    read(fh, '(A)') lineread
    read(fh, '(/)')
    read(fh, '(A)') gridfile
    close(fh)

    if (trim(flowtype) /= 'tgv' .and. trim(flowtype) /= '2dvort') call input_stop('flowtype '//trim(flowtype))
    if (.not. (lihomo .and. ljhomo .and. lkhomo)) call input_stop('non-homogeneous directions')
    if (any(bctype /= 1)) call input_stop('boundary types other than periodic (1)')
    if (.not. lfilter) call input_stop('lfilter = f')
    if (lreadgrid .or. lfftk .or. limmbou) call input_stop('lreadgrid, lfftz or limmbou = t')
    ! rhscal's choice (solver.F90:210-236): an even first digit is the
    ! central scheme, an odd one the upwind-biased scheme, explicit for
    ! suffix 'e', compact for suffix 'c' (scheme 543 only,
    ! compact_flux_initiate, flux.F90:93-110)
    read(conschm(1:1), '(I1)', iostat=ios) nconv
    if (ios /= 0 .or. len_trim(conschm) /= 4) call input_stop('convective scheme '//trim(conschm))
    if (mod(nconv, 2) == 0) then
      if (trim(conschm) /= '643c') call input_stop('central scheme '//trim(conschm))
      lupwind = .false.
    else if (conschm(4:4) == 'e') then
      lupwind = .true.
    else if (trim(conschm) == '543c') then
      lupwind = .true.
      lcompact = .true.
    else
      call input_stop('upwind-biased scheme '//trim(conschm))
    end if
    if (trim(dschm) /= '643c') call input_stop('diffusion scheme '//trim(dschm))
    if (trim(rkscheme) /= 'rk3') call input_stop('time scheme '//trim(rkscheme))
    if (num_species /= 0) call input_stop('species (num_species > 0)')
    if (trim(turbmode) /= 'none') call input_stop('turbmode '//trim(turbmode))
    if (ninit /= 0) call input_stop('ninit /= 0')
    if (ja == 0) call input_stop('one-dimensional grid (ja = 0)')

    ! one rank: the local block is the whole domain (parallel.F90:389-399)
    im = ia; jm = ja; km = ka
    ndims = merge(2, 3, ka == 0)
    numq = 5
    write(*, '(3(A,I0))') '  ** dimension set: ', im, ' x ', jm, ' x ', km
    write(*, '(A,4(1X,ES22.15))') '  ** ref_tem, reynolds, mach, alfa_filter:', ref_tem, reynolds, &
      mach, alfa_filter
    write(*, '(A,L1,A,L1)') '  ** diffterm ', diffterm, ', lrestart ', lrestart
    print *, ' >> ', trim(inputfile), ' ... done'
  end subroutine readinput

  ! readcont (src/readwrite.F90:729-769): datin/controller
  subroutine readcont
    logical :: lwsequ, lwslic, lavg, lcracon
    integer :: fh
    open(newunit=fh, file=trim(contfile), action='read', status='old')
    read(fh, '(////)')
    read(fh, *) lwsequ, lwslic, lavg, lcracon
    read(fh, '(/)')
    read(fh, *) maxstep, feqchkpt, feqwsequ, feqslice, feqlist, feqavg
    read(fh, '(/)')
    read(fh, *) deltat
    close(fh)
    write(*, '(A,I0,A,I0,A,ES22.15)') '  ** maxstep ', maxstep, ', feqchkpt ', feqchkpt, ', deltat', deltat
    print *, ' >> ', trim(contfile), ' ... done'
  end subroutine readcont

  ! refcal (src/solver.F90:104-128, nondimen): the flow constants and the
  ! free-stream state
  subroutine refcal
    const1 = 1.d0 / (gamma * (gamma - 1.d0) * mach**2)
    const2 = gamma * mach**2
    const5 = (gamma - 1.d0) * mach**2
    const6 = 1.d0 / (gamma - 1.d0)
    roinf = 1.d0; uinf = 1.d0; vinf = 0.d0; winf = 0.d0; tinf = 1.d0
    pinf = roinf * tinf / const2
    tempconst = 110.3d0 / ref_tem
  end subroutine refcal

  ! allocommarray (src/commarray.F90:315-367): the arrays this run uses
  subroutine allocommarray
    allocate(x(0:im, 0:jm, 0:km, 1:3), jacob(0:im, 0:jm, 0:km), dxi(0:im, 0:jm, 0:km, 1:3, 1:3))
    allocate(q(0:im, 0:jm, 0:km, 1:numq), rho(0:im, 0:jm, 0:km), vel(0:im, 0:jm, 0:km, 1:3))
    allocate(prs(0:im, 0:jm, 0:km), tmp(0:im, 0:jm, 0:km))
    vel = 0.d0
    print *, ' ** global arrays allocated'
  end subroutine allocommarray

  ! gridgen (src/gridgeneration.F90:25-45) for the flow types run here:
  ! gridcube (:233-267) of the flow type's box
  subroutine gridgen
    integer :: i, j, k
    real(8) :: lx, ly, lz
    if (trim(flowtype) == 'tgv') then
      lx = ref_len * 2.d0 * pi; ly = lx; lz = lx
    else
      lx = 20.d0; ly = 10.d0; lz = 1.d0
    end if
    do k = 0, km
      do j = 0, jm
        do i = 0, im
          x(i, j, k, 1) = lx / real(ia, 8) * real(i, 8)
          x(i, j, k, 2) = ly / real(ja, 8) * real(j, 8)
          if (ka == 0) then
            x(i, j, k, 3) = 0.d0
          else
            x(i, j, k, 3) = lz / real(ka, 8) * real(k, 8)
          end if
        end do
      end do
    end do
    print *, ' ** cubic grid generated'
  end subroutine gridgen

  ! solvrinit (src/solver.F90:28-173): schemes, filter and constants on the
  ! device, with the main path's stage (ASTR_SEM_MAIN) and Sutherland
  ! constant 110.3 K.  diffterm = f runs with 1/Re = 0, which makes every
  ! viscous flux and its derivative an exact zero.
  subroutine solvrinit
    type(astr_hip_config) :: cfg
    cfg%im = im; cfg%jm = jm; cfg%km = km
    cfg%semantics = ASTR_SEM_MAIN
    cfg%alfa_filter = alfa_filter
    cfg%beta_halo = 1.11d0
    cfg%beta_bound = 0.98d0   ! filter_coefficient_cal(beter_bouond=0.98d0), comsolver.F90:121
    cfg%gamma = gamma; cfg%mach = mach; cfg%prandtl = prandtl
    if (diffterm) then
      cfg%reynolds = reynolds
    else
      cfg%reynolds = ieee_value(1.d0, ieee_positive_inf)
    end if
    cfg%ref_tem = ref_tem
    cfg%sutherland = 110.3d0
    cfg%lx = x(im, 0, 0, 1) - x(0, 0, 0, 1)
    cfg%ly = x(0, jm, 0, 2) - x(0, 0, 0, 2)
    cfg%lz = merge(x(0, 0, km, 3) - x(0, 0, 0, 3), 1.d0, km > 0)
    cfg%deltat = deltat
    call hipcheck(astr_hip_create(cfg, ctx), 'solvrinit')
    if (lupwind) then
      call hipcheck(astr_hip_set_convection(ctx, merge(2, 1, lcompact), recon_schem, &
                                            merge(1, 0, lchardecomp), bfacmpld, shkcrt), &
                    'solvrinit: convective scheme')
      if (lcompact) then
        write(*, '(A,L1)') '  ** compact upwind-biased convection 543c, lchardecomp ', lchardecomp
      else
        write(*, '(A,I0,A,L1)') '  ** upwind-biased convection, recon_schem ', recon_schem, &
          ', lchardecomp ', lchardecomp
      end if
    end if
    print *, ' ** solver initialised on the device'
  end subroutine solvrinit

  ! geomcal (src/geom.F90:99-692): the metrics of the grid, computed on the
  ! device, then handed to the passes
  subroutine geomcal
    call hipcheck(astr_hip_gridgeom(ctx, c_loc(x), c_loc(jacob), c_loc(dxi)), 'gridgeom')
    call hipcheck(astr_hip_set_metrics(ctx, c_loc(jacob), c_loc(dxi)), 'geomcal')
    print *, ' ** geometrical variables calculated'
  end subroutine geomcal

  ! tgvini (src/initialisation.F90:621-691, nondimen)
  subroutine tgvini
    integer :: i, j, k
    real(8) :: l_0
    l_0 = 1.d0
    do k = 0, km
      do j = 0, jm
        do i = 0, im
          rho(i, j, k) = roinf
          vel(i, j, k, 1) = uinf * sin(x(i, j, k, 1) / l_0) * cos(x(i, j, k, 2) / l_0) * cos(x(i, j, k, 3) / l_0)
          vel(i, j, k, 2) = -uinf * cos(x(i, j, k, 1) / l_0) * sin(x(i, j, k, 2) / l_0) * cos(x(i, j, k, 3) / l_0)
          vel(i, j, k, 3) = 0.d0
          prs(i, j, k) = pinf + rho(i, j, k) / 16.d0 * (uinf**2) &
                         * (cos(2.d0 * x(i, j, k, 1) / l_0) + cos(2.d0 * x(i, j, k, 2) / l_0)) &
                         * (cos(2.d0 * x(i, j, k, 3) / l_0) + 2.d0)
          tmp(i, j, k) = prs(i, j, k) / rho(i, j, k) * const2
        end do
      end do
    end do
  end subroutine tgvini

  ! vortini (src/initialisation.F90:480-533, 2dvort)
  subroutine vortini
    integer :: i, j, k
    real(8) :: xc, yc, radi2, rvor, cvor, var1
    xc = 10.d0
    yc = 5.d0
    rvor = 0.7d0
    cvor = 0.1d0 * rvor
    do k = 0, km
      do j = 0, jm
        do i = 0, im
          radi2 = ((x(i, j, k, 1) - xc)**2 + (x(i, j, k, 2) - yc)**2) / rvor / rvor
          var1 = cvor / rvor / rvor * exp(-0.5d0 * radi2)
          rho(i, j, k) = roinf
          vel(i, j, k, 1) = uinf - var1 * (x(i, j, k, 2) - yc)
          if (ndims >= 2) vel(i, j, k, 2) = vinf + var1 * (x(i, j, k, 1) - xc)
          if (ndims == 3) vel(i, j, k, 3) = 0.d0
          prs(i, j, k) = pinf - 0.5d0 * rho(i, j, k) * cvor**2 / rvor**2 * exp(-radi2)
          tmp(i, j, k) = prs(i, j, k) / rho(i, j, k) * const2
        end do
      end do
    end do
  end subroutine vortini

  ! updateq (src/fludyna.F90:254-300): q = fvar2q of the primitives
  ! (fvar2q_sca :312-360 with the temperature: q5 = rho (T const1 + |u|^2/2))
  subroutine updateq
    integer :: i, j, k
    real(8) :: var1
    do k = 0, km
      do j = 0, jm
        do i = 0, im
          q(i, j, k, 1) = rho(i, j, k)
          q(i, j, k, 2) = rho(i, j, k) * vel(i, j, k, 1)
          q(i, j, k, 3) = rho(i, j, k) * vel(i, j, k, 2)
          q(i, j, k, 4) = rho(i, j, k) * vel(i, j, k, 3)
          var1 = 0.5d0 * sum(vel(i, j, k, :) * vel(i, j, k, :))
          q(i, j, k, 5) = rho(i, j, k) * (tmp(i, j, k) * const1 + var1)
        end do
      end do
    end do
  end subroutine updateq

  ! flowinit (src/initialisation.F90:28-140): the flow type's field, or a
  ! checkpoint when lrestart; q and the primitives go to the device
  subroutine flowinit
    integer(c_long_long) :: ns
    real(c_double) :: t
    if (lrestart) then
      call readcheckpoint('checkpoint', ns, t)
      nstep = int(ns)
      time = t
    else
      if (trim(flowtype) == 'tgv') then
        call tgvini
      else
        call vortini
      end if
      call updateq
      call hipcheck(astr_hip_upload(ctx, q, 0_c_int), 'flowinit')
      call hipcheck(astr_hip_set_fvar(ctx, c_loc(rho), c_loc(vel), c_loc(prs), c_loc(tmp), 0_c_int), &
                    'flowinit')
      nstep = 0
      time = 0.d0
    end if
    call hipcheck(astr_hip_set_deltat(ctx, deltat), 'flowinit')
    print *, ' ** flow field initialised'
  end subroutine flowinit

  ! updatefvar (src/fludyna.F90:191-242): rho, vel, prs, tmp of the device
  ! state into the host arrays
  subroutine updatefvar
    call hipcheck(astr_hip_updatefvar(ctx, c_loc(rho), c_loc(vel), c_loc(prs), c_loc(tmp), 0_c_int), &
                  'updatefvar')
  end subroutine updatefvar

  ! writechkpt (src/readwrite.F90:2390-2459): checkpoint/flowfield.bin
  subroutine writechkpt
    call execute_command_line('mkdir -p checkpoint')
    call hipcheck(astr_hip_write_checkpoint(ctx, 'checkpoint/flowfield.bin'//c_null_char, &
                                            int(nstep, c_long_long), time), 'writechkpt')
    print *, ' << checkpoint/flowfield.bin at step ', nstep
  end subroutine writechkpt

  ! readcheckpoint (src/readwrite.F90:1381-1531)
  subroutine readcheckpoint(folder, ns, t)
    character(len=*), intent(in) :: folder
    integer(c_long_long), intent(out) :: ns
    real(c_double), intent(out) :: t
    call hipcheck(astr_hip_read_checkpoint(ctx, trim(folder)//'/flowfield.bin'//c_null_char, ns, t), &
                  'readcheckpoint')
    print *, ' >> ', trim(folder), '/flowfield.bin at step ', ns
  end subroutine readcheckpoint

  ! statout for tgv (src/statistic.F90:777-859) with listinit/listwrite
  ! (src/utility.F90:310-434): nstep, time, kenergy, enstophy, dissipation
  subroutine statout(kenergy, enstophy, dissipation)
    real(8), intent(in) :: kenergy, enstophy, dissipation
    if (.not. fsopen) then
      open(newunit=fsfile, file='flowstate.dat')
      fsopen = .true.
      write(fsfile, '(5(1X,A20))') 'nstep', 'time', 'kenergy', 'enstophy', 'dissipation'
    end if
    write(fsfile, '(1X,I20,4(1X,E20.13E2))') nstep, time, kenergy, enstophy, dissipation
  end subroutine statout

  ! time_integration_rk (src/mainloop.F90:301-601) for n steps in one device
  ! call: per step the three stages of filterq, qswap, gradcal, rhscal, the
  ! RK3 update and updatefvar; rkfirst's statcal of each step's first stage
  ! comes back in ke, en, di
  subroutine time_integration_rk(n, ke, en, di)
    integer, intent(in) :: n
    real(8), intent(out), target :: ke(n), en(n), di(n)
    call hipcheck(astr_hip_rk3_steps_stats(ctx, int(n, c_int), c_loc(ke), c_loc(en), c_loc(di)), &
                  'time_integration_rk')
  end subroutine time_integration_rk

  ! steploop (src/mainloop.F90:22-150): do while (nstep <= maxstep) one
  ! time_integration_rk, then nstep and time advance.  rkfirst writes the
  ! statistics of each step's first stage except on its very first call
  ! (:632-640).  After every step divisible by feqchkpt steploop re-reads the
  ! controller (:111), so maxstep and deltat may change, and (after the
  ! first step) a checkpoint is written.  The device runs the steps up to the
  ! next such step in one call.  The checkpoint holds the state at the start
  ! of the following step (after the RK update and updatefvar), where the
  ! reference's writechkpt runs inside the first stage after filterq; a run
  ! resumed from it continues bit for bit.
  subroutine steploop
    real(8), allocatable :: ke(:), en(:), di(:)
    integer :: n, it, nend
    logical :: firstcall
    firstcall = .true.
    do while (nstep <= maxstep)
      nend = min(maxstep, ((nstep + feqchkpt - 1) / feqchkpt) * feqchkpt)
      n = nend - nstep + 1
      allocate(ke(n), en(n), di(n))
      call time_integration_rk(n, ke, en, di)
      do it = 1, n
        if (.not. firstcall .and. trim(flowtype) == 'tgv') call statout(ke(it), en(it), di(it))
        firstcall = .false.
        nstep = nstep + 1
        time = time + deltat
      end do
      deallocate(ke, en, di)
      if (mod(nend, feqchkpt) == 0) then
        if (nend > 0) call writechkpt
        call readcont
        call hipcheck(astr_hip_set_deltat(ctx, deltat), 'steploop')
      end if
    end do
    if (fsopen) flush(fsfile)
  end subroutine steploop

  subroutine solver_end
    if (c_associated(ctx)) call hipcheck(astr_hip_destroy(ctx), 'solver_end')
    ctx = c_null_ptr
    if (fsopen) then
      close(fsfile)
      print *, ' << flowstate.dat'
    end if
    fsopen = .false.
  end subroutine solver_end

end module astr_hip_main

! astr run (src/astr.F90:39-93) on the device
program astr_hip
  use astr_hip_main
  implicit none
  character(len=512) :: cmd, inputfile
  integer :: nargs, i
  nargs = command_argument_count()
  cmd = 'run'
  if (nargs >= 1) call get_command_argument(1, cmd)
  if (trim(cmd) /= 'run') then
    print *, ' usage: astr_hip run [inputfile]'
    stop 1
  end if
  inputfile = 'datin/input.dat'
  if (nargs >= 2) call get_command_argument(2, inputfile)
  ! the controller sits beside the input file (datin/controller)
  i = index(inputfile, '/', back=.true.)
  if (i > 0) then
    contfile = inputfile(1:i)//'controller'
  else
    contfile = 'controller'
  end if
  call readinput(trim(inputfile))
  call readcont
  call refcal
  call allocommarray
  call gridgen
  call solvrinit
  call geomcal
  call flowinit
  call steploop
  call solver_end
end program astr_hip
