!+---------------------------------------------------------------------+
!| ref_driver.F90 -- C-callable harness around the reference's own     |
!| hot-path Fortran (TEST INFRASTRUCTURE ONLY).                        |
!|                                                                     |
!| It is compiled together with the unmodified reference sources       |
!| src/{constdef,commvar,commtype,commfunc,singleton,derivative,       |
!| filter}.F90 and miniapps/tgv_solver_3d/tgvsolver.F90 into           |
!| oracle/_ref/libastr_ref.so (see oracle/Makefile).  Nothing here     |
!| re-implements reference arithmetic: every routine only sets the     |
!| reference's module variables, calls the reference's own procedures  |
!| and copies arrays in or out.                                        |
!|                                                                     |
!| Module-level SAVE state in the reference (rk3's qsave, stacal's     |
!| file unit, fds allocation) means one grid per process: callers run  |
!| each configuration in a fresh process.                              |
!+---------------------------------------------------------------------+
module astr_ref_driver

  use iso_c_binding
  implicit none

contains

  ! one line of df_compact (derivative.F90:171-198), ntype=3
  subroutine ref_line_deriv(n, f, df) bind(C, name="ref_line_deriv")
    use commtype,   only : compact_scheme
    use derivative, only : fd_scheme_initiate, fds, compact_central
    integer(c_int), value :: n
    real(c_double), intent(in) :: f(-5:n+5)
    real(c_double), intent(out) :: df(0:n)
    type(compact_scheme) :: s
    if (.not. allocated(fds)) allocate(compact_central :: fds)
    call fd_scheme_initiate(asolver=s, scheme='643c', ntype=3, dim=n, dir=1)
    df = fds%central(s, f=f, dim=n)
    deallocate(s%a, s%c, s%ac)
  end subroutine ref_line_deriv

  ! one line of compact_filter (filter.F90:112-144), ntype=3
  subroutine ref_line_filter(n, alfa, f, ff) bind(C, name="ref_line_filter")
    use commtype, only : compact_scheme
    use commvar,  only : alfa_filter
    use filter,   only : compact_filter_initiate, filter_coefficient_cal,     &
                         compact_filter, coef10i
    integer(c_int), value :: n
    real(c_double), value :: alfa
    real(c_double), intent(in) :: f(-5:n+5)
    real(c_double), intent(out) :: ff(0:n)
    type(compact_scheme) :: s
    alfa_filter = alfa
    if (.not. allocated(coef10i)) then
      call filter_coefficient_cal(alfa=alfa, beter_halo=1.11d0, beter_bouond=1.09d0)
    endif
    call compact_filter_initiate(afilter=s, ntype=3, dim=n)
    ff = compact_filter(afilter=s, f=f, dim=n)
    deallocate(s%a, s%c, s%ac)
  end subroutine ref_line_filter

  ! one line of df_compact (derivative.F90:63-306) with boundary type ntype
  ! (1: physical boundary at 0, 2: at n, 3: interfaces, 4: both physical)
  subroutine ref_line_deriv_nt(n, ntype, f, df) bind(C, name="ref_line_deriv_nt")
    use commtype,   only : compact_scheme
    use derivative, only : fd_scheme_initiate, fds, compact_central
    integer(c_int), value :: n, ntype
    real(c_double), intent(in) :: f(-5:n+5)
    real(c_double), intent(out) :: df(0:n)
    type(compact_scheme) :: s
    if (.not. allocated(fds)) allocate(compact_central :: fds)
    call fd_scheme_initiate(asolver=s, scheme='643c', ntype=ntype, dim=n, dir=1)
    df = fds%central(s, f=f, dim=n)
    deallocate(s%a, s%c, s%ac)
  end subroutine ref_line_deriv_nt

  ! one line of compact_filter (filter.F90:31-285) with boundary type ntype;
  ! the coefficients are set once per process from (alfa, beta_halo, beta_bound)
  subroutine ref_line_filter_nt(n, ntype, alfa, bh, bb, f, ff) bind(C, name="ref_line_filter_nt")
    use commtype, only : compact_scheme
    use commvar,  only : alfa_filter
    use filter,   only : compact_filter_initiate, filter_coefficient_cal,     &
                         compact_filter, coef10i
    integer(c_int), value :: n, ntype
    real(c_double), value :: alfa, bh, bb
    real(c_double), intent(in) :: f(-5:n+5)
    real(c_double), intent(out) :: ff(0:n)
    type(compact_scheme) :: s
    alfa_filter = alfa
    if (.not. allocated(coef10i)) then
      call filter_coefficient_cal(alfa=alfa, beter_halo=bh, beter_bouond=bb)
    endif
    call compact_filter_initiate(afilter=s, ntype=ntype, dim=n)
    ff = compact_filter(afilter=s, f=f, dim=n)
    deallocate(s%a, s%c, s%ac)
  end subroutine ref_line_filter_nt

  ! set up the miniapp for an im x jm x km TGV box (tgvsolver.F90:152-261)
  subroutine ref_field_init(im_, jm_, km_, dt) bind(C, name="ref_field_init")
    use commvar,   only : im, jm, km, numq, difschm, alfa_filter, deltat
    use commarray, only : allocommarray
    use initilise, only : solver_init, flowfield_init, maxstep
    integer(c_int), value :: im_, jm_, km_
    real(c_double), value :: dt
    numq = 5
    im = im_
    jm = jm_
    km = km_
    difschm = '643c'
    alfa_filter = 0.49d0
    maxstep = 0
    call solver_init
    call allocommarray
    call flowfield_init
    deltat = dt
  end subroutine ref_field_init

  ! replace q on 0:im and recompute primitives with the reference's q2fvar
  subroutine ref_set_q(qin) bind(C, name="ref_set_q")
    use commvar,   only : im, jm, km
    use commarray, only : q, rho, vel, prs, tmp
    use fludyna,   only : q2fvar
    real(c_double), intent(in) :: qin(0:im, 0:jm, 0:km, 1:5)
    integer :: i, j, k
    q(0:im, 0:jm, 0:km, 1:5) = qin
    do k = 0, km
    do j = 0, jm
    do i = 0, im
      call q2fvar(q=q(i,j,k,:), density=rho(i,j,k), velocity=vel(i,j,k,:),  &
                  pressure=prs(i,j,k), temperature=tmp(i,j,k))
    enddo
    enddo
    enddo
  end subroutine ref_set_q

  subroutine ref_get_q(qout) bind(C, name="ref_get_q")
    use commvar,   only : im, jm, km
    use commarray, only : q
    real(c_double), intent(out) :: qout(0:im, 0:jm, 0:km, 1:5)
    qout = q(0:im, 0:jm, 0:km, 1:5)
  end subroutine ref_get_q

  ! rho,u,v,w,p,T on 0:im
  subroutine ref_get_prims(pout) bind(C, name="ref_get_prims")
    use commvar,   only : im, jm, km
    use commarray, only : rho, vel, prs, tmp
    real(c_double), intent(out) :: pout(0:im, 0:jm, 0:km, 1:6)
    pout(:,:,:,1) = rho(0:im, 0:jm, 0:km)
    pout(:,:,:,2) = vel(0:im, 0:jm, 0:km, 1)
    pout(:,:,:,3) = vel(0:im, 0:jm, 0:km, 2)
    pout(:,:,:,4) = vel(0:im, 0:jm, 0:km, 3)
    pout(:,:,:,5) = prs(0:im, 0:jm, 0:km)
    pout(:,:,:,6) = tmp(0:im, 0:jm, 0:km)
  end subroutine ref_get_prims

  ! one RHS evaluation as a stage does it: bchomo then rhscal
  ! (tgvsolver.F90:1039-1041)
  subroutine ref_rhscal(qrhs_out, dvel_out, dtmp_out) bind(C, name="ref_rhscal")
    use commvar,   only : im, jm, km
    use commarray, only : qrhs, dvel, dtmp
    use bc,        only : bchomo
    use mainloop,  only : rhscal
    real(c_double), intent(out) :: qrhs_out(0:im, 0:jm, 0:km, 1:5)
    real(c_double), intent(out) :: dvel_out(0:im, 0:jm, 0:km, 1:3, 1:3)
    real(c_double), intent(out) :: dtmp_out(0:im, 0:jm, 0:km, 1:3)
    call bchomo
    call rhscal
    qrhs_out = qrhs
    dvel_out = dvel
    dtmp_out = dtmp
  end subroutine ref_rhscal

  ! the reference's filterq on the current q (tgvsolver.F90:791-883)
  subroutine ref_filterq() bind(C, name="ref_filterq")
    use solver, only : filterq
    call filterq
  end subroutine ref_filterq

  ! nsteps of rk3 (tgvsolver.F90:944-960); stacal appends to ./state.dat
  subroutine ref_rk3_steps(nsteps) bind(C, name="ref_rk3_steps")
    use commvar,  only : nstep, time, deltat
    use mainloop, only : rk3
    integer(c_int), value :: nsteps
    integer :: it
    do it = 1, nsteps
      call rk3
      nstep = nstep + 1
      time = time + deltat
    enddo
    flush(13)
  end subroutine ref_rk3_steps

end module astr_ref_driver
