!-----------------------------------------------------------------------
! astr_hip_mod.F90 -- Fortran host binding of libastr_hip (ISO_C_BINDING).
!
! module astr_hip_capi   the C ABI of include/astr_hip.h, one interface
!                        per entry point, and hipcheck (stop on error, the
!                        reference's termination convention)
! module astr_hip_tgv    the reference TGV solver's module state and
!                        subroutines under their own names
!                        (miniapps/tgv_solver_3d/tgvsolver.F90): im,jm,km,
!                        q(-hm:..,1:5), qrhs, dvel, dtmp, rho, vel, prs, tmp;
!                        program_init, solver_init, allocommarray,
!                        flowfield_init, filterq, gradcal, rhscal, stacal,
!                        rk3, steploop.  The fields live in HBM; the host
!                        arrays are filled when a subroutine's outputs are
!                        the reference's module arrays.
! The main solver's run (astr run) is astr_hip_run.F90 (module
! astr_hip_main, program astr_hip).
! module astr_hip_thermchem  chemread, chemrate, cpeval, enthpy,
!                        temperature_calc, rgcmix_scar, convertxiyi with the
!                        reference's argument lists (src/thermchem.F90:59,
!                        :1553, :1342, :1498, :1775, :1281, :1639) plus an
!                        array form of chemrate over many cells
!-----------------------------------------------------------------------
module astr_hip_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ASTR_SEM_MINIAPP = 0, ASTR_SEM_MAIN = 1

  type, bind(C) :: astr_hip_config
    integer(c_int) :: im, jm, km, semantics
    real(c_double) :: alfa_filter, beta_halo, beta_bound
    real(c_double) :: gamma, mach, reynolds, prandtl
    real(c_double) :: ref_tem, sutherland
    real(c_double) :: lx, ly, lz
    real(c_double) :: deltat
  end type astr_hip_config

  interface
    function astr_hip_last_error() bind(C, name='astr_hip_last_error')
      import :: c_ptr
      type(c_ptr) :: astr_hip_last_error
    end function
    function astr_hip_chem_last_error() bind(C, name='astr_hip_chem_last_error')
      import :: c_ptr
      type(c_ptr) :: astr_hip_chem_last_error
    end function
    function astr_hip_create(cfg, ctx) bind(C, name='astr_hip_create')
      import :: c_int, c_ptr, astr_hip_config
      type(astr_hip_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: astr_hip_create
    end function
    function astr_hip_destroy(ctx) bind(C, name='astr_hip_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: astr_hip_destroy
    end function
    function astr_hip_set_deltat(ctx, dt) bind(C, name='astr_hip_set_deltat')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int) :: astr_hip_set_deltat
    end function
    function astr_hip_init_tgv(ctx, amp, seed) bind(C, name='astr_hip_init_tgv')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), value :: amp
      integer(c_long_long), value :: seed
      integer(c_int) :: astr_hip_init_tgv
    end function
    function astr_hip_upload(ctx, q, hm) bind(C, name='astr_hip_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
      integer(c_int), value :: hm
      integer(c_int) :: astr_hip_upload
    end function
    function astr_hip_download(ctx, q, prims, hm) bind(C, name='astr_hip_download')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, q, prims
      integer(c_int), value :: hm
      integer(c_int) :: astr_hip_download
    end function
    function astr_hip_filterq(ctx) bind(C, name='astr_hip_filterq')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: astr_hip_filterq
    end function
    function astr_hip_gradcal(ctx, dvel, dtmp) bind(C, name='astr_hip_gradcal')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dvel, dtmp
      integer(c_int) :: astr_hip_gradcal
    end function
    function astr_hip_rhscal(ctx, qrhs) bind(C, name='astr_hip_rhscal')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, qrhs
      integer(c_int) :: astr_hip_rhscal
    end function
    function astr_hip_rk3_steps(ctx, nsteps, ke, ens) bind(C, name='astr_hip_rk3_steps')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(c_ptr), value :: ke, ens
      integer(c_int) :: astr_hip_rk3_steps
    end function
    ! rk3_steps with the main path's statcal statistics (kenergycal,
    ! enstophycal, diss_rate_cal, statistic.F90:593-595) of each step
    function astr_hip_rk3_steps_stats(ctx, nsteps, ke, ens, diss) &
        bind(C, name='astr_hip_rk3_steps_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(c_ptr), value :: ke, ens, diss
      integer(c_int) :: astr_hip_rk3_steps_stats
    end function
    ! qswap, single rank (parallel.F90:4876-4884, 5018-5030, 5159-5170)
    function astr_hip_qswap(ctx) bind(C, name='astr_hip_qswap')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: astr_hip_qswap
    end function
    ! gridgeom (geom.F90:99-692): x(0:im,0:jm,0:km,1:3) -> jacob, dxi
    function astr_hip_gridgeom(ctx, x, jacob, dxi) bind(C, name='astr_hip_gridgeom')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, x, jacob, dxi
      integer(c_int) :: astr_hip_gridgeom
    end function
    ! npdc(1:3) = npdci, npdcj, npdck: 3 periodic, 4 physical at both ends
    function astr_hip_gridgeom_nt(ctx, npdc, x, jacob, dxi) bind(C, name='astr_hip_gridgeom_nt')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: npdc(3)
      type(c_ptr), value :: x, jacob, dxi
      integer(c_int) :: astr_hip_gridgeom_nt
    end function
    ! fds%central (op 0) / compact_filter (op 1) on nlines lines f(-5:n+5)
    ! with the boundary type ntype of their ends
    function astr_hip_line_batch_nt(op, ntype, n, nlines, f, out, alfa_filter, beta_halo, &
                                    beta_bound) bind(C, name='astr_hip_line_batch_nt')
      import :: c_int, c_long_long, c_double, c_ptr
      integer(c_int), value :: op, ntype, n
      integer(c_long_long), value :: nlines
      type(c_ptr), value :: f, out
      real(c_double), value :: alfa_filter, beta_halo, beta_bound
      integer(c_int) :: astr_hip_line_batch_nt
    end function
    function astr_hip_updatefvar(ctx, rho, vel, prs, tmp, hm) bind(C, name='astr_hip_updatefvar')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho, vel, prs, tmp
      integer(c_int), value :: hm
      integer(c_int) :: astr_hip_updatefvar
    end function
    function astr_hip_set_metrics(ctx, jacob, dxi) bind(C, name='astr_hip_set_metrics')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, jacob, dxi
      integer(c_int) :: astr_hip_set_metrics
    end function
    ! the convective scheme: central 643c (upwind = 0), the explicit
    ! upwind-biased one (1; convrsduwd, src/solver.F90:548-1201) or the
    ! compact upwind-biased 543c (2; convrsdcmp, src/solver.F90:1271-1937)
    function astr_hip_set_convection(ctx, upwind, recon_schem, lchardecomp, bfacmpld, shkcrt) &
        bind(C, name='astr_hip_set_convection')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: upwind, recon_schem, lchardecomp
      real(c_double), value :: bfacmpld, shkcrt
      integer(c_int) :: astr_hip_set_convection
    end function
    ! the main path's stored primitives rho, vel(...,1:3), prs, tmp
    function astr_hip_set_fvar(ctx, rho, vel, prs, tmp, hm) bind(C, name='astr_hip_set_fvar')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho, vel, prs, tmp
      integer(c_int), value :: hm
      integer(c_int) :: astr_hip_set_fvar
    end function
    ! writechkpt / readcheckpoint (src/readwrite.F90:2390, 1381)
    function astr_hip_write_checkpoint(ctx, path, nstep, time) bind(C, name='astr_hip_write_checkpoint')
      import :: c_int, c_ptr, c_char, c_long_long, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_long_long), value :: nstep
      real(c_double), value :: time
      integer(c_int) :: astr_hip_write_checkpoint
    end function
    function astr_hip_read_checkpoint(ctx, path, nstep, time) bind(C, name='astr_hip_read_checkpoint')
      import :: c_int, c_ptr, c_char, c_long_long, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_long_long), intent(out) :: nstep
      real(c_double), intent(out) :: time
      integer(c_int) :: astr_hip_read_checkpoint
    end function
    function astr_hip_sync(ctx) bind(C, name='astr_hip_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: astr_hip_sync
    end function
    function astr_hip_chem_create(path, phase, chem) bind(C, name='astr_hip_chem_create')
      import :: c_int, c_ptr, c_char
      character(kind=c_char), intent(in) :: path(*), phase(*)
      type(c_ptr), intent(out) :: chem
      integer(c_int) :: astr_hip_chem_create
    end function
    function astr_hip_chem_destroy(chem) bind(C, name='astr_hip_chem_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: chem
      integer(c_int) :: astr_hip_chem_destroy
    end function
    function astr_hip_chem_info(chem, nsp, nrx, mw, names, name_len) bind(C, name='astr_hip_chem_info')
      import :: c_int, c_ptr
      type(c_ptr), value :: chem, nsp, nrx, mw, names
      integer(c_int), value :: name_len
      integer(c_int) :: astr_hip_chem_info
    end function
    function astr_hip_chem_temperature(chem, n, y, e, t) bind(C, name='astr_hip_chem_temperature')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: chem
      integer(c_long_long), value :: n
      type(c_ptr), value :: y, e, t
      integer(c_int) :: astr_hip_chem_temperature
    end function
    function astr_hip_chem_rates(chem, n, t, rp, input_is_p, y, wdot, hmix, cpmix, hk) &
        bind(C, name='astr_hip_chem_rates')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: chem
      integer(c_long_long), value :: n
      type(c_ptr), value :: t, rp
      integer(c_int), value :: input_is_p
      type(c_ptr), value :: y, wdot, hmix, cpmix, hk
      integer(c_int) :: astr_hip_chem_rates
    end function
    function astr_hip_srccomb(chem, im, jm, km, hm, numq, rho, tmp, spc, jacob, qrhs) &
        bind(C, name='astr_hip_srccomb')
      import :: c_int, c_ptr
      type(c_ptr), value :: chem
      integer(c_int), value :: im, jm, km, hm, numq
      type(c_ptr), value :: rho, tmp, spc, jacob, qrhs
      integer(c_int) :: astr_hip_srccomb
    end function
  end interface

contains

  ! the text of a NUL-terminated C string
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    call c_f_pointer(p, a, [4096])
    n = 0
    do while (n < 4096)
      if (a(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(a(1:n), s)
  end function c_string

  ! nonzero status: print the library's message and stop, as the reference
  ! stops on its own errors (e.g. src/commarray.F90:323)
  subroutine hipcheck(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(4A)') ' !! ', where, ': ', c_string(astr_hip_last_error())
      stop 1
    end if
  end subroutine hipcheck

  subroutine chemcheck(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(4A)') ' !! ', where, ': ', c_string(astr_hip_chem_last_error())
      stop 1
    end if
  end subroutine chemcheck

end module astr_hip_capi

!-----------------------------------------------------------------------
! the TGV solver (tgvsolver.F90) on the device
!-----------------------------------------------------------------------
module astr_hip_tgv
  use iso_c_binding
  use astr_hip_capi
  implicit none

  ! commvar (src/commvar.F90:18-98), the values tgvsolver.F90 sets
  integer :: im = 128, jm = 128, km = 128, hm = 5, numq = 5
  integer :: nstep = 0, maxstep = 100
  real(8) :: time = 0.d0, deltat = 1.d-3
  real(8) :: gamma = 1.4d0, mach = 0.1d0, reynolds = 1600.d0, prandtl = 0.72d0
  real(8) :: ref_tem = 273.15d0, alfa_filter = 0.49d0
  character(len=4) :: difschm = '643c'
  ! seeded Gaussian velocity noise on the initial field (0: the reference's field)
  real(8) :: noise_amp = 0.d0
  integer(c_long_long) :: noise_seed = 0

  ! commarray (tgvsolver.F90:1-40)
  real(8), allocatable, target :: q(:,:,:,:), qrhs(:,:,:,:), vel(:,:,:,:), dtmp(:,:,:,:)
  real(8), allocatable, target :: dvel(:,:,:,:,:)
  real(8), allocatable, target :: rho(:,:,:), prs(:,:,:), tmp(:,:,:)

  type(c_ptr) :: ctx = c_null_ptr
  integer :: statfile = 13
  logical :: statopen = .false.

contains

  ! program_init (tgvsolver.F90:152-170): grid and step count
  subroutine program_init(nim, njm, nkm, nmax)
    integer, intent(in), optional :: nim, njm, nkm, nmax
    numq = 5
    if (present(nim)) im = nim
    if (present(njm)) jm = njm
    if (present(nkm)) km = nkm
    if (present(nmax)) maxstep = nmax
    difschm = '643c'
    alfa_filter = 0.49d0
    write(*, '(3(A,I0))') '  ** dimension set: ', im, ' x ', jm, ' x ', km
  end subroutine program_init

  ! solver_init (tgvsolver.F90:236-261): scheme and filter coefficients,
  ! built on the device with the flow constants (flowfield_init :181-196)
  subroutine solver_init
    type(astr_hip_config) :: cfg
    real(8), parameter :: pi = 4.d0 * atan(1.d0)
    cfg%im = im; cfg%jm = jm; cfg%km = km
    cfg%semantics = ASTR_SEM_MINIAPP
    cfg%alfa_filter = alfa_filter
    cfg%beta_halo = 1.11d0
    cfg%beta_bound = 1.09d0
    cfg%gamma = gamma; cfg%mach = mach; cfg%reynolds = reynolds; cfg%prandtl = prandtl
    cfg%ref_tem = ref_tem
    cfg%sutherland = 110.4d0
    cfg%lx = 2.d0 * pi; cfg%ly = 2.d0 * pi; cfg%lz = 2.d0 * pi
    cfg%deltat = deltat
    call hipcheck(astr_hip_create(cfg, ctx), 'solver_init')
    print *, ' ** solver initialised on the device'
  end subroutine solver_init

  ! allocommarray (tgvsolver.F90:13-38)
  subroutine allocommarray
    allocate(q(-hm:im+hm, -hm:jm+hm, -hm:km+hm, 1:numq))
    allocate(rho(-hm:im+hm, -hm:jm+hm, -hm:km+hm))
    allocate(prs(-hm:im+hm, -hm:jm+hm, -hm:km+hm))
    allocate(tmp(-hm:im+hm, -hm:jm+hm, -hm:km+hm))
    allocate(vel(-hm:im+hm, -hm:jm+hm, -hm:km+hm, 1:3))
    allocate(qrhs(0:im, 0:jm, 0:km, 1:numq))
    allocate(dvel(0:im, 0:jm, 0:km, 1:3, 1:3))
    allocate(dtmp(0:im, 0:jm, 0:km, 1:3))
    print *, ' ** global arrays allocated'
  end subroutine allocommarray

  ! flowfield_init (tgvsolver.F90:172-234), evaluated on the device
  subroutine flowfield_init
    call hipcheck(astr_hip_init_tgv(ctx, noise_amp, noise_seed), 'flowfield_init')
    nstep = 0
    time = 0.d0
    call hipcheck(astr_hip_set_deltat(ctx, deltat), 'flowfield_init')
    call download_state
    print *, ' ** flow field initilised'
  end subroutine flowfield_init

  ! host copies of q and the primitives (q2fvar, tgvsolver.F90:92-121),
  ! halo cells included
  subroutine download_state
    real(8), allocatable, target :: pr(:,:,:,:)
    allocate(pr(-hm:im+hm, -hm:jm+hm, -hm:km+hm, 1:6))
    call hipcheck(astr_hip_download(ctx, c_loc(q), c_loc(pr), int(hm, c_int)), 'download')
    rho = pr(:,:,:,1)
    vel(:,:,:,1) = pr(:,:,:,2)
    vel(:,:,:,2) = pr(:,:,:,3)
    vel(:,:,:,3) = pr(:,:,:,4)
    prs = pr(:,:,:,5)
    tmp = pr(:,:,:,6)
  end subroutine download_state

  ! the host q (halo cells are not read) becomes the device state
  subroutine upload_state
    call hipcheck(astr_hip_upload(ctx, q, int(hm, c_int)), 'upload')
  end subroutine upload_state

  ! filterq (tgvsolver.F90:791-883), in place on the device state
  subroutine filterq
    call hipcheck(astr_hip_filterq(ctx), 'filterq')
    call download_state
  end subroutine filterq

  ! gradcal (tgvsolver.F90:450-537) -> dvel, dtmp
  subroutine gradcal
    call hipcheck(astr_hip_gradcal(ctx, c_loc(dvel), c_loc(dtmp)), 'gradcal')
  end subroutine gradcal

  ! bchomo + rhscal (tgvsolver.F90:1039-1041, :962-992) -> qrhs
  subroutine rhscal
    call hipcheck(astr_hip_rhscal(ctx, c_loc(qrhs)), 'rhscal')
  end subroutine rhscal

  ! stacal (tgvsolver.F90:893-942): one line per step into state.dat
  subroutine stacal(tke, enst)
    real(8), intent(in) :: tke, enst
    if (.not. statopen) then
      open(statfile, file='state.dat')
      statopen = .true.
    end if
    write(statfile, *) nstep, time, tke, enst
    if (mod(nstep, 10) == 0) flush(statfile)
  end subroutine stacal

  ! rk3 (tgvsolver.F90:994-1082): one step, statistics written by stacal
  subroutine rk3
    real(8), target :: ke(1), en(1)
    call hipcheck(astr_hip_rk3_steps(ctx, 1_c_int, c_loc(ke), c_loc(en)), 'rk3')
    call stacal(ke(1), en(1))
  end subroutine rk3

  ! steploop (tgvsolver.F90:944-960): the whole run in one device call
  ! (statistics of every step come back at the end), then the stacal lines
  subroutine steploop
    real(8), allocatable, target :: ke(:), en(:), di(:)
    integer :: n, it
    n = maxstep - nstep
    if (n <= 0) return
    allocate(ke(n), en(n), di(n))
    call hipcheck(astr_hip_rk3_steps(ctx, int(n, c_int), c_loc(ke), c_loc(en)), 'steploop')
    do it = 1, n
      call stacal(ke(it), en(it))
      nstep = nstep + 1
      time = time + deltat
    end do
    if (statopen) flush(statfile)
  end subroutine steploop

  subroutine solver_end
    if (c_associated(ctx)) call hipcheck(astr_hip_destroy(ctx), 'solver_end')
    ctx = c_null_ptr
    if (statopen) close(statfile)
    statopen = .false.
  end subroutine solver_end

end module astr_hip_tgv

!-----------------------------------------------------------------------
! chemistry (src/thermchem.F90) on the device
!-----------------------------------------------------------------------
module astr_hip_thermchem
  use iso_c_binding
  use astr_hip_capi
  implicit none

  ! Cantera 2.5's gas constant (inside cp and the rates) and the reference's
  ! own rguniv (thermchem.F90:26, used by enthpy)
  real(8), parameter :: r_cantera = 8314.46261815324d0, rguniv = 8314.2d0
  real(8), parameter :: prefgb_default = 1.d5

  integer :: num_species = 0, ncstep = 0
  real(8), allocatable :: wmolar(:)
  character(len=16), allocatable :: spcsym(:)
  real(8) :: prefgb = prefgb_default
  type(c_ptr) :: mixture = c_null_ptr

contains

  ! chemread (thermchem.F90:59-98): mechanism from a CTML file
  subroutine chemread(cheminfile)
    character(len=*), intent(in) :: cheminfile
    integer(c_int), target :: ns, nr
    real(8), allocatable, target :: w(:)
    character(kind=c_char), allocatable, target :: names(:)
    integer :: k, j
    call chemcheck(astr_hip_chem_create(trim(cheminfile) // c_null_char, 'gas' // c_null_char, &
                                        mixture), 'chemread')
    call chemcheck(astr_hip_chem_info(mixture, c_loc(ns), c_loc(nr), c_null_ptr, c_null_ptr, 0), &
                   'chemread')
    num_species = ns
    ncstep = nr
    allocate(w(ns), names(16 * ns))
    call chemcheck(astr_hip_chem_info(mixture, c_null_ptr, c_null_ptr, c_loc(w), c_loc(names), 16), &
                   'chemread')
    if (allocated(wmolar)) deallocate(wmolar, spcsym)
    allocate(wmolar(ns), spcsym(ns))
    wmolar = w
    do k = 1, ns
      spcsym(k) = ''
      do j = 1, 16
        if (names(16 * (k - 1) + j) == c_null_char) exit
        spcsym(k)(j:j) = names(16 * (k - 1) + j)
      end do
    end do
  end subroutine chemread

  ! chemrate (thermchem.F90:1553-1575) for one cell: wi = wdot_k W_k (kg/m^3/s)
  subroutine chemrate(den, tmp, spc, wi)
    real(8), intent(in) :: den, tmp, spc(:)
    real(8), intent(out) :: wi(:)
    real(8) :: d(1), t(1), y(num_species), w(num_species)
    d(1) = den; t(1) = tmp; y = spc(1:num_species)
    call chemrate_cells(1, d, t, y, w)
    wi(1:num_species) = w
  end subroutine chemrate

  ! chemrate over n cells: den(n), tmp(n), spc(n,num_species) -> wi(n,num_species)
  subroutine chemrate_cells(n, den, tmp, spc, wi)
    integer, intent(in) :: n
    real(8), intent(in), target :: den(n), tmp(n), spc(n, num_species)
    real(8), intent(out), target :: wi(n, num_species)
    call chemcheck(astr_hip_chem_rates(mixture, int(n, c_long_long), c_loc(tmp), c_loc(den), 0_c_int, &
                                       c_loc(spc), c_loc(wi), c_null_ptr, c_null_ptr, c_null_ptr), &
                   'chemrate')
  end subroutine chemrate_cells

  ! srccomb (solver.F90:500-537) over the reference's arrays, passed
  ! explicitly: qrhs(i,j,k,6:numq) += chemrate(rho, tmp, spc) * jacob at every
  ! point of 0:im x 0:jm x 0:km, in one launch
  subroutine srccomb(im, jm, km, hm, numq, rho, tmp, spc, jacob, qrhs)
    integer, intent(in) :: im, jm, km, hm, numq
    real(8), intent(in), target :: rho(-hm:im+hm, -hm:jm+hm, -hm:km+hm), &
                                   tmp(-hm:im+hm, -hm:jm+hm, -hm:km+hm), &
                                   spc(-hm:im+hm, -hm:jm+hm, -hm:km+hm, 1:num_species), &
                                   jacob(-hm:im+hm, -hm:jm+hm, -hm:km+hm)
    real(8), intent(inout), target :: qrhs(0:im, 0:jm, 0:km, 1:numq)
    call chemcheck(astr_hip_srccomb(mixture, int(im, c_int), int(jm, c_int), int(km, c_int), &
                                    int(hm, c_int), int(numq, c_int), c_loc(rho), c_loc(tmp), &
                                    c_loc(spc), c_loc(jacob), c_loc(qrhs)), 'srccomb')
  end subroutine srccomb

  ! cpeval (thermchem.F90:1342-1428): cp_mass or intEnergy_mass at (tmp,
  ! prefgb, spc), exactly one of cp, eng present
  subroutine cpeval(tmp, spc, eng, cp, ke)
    real(8), intent(in) :: tmp, spc(:)
    real(8), intent(in), optional :: ke
    real(8), intent(out), optional :: cp, eng
    real(8), target :: t(1), p(1), y(num_species), w(num_species), hk(num_species), c(1), h(1)
    if (present(cp) .and. present(eng)) stop ' !! Conflict - both cp and e are given!!'
    t(1) = tmp; p(1) = prefgb; y = spc(1:num_species)
    call chemcheck(astr_hip_chem_rates(mixture, 1_c_long_long, c_loc(t), c_loc(p), 1_c_int, c_loc(y), &
                                       c_loc(w), c_loc(h), c_loc(c), c_loc(hk)), 'cpeval')
    if (present(cp)) cp = c(1)
    if (present(eng)) then
      ! e = sum Y_k (h_k - R T / W_k) with h_k at Cantera's gas constant
      eng = sum(y * (hk * (r_cantera / rguniv) - r_cantera * tmp / wmolar))
    end if
  end subroutine cpeval

  ! enthpy (thermchem.F90:1498-1541): h_k = (H/RT)_k (8314.2/W_k) T
  subroutine enthpy(tmp, hi)
    real(8), intent(in) :: tmp
    real(8), intent(out) :: hi(num_species)
    real(8), target :: t(1), p(1), y(num_species), w(num_species), hk(num_species)
    t(1) = tmp; p(1) = prefgb
    y = 0.d0
    y(num_species) = 1.d0
    call chemcheck(astr_hip_chem_rates(mixture, 1_c_long_long, c_loc(t), c_loc(p), 1_c_int, c_loc(y), &
                                       c_loc(w), c_null_ptr, c_null_ptr, c_loc(hk)), 'enthpy')
    hi = hk
  end subroutine enthpy

  ! temperature_calc (thermchem.F90:1775-1786): T from (den, spc, eint);
  ! tmp holds the initial guess on entry
  subroutine temperature_calc(tmp, den, spc, eint)
    real(8), intent(in) :: den, spc(:), eint
    real(8), intent(inout) :: tmp
    real(8), target :: y(num_species), e(1), t(1)
    y = spc(1:num_species); e(1) = eint; t(1) = tmp
    call chemcheck(astr_hip_chem_temperature(mixture, 1_c_long_long, c_loc(y), c_loc(e), c_loc(t)), &
                   'temperature_calc')
    tmp = t(1)
  end subroutine temperature_calc

  ! rgcmix (thermchem.F90:1281-1289): mixture gas constant, R = 8314.2
  function rgcmix_scar(spc) result(vout)
    real(8), intent(in) :: spc(:)
    real(8) :: vout
    vout = sum(spc(1:num_species) * rguniv / wmolar)
  end function rgcmix_scar

  ! convertxiyi (thermchem.F90:1639-1657): mole <-> mass fractions
  subroutine convertxiyi(fracin, fracout, mode)
    character(len=*), intent(in) :: mode
    real(8), intent(in) :: fracin(:)
    real(8), intent(out) :: fracout(:)
    if (mode == 'X2Y') then
      fracout(1:num_species) = (fracin(1:num_species) * wmolar) / sum(fracin(1:num_species) * wmolar)
    elseif (mode == 'Y2X') then
      fracout(1:num_species) = (fracin(1:num_species) / wmolar) / sum(fracin(1:num_species) / wmolar)
    else
      stop ' !!Error - wrong mode given in convertxiyi!!'
    end if
  end subroutine convertxiyi

  subroutine chemend
    if (c_associated(mixture)) call chemcheck(astr_hip_chem_destroy(mixture), 'chemend')
    mixture = c_null_ptr
  end subroutine chemend

end module astr_hip_thermchem
