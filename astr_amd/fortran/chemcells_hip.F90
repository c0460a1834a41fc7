!-----------------------------------------------------------------------
! chemcells_hip -- source terms and thermodynamics of a list of cells
! through the reference's chemistry interface (src/thermchem.F90: chemread,
! chemrate, cpeval, enthpy), evaluated on the GPU.
!
! usage: chemcells_hip MECHANISM.xml CELLS.txt OUT.txt
!   CELLS.txt: first line n; then n lines "den tmp Y(1..num_species)"
!   OUT.txt:   per cell one line
!              "wi(1..num_species) cp eng hi(1..num_species) T_e rgc Xmole(1..)"
!              wi from one chemrate_cells call over all cells; cp, eng and hi
!              from per-cell cpeval / enthpy calls; T_e the temperature that
!              temperature_calc recovers from eng starting at 1000 K; rgc
!              from rgcmix_scar; Xmole from convertxiyi(Y2X)
!
!        chemcells_hip MECHANISM.xml CELLS.txt OUT.txt GRID.txt QRHS.txt
!   also srccomb over a block: GRID.txt holds "im jm km hm", then rho, tmp,
!   spc, jacob and qrhs in Fortran array order (the reference's shapes);
!   QRHS.txt receives qrhs after srccomb, one value a line, in array order
!-----------------------------------------------------------------------
program chemcells_hip
  use astr_hip_thermchem
  implicit none
  character(len=512) :: mech, fin, fout
  integer :: n, i, k
  real(8), allocatable :: den(:), tmp(:), spc(:,:), wi(:,:), hi(:), row(:)
  real(8) :: cp, eng, te
  real(8), allocatable :: xm(:)
  character(len=512) :: fgrid, fqrhs
  integer :: im, jm, km, hm
  real(8), allocatable :: grho(:,:,:), gtmp(:,:,:), gspc(:,:,:,:), gjac(:,:,:), gq(:,:,:,:)

  call get_command_argument(1, mech)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call chemread(trim(mech))

  open(21, file=trim(fin), status='old')
  read(21, *) n
  allocate(den(n), tmp(n), spc(n, num_species), wi(n, num_species), hi(num_species), &
           row(2 + num_species))
  do i = 1, n
    read(21, *) row
    den(i) = row(1)
    tmp(i) = row(2)
    spc(i, :) = row(3:)
  end do
  close(21)

  call chemrate_cells(n, den, tmp, spc, wi)
  allocate(xm(num_species))

  open(22, file=trim(fout), status='replace')
  do i = 1, n
    call cpeval(tmp(i), spc(i, :), cp=cp)
    call cpeval(tmp(i), spc(i, :), eng=eng)
    call enthpy(tmp(i), hi)
    te = 1000.d0
    call temperature_calc(te, den(i), spc(i, :), eng)
    call convertxiyi(spc(i, :), xm, 'Y2X')
    write(22, '(*(ES25.17E3,1X))') (wi(i, k), k = 1, num_species), cp, eng, (hi(k), k = 1, num_species), &
        te, rgcmix_scar(spc(i, :)), (xm(k), k = 1, num_species)
  end do
  close(22)

  if (command_argument_count() >= 5) then
    call get_command_argument(4, fgrid)
    call get_command_argument(5, fqrhs)
    open(23, file=trim(fgrid), status='old')
    read(23, *) im, jm, km, hm
    allocate(grho(-hm:im+hm, -hm:jm+hm, -hm:km+hm), gtmp(-hm:im+hm, -hm:jm+hm, -hm:km+hm), &
             gspc(-hm:im+hm, -hm:jm+hm, -hm:km+hm, num_species), &
             gjac(-hm:im+hm, -hm:jm+hm, -hm:km+hm), gq(0:im, 0:jm, 0:km, 5 + num_species))
    read(23, *) grho, gtmp, gspc, gjac, gq
    close(23)
    call srccomb(im, jm, km, hm, 5 + num_species, grho, gtmp, gspc, gjac, gq)
    open(24, file=trim(fqrhs), status='replace')
    write(24, '(ES25.17E3)') gq
    close(24)
  end if
  call chemend
end program chemcells_hip
