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
!-----------------------------------------------------------------------
program chemcells_hip
  use astr_hip_thermchem
  implicit none
  character(len=512) :: mech, fin, fout
  integer :: n, i, k
  real(8), allocatable :: den(:), tmp(:), spc(:,:), wi(:,:), hi(:), row(:)
  real(8) :: cp, eng, te
  real(8), allocatable :: xm(:)

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
  call chemend
end program chemcells_hip
