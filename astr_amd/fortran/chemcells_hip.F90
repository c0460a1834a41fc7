!-----------------------------------------------------------------------
! chemcells_hip -- source terms and thermodynamics of a list of cells
! through the reference's chemistry interface (src/thermchem.F90: chemread,
! chemrate, cpeval, enthpy), evaluated on the GPU.
!
! usage: chemcells_hip MECHANISM.xml CELLS.txt OUT.txt
!   CELLS.txt: first line n; then n lines "den tmp Y(1..num_species)"
!   OUT.txt:   per cell one line
!              "wi(1..num_species) cp eng hi(1..num_species)"
!              wi from one chemrate_cells call over all cells; cp, eng and hi
!              from per-cell cpeval / enthpy calls
!-----------------------------------------------------------------------
program chemcells_hip
  use astr_hip_thermchem
  implicit none
  character(len=512) :: mech, fin, fout
  integer :: n, i, k
  real(8), allocatable :: den(:), tmp(:), spc(:,:), wi(:,:), hi(:), row(:)
  real(8) :: cp, eng

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

  open(22, file=trim(fout), status='replace')
  do i = 1, n
    call cpeval(tmp(i), spc(i, :), cp=cp)
    call cpeval(tmp(i), spc(i, :), eng=eng)
    call enthpy(tmp(i), hi)
    write(22, '(*(ES25.17E3,1X))') (wi(i, k), k = 1, num_species), cp, eng, (hi(k), k = 1, num_species)
  end do
  close(22)
  call chemend
end program chemcells_hip
