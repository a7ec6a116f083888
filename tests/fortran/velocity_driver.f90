! Fortran host driver: reads a cell configuration and a target list written by the
! test harness (raw stream file), evaluates CalcVelocityField through the rbc3d_bind
! ISO_C_BINDING module, and writes the velocities and pair statistics.
!
! Input (stream, native endianness):
!   int32 nrbc, np, nT
!   rbc3d_conf
!   int32 nlat0(nrbc), nlat(nrbc), celltype(nrbc)
!   real64 x, a3, detj, f, g: per-cell (nlat,nlon,3) and (nlat,nlon) blocks, concatenated
!   real64 tx(nT,3); int32 tindx(nT,3), active(nT); real64 Acoef(nT)
! Output: real64 v(nT,3); int64 npairs, checksum
program velocity_driver
  use iso_c_binding
  use rbc3d_bind
  implicit none
  character(len=4096) :: fin, fout
  integer(c_int) :: nrbc, np, nT
  type(rbc3d_conf) :: conf
  integer(c_int), allocatable :: nlat0(:), nlat(:), celltype(:), tindx(:, :), active(:)
  real(c_double), allocatable :: x(:, :), a3(:, :), detj(:), f(:, :), g(:, :)
  real(c_double), allocatable :: tx(:, :), Acoef(:), v(:, :)
  integer(c_long_long) :: npairs, checksum
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) nrbc, np, nT
  read (u) conf
  allocate (nlat0(nrbc), nlat(nrbc), celltype(nrbc))
  allocate (x(np, 3), a3(np, 3), detj(np), f(np, 3), g(np, 3))
  allocate (tx(nT, 3), tindx(nT, 3), active(nT), Acoef(nT), v(nT, 3))
  read (u) nlat0, nlat, celltype
  read (u) x, a3, detj, f, g
  read (u) tx, tindx, active, Acoef
  close (u)

  call rbc3d_check(rbc3d_create(conf, ctx), 'rbc3d_create')
  call rbc3d_check(rbc3d_set_cells(ctx, nrbc, nlat0, nlat, celltype, x, a3, detj), 'set_cells')
  call rbc3d_check(rbc3d_set_cell_density(ctx, 1_c_int, f), 'set_cell_density(f)')
  call rbc3d_check(rbc3d_set_cell_density(ctx, 2_c_int, g), 'set_cell_density(g)')
  v = 0
  call rbc3d_check(rbc3d_calc_velocity_field(ctx, nT, tx, tindx, active, Acoef, v), &
       'calc_velocity_field')
  call rbc3d_check(rbc3d_pair_stats(ctx, npairs, checksum), 'pair_stats')
  call rbc3d_check(rbc3d_destroy(ctx), 'rbc3d_destroy')

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) v, npairs, checksum
  close (u)
end program velocity_driver
