! ISO_C_BINDING interface to librbc3d_gpu.so (include/rbc3d.h).
!
! A Fortran host (the RBC3D reference is one) uses this module to call the GPU
! implementation of the Ewald-split boundary-integral operator.  Arrays are passed by
! base address exactly as the reference stores them (column-major, indices 1-based),
! Fortran logicals cross as integer(c_int) 0/1, and every routine returns a status
! that rbc3d_check turns into an error stop with the library's message.
module rbc3d_bind
  use iso_c_binding
  implicit none
  private
  public :: rbc3d_conf, rbc3d_check, c_ptr, c_null_ptr

  integer(c_int), parameter, public :: RBC3D_OK = 0, RBC3D_EINVAL = 1, RBC3D_EHIP = 2, &
       RBC3D_ENBR = 3, RBC3D_ESTATE = 4, RBC3D_ENONFINITE = 5, RBC3D_ENOCONV = 6

  ! ModConf parameters (common/ModConf.F90:14-47) as the library's rbc3d_conf
  type, bind(C) :: rbc3d_conf
    real(c_double) :: Lb(3)
    real(c_double) :: alpha
    real(c_double) :: rc
    integer(c_int) :: Nb(3)
    integer(c_int) :: P
    integer(c_int) :: ntypes
    real(c_double) :: viscRat(8)
    real(c_double) :: vBkg(3)
  end type rbc3d_conf

  public :: rbc3d_create, rbc3d_destroy, rbc3d_set_cells, rbc3d_set_cell_density, &
       rbc3d_set_wall, rbc3d_set_wall_density, rbc3d_prepare_sing_int_on_wall, &
       rbc3d_hashtable_build, rbc3d_add_int_on_rbcs, rbc3d_add_int_on_walls, &
       rbc3d_pme_distrib_source, rbc3d_pme_transform, rbc3d_pme_add_interp_vel, &
       rbc3d_sh_anal_gau, rbc3d_sh_synth_gau, rbc3d_sh_synth_equ, rbc3d_sh_grad_gau, &
       rbc3d_glob_sph_trans, &
       rbc3d_calc_velocity_field, rbc3d_cell_dof, rbc3d_cell_matmult, rbc3d_cell_rhs, &
       rbc3d_wall_dof, rbc3d_wall_matmult, rbc3d_wall_residual_vel, rbc3d_get_wall_density, &
       rbc3d_noslip_wall, rbc3d_solve_rbc_vel, rbc3d_pair_stats, rbc3d_set_precision, &
       rbc3d_mlp_create, rbc3d_mlp_destroy, rbc3d_mlp_set_params, rbc3d_mlp_set_norm, &
       rbc3d_noslip_wall_learned, rbc3d_set_gmres_restart, &
       rbc3d_rbc_compute_geometry, rbc3d_shell_res_force, rbc3d_rbc_sph_project, &
       rbc3d_set_ref_cells, rbc3d_compute_rbc_vel, rbc3d_move_rbcs, rbc3d_rebox_rbcs, &
       rbc3d_vol_constrain_rbcs, rbc3d_filter_rbcs, rbc3d_get_rbc_coords, rbc3d_timeint_euler, &
       rbc3d_set_repulsion, rbc3d_intra_cell_repulsion, rbc3d_inter_cell_repulsion, &
       rbc3d_wall_mesh_info, rbc3d_read_wall_mesh, rbc3d_timeint_rk2, rbc3d_timeint_ab2

  ! rbc3d_set_precision modes
  integer(c_int), parameter, public :: RBC3D_PREC_FP64 = 0, RBC3D_PREC_SPLIT = 1

  interface
    integer(c_int) function rbc3d_create(conf, ctx) bind(C, name="rbc3d_create")
      import :: c_int, c_ptr, rbc3d_conf
      type(rbc3d_conf), intent(in) :: conf
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function rbc3d_destroy(ctx) bind(C, name="rbc3d_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function rbc3d_last_error() bind(C, name="rbc3d_last_error")
      import :: c_ptr
    end function
    integer(c_int) function rbc3d_set_cells(ctx, nrbc, nlat0, nlat, celltype, x, a3, detj) &
         bind(C, name="rbc3d_set_cells")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrbc
      integer(c_int), intent(in) :: nlat0(*), nlat(*), celltype(*)
      real(c_double), intent(in) :: x(*), a3(*), detj(*)
    end function
    integer(c_int) function rbc3d_set_cell_density(ctx, which, dens) &
         bind(C, name="rbc3d_set_cell_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(in) :: dens(*)
    end function
    integer(c_int) function rbc3d_set_wall(ctx, nvert, nele, x, e2v) bind(C, name="rbc3d_set_wall")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nvert, nele
      real(c_double), intent(in) :: x(*)
      integer(c_int), intent(in) :: e2v(*)
    end function
    integer(c_int) function rbc3d_set_wall_density(ctx, f) bind(C, name="rbc3d_set_wall_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: f(*)
    end function
    integer(c_int) function rbc3d_prepare_sing_int_on_wall(ctx) &
         bind(C, name="rbc3d_prepare_sing_int_on_wall")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rbc3d_hashtable_build(Nc, iLbNc, nPoint, x, ldx, hoc, next) &
         bind(C, name="rbc3d_hashtable_build")
      import :: c_int, c_double
      integer(c_int), intent(in) :: Nc(3)
      real(c_double), intent(in) :: iLbNc(3)
      integer(c_int), value :: nPoint, ldx
      real(c_double), intent(in) :: x(*)
      integer(c_int), intent(out) :: hoc(*), next(*)
    end function
    integer(c_int) function rbc3d_add_int_on_rbcs(ctx, c1, c2, nT, tx, tindx, active, Acoef, v) &
         bind(C, name="rbc3d_add_int_on_rbcs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: c1, c2
      integer(c_int), value :: nT
      real(c_double), intent(in) :: tx(*), Acoef(*)
      integer(c_int), intent(in) :: tindx(*), active(*)
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function rbc3d_add_int_on_walls(ctx, c1, nT, tx, tindx, active, Acoef, v) &
         bind(C, name="rbc3d_add_int_on_walls")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: c1
      integer(c_int), value :: nT
      real(c_double), intent(in) :: tx(*), Acoef(*)
      integer(c_int), intent(in) :: tindx(*), active(*)
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function rbc3d_pair_stats(ctx, npairs, checksum) bind(C, name="rbc3d_pair_stats")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: npairs, checksum
    end function
    integer(c_int) function rbc3d_pme_distrib_source(ctx, c1, c2, cells, walls) &
         bind(C, name="rbc3d_pme_distrib_source")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: c1, c2
      integer(c_int), value :: cells, walls
    end function
    integer(c_int) function rbc3d_pme_transform(ctx) bind(C, name="rbc3d_pme_transform")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rbc3d_pme_add_interp_vel(ctx, nT, tx, active, Acoef, v) &
         bind(C, name="rbc3d_pme_add_interp_vel")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nT
      real(c_double), intent(in) :: tx(*), Acoef(*)
      integer(c_int), intent(in) :: active(*)
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function rbc3d_sh_anal_gau(nlat, nlon, nt, g, idg, jdg, a, b, ida, jda) &
         bind(C, name="rbc3d_sh_anal_gau")
      import :: c_int, c_double
      integer(c_int), value :: nlat, nlon, nt, idg, jdg, ida, jda
      real(c_double), intent(in) :: g(*)
      real(c_double), intent(out) :: a(*), b(*)
    end function
    integer(c_int) function rbc3d_sh_synth_gau(nlat, nlon, nt, g, idg, jdg, a, b, ida, jda) &
         bind(C, name="rbc3d_sh_synth_gau")
      import :: c_int, c_double
      integer(c_int), value :: nlat, nlon, nt, idg, jdg, ida, jda
      real(c_double), intent(out) :: g(*)
      real(c_double), intent(in) :: a(*), b(*)
    end function
    integer(c_int) function rbc3d_sh_synth_equ(nlat, nlon, nt, g, idg, jdg, a, b, ida, jda) &
         bind(C, name="rbc3d_sh_synth_equ")
      import :: c_int, c_double
      integer(c_int), value :: nlat, nlon, nt, idg, jdg, ida, jda
      real(c_double), intent(out) :: g(*)
      real(c_double), intent(in) :: a(*), b(*)
    end function
    integer(c_int) function rbc3d_sh_grad_gau(nlat, nlon, nt, v, w, idv, jdv, a, b, ida, jda) &
         bind(C, name="rbc3d_sh_grad_gau")
      import :: c_int, c_double
      integer(c_int), value :: nlat, nlon, nt, idv, jdv, ida, jda
      real(c_double), intent(out) :: v(*), w(*)
      real(c_double), intent(in) :: a(*), b(*)
    end function
    integer(c_int) function rbc3d_glob_sph_trans(ctx, v, c, direction) &
         bind(C, name="rbc3d_glob_sph_trans")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: v(*), c(*)
      integer(c_int), value :: direction
    end function
    integer(c_int) function rbc3d_calc_velocity_field(ctx, nT, tx, tindx, active, Acoef, v) &
         bind(C, name="rbc3d_calc_velocity_field")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nT
      real(c_double), intent(in) :: tx(*), Acoef(*)
      integer(c_int), intent(in) :: tindx(*), active(*)
      real(c_double), intent(inout) :: v(*)
    end function
    integer(c_int) function rbc3d_cell_dof(ctx, dof) bind(C, name="rbc3d_cell_dof")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: dof
    end function
    integer(c_int) function rbc3d_cell_matmult(ctx, u, b) bind(C, name="rbc3d_cell_matmult")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(out) :: b(*)
    end function
    integer(c_int) function rbc3d_cell_rhs(ctx, rhs) bind(C, name="rbc3d_cell_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhs(*)
    end function
    integer(c_int) function rbc3d_wall_dof(ctx, dof) bind(C, name="rbc3d_wall_dof")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: dof
    end function
    integer(c_int) function rbc3d_wall_matmult(ctx, u, b) bind(C, name="rbc3d_wall_matmult")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(out) :: b(*)
    end function
    integer(c_int) function rbc3d_wall_residual_vel(ctx, v) bind(C, name="rbc3d_wall_residual_vel")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v(*)
    end function
    integer(c_int) function rbc3d_get_wall_density(ctx, f) bind(C, name="rbc3d_get_wall_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: f(*)
    end function
    integer(c_int) function rbc3d_noslip_wall(ctx, rtol, maxit, finc0, niter, rnorm, converged, &
         vres_max) bind(C, name="rbc3d_noslip_wall")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: rtol
      integer(c_int), value :: maxit
      type(c_ptr), value :: finc0
      integer(c_int), intent(out) :: niter, converged
      real(c_double), intent(out) :: rnorm, vres_max
    end function
    integer(c_int) function rbc3d_solve_rbc_vel(ctx, v, niter, rnorm) bind(C, name="rbc3d_solve_rbc_vel")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v(*)
      integer(c_int), intent(out) :: niter
      real(c_double), intent(out) :: rnorm
    end function
    integer(c_int) function rbc3d_set_ref_cells(ctx, ntypes, nlat, x, ES, ED, EB) &
        bind(C, name="rbc3d_set_ref_cells")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntypes
      integer(c_int), intent(in) :: nlat(*)
      real(c_double), intent(in) :: x(*), ES(*), ED(*), EB(*)
    end function
    integer(c_int) function rbc3d_compute_rbc_vel(ctx, v, niter, rnorm) bind(C, name="rbc3d_compute_rbc_vel")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v(*)
      integer(c_int), intent(out) :: niter
      real(c_double), intent(out) :: rnorm
    end function
    integer(c_int) function rbc3d_move_rbcs(ctx, Ts) bind(C, name="rbc3d_move_rbcs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: Ts
    end function
    integer(c_int) function rbc3d_rebox_rbcs(ctx) bind(C, name="rbc3d_rebox_rbcs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rbc3d_vol_constrain_rbcs(ctx, eps_dist) bind(C, name="rbc3d_vol_constrain_rbcs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: eps_dist
    end function
    integer(c_int) function rbc3d_filter_rbcs(ctx) bind(C, name="rbc3d_filter_rbcs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function rbc3d_get_rbc_coords(ctx, x) bind(C, name="rbc3d_get_rbc_coords")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: x(*)
    end function
    integer(c_int) function rbc3d_timeint_euler(ctx, Ts, nsteps, eps_dist, noslip_rtol, noslip_maxit, &
        time, its_cell, its_wall) bind(C, name="rbc3d_timeint_euler")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: Ts, eps_dist, noslip_rtol
      integer(c_int), value :: nsteps, noslip_maxit
      real(c_double), intent(inout) :: time
      integer(c_int), intent(out) :: its_cell(*), its_wall(*)
    end function
    integer(c_int) function rbc3d_set_repulsion(ctx, on, force_coef, visc_rat_thresh, rigidsep) &
        bind(C, name="rbc3d_set_repulsion")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: on, rigidsep
      real(c_double), value :: force_coef, visc_rat_thresh
    end function
    integer(c_int) function rbc3d_intra_cell_repulsion(ctx, df, npoints) &
        bind(C, name="rbc3d_intra_cell_repulsion")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: df(*)
      integer(c_int), intent(out) :: npoints
    end function
    integer(c_int) function rbc3d_inter_cell_repulsion(ctx, eps_dist, npoints, dist_min) &
        bind(C, name="rbc3d_inter_cell_repulsion")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: eps_dist
      integer(c_int), intent(out) :: npoints
      real(c_double), intent(out) :: dist_min
    end function
    integer(c_int) function rbc3d_timeint_rk2(ctx, alpha, Ts, nsteps, noslip_rtol, noslip_maxit, &
        time, its_cell, its_wall) bind(C, name="rbc3d_timeint_rk2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: alpha, Ts, noslip_rtol
      integer(c_int), value :: nsteps, noslip_maxit
      real(c_double), intent(inout) :: time
      integer(c_int), intent(out) :: its_cell(*), its_wall(*)
    end function
    integer(c_int) function rbc3d_timeint_ab2(ctx, Ts, nsteps, noslip_rtol, noslip_maxit, &
        time, its_cell, its_wall) bind(C, name="rbc3d_timeint_ab2")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: Ts, noslip_rtol
      integer(c_int), value :: nsteps, noslip_maxit
      real(c_double), intent(inout) :: time
      integer(c_int), intent(out) :: its_cell(*), its_wall(*)
    end function
    ! file names are NUL-terminated: pass trim(fn)//c_null_char
    integer(c_int) function rbc3d_wall_mesh_info(fn, nvert, nele) bind(C, name="rbc3d_wall_mesh_info")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: fn(*)
      integer(c_int), intent(out) :: nvert, nele
    end function
    integer(c_int) function rbc3d_read_wall_mesh(fn, x, e2v) bind(C, name="rbc3d_read_wall_mesh")
      import :: c_int, c_char, c_double
      character(kind=c_char), intent(in) :: fn(*)
      real(c_double), intent(out) :: x(*)
      integer(c_int), intent(out) :: e2v(*)
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
    ! per-cell geometry and membrane mechanics (RBC_ComputeGeometry, Shell_ResForce,
    ! RBC_SphProject of common/ModRbc.F90), batched over cells
    integer(c_int) function rbc3d_rbc_compute_geometry(ncell, nlat, x, a1, a2, a3, a, a_rcp, &
         b, detj, xc, area, vol, mesh_size) bind(C, name="rbc3d_rbc_compute_geometry")
      import :: c_int, c_double
      integer(c_int), value :: ncell
      integer(c_int), intent(in) :: nlat(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: a1(*), a2(*), a3(*), a(*), a_rcp(*), b(*), detj(*)
      real(c_double), intent(out) :: xc(*), area(*), vol(*), mesh_size(*)
    end function
    integer(c_int) function rbc3d_shell_res_force(ncell, nlat, celltype, a1, a2, a3, a, a_rcp, &
         b, detj, area, ntypes, nlat_ref, a_rcp_ref, b_ref, es, ed, eb, f) &
         bind(C, name="rbc3d_shell_res_force")
      import :: c_int, c_double
      integer(c_int), value :: ncell, ntypes
      integer(c_int), intent(in) :: nlat(*), celltype(*), nlat_ref(*)
      real(c_double), intent(in) :: a1(*), a2(*), a3(*), a(*), a_rcp(*), b(*), detj(*), area(*)
      real(c_double), intent(in) :: a_rcp_ref(*), b_ref(*), es(*), ed(*), eb(*)
      real(c_double), intent(out) :: f(*)
    end function
    integer(c_int) function rbc3d_rbc_sph_project(ncell, nlat, nlat0, nvar, u) &
         bind(C, name="rbc3d_rbc_sph_project")
      import :: c_int, c_double
      integer(c_int), value :: ncell, nvar
      integer(c_int), intent(in) :: nlat(*), nlat0(*)
      real(c_double), intent(inout) :: u(*)
    end function
    ! precision of the hot path (no counterpart in the reference)
    integer(c_int) function rbc3d_set_precision(ctx, mode) bind(C, name="rbc3d_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    ! GMRES restart length of both solves (the reference keeps PETSc's 30)
    integer(c_int) function rbc3d_set_gmres_restart(ctx, restart) &
        bind(C, name="rbc3d_set_gmres_restart")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: restart
    end function
    ! the learned warm start of NoSlipWall (no counterpart in the reference)
    integer(c_int) function rbc3d_mlp_create(m) bind(C, name="rbc3d_mlp_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: m
    end function
    integer(c_int) function rbc3d_mlp_destroy(m) bind(C, name="rbc3d_mlp_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: m
    end function
    integer(c_int) function rbc3d_mlp_set_params(m, params) bind(C, name="rbc3d_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: params(*)
    end function
    integer(c_int) function rbc3d_mlp_set_norm(m, mean_in, std_in, scale_out) &
         bind(C, name="rbc3d_mlp_set_norm")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: mean_in(9), std_in(9), scale_out(3)
    end function
    integer(c_int) function rbc3d_noslip_wall_learned(ctx, m, rtol, maxit, niter, rnorm, &
         converged, vres_max, fallback) bind(C, name="rbc3d_noslip_wall_learned")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, m
      real(c_double), value :: rtol
      integer(c_int), value :: maxit
      integer(c_int), intent(out) :: niter, converged, fallback
      real(c_double), intent(out) :: rnorm, vres_max
    end function
  end interface

contains

  ! A nonzero status stops the program with the library's message, as the reference's
  ! SPHEREPACK wrappers stop on failure (common/ModSphpk.F90:51-56).
  subroutine rbc3d_check(status, where)
    integer(c_int), intent(in) :: status
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (status == RBC3D_OK) return
    p = rbc3d_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write (*, '(a, a, i0, a, *(a))') where, ': rbc3d status ', status, ': ', msg(1:n)
    error stop 1
  end subroutine rbc3d_check

end module rbc3d_bind
