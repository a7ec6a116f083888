/*
 * rbc3d.h -- C ABI of the MI355X (gfx950) implementation of RBC3D's Ewald-split
 * boundary-integral matrix-vector product.
 *
 * Every entry point returns an int status (RBC3D_OK = 0).  Arrays are the
 * reference's Fortran arrays passed by base address (column-major, indices kept
 * 1-based where the reference stores indices), so a Fortran ISO_C_BINDING shim can
 * forward c_loc(array(1,1)) unchanged.  Host pointers unless the name ends in _dev.
 *
 * Each declaration names the reference routine it replaces (paths relative to the
 * RBC3D source tree).  The reference reads module globals (ModConf, ModData); here
 * they live in an opaque context created once per process/rank and GPU.
 */
#ifndef RBC3D_H
#define RBC3D_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RBC3D_OK = 0,
  RBC3D_EINVAL = 1,      /* invalid argument or shape */
  RBC3D_EHIP = 2,        /* HIP runtime error */
  RBC3D_ENBR = 3,        /* more than 32 neighbouring cells within rc of a target */
  RBC3D_ESTATE = 4,      /* call out of order (e.g. no cells set) */
  RBC3D_ENONFINITE = 5,  /* non-finite input */
  RBC3D_ENOCONV = 6,     /* GMRES did not reach the tolerance within maxit */
  RBC3D_EIO = 7          /* a file is missing, unreadable or not of the expected format */
};

typedef struct rbc3d_ctx rbc3d_ctx;

/* Module-global parameters (common/ModConf.F90:14-47).  alpha is the reference's
 * Ewald parameter (erfc(sqrt(pi/alpha) r), exp(-pi alpha q^2)); rc, Nb and P are
 * given explicitly (what SetEwaldPrms, common/ModConf.F90:348-408, derives). */
typedef struct {
  double Lb[3];
  double alpha;
  double rc;
  int Nb[3];
  int P;
  int ntypes;
  double viscRat[8]; /* initCOEFs, common/ModConf.F90:327-344 -> Acoef, Bcoef */
  double vBkg[3];
} rbc3d_conf;

/* --- lifetime: PME_Init/PME_Finalize (common/ModPME.F90:252-345) + the Ewald
 *     tables of common/ModEwaldFunc.F90:97-107,152-162 built eagerly ---------- */
int rbc3d_create(const rbc3d_conf *conf, rbc3d_ctx **out);
int rbc3d_destroy(rbc3d_ctx *ctx);
const char *rbc3d_last_error(void);

/* --- module data: rbcs(:) and walls(:) of common/ModData.F90:13-20 ---------- */
/* Cells: RBC_Create grid (common/ModRbc.F90:44-110) for per-cell nlat (nlon = 2 nlat),
 * geometry x(nlat,nlon,3), a3(nlat,nlon,3), detj(nlat,nlon) concatenated over cells.
 * Builds the x/a3/detJ surface splines (Rbc_BuildSurfaceSource(xFlag),
 * common/ModRbc.F90:737-767), the polar patch (RbcPolarPatch_Create,
 * common/ModPolarPatch.F90:27-75), and the source list + cell list
 * (SourceList_UpdateCoord + HashTable_Build, common/ModSourceList.F90:92-151). */
int rbc3d_set_cells(rbc3d_ctx *ctx, int nrbc, const int *nlat0, const int *nlat,
                    const int *celltype, const double *x, const double *a3, const double *detj);
/* which = 1: f, 2: g.  Rbc_BuildSurfaceSource(fFlag|gFlag) (common/ModRbc.F90:770-802)
 * + SourceList_UpdateDensity (common/ModSourceList.F90:160-187). */
int rbc3d_set_cell_density(rbc3d_ctx *ctx, int which, const double *dens);
/* One wall: x(nvert,3), e2v(nele,3) 1-based.  Wall_Build_V2V + Wall_ComputeGeometry
 * (common/ModWall.F90:66-145), global vertex numbering (common/ModData.F90:50-63),
 * wall source list (common/ModSourceList.F90:130-147). */
int rbc3d_set_wall(rbc3d_ctx *ctx, int nvert, int nele, const double *x, const int *e2v);
int rbc3d_set_wall_density(rbc3d_ctx *ctx, const double *f);
/* PrepareSingIntOnWall (common/ModIntOnWalls.F90:181-308): sparse wall self-matrix */
int rbc3d_prepare_sing_int_on_wall(rbc3d_ctx *ctx);
int rbc3d_wall_self_nnz(rbc3d_ctx *ctx, long long *nnz);

/* --- cell list ----------------------------------------------------------- */
/* HashTable_Build (common/ModHashTable.F90:23-58): hoc(0:Nc1+1,0:Nc2+1,0:Nc3+1),
 * next(nPoint), 1-based, -1 sentinel; x(ldx,3). Bit-identical to the reference. */
int rbc3d_hashtable_build(const int Nc[3], const double iLbNc[3], int nPoint, const double *x,
                          int ldx, int *hoc, int *next);

/* No counterpart in the reference: rbc3d_hashtable_build for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result, bit for bit, on device pointers: d_x(ldx,3), d_hoc
 * ((Nc1+2)(Nc2+2)(Nc3+2)) and d_next(nPoint) in HBM; queued asynchronously on `stream`
 * (a hipStream_t; NULL = the null stream), results valid once the stream reaches them.
 * Working storage is one device array of RBC3D_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends: nPoint <= RBC3D_NOALLOC_MAX_POINTS and Nc1*Nc2*Nc3 <=
 * RBC3D_NOALLOC_MAX_BINS, else RBC3D_EINVAL.  Every call on one GPU shares that array,
 * so calls that overlap on one GPU (two streams, two threads) race. */
#define RBC3D_NOALLOC_MAX_POINTS (1 << 20)
#define RBC3D_NOALLOC_MAX_BINS (1 << 18)
#define RBC3D_NOALLOC_SCRATCH_BYTES (40u << 20)
int rbc3d_hashtable_build_noalloc(const int Nc[3], const double iLbNc[3], int nPoint,
                                  const double *d_x, int ldx, int *d_hoc, int *d_next,
                                  void *stream);

/* --- hot path: real space ------------------------------------------------- */
/* AddIntOnRbcs(c1,c2,tlist,v) (common/ModIntOnRbcs.F90:25-158): pair sums over the
 * cell list, same-surface masking, self-patch singular quadrature, near-singular
 * corrections, AddLinearInt.  tlist = (nT, x(nT,3), indx(nT,0:2), active(nT) as
 * 4-byte logicals, Acoef(nT)); v(nT,3) accumulated in place. */
int rbc3d_add_int_on_rbcs(rbc3d_ctx *ctx, double c1, double c2, int nT, const double *tx,
                          const int *tindx, const int *active, const double *Acoef, double *v);
/* AddIntOnWalls(c1,tlist,v) (common/ModIntOnWalls.F90:37-128) */
int rbc3d_add_int_on_walls(rbc3d_ctx *ctx, double c1, int nT, const double *tx, const int *tindx,
                           const int *active, const double *Acoef, double *v);
/* Diagnostics of the last add_int_on_rbcs: in-cutoff pair count and a checksum of
 * the pair set (sum over pairs of (i+1)*(j+1) mod 2^64, 1-based target/source). */
int rbc3d_pair_stats(rbc3d_ctx *ctx, long long *npairs, unsigned long long *checksum);
/* Diagnostics of the last add_int_on_rbcs (no counterpart in the reference): the
 * number of (target, neighbouring cell) projections of RBC_NearSingInt
 * (common/ModRbcSingInt.F90:103-166) and how many passed its distance test. */
int rbc3d_near_stats(rbc3d_ctx *ctx, long long *nitems, long long *nwork);

/* --- hot path: PME (common/ModPME.F90) ------------------------------------ */
/* PME_Distrib_Source(c1,c2,cells,walls) :58-133 (cells/walls = present flags) */
int rbc3d_pme_distrib_source(rbc3d_ctx *ctx, double c1, double c2, int cells, int walls);
/* PME_Transform :137-222 (forward FFTs, Green's function, B-factor, inverse FFTs) */
int rbc3d_pme_transform(rbc3d_ctx *ctx);
/* PME_Add_Interp_Vel(tlist,v) :228-248 */
int rbc3d_pme_add_interp_vel(rbc3d_ctx *ctx, int nT, const double *tx, const int *active,
                             const double *Acoef, double *v);
/* copy a PME grid to the host: which = 0 ff (3 fields), 1 tt (9), 2 vv (3);
 * layout ff(0:N1-1,0:N2-1,0:N3-1,field) */
int rbc3d_pme_grid(rbc3d_ctx *ctx, int which, double *out);

/* --- hot path: spherical harmonics (common/ModSphpk.F90) ------------------ */
/* ShAnalGau (:74-105, SPHEREPACK shags), ShSynthGau (:240-270, shsgs),
 * ShSynthEqu (:308-338, shses) with the reference's argument lists. */
int rbc3d_sh_anal_gau(int nlat, int nlon, int nt, const double *g, int idg, int jdg, double *a,
                      double *b, int ida, int jda);
int rbc3d_sh_synth_gau(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                       const double *b, int ida, int jda);
int rbc3d_sh_synth_equ(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                       const double *b, int ida, int jda);
/* ShGradGau (common/ModSphpk.F90:377-417, SPHEREPACK gradgs plus the sin(th) factor the
 * routine multiplies back): v(idv,jdv,nt) = dg/dth and w = dg/dphi on the Gauss grid of
 * the scalar field whose shags coefficients are a, b. */
int rbc3d_sh_grad_gau(int nlat, int nlon, int nt, double *v, double *w, int idv, int jdv,
                      const double *a, const double *b, int ida, int jda);
/* --- next to the hot path: per-cell geometry and membrane mechanics ---------------
 * Batched over ncell cells, each on its Gauss grid nlat(i) x 2 nlat(i) (2 <= nlat <= 28),
 * arrays of the reference's cell type concatenated over the cells in cell order:
 * x, a1, a2, a3, f (nlat,nlon,3); a, a_rcp, b (nlat,nlon,2,2); detj (nlat,nlon);
 * xc (3,ncell); area, vol, mesh_size (ncell). */
/* RBC_ComputeGeometry (common/ModRbc.F90:404-509): from x, the tangents a1 = dx/dth,
 * a2 = dx/dphi, unit normal a3, metric a and its inverse a_rcp, curvature tensor b,
 * detJ = sqrt(det a)/sin th, centroid, area, volume and meshSize. */
int rbc3d_rbc_compute_geometry(int ncell, const int *nlat, const double *x, double *a1,
                               double *a2, double *a3, double *a, double *a_rcp, double *b,
                               double *detj, double *xc, double *area, double *vol,
                               double *mesh_size);
/* Shell_ResForce (common/ModRbc.F90:815-900) with Shell_ElasStrs (Skalak law, :908-944)
 * and Shell_Bend (:951-969): the membrane's residual force f (Cartesian) of each cell from
 * its geometry (as rbc3d_rbc_compute_geometry returns it) against the reference cell of
 * its type, rbcRefs(celltype): per type k (1-based celltype) nlat_ref(k), its a_rcp and b
 * concatenated over the types, and the moduli ES(k), ED(k), EB(k).  Compute_Rbc_Vel
 * (common/ModTimeInt.F90:746-753) uses -f as the surface force density. */
int rbc3d_shell_res_force(int ncell, const int *nlat, const int *celltype, const double *a1,
                          const double *a2, const double *a3, const double *a,
                          const double *a_rcp, const double *b, const double *detj,
                          const double *area, int ntypes, const int *nlat_ref,
                          const double *a_rcp_ref, const double *b_ref, const double *ES,
                          const double *ED, const double *EB, double *f);
/* RBC_SphProject (common/ModRbc.F90:659-686): each cell's nvar fields u(nlat,nlon,nvar)
 * (concatenated over cells) replaced by ShSynthGau(ShFilter(ShAnalGau(u))), modes
 * n >= nlat0(i) dropped; with nvar = 3 on the coordinates this is FilterRbcs
 * (common/ModTimeInt.F90:908-934). */
int rbc3d_rbc_sph_project(int ncell, const int *nlat, const int *nlat0, int nvar, double *u);

/* --- time stepping of the context's cells (common/ModTimeInt.F90) ---------------
 * The cells' coordinates stay on the device between steps; rbc3d_set_cells gives the
 * initial state, rbc3d_get_rbc_coords reads it back in the same per-cell layout.  After
 * an entry that moves the cells (move, rebox, volume constraint, filter, inter-cell
 * repulsion), the next operator call (AddIntOnRbcs, the cell matvec and right-hand side,
 * the solves, the wall residual, CalcVelocityField, a PME spread of the cells) first
 * rebuilds the surface sources of the new coordinates: geometry, splines, source lists
 * and the f and g densities on the new dS. */
/* rbcRefs (common/ModTimeInt.F90:63-76, common/ModRbc.F90 RBC_ComputeGeometry): one
 * reference cell per cell type k (1..ntypes), its coordinates x(nlat(k), 2 nlat(k), 3)
 * concatenated over types, and the membrane moduli ES(k), ED(k), EB(k).  A cell's grid
 * must be that of its type's reference cell. */
int rbc3d_set_ref_cells(rbc3d_ctx *ctx, int ntypes, const int *nlat, const double *x,
                        const double *ES, const double *ED, const double *EB);
/* Compute_Rbc_Vel (common/ModTimeInt.F90:716-808): RBC_ComputeGeometry and the surface
 * sources of every cell, the force density -Shell_ResForce (plus the intra-cell repulsion
 * when rbc3d_set_repulsion turned it on) projected by RBC_SphProject, Solve_RBC_Vel.  The
 * node velocities v(3, np) (as rbc3d_solve_rbc_vel returns them) stay on the device for
 * rbc3d_move_rbcs; v may be NULL. */
int rbc3d_compute_rbc_vel(rbc3d_ctx *ctx, double *v, int *niter, double *rnorm);
/* rbc%x = rbc%x + Ts*rbc%v for every cell (common/ModTimeInt.F90:131-136), with the
 * velocities of the last rbc3d_compute_rbc_vel. */
int rbc3d_move_rbcs(rbc3d_ctx *ctx, double Ts);
/* ReboxRbcs (common/ModTimeInt.F90:1006-1020): each cell shifted by whole periods so the
 * centre of its coordinate range lies in [0, Lb). */
int rbc3d_rebox_rbcs(rbc3d_ctx *ctx);
/* VolConstrainRbcs (common/ModTimeInt.F90:938-963): x += fac a3 per cell, fac =
 * (vol_ref - vol)/area limited to eps_dist/20 in magnitude (double precision: the
 * reference's default REAL under its -freal-4-real-8 build). */
int rbc3d_vol_constrain_rbcs(rbc3d_ctx *ctx, double eps_dist);
/* FilterRbcs (common/ModTimeInt.F90:908-934): every cell's coordinates projected onto
 * spherical-harmonic modes below nlat0. */
int rbc3d_filter_rbcs(rbc3d_ctx *ctx);
/* The cells' coordinates, x(nlat, nlon, 3) per cell, concatenated (rbc3d_set_cells' layout). */
int rbc3d_get_rbc_coords(rbc3d_ctx *ctx, double *x);
/* The repulsion (common/ModRepulsion.F90, parameters ForceCoef, viscRatThresh and
 * rigidsep of common/ModConf.F90:52-59).  When on, rbc3d_compute_rbc_vel adds
 * AddIntraCellRepulsionForce (:126-266) to the membrane force density before its
 * projection, and rbc3d_timeint_euler runs InterCellRepulsion after every
 * VolConstrainRbcs.  Off by default. */
int rbc3d_set_repulsion(rbc3d_ctx *ctx, int on, double force_coef, double visc_rat_thresh,
                        int rigidsep);
/* AddIntraCellRepulsionForce on the cells' current coordinates: df(nlat, nlon, 3) per cell,
 * concatenated (may be NULL), and the number of nodes that receive a force. */
int rbc3d_intra_cell_repulsion(rbc3d_ctx *ctx, double *df, int *npoints);
/* InterCellRepulsion (common/ModRepulsion.F90:270-402, with Closest_Neighbor_Cell :480-546
 * and Closest_Neighbor_Wall :556-613) at epsDist = eps_dist: the cells' coordinates are
 * moved in place.  npoints: nodes displaced; dist_min: the smallest node-surface
 * separation below 2 eps_dist (infinity when none).  Optional outputs. */
int rbc3d_inter_cell_repulsion(rbc3d_ctx *ctx, double eps_dist, int *npoints, double *dist_min);
/* TimeInt_Euler (common/ModTimeInt.F90:109-178), nsteps steps of Ts: Compute_Rbc_Vel,
 * NoSlipWall when a wall is set (rtol, maxit as rbc3d_noslip_wall), x += Ts v,
 * ReboxRbcs, four rounds of VolConstrainRbcs, InterCellRepulsion (when the repulsion is
 * on, at eps_dist) and FilterRbcs; *time += Ts per step.  its_cell, its_wall (length
 * nsteps, optional): the GMRES iterations of each step's two solves.  In all three
 * integrators a solve that reaches its maxit does not stop the run (the reference prints
 * the residual and goes on); its count then equals maxit. */
int rbc3d_timeint_euler(rbc3d_ctx *ctx, double Ts, int nsteps, double eps_dist,
                        double noslip_rtol, int noslip_maxit, double *time, int *its_cell,
                        int *its_wall);
/* TimeInt_RK2(alpha) (common/ModTimeInt.F90:183-271), nsteps steps: k1 = Compute_Rbc_Vel +
 * NoSlipWall, the cells moved by alpha Ts k1 for k2 and moved back, x += Ts (b1 k1 + b2 k2)
 * with b2 = 1/(2 alpha), b1 = 1 - b2, ReboxRbcs.  its_cell, its_wall (optional): 2 nsteps
 * counts, the two evaluations of every step. */
int rbc3d_timeint_rk2(rbc3d_ctx *ctx, double alpha, double Ts, int nsteps, double noslip_rtol,
                      int noslip_maxit, double *time, int *its_cell, int *its_wall);
/* TimeInt_AB2 (common/ModTimeInt.F90:275-341): one Euler step from Compute_Rbc_Vel alone
 * (no NoSlipWall or rebox, time not advanced, as the reference), then nsteps - 1 steps of
 * x += Ts (3 v - v_prev)/2 after Compute_Rbc_Vel + NoSlipWall, ReboxRbcs, time + Ts.
 * its_cell, its_wall (optional): nsteps counts. */
int rbc3d_timeint_ab2(rbc3d_ctx *ctx, double Ts, int nsteps, double noslip_rtol, int noslip_maxit,
                      double *time, int *its_cell, int *its_wall);

/* Glob_Sph_Trans(v,c,direction) (common/ModVelSolver.F90:641-719); direction
 * 1 = PHYS_TO_FOUR, -1 = FOUR_TO_PHYS.  v(nPoint,3), c(3*sum nlat0^2). */
int rbc3d_glob_sph_trans(rbc3d_ctx *ctx, double *v, double *c, int direction);
/* spline tables of cell irbc (0-based): which 0 x, 1 a3, 2 detJ, 3 f*detJ, 4 g*detJ;
 * each u,u1,u2,u12 (2 nlat, nlon, nvar) as Spline_Build_on_Sphere leaves them */
int rbc3d_cell_spline(rbc3d_ctx *ctx, int irbc, int which, double *u, double *u1, double *u2,
                      double *u12);

/* --- composite operators --------------------------------------------------- */
/* CalcVelocityField(tlist,v) (common/ModPostProcess.F90:29-61) */
int rbc3d_calc_velocity_field(rbc3d_ctx *ctx, int nT, const double *tx, const int *tindx,
                              const int *active, const double *Acoef, double *v);
/* MyMatMult of the cell BIE (common/ModVelSolver.F90:523-601): b = A u, u and b of
 * length 3*sum(nlat0^2) in Glob_Sph_Trans packing.  The _dev form takes device
 * pointers and leaves the result on the device (stream-ordered). */
int rbc3d_cell_dof(rbc3d_ctx *ctx, int *dof);
int rbc3d_cell_matmult(rbc3d_ctx *ctx, const double *u, double *b);
int rbc3d_cell_matmult_dev(rbc3d_ctx *ctx, const double *d_u, double *d_b);
/* Compute_Rhs (common/ModVelSolver.F90:455-515) */
int rbc3d_cell_rhs(rbc3d_ctx *ctx, double *rhs);
/* Wall no-slip operator (common/ModNoSlip.F90): MyMatMult :255-308 on the
 * de-duplicated traction vector (3 per unique vertex), Compute_Wall_Residual_Vel
 * :153-193 on all wall vertices v(nvert,3). */
int rbc3d_wall_dof(rbc3d_ctx *ctx, int *dof);
int rbc3d_wall_matmult(rbc3d_ctx *ctx, const double *u, double *b);
int rbc3d_wall_residual_vel(rbc3d_ctx *ctx, double *v);
int rbc3d_get_wall_density(rbc3d_ctx *ctx, double *f);
/* Compute_Rhs into a device vector (common/ModVelSolver.F90:455-515) */
int rbc3d_cell_rhs_dev(rbc3d_ctx *ctx, double *d_rhs);

/* Geometry reuse (no counterpart in the reference, which recomputes everything in
 * each MyMatMult): the cell BIE matvec and right-hand side keep their geometry-only
 * intermediate results -- target binning, neighbouring-cell lists (NbrRbcList), the
 * closest-point projections of RBC_NearSingInt and its distance test -- until the
 * cells are set again, so the matvecs of one GMRES solve compute them once.  Results
 * are identical either way.  enable: 1 (default) or 0. */
int rbc3d_set_geometry_reuse(rbc3d_ctx *ctx, int enable);

/* Deflation of the cell BIE (the reference's Deflate switch, common/ModVelSolver.F90:
 * 103-106, 122-125 and 144-449; Pozrikidis): rbc3d_solve_rbc_vel builds each cell's six
 * solid-body modes after the right-hand side (DeflateGetSBModes), the right-hand side and
 * the matvec carry the deflation terms (DeflateRHS, DeflateMatMult) and the solution is
 * recovered from the deflated one (DeflateRecoverSol).  Each cell is addressed at its own
 * node offset (the reference assumes one grid for all cells).  on: 1 or 0 (default). */
int rbc3d_set_deflation(rbc3d_ctx *ctx, int on);
/* the solid-body modes of the last deflated solve, modes[(m * 3 + d) * nPoint + node] */
int rbc3d_sb_modes(rbc3d_ctx *ctx, double *modes);

/* Graph form of the cell BIE matvec (no counterpart in the reference): with on = 1 and
 * geometry reuse on, rbc3d_cell_matmult_dev records the matvec at the current geometry
 * once into a hipGraph (the first call of a geometry runs eagerly, then records) and
 * replays it in later calls, one launch per matvec instead of one per kernel.  A new
 * geometry (rbc3d_set_cells), precision form or grown work buffer records a new graph.
 * Results are identical to the eager matvec.  on = 0 (default) frees the graph. */
int rbc3d_set_graph(rbc3d_ctx *ctx, int on);
/* 1 when a recorded cell-matvec graph is ready to replay, else 0 */
int rbc3d_graph_ready(rbc3d_ctx *ctx, int *ready);

/* Precision of the hot path (no counterpart in the reference, which is FP64
 * throughout).  RBC3D_PREC_FP64, the default, is the plain form: every stage in FP64,
 * the reference's arithmetic.  RBC3D_PREC_SPLIT is the split form: real-space pair
 * sums in FP64 or FP32 (rbc3d_set_split_pairs); PME
 * spread, FFTs, Green's function and gather on FP32 grids; the Legendre GEMMs of the
 * global SH transforms as three-way split BF16 products on the matrix cores with
 * FP32 accumulation; the wall self-matrix stored in FP32.  The self-patch and
 * near-singular quadratures, the neighbour lists and projections, the periodic
 * linear term, the surface splines and the GMRES Arnoldi process stay FP64. */
enum { RBC3D_PREC_FP64 = 0, RBC3D_PREC_SPLIT = 1 };
int rbc3d_set_precision(rbc3d_ctx *ctx, int mode);
/* The real-space pair sums of the split form (no counterpart in the reference): FP64, the
 * default (fp32 = 0), or FP32 on target-bin-relative coordinates with FP32 accumulation
 * and the pairs closer than rc/4 in FP64 (fp32 = 1).  On gfx950 the pair kernel is bound
 * by instruction issue, and FP32 arithmetic without packed instructions issues at the FP64
 * rate, so the FP32 kernel is not faster and the split form keeps FP64 pairs unless asked. */
int rbc3d_set_split_pairs(rbc3d_ctx *ctx, int fp32);

/* --- the two GMRES solves (PETSc KSPGMRES semantics: restart 30, classical
 *     Gram-Schmidt, no preconditioner, x += V y at restarts) ------------------ */
/* Restart length of both solves (and of the learned warm start's), 1..255.  The
 * reference's solves keep PETSc's default, 30 (KSPSetType(KSPGMRES) with no
 * KSPGMRESSetRestart, common/ModVelSolver.F90:82, common/ModNoSlip.F90:79); this is the
 * KSPGMRESSetRestart call a maintainer would add there.  The Krylov basis takes
 * (restart + 1) x ndof doubles of HBM. */
#define RBC3D_GMRES_MAX_RESTART 255
int rbc3d_set_gmres_restart(rbc3d_ctx *ctx, int restart);
/* NoSlipWall (common/ModNoSlip.F90:45-150): residual wall velocity, GMRES for the
 * traction increment (rtol relative to the initial residual, maxit), f = f0 + finc.
 * finc0: initial guess of the increment on the de-duplicated unknowns (NULL: zero,
 * the reference's VecZeroEntries at :117).  vres_max (optional): max |v| of the
 * residual velocity after the update (the monitor of :140-145). */
int rbc3d_noslip_wall(rbc3d_ctx *ctx, double rtol, int maxit, const double *finc0, int *niter,
                      double *rnorm, int *converged, double *vres_max);
/* Residual history of the last NoSlipWall solve (no counterpart in the reference,
 * which prints only the final niter and residual, common/ModNoSlip.F90:120-126):
 * the initial residual norm, then the GMRES estimate after every iteration.  *n is
 * the length of the history; at most maxn values are copied into h. */
int rbc3d_gmres_history(rbc3d_ctx *ctx, double *h, int maxn, int *n);
/* Solve_RBC_Vel (common/ModVelSolver.F90:44-132): v(nPoint,3) cell velocities; with
 * all viscosity ratios 1 the solution is the right-hand side (no GMRES); otherwise
 * GMRES to 1e-11 (maxit 200) from the previous solution.  The cells' double-layer
 * density is then set to v (common/ModTimeInt.F90:773-790).  v may be NULL. */
int rbc3d_solve_rbc_vel(rbc3d_ctx *ctx, double *v, int *niter, double *rnorm);

/* --- the learned warm start of NoSlipWall (no counterpart in the reference) -----
 * A per-wall-vertex MLP, 9 -> 64 -> 64 -> 64 -> 3 with GELU, BF16 weights and
 * activations, FP32 accumulation.  Inputs per vertex: the cells' velocity there
 * (rbc3d_wall_cell_velocity), the previous traction and the vertex normal, each
 * standardised (x - mean) / std; the output times scale_out is the predicted traction
 * increment.  Parameters: an FP32 vector of rbc3d_mlp_nparams() values,
 * W1[64][9] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3] (row-major).
 * Training: mean squared error of the scaled increments, Adam (beta 0.9, 0.999,
 * eps 1e-8).  Device arrays: x(n,9) and t(n,3) row-major FP32. */
typedef struct rbc3d_mlp rbc3d_mlp;
int rbc3d_mlp_create(rbc3d_mlp **out);
int rbc3d_mlp_destroy(rbc3d_mlp *m);
int rbc3d_mlp_nparams(void);
/* The HIP stream the device-pointer entries (_dev) run on from now on (NULL names
 * the null stream); until this is called they run on the network's own stream.  A
 * caller whose device arrays are written on a stream of its own passes that stream,
 * so the network's kernels are ordered after that work. */
int rbc3d_mlp_set_stream(rbc3d_mlp *m, void *hip_stream);
int rbc3d_mlp_set_params(rbc3d_mlp *m, const float *params);
int rbc3d_mlp_get_params(rbc3d_mlp *m, float *params);
int rbc3d_mlp_set_norm(rbc3d_mlp *m, const float *mean_in, const float *std_in,
                       const float *scale_out);
/* y(n,3) = the predicted increments */
int rbc3d_mlp_forward_dev(rbc3d_mlp *m, int n, const float *d_x, float *d_y);
/* gradient of the loss w.r.t. the parameters into d_grad (may be NULL), and the loss */
int rbc3d_mlp_grad_dev(rbc3d_mlp *m, int n, const float *d_x, const float *d_t, float *d_grad,
                       double *loss);
/* one Adam step on the batch; loss (may be NULL) is the batch loss before the step */
int rbc3d_mlp_adam_step_dev(rbc3d_mlp *m, int n, const float *d_x, const float *d_t, float lr,
                            double *loss);
/* The cells' part of Compute_Wall_Residual_Vel (common/ModNoSlip.F90:153-193) at the
 * wall vertices, v(nvert,3), and the area-weighted vertex normals n(nvert,3) (the
 * reference keeps element normals only, common/ModWall.F90:134-138). */
int rbc3d_wall_cell_velocity(rbc3d_ctx *ctx, double *v);
int rbc3d_wall_vertex_normals(rbc3d_ctx *ctx, double *n);
/* NoSlipWall (common/ModNoSlip.F90:45-150) with the network's prediction as GMRES's
 * initial guess (tolerance rtol * ||b||, as for a zero guess); plain GMRES from zero
 * is the fallback (*fallback = 1) when the prediction's residual is not below the
 * right-hand side's or the warm-started solve does not converge.  *niter counts the
 * iterations of both solves. */
int rbc3d_noslip_wall_learned(rbc3d_ctx *ctx, rbc3d_mlp *m, double rtol, int maxit, int *niter,
                              double *rnorm, int *converged, double *vres_max, int *fallback);

/* --- diagnostics (no counterpart in the reference, which only prints a per-step
 *     MPI_WTime, common/ModTimeInt.F90:122,167) ---------------------------------
 * Per-stage GPU time from HIP events recorded on the context's stream around each
 * kernel stage.  enable: 1 on, 0 off, -1 reset and off.  Stages are numbered
 * 0..rbc3d_stage_time() returns RBC3D_EINVAL past the last one. */
int rbc3d_timing(rbc3d_ctx *ctx, int enable);
int rbc3d_stage_time(rbc3d_ctx *ctx, int stage, double *ms, long long *count, const char **name);
/* --- timing hooks (HIP events on the context's stream) --------------------- */
int rbc3d_stream(rbc3d_ctx *ctx, void **hip_stream);
int rbc3d_sync(rbc3d_ctx *ctx);

/* --- files (common/ModIO.F90), host only ---------------------------------------- */
/* ReadWallMesh (:547-605): an Exodus II Tri3 wall mesh (netCDF, classic, 64-bit-offset or
 * 64-bit-data encoding).  rbc3d_wall_mesh_info gives num_nodes and num_elem; then
 * rbc3d_read_wall_mesh fills x(nvert,3) from coordx/coordy/coordz and e2v(nele,3),
 * 1-based, from connect1 (what rbc3d_set_wall takes).  A mesh that is not Tri3, has more
 * than one element block or names nodes outside 1..num_nodes is RBC3D_EIO. */
int rbc3d_wall_mesh_info(const char *fn, int *nvert, int *nele);
int rbc3d_read_wall_mesh(const char *fn, double *x, int *e2v);

/* The restart file of WriteRestart (:742-786) and ReadRestart (:792-921), byte for byte:
 * Fortran sequential unformatted records (a 4-byte length before and after each), 4-byte
 * integers, 8-byte reals.  Records: Lb(3); lt; time; vBkg(3); nrbc; per cell (nlat0,
 * nlon0), (nlat, nlon), celltype, starting_area, x(nlat,nlon,3); nwall; per wall (nvert,
 * nele), x(nvert,3), f(nvert,3), e2v(nele,3).  Arrays are concatenated over cells and over
 * walls.  npoint, nvert, nele of the info are totals (sum of nlat*nlon, of nvert, of nele). */
typedef struct {
  double Lb[3];
  int lt; /* time step (the reader's Nt0) */
  double time;
  double vBkg[3];
  int nrbc, nwall;
  long long npoint, nvert, nele;
} rbc3d_restart_info;
int rbc3d_restart_write(const char *fn, const rbc3d_restart_info *info, const int *nlat0,
                        const int *nlon0, const int *nlat, const int *nlon, const int *celltype,
                        const double *starting_area, const double *x, const int *nvert,
                        const int *nele, const double *wx, const double *wf, const int *e2v);
/* The header and counts, to size the arrays of rbc3d_restart_read; that call checks the
 * file still has the counts given in *info, then fills *info and the arrays.  Both refuse
 * (RBC3D_EIO), as ReadRestart does (common/ModIO.F90:858-875), a cell whose grid
 * RBC_Create would not re-create from its nlat0: nlat must be nlat0 + 2 when
 * nlat/nlat0 < 1.5, else k nlat0 with k = nlat / nlat0 (integer quotient) > 1, and
 * nlon = 2 nlat. */
int rbc3d_restart_info_read(const char *fn, rbc3d_restart_info *info);
int rbc3d_restart_read(const char *fn, rbc3d_restart_info *info, int *nlat0, int *nlon0, int *nlat,
                       int *nlon, int *celltype, double *starting_area, double *x, int *nvert,
                       int *nele, double *wx, double *wf, int *e2v);

#ifdef __cplusplus
}
#endif
#endif
