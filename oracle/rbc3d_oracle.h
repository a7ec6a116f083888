/*
 * rbc3d_oracle.h -- CPU FP64 restatement of RBC3D's Ewald-split boundary-integral
 * matrix-vector product.  TEST INFRASTRUCTURE ONLY: this library is the parity
 * checker and the host-CPU baseline; the product path (rbc3d_amd/, include/rbc3d.h)
 * never links, loads or calls it.
 *
 * Array conventions follow the reference's Fortran layouts (column-major, stored
 * 0-based here):
 *   cell field  x(nlat,nlon,3)        -> x[ilat + nlat*(ilon + nlon*d)]
 *   list field  x(Np,3)               -> x[p + Np*d]
 *   list index  indx(Np,0:2)          -> indx[p + Np*c]  (values stay 1-based)
 *   hash heads  hoc(0:Nc1+1,0:Nc2+1,0:Nc3+1), next(Np): 1-based, -1 sentinel
 *   PME grids   ff(0:N1-1,0:N2-1,0:N3-1,f)
 */
#ifndef RBC3D_ORACLE_H
#define RBC3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NTAB 8192
#define ORC_MAXTYPES 8

typedef struct {
  double Lb[3];
  double alpha;     /* reference's Ewald parameter (length^2): erfc(sqrt(pi/alpha) r) */
  double rc;        /* real-space cut-off */
  int Nb[3];        /* PME grid */
  int P;            /* B-spline order (PBspln) */
  int ntypes;
  double viscRat[ORC_MAXTYPES];
  double vBkg[3];
} orc_conf;

typedef struct orc_sys orc_sys;

/* lifetime */
orc_sys *orc_create(const orc_conf *conf);
void orc_destroy(orc_sys *s);
int orc_set_threads(int n);

/* cells: nrbc cells, per-cell nlat0/nlat/celltype (1-based); nlon = 2*nlat.
 * x, a3, detj concatenated per cell in (nlat,nlon,3) / (nlat,nlon) order. */
int orc_set_cells(orc_sys *s, int nrbc, const int *nlat0, const int *nlat, const int *celltype,
                  const double *x, const double *a3, const double *detj);
/* which = 1: single-layer density f, 2: double-layer density g (concatenated (nlat,nlon,3)) */
int orc_set_cell_density(orc_sys *s, int which, const double *dens);
/* one wall (ModWall): x(nvert,3), e2v(nele,3) 1-based, f(nvert,3) */
int orc_set_wall(orc_sys *s, int nvert, int nele, const double *x, const int *e2v);
int orc_set_wall_density(orc_sys *s, const double *f);
/* GMRES restart length of the NoSlipWall and Solve_RBC_Vel solves, 1..255 (30) */
int orc_set_gmres_restart(orc_sys *s, int m);
int orc_prepare_sing_int_on_wall(orc_sys *s);

/* target-list helpers: the reference's tlist_rbc / tlist_wall */
int orc_ntargets_rbc(const orc_sys *s);
int orc_ntargets_wall(const orc_sys *s);
int orc_tlist_rbc(const orc_sys *s, double *tx, int *tindx, double *Acoef);
int orc_tlist_wall(const orc_sys *s, double *tx, int *tindx, double *Acoef);

/* hot-path routines (v accumulated in place, v(nT,3)) */
int orc_add_int_on_rbcs(orc_sys *s, double c1, double c2, int nT, const double *tx,
                        const int *tindx, const int *active, const double *Acoef, double *v,
                        long long *npairs_out);
int orc_add_int_on_walls(orc_sys *s, double c1, int nT, const double *tx, const int *tindx,
                         const int *active, const double *Acoef, double *v);
int orc_pme_distrib_source(orc_sys *s, double c1, double c2, int cells, int walls);
int orc_pme_transform(orc_sys *s);
int orc_pme_add_interp_vel(orc_sys *s, int nT, const double *tx, const int *active,
                           const double *Acoef, double *v);
int orc_calc_velocity_field(orc_sys *s, int nT, const double *tx, const int *tindx,
                            const int *active, const double *Acoef, double *v);
int orc_glob_sph_trans(orc_sys *s, double *v, double *c, int direction);
int orc_cell_dof(const orc_sys *s);
int orc_cell_matmult(orc_sys *s, const double *u, double *b);
/* the same on the nodes of the cells with cell_active[irbc] != 0 only (other cells'
 * coefficient blocks of b are not the reference's) */
int orc_cell_matmult_sample(orc_sys *s, const double *u, double *b, const int *cell_active);
int orc_cell_rhs(orc_sys *s, double *rhs);
int orc_wall_dof(const orc_sys *s);
int orc_wall_matmult(orc_sys *s, const double *u, double *b);
int orc_wall_residual_vel(orc_sys *s, double *v);

/* pieces, exposed for pinning */
void orc_ewald_sl(const orc_sys *s, double r, double *A, double *B);
void orc_ewald_dl(const orc_sys *s, double r, double *A);
double orc_mask(const orc_sys *s, double x);
void orc_bspline(double xc, int P, int *imin, double *w);
void orc_gauleg(double x1, double x2, int n, double *x, double *w);
void orc_gauleg_sinh(double xmin, double xmax, double a, double b, int n, double *x, double *w);
void orc_gaqd(int nlat, double *th, double *wts);
void orc_pbar_table(int nlat, const double *th, int mmax, double *pbar);
int orc_hashtable_build(const int Nc[3], const double iLbNc[3], int nPoint, const double *x,
                        int ldx, int *hoc, int *next);
int orc_sh_anal_gau(int nlat, int nlon, int nt, const double *g, int idg, int jdg, double *a,
                    double *b, int ida, int jda);
int orc_sh_synth_gau(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                     const double *b, int ida, int jda);
int orc_sh_synth_equ(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                     const double *b, int ida, int jda);
/* ShGradGau: v = dg/dth, w = dg/dphi on the Gauss grid from the shags coefficients */
int orc_sh_grad_gau(int nlat, int nlon, int nt, double *v, double *w, int idv, int jdv,
                    const double *a, const double *b, int ida, int jda);
void orc_fft_diff(int m, int n, int nvar, const double *u, double *u1, double *u2);
double orc_min_dist_to_tri(const double *xtar, const double *x /* x(3,3) */, double *s0, double *t0);
void orc_tri_int_regular(const orc_sys *s, const double *x, const double *f, const double *xtar,
                         double *rhs);
void orc_tri_int_duffy(const orc_sys *s, const double *x, const double *f, const double *xtar,
                       double s0, double t0, double *rhs);
int orc_wall_self_nnz(const orc_sys *s);
int orc_wall_self_csr(const orc_sys *s, int *rowptr, int *col, double *val);
int orc_pme_grid(const orc_sys *s, int which, double *out);
void orc_fft3d_forward(int N1, int N2, int N3, const double *in, double *out_cplx);
void orc_fft3d_inverse(int N1, int N2, int N3, const double *in_cplx, double *out);
int orc_cell_spline(const orc_sys *s, int irbc, int which, double *u, double *u1, double *u2,
                    double *u12);
unsigned long long orc_pair_checksum(const orc_sys *s);
int orc_noslip_wall(orc_sys *s, double rtol, int maxit, const double *finc0, int *niter,
                    double *rnorm, double *vres_max);
int orc_get_wall_density(const orc_sys *s, double *f);
int orc_solve_rbc_vel(orc_sys *s, double *sol, double *v, int *niter, double *rnorm);
int orc_near_list(orc_sys *s, int nT, const double *tx, const int *tindx, int i, int *cells,
                  int *ilat, int *ilon, double *dist, int cap);

/* test hooks: the parts of AddIntOnRbcs evaluated (all by default), a cell's surface
 * splines (which: 0 x, 1 a3, 2 detJ, 3 f detJ, 4 g detJ) at (th, phi), the closest-point
 * projection onto cell irbc (0-based) starting from (th0, phi0), the bicubic Hermite
 * interpolation of given tables on the doubly periodic [0,2pi)^2 grid (m x n), and a
 * cell's mesh size */
#define ORC_TERM_PAIR 1
#define ORC_TERM_SELF 2
#define ORC_TERM_NEAR 4
#define ORC_TERM_LINEAR 8
#define ORC_TERM_ALL 15
int orc_set_terms(orc_sys *s, int terms);
int orc_spline_eval(const orc_sys *s, int irbc, int which, double th, double phi, double *out);
int orc_find_projection(const orc_sys *s, int irbc, const double *xtar, double *th0,
                        double *phi0, double *x0);
void orc_spline_interp_raw(int m, int n, int nvar, const double *u, const double *u1,
                           const double *u2, const double *u12, double x, double y, double *f);
double orc_cell_mesh_size(const orc_sys *s, int irbc);

/* Per-cell surface geometry and membrane mechanics (mech_oracle.c), one cell on its
 * nlat x 2 nlat Gauss grid in the reference's layouts: RBC_ComputeGeometry
 * (common/ModRbc.F90:404-509), Shell_ResForce with Shell_ElasStrs and Shell_Bend
 * (:815-969), RBC_SphProject (:659-686). */
int orc_rbc_geometry(int nlat, const double *x, double *a1, double *a2, double *a3, double *a,
                     double *a_rcp, double *b, double *detj, double *xc, double *area,
                     double *vol, double *mesh);
int orc_shell_res_force(int nlat, const double *a1, const double *a2, const double *a3,
                        const double *a, const double *a_rcp, const double *b, const double *detj,
                        double area, const double *a_rcp_ref, const double *b_ref, double ES,
                        double ED, double EB, double *f);
int orc_rbc_sph_project(int nlat, int nlat0, int nvar, double *u);
/* the reference's Deflate switch of the cell BIE (common/ModVelSolver.F90:103-106,
 * 144-449): DeflateGetSBModes in Solve_RBC_Vel, DeflateRHS in Compute_Rhs,
 * DeflateMatMult in MyMatMult, DeflateRecoverSol after the solve */
int orc_set_deflation(orc_sys *s, int on);
/* the solid-body modes of the last deflated solve: modes[(m * 3 + d) * nP + node] */
int orc_sb_modes(const orc_sys *s, double *modes);

#ifdef __cplusplus
}
#endif
#endif
