// Internal declarations of the gfx950 implementation of RBC3D's Ewald-split
// boundary-integral matvec.  The plain form is FP64 throughout; the split form
// (rbc3d_set_precision) swaps the pair sums, PME, SH Legendre GEMMs and the wall
// self-matrix for FP32 / split-BF16 kernels over the same data.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/rbc3d.h"

#define RBC_NTAB 8192
#define RBC_NBR_MAX 32
// closest-point projection tables per grid type: least-squares weights W[3][17][6]
// then local patch directions [3][16][3]
#define RBC_PJ_W 0
#define RBC_PJ_DIR 306
#define RBC_PJ_SIZE 450
#define RBC_PI 3.14159265358979323846
#define RBC_TWO_PI (2 * RBC_PI)

void rbc_set_error(const std::string &msg);

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess) {                                                              \
      rbc_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
      return RBC3D_EHIP;                                                                 \
    }                                                                                    \
  } while (0)

#define RBC_TRY(expr)            \
  do {                           \
    int _r = (expr);             \
    if (_r != RBC3D_OK) return _r; \
  } while (0)

// Count of device-buffer (re)allocations in the process: a captured graph holds buffer
// addresses, so it is valid only while this count is unchanged.
extern unsigned long long g_dbuf_allocs;

// Device buffer with explicit lifetime (hipMalloc / hipFree); grows on demand.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  int ensure(size_t count) {
    if (count <= n && p) return RBC3D_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return RBC3D_OK;
    HIP_TRY(hipMalloc(&p, count * sizeof(T)));
    n = count;
    g_dbuf_allocs++;
    return RBC3D_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Tables shared by every cell with the same grid (nlat, nlon = 2 nlat).
struct GridType {
  int nlat, nlon, mmax;  // mmax = min(nlat, nlon/2 + 1) (SPHEREPACK Gauss grids)
  double radius;         // polar patch radius pi/sqrt(nlat)
  int nrad, nazm;
  std::vector<double> th, phi, w;  // w includes 2 pi / nlon (RBC_Create)
  DBuf<double> d_th, d_phi, d_w;
  DBuf<double> d_pw;     // analysis  [m][n][i] = (2/nlon) wgauss_i Pbar(m,n,th_i)
  DBuf<double> d_pws;    // analysis of the spline build [m][n][i] = Pbar(m,n,th_i) wgauss_i
  DBuf<double> d_pg;     // synthesis [m][n][i] = Pbar(m,n,th_i) (x 1/2 for m = 0)
  DBuf<double> d_pe;     // equi synthesis, nlat+1 rows incl. poles, [m][n][ie]
  DBuf<double> d_cs;     // cos(m phi_j) [m][j] then sin(m phi_j) [m][j], m <= nlat
  DBuf<double> d_Dm;     // periodic derivative matrix, length 2 nlat, [p][q]
  DBuf<double> d_Dn;     // periodic derivative matrix, length nlon, [p][q]
  DBuf<double> d_thG, d_phiG, d_pwt;  // polar patch: [node][irad + nrad*iazm], weights
  DBuf<double> d_mask;   // same-surface mask [ilat_i][ilat_j][dlon]
  DBuf<float> d_maskf;   // the same in FP32
  DBuf<double> d_proj;   // projection tables (RBC_PJ_*)
  // split form: the Legendre tables as three BF16 terms, [3][mmax][nlat][nlat] each
  DBuf<unsigned short> d_pw3, d_pg3;
};

// GMRES work space (Krylov basis, partial dot products), kept across solves
struct GmresWork {
  DBuf<double> V, w, part, part2, hd;
  // vectors of the solves around the GMRES (previous traction, right-hand side,
  // solution, the learned warm start's inputs and residual images), kept across calls
  DBuf<double> f0, rhs, x, r0, uc, LV, tmp;
  DBuf<double> ir_r, ir_d;  // iterative refinement of the split form: FP64 residual, correction
  DBuf<float> X, Y;
  std::vector<double> history;  // residual estimates of the last solve
  // pinned host slots for the Hessenberg columns of two Arnoldi steps in flight and
  // the update coefficients, with an event per column slot (created on first use)
  double *hpin = nullptr;
  // coherent pinned slots the Arnoldi step's last kernel writes its Hessenberg column
  // to (host and device addresses)
  double *hcoh = nullptr, *hcoh_dev = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

// Stage timing (HIP events on the context's stream, enabled on request)
enum {
  ST_TSORT = 0, ST_PAIR, ST_NBR, ST_SELF, ST_PROJ, ST_NEAR, ST_LINEAR, ST_COMBINE, ST_SH_GLOB,
  ST_SPLINE, ST_SPREAD_PREP, ST_SPREAD, ST_FFT, ST_GREEN, ST_GATHER, ST_WALL, ST_NSTAGE
};
struct StageTimer {
  bool on = false;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  double ms[ST_NSTAGE] = {0};
  long long cnt[ST_NSTAGE] = {0};
};

// Per-cell record (device array of these).
struct CellRec {
  int off;      // first node in the concatenated node arrays
  int gtype;    // GridType index
  int nlat, nlon, nlat0;
  int celltype;  // 1-based
  int coff;      // offset of the cell's block in the Glob_Sph_Trans vector
  int choff;     // first 64-node chunk of the cell (chunk bounding spheres)
  long long spl;  // offset (doubles) of the cell's spline block
  double Acoef, Bcoef, area, meshSize, radius;
};

// Spline block of a cell: 13 variables (x:3, a3:3, detJ:1, f detJ:3, g detJ:3), each
// (2 nlat) x nlon points of 4 doubles (u, u1, u2, u12), point (i,j) at i + m*j.
enum { SPL_X = 0, SPL_A3 = 3, SPL_DETJ = 6, SPL_F = 7, SPL_G = 10, SPL_NVAR = 13 };

struct WallData {
  int nvert = 0, nele = 0, id = 0, ndof = 0;
  std::vector<double> x;
  std::vector<int> e2v, v2v, glb;
  std::vector<double> vnorm;  // area-weighted vertex normals (nvert,3)
  double rmax = 0;            // largest triangle centroid-to-vertex distance
  DBuf<double> d_x, d_f, d_area, d_eps, d_cen, d_vnorm;
  DBuf<int> d_e2v, d_glb;
  // cell list of wall sources (triangle centroids)
  DBuf<int> d_bin_start, d_order;
  // targets of AddIntOnWalls sorted by the same coarse bins, and their 64-target chunks
  DBuf<int> d_tg_bin, d_tg_order, d_tg_start, d_tg_chunk;
  // self matrix CSR (rows i + k nvert)
  DBuf<int> d_rowptr, d_col;
  DBuf<double> d_val;   // symmetric 3x3 blocks: B00, B11, B22, B01, B02, B12 arrays
  DBuf<float> d_valf;  // split form: the same in FP32
  long long nnz = 0;
  bool prepared = false, valf_ready = false;
  // wall-only spread: per 4^3 brick the sources reaching it and their brick-local
  // stencil origins (CSR), per source its B-spline weights (fixed for a wall and
  // grid), and the per-source values of the current spread
  DBuf<int> d_wsb_start, d_wsb_src, d_wsb_rel;
  DBuf<double> d_wsb_w, d_wsb_f;
  bool wsp_ready = false;
  // the gather at the vertices: per vertex its stencil origin and B-spline weights
  DBuf<int> d_wg_org;
  DBuf<double> d_wg_w;
  DBuf<float> d_wg_wf;  // the same weights in FP32 (the split form's gather)
  bool wgt_ready = false;
  // the operator on the de-duplicated vector: self-matrix columns and triangle
  // vertices as unique indices, the targets' A (2 at every vertex)
  DBuf<int> d_colu, d_e2u;
  DBuf<double> d_A2;
  bool op_ready = false;
  // images of the three uniform tractions under the no-slip operator (learned warm
  // start), for the precision mode they were computed in
  DBuf<double> d_lmodes;
  int lmodes_prec = -1;
  bool lmodes_ready = false;
};

struct rbc3d_ctx {
  rbc3d_conf conf;
  double Lb[3], iLb[3];
  double Acoef[8], Bcoef[8];
  hipStream_t stream = nullptr;
  int prec = RBC3D_PREC_FP64;  // rbc3d_set_precision
  bool split_pair32 = false;   // rbc3d_set_split_pairs: FP32 pair sums in the split form
  int gmres_restart = 30;      // rbc3d_set_gmres_restart (PETSc's default length)
  int ncu = 256;               // compute units of the device
  int lds_max = 65536;         // LDS bytes one workgroup may allocate on the device
  // tables
  DBuf<double> d_sl1, d_sl2, d_dl1, d_mtab;
  DBuf<float> d_tabf;  // split form: sl1, sl2, dl1 in FP32, (RBC_NTAB + 1) each
  double r_eps;
  // cell list geometry
  int Nc[3];
  double iLbNc[3];
  int nbins;
  // finer grid of the pair kernel (width >= rc/2) and its pruned source-bin offsets
  int Nf[3];
  // every dimension has at least 2R+3 fine bins (R: bins within rc): the visited
  // source bins then determine the periodic image of every in-cutoff pair
  int pair_img = 0;
  double iLbNf[3];
  int nfbins;
  size_t noffs = 0;
  DBuf<int> d_offs, d_tmp_f, d_fbin_start;
  // cells
  int nrbc = 0, np = 0;
  std::vector<GridType *> gtypes;
  std::vector<CellRec> cells;
  DBuf<CellRec> d_cells;
  DBuf<double> d_x, d_a3, d_detj, d_ds, d_f, d_g, d_fs, d_gs;  // node arrays, [d][np]
  DBuf<int> d_node_cell, d_node_ll;  // cell index; ilat | ilon << 16
  DBuf<double> d_spl;
  long long spl_total = 0;
  // source cell list (sorted by bin, ascending index inside a bin)
  DBuf<int> d_src_bin, d_src_order, d_bin_start, d_tmp_i;
  DBuf<double> d_src_rec;  // sorted source records
  DBuf<unsigned char> d_cub_tmp;
  // target scratch
  DBuf<double> d_tx, d_tA, d_v, d_vacc;
  DBuf<int> d_tindx, d_tact, d_tbin, d_torder, d_tbin_start, d_chunk_off;
  DBuf<int> d_nbr_cnt, d_nbr_cell, d_nbr_node;
  DBuf<int> d_cell_cand;  // per cell: candidate neighbouring cells, then their counts
  DBuf<double> d_nbr_dist;
  DBuf<int> d_work, d_work_cnt;
  DBuf<double> d_dvitem, d_scratch;
  DBuf<int> d_flag;
  DBuf<unsigned long long> d_pairstat;
  // cell bounding spheres binned by centre (neighbour-cell candidates)
  DBuf<double> d_cell_sph;
  DBuf<double> d_chunk_sph;  // per 64-node chunk of every cell: centre(3), radius
  DBuf<int> d_src_cc;        // coarse cell-list coordinates per node (packed)
  DBuf<double> d_shx;        // SH Legendre GEMM operands (m, row, col) x 2
  int nchunks = 0;
  DBuf<int> d_cellbin_start, d_cellbin_order;
  double max_cell_radius = 0;
  // PME
  bool pme_ready = false;
  int flag_sl = 0, flag_dl = 0;
  int pme_grid_prec = RBC3D_PREC_FP64;  // scalar of the grids the last transform wrote
  DBuf<double> d_grid;   // real grids: up to 12 fields (ff 3 | tt 9), later vv 3
  DBuf<double> d_cgrid;  // complex grids (interleaved) up to 12 fields
  DBuf<double> d_vgrid;  // velocity grid 3 fields
  DBuf<double> d_bb;     // B-factor (kx,ky,kz)
  DBuf<double> d_tw;     // FFT twiddles for the three lengths
  DBuf<float> d_twf;     // the same in FP32 (split form)
  int tw_off[3];
  DBuf<int> d_pme_key, d_pme_order, d_pme_start;
  DBuf<double> d_pme_src;  // sorted spread records
  DBuf<double> d_spread_tiles;  // per-brick tiles of the two-pass single-layer spread
  // SH scratch
  DBuf<double> d_shbuf, d_shbuf2;
  // wall
  WallData wall;
  // per-grid-type table pointers for the kernels
  DBuf<const double *> d_tp_ptr;
  DBuf<const float *> d_tp_maskf;  // FP32 mask table per grid type
  DBuf<int> d_tp_int;
  size_t tp_ntypes = 0;
  StageTimer timer;
  // persistent cell-BIE solution (KSPSetInitialGuessNonzero, common/ModVelSolver.F90:87)
  DBuf<double> d_cell_sol;
  bool cell_sol_init = false;
  // last pair statistics
  long long last_npairs = 0, last_nitems = 0, last_nwork = 0;
  // geometry epoch (advanced by set_cells) and the epoch of the geometry-only results
  // kept by the last cacheable add_int_on_rbcs (-1: none)
  int geom_epoch = 0, cache_epoch = -1, cache_nT = 0, cache_nitems = 0, cache_nwork = 0;
  bool geom_reuse = true;  // rbc3d_set_geometry_reuse
  DBuf<int> d_worklist, d_ikey, d_ival, d_islot;
  GmresWork gmres;
  DBuf<double> d_linpart;
  unsigned long long last_checksum = 0;
  // the PME spread's sort of the cell sources by brick depends on the geometry only:
  // kept for the geometry epoch it was made at (with geometry reuse on)
  int spread_sort_epoch = -1, spread_sort_ns = 0;
  // hipGraph of the cell matvec at fixed geometry (rbc3d_set_graph): captured on the
  // matvec's own input and output buffers, valid for one geometry epoch, precision
  // form and buffer-allocation count
  bool use_graph = false;
  bool capturing = false;  // a capture is being recorded: no host synchronisation
  hipGraph_t mv_graph = nullptr;
  hipGraphExec_t mv_exec = nullptr;
  int mv_epoch = -1, mv_prec = -1;
  unsigned long long mv_allocs = 0;
  DBuf<double> d_gin, d_gout;
  // deflation of the cell BIE (rbc3d_set_deflation): the solid-body modes of the last
  // Solve_RBC_Vel, [6][3][np]
  bool deflate = false, sbq_ready = false;
  DBuf<double> d_sbq, d_visc;
  // time stepping on the device (timeint.hip, mech.hip): reference cells per type
  // (rbcRefs: a_rcp, b of their grid, volume), membrane moduli, and the cells' geometry
  // in per-cell blocks (x, a1, a2, a3, a, a_rcp, b, detJ, ds; xc, area, vol, mesh)
  int nref = 0;
  std::vector<int> ref_nlat, ref_off;
  std::vector<double> ref_vol, ES, ED, EB;
  DBuf<double> d_ref_ar, d_ref_b;
  DBuf<double> d_mx, d_ma1, d_ma2, d_ma3, d_ma, d_mar, d_mb, d_mdetj, d_mscr, d_mf;
  DBuf<double> d_mxc, d_marea, d_mvol, d_mmesh, d_mmod;
  DBuf<int> d_moff, d_mnlat, d_mnlat0, d_mref;
  DBuf<double> d_vel;  // the cells' node velocities of the last Compute_Rbc_Vel, [3][np]
  DBuf<double> d_vel_k1;  // an earlier velocity of the multi-stage integrators (RK2 k1, AB2 v_1)
  bool vel_ready = false;
  int mech_epoch = -1;  // geom_epoch the per-cell tables above were built for
  bool rbc_geom_ready = false;  // the block geometry above is that of d_x
  // d_x moved by a cell update (move, rebox, volume constraint, filter, repulsion) since
  // the surface sources were last built: the operator entries rebuild them first
  bool x_moved = false;
  // repulsion (rbc3d_set_repulsion, repulsion.hip): ForceCoef of the intra-cell force,
  // viscRatThresh and rigidsep of the inter-cell displacement; per-cell bounding spheres,
  // candidate neighbour cells and wall triangles (capacities grown on overflow), the
  // displacement of every node, and (moved count, smallest separation)
  bool rep_on = false, rigidsep = false;
  double force_coef = 0, visc_rat_thresh = 0;
  int rep_cap_c = 32, rep_cap_e = 256;
  DBuf<double> d_rep_sph, d_rep_dx, d_rep_stat;
  DBuf<int> d_rep_cnt, d_rep_cand_c, d_rep_cand_e;
};

// ---------------------------------------------------------------------------
// Device helpers shared by several kernels.

// Reference distance, evaluated without FMA contraction so the in-cutoff test
// matches the reference bit for bit: xx = xj - xi; xx -= nint(xx/Lb) Lb;
// rr = sqrt((x1*x1 + x2*x2) + x3*x3)   (common/ModIntOnRbcs.F90:68-74)
__device__ __forceinline__ double ref_min_image_dist(const double xi[3], const double xj[3],
                                                     const double Lb[3], const double iLb[3],
                                                     double xx[3]) {
#pragma clang fp contract(off)
  for (int d = 0; d < 3; d++) {
    double t = xj[d] - xi[d];
    double sh = round(t * iLb[d]) * Lb[d];
    xx[d] = t - sh;
  }
  double s = xx[0] * xx[0] + xx[1] * xx[1];
  s = s + xx[2] * xx[2];
  return sqrt(s);
}

__device__ __forceinline__ int imod_d(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}
// imod_d for indices that usually lie within one period of [0, n) (stencil points of a
// point near the box): compares and one add or subtract, the division only beyond that
__device__ __forceinline__ int wrap_idx(int a, int n) {
  if ((unsigned)a < (unsigned)n) return a;
  const int b = a < 0 ? a + n : a - n;
  return (unsigned)b < (unsigned)n ? b : imod_d(a, n);
}

// XCD-aware block order: workgroups are dealt to the 8 XCDs round-robin
// (blockIdx % 8), so logical block (blockIdx % 8) * per + blockIdx / 8 gives each XCD
// one contiguous range of the logical order and keeps that range's data in one L2.
// Launch 8 * per blocks, per = ceil(nblk / 8); returns -1 for the padding blocks.
#define RBC_NXCD 8
__device__ __forceinline__ int xcd_block(int nblk) {
  const int per = (nblk + RBC_NXCD - 1) / RBC_NXCD;
  const int b = (int)(blockIdx.x % RBC_NXCD) * per + (int)(blockIdx.x / RBC_NXCD);
  return b < nblk ? b : -1;
}
inline unsigned xcd_grid(size_t nblk) {
  return (unsigned)(RBC_NXCD * ((nblk + RBC_NXCD - 1) / RBC_NXCD));
}

// 1/x from the hardware reciprocal estimate and two Newton steps (x normal, finite)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) from the hardware estimate and two Newton steps (x > 0 normal)
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// EwaldCoeff_SL / EwaldCoeff_DL (common/ModEwaldFunc.F90:86-178) by table interpolation.
// Ewald coefficients at r with 1/r supplied (r >= reps already established)
__device__ __forceinline__ double ewald_dl_ir(const double *__restrict__ t, double r, double ir,
                                              double tabscale) {
  double s = r * tabscale;
  double fi = floor(s);
  if (fi >= RBC_NTAB) return 0.;
  int i = (int)fi;
  double c1 = t[i] * (i + 1 - s) + t[i + 1] * (s - i);
  double ir2 = ir * ir;
  return c1 * (ir2 * ir2 * ir);
}
__device__ __forceinline__ void ewald_sl_ir(const double *__restrict__ t1,
                                            const double *__restrict__ t2, double r, double ir,
                                            double tabscale, double &A, double &B) {
  double s = r * tabscale;
  double fi = floor(s);
  if (fi >= RBC_NTAB) {
    A = 0;
    B = 0;
    return;
  }
  int i = (int)fi;
  double c1 = t1[i] * (i + 1 - s) + t1[i + 1] * (s - i);
  double c2 = t2[i] * (i + 1 - s) + t2[i + 1] * (s - i);
  double ir2 = ir * ir;
  A = c1 * ir * ir2 + c2 * ir2;
  B = c1 * ir - c2;
}

// The table position is r * (N / rc): the factor is loop-invariant for the callers'
// loops, and the interpolated value is continuous in it, so this differs from the
// reference's N * r / rc only at the level of rounding.
__device__ __forceinline__ void ewald_sl(const double *__restrict__ t1, const double *__restrict__ t2,
                                         double r, double reps, double rc, double &A, double &B) {
  if (r < reps) {
    A = 0;
    B = 0;
    return;
  }
  double s = r * (RBC_NTAB / rc);
  double fi = floor(s);
  if (fi >= RBC_NTAB) {
    A = 0;
    B = 0;
    return;
  }
  int i = (int)fi;
  double c1 = t1[i] * (i + 1 - s) + t1[i + 1] * (s - i);
  double c2 = t2[i] * (i + 1 - s) + t2[i + 1] * (s - i);
  double ir = rcp_nr(r), ir2 = ir * ir;
  A = c1 * ir * ir2 + c2 * ir2;
  B = c1 * ir - c2;
}

__device__ __forceinline__ double ewald_dl(const double *__restrict__ t, double r, double reps,
                                           double rc) {
  if (r < reps) return 0.;
  double s = r * (RBC_NTAB / rc);
  double fi = floor(s);
  if (fi >= RBC_NTAB) return 0.;
  int i = (int)fi;
  double c1 = t[i] * (i + 1 - s) + t[i + 1] * (s - i);
  double ir = rcp_nr(r), ir2 = ir * ir;
  return c1 * (ir2 * ir2 * ir);
}

// MaskFunc (common/ModBasicMath.F90:351-379)
__device__ __forceinline__ double mask_fn(const double *__restrict__ mt, double x) {
  double s = fabs(x) * RBC_NTAB;
  double fi = floor(s);
  if (fi >= RBC_NTAB) return 0.;
  int i = (int)fi;
  return mt[i] * (i + 1 - s) + mt[i + 1] * (s - i);
}

// Division of a 32-bit unsigned n by a divisor fixed at launch, by multiply-high and
// shift (Granlund and Montgomery, "Division by invariant integers using
// multiplication", 1994, round-up method): q = (umulhi(n, m) + n) >> s with
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1; exact for every n < 2^32.
struct FastDiv {
  unsigned d, m, s;
};
inline FastDiv make_fastdiv(unsigned d) {
  FastDiv f;
  f.d = d;
  f.s = 0;
  while ((1ull << f.s) < d) f.s++;
  f.m = (unsigned)(((1ull << 32) * ((1ull << f.s) - d)) / d + 1);
  return f;
}
__host__ __device__ __forceinline__ unsigned fd_div(unsigned n, const FastDiv &f) {
  unsigned t = (unsigned)(((unsigned long long)n * f.m) >> 32);  // multiply-high
  return (unsigned)(((unsigned long long)t + n) >> f.s);
}
// the same for n < 2^31 in 32-bit arithmetic: the multiply-high is below n, so the sum
// does not overflow
__device__ __forceinline__ unsigned fd_div32(unsigned n, const FastDiv &f) {
  return (__umulhi(n, f.m) + n) >> f.s;
}

// Spline_Interp (common/ModSpline.F90:150-191) on a cell's spline block.
// blk points at variable v0; nvar consecutive variables are interpolated.
__device__ __forceinline__ void spline_interp(const double *__restrict__ blk, int m, int n,
                                              double x, double y, int nvar, double *f) {
  // evaluated without FMA contraction, operation for operation as the reference's
  // compiled Fortran: the double-layer near-singular correction amplifies ulp-level
  // differences of these values about a thousandfold
#pragma clang fp contract(off)
  double hx = RBC_TWO_PI / m, hy = RBC_TWO_PI / n;
  double ihx = 1. / hx, ihy = 1. / hy;
  double fi1 = floor(x * ihx), fj1 = floor(y * ihy);
  double s = x * ihx - fi1, t = y * ihy - fj1;
  int i1 = imod_d((int)fi1, m), j1 = imod_d((int)fj1, n);
  int i2 = i1 + 1 == m ? 0 : i1 + 1, j2 = j1 + 1 == n ? 0 : j1 + 1;
  double cx0 = 1 + s * s * (-3. + 2. * s), cx1 = s * s * (3. - 2. * s);
  double cx2 = hx * s * (1. + s * (-2. + s)), cx3 = hx * s * s * (-1. + s);
  double cy0 = 1 + t * t * (-3. + 2. * t), cy1 = t * t * (3. - 2. * t);
  double cy2 = hy * t * (1. + t * (-2. + t)), cy3 = hy * t * t * (-1. + t);
  size_t stride = (size_t)m * n * 4;
  size_t p11 = 4 * (size_t)(i1 + m * j1), p12 = 4 * (size_t)(i1 + m * j2);
  size_t p21 = 4 * (size_t)(i2 + m * j1), p22 = 4 * (size_t)(i2 + m * j2);
#pragma nounroll
  for (int l = 0; l < nvar; l++) {
    const double *b = blk + l * stride;
    double4 a11 = *(const double4 *)(b + p11), a12 = *(const double4 *)(b + p12);
    double4 a21 = *(const double4 *)(b + p21), a22 = *(const double4 *)(b + p22);
    // rows of u(4,4): (u, u, u2, u2)_(i1), (.)_(i2), (u1, u1, u12, u12)_(i1), (.)_(i2)
    double r0 = a11.x * cy0 + a12.x * cy1 + a11.z * cy2 + a12.z * cy3;
    double r1 = a21.x * cy0 + a22.x * cy1 + a21.z * cy2 + a22.z * cy3;
    double r2 = a11.y * cy0 + a12.y * cy1 + a11.w * cy2 + a12.w * cy3;
    double r3 = a21.y * cy0 + a22.y * cy1 + a21.w * cy2 + a22.w * cy3;
    f[l] = cx0 * r0 + cx1 * r1 + cx2 * r2 + cx3 * r3;
  }
}

// DistOnSphere (common/ModPolarPatch.F90:249-257)
__device__ __forceinline__ double dist_on_sphere(double th0, double phi0, double th1, double phi1) {
  double d = cos(th0 - th1) - sin(th0) * sin(th1) * (1. - cos(phi0 - phi1));
  d = fmin(1., fmax(-1., d));
  return acos(d);
}

void rbc_stage_begin(rbc3d_ctx *c, int stage);
void rbc_stage_end(rbc3d_ctx *c, int stage);
void rbc_stage_flush(rbc3d_ctx *c);

// Gauss colatitudes and weights (gaqd), normalised associated Legendre table
// [m][n][i] on the given colatitudes (sh.hip)
void host_gauss(int n, std::vector<double> &th, std::vector<double> &w);
void host_pbar(int nl, int mmax, const std::vector<double> &th, std::vector<double> &out);

// Kernel launchers (defined in the .hip files)
int rbc_sh_build_splines(rbc3d_ctx *c, int which);
int rbc_sh_glob(rbc3d_ctx *c, const double *d_in, double *d_out, int direction);
int rbc_sh_anal_batch(int nlat, int nlon, int nt, const double *d_g, int idg, int jdg, double *d_a,
                      double *d_b, int ida, int jda, hipStream_t st);
int rbc_sh_synth_batch(int nlat, int nlon, int nt, double *d_g, int idg, int jdg, const double *d_a,
                       const double *d_b, int ida, int jda, int equi, hipStream_t st);
int rbc_build_source_list(rbc3d_ctx *c);
// device arrays of per-grid-type table pointers, indexed by CellRec::gtype (pair.hip):
// same-surface mask, grid colatitudes and longitudes, patch points and weights, projection
// tables (patch directions and least-squares weights)
struct TypePtrs {
  const double *const *mask, *const *th, *const *phi, *const *thG, *const *phiG, *const *pwt,
      *const *proj;
  const float *const *maskf;
  const int *nrad, *nazm;
};
int rbc_type_ptrs(rbc3d_ctx *c, TypePtrs *tp);
int rbc_update_density(rbc3d_ctx *c, int which);
// per-cell mechanics on the context's cells (mech.hip): RBC_ComputeGeometry of every
// cell from d_x (refreshing a3, detJ, dS, the cell records, splines and source lists),
// and the surface force density -Shell_ResForce, projected (RBC_SphProject), into d_f
int rbc_cells_geometry(rbc3d_ctx *c, bool sources);
int rbc_cells_force(rbc3d_ctx *c);
int rbc_cells_project_x(rbc3d_ctx *c);
// the surface sources of moved cells rebuilt (geometry, splines, source lists, both
// densities on the new dS); nothing when the cells have not moved (timeint.hip)
int rbc_refresh_moved(rbc3d_ctx *c);
// AddIntraCellRepulsionForce on the block geometry (repulsion.hip): mode 0 writes df to
// out, mode 1 turns the block force out = f into -f + df; count (optional, device) is
// increased by the number of nodes that received a force
int rbc_intra_repulsion(rbc3d_ctx *c, int mode, double *out, int *count);
// deflation corrections (gmres.hip); no-ops unless rbc3d_set_deflation is on
int rbc_defl_rhs(rbc3d_ctx *c, double *d_v);
int rbc_defl_matmult(rbc3d_ctx *c, const double *d_g, double *d_v);
int rbc_add_int_on_rbcs_dev(rbc3d_ctx *c, double c1, double c2, int nT, const double *d_tx,
                            const int *d_tindx, const int *d_act, const double *d_A, double *d_v,
                            bool cacheable);
int rbc_pme_init(rbc3d_ctx *c);
int rbc_pme_spread(rbc3d_ctx *c, double c1, double c2, int cells, int walls);
int rbc_pme_transform(rbc3d_ctx *c);
// the gather at the wall's vertices (targets as set by the wall target list: all active)
int rbc_pme_gather_wall(rbc3d_ctx *c, const double *d_A, double *d_v);
// the same for the de-duplicated output: b(3 k + d) = v(i, d) + pme(i, d) / A_i for the
// vertex i = last[k] of each unique index k < nu
int rbc_pme_gather_wall_out(rbc3d_ctx *c, const double *d_A, const double *d_v, const int *d_last,
                            int nu, double *d_b);
// the wall-only single-layer spread with the sources' tractions read from the
// de-duplicated vector u through the triangles' unique vertex indices e2u(nele, 3)
int rbc_pme_spread_wall_u(rbc3d_ctx *c, double c1, const double *d_u, const int *d_e2u);
// whether the wall-only spread kernels apply to this grid
bool rbc_wall_spread_applies(const rbc3d_ctx *c);
int rbc_pme_gather(rbc3d_ctx *c, int nT, const double *d_tx, const int *d_act, const double *d_A,
                   double *d_v);
int rbc_wall_setup(rbc3d_ctx *c);
int rbc_wall_prepare(rbc3d_ctx *c);
int rbc_add_int_on_walls_dev(rbc3d_ctx *c, double c1, int nT, const double *d_tx, const int *d_tindx,
                             const int *d_act, const double *d_A, double *d_v, bool wall_targets);
typedef std::function<int(const double *, double *)> MatvecFn;
// Restarted GMRES (PETSc KSPGMRES semantics).  matvec_lo (optional): a cheaper, inexact
// operator that Arnoldi steps use once the residual estimate is below lo_below times
// the initial residual.
int rbc_gmres(rbc3d_ctx *c, int n, const MatvecFn &matvec, const double *d_b, double *d_x,
              bool x0_nonzero, double rtol, double atol, double dtol, int maxit, int restart,
              int *niter, double *rnorm, int *reason, std::vector<double> *history,
              const MatvecFn *matvec_lo = nullptr, double lo_below = 0.0);
int rbc_hash_sort(rbc3d_ctx *c, int n, const double *d_x, int ldx, int *d_bin, int *d_order,
                  int *d_bin_start);
