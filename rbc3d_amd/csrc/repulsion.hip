// Repulsion of the cells (common/ModRepulsion.F90) on the device:
//   AddIntraCellRepulsionForce (:126-266): a force along the outward normal at nodes whose
//     own surface folds back within 2 sqrt(area)/nlat of them, added to the membrane force
//     density before its projection (Compute_Rbc_Vel, common/ModTimeInt.F90:755-756);
//   InterCellRepulsion (:270-402) with Closest_Neighbor_Cell (:480-546) and
//     Closest_Neighbor_Wall (:556-613): nodes closer than epsDist to another cell's surface
//     or to the wall are moved half the shortfall away from it.
//
// The reference walks its cell list (27 bins of width >= rc around each node) and keeps the
// minimum distance.  Every candidate that can change the result lies within 2 epsDist of
// the node (cells) or epsDist (wall); here the candidates are every node of a cell whose
// bounding sphere comes that close (and the wall triangles whose centroid sphere does), a
// superset of what matters, so the minima agree.  Exact ties between candidates are broken
// by (distance, cell, node) or (distance, triangle) instead of the bin-walk order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "surf_core.h"

#define REP_T 256

// AddIntraCellRepulsionForce for one cell per workgroup, on the per-cell (nlat, nlon, 3)
// blocks of the mechanics buffers.  For node i, over the nodes j of the same cell:
// skipped if DistOnSphere(i, j) < patch radius, |xi - xj| > epsDist or
// -(xi - xj).a3j / |xi - xj| < 0.707; dist = min |xi - xj| over the rest; if
// dist < epsDist, df_i = ForceCoef exp(1 - dist/epsDist) a3i.  epsDist = 2 sqrt(area)/nlat.
// mode 0: df written to out; mode 1: f := -f + df in place (rbc%f = -rbc%f, then the
// repulsion added, common/ModTimeInt.F90:752,756).
__global__ void __launch_bounds__(REP_T) k_intra_repulsion(int nrbc, const CellRec *__restrict__ cells,
                                                           const double *const *__restrict__ th_t,
                                                           const double *const *__restrict__ phi_t,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ a3,
                                                           const double *__restrict__ area,
                                                           double force_coef, int mode, double *out,
                                                           int *count) {
  const int c = blockIdx.x;
  if (c >= nrbc) return;
  const CellRec cr = cells[c];
  const int nl = cr.nlat, npt = nl * cr.nlon;
  const double *th = th_t[cr.gtype], *phi = phi_t[cr.gtype];
  const double *X = x + 3 * (size_t)cr.off, *A = a3 + 3 * (size_t)cr.off;
  double *F = out + 3 * (size_t)cr.off;
  const double eps = 2 * sqrt(area[c]) / nl, epsRef = cr.radius;
  int hits = 0;
  for (int i = threadIdx.x; i < npt; i += REP_T) {
    const double xi0 = X[i], xi1 = X[i + npt], xi2 = X[i + 2 * npt];
    const double thi = th[i % nl], phii = phi[i / nl];
    double dist = HUGE_VAL;
    for (int j = 0; j < npt; j++) {
#pragma clang fp contract(off)
      const double xx0 = xi0 - X[j], xx1 = xi1 - X[j + npt], xx2 = xi2 - X[j + 2 * npt];
      const double rr = sqrt(xx0 * xx0 + xx1 * xx1 + xx2 * xx2);
      if (rr > eps) continue;
      if (dist_on_sphere(thi, phii, th[j % nl], phi[j / nl]) < epsRef) continue;
      const double xxn = -(xx0 * A[j] + xx1 * A[j + npt] + xx2 * A[j + 2 * npt]);
      if (xxn / rr < 0.707) continue;
      dist = fmin(dist, rr);
    }
    double df[3] = {0., 0., 0.};
    if (dist < eps) {
      const double FF = force_coef * exp(1 - dist / eps);
      for (int d = 0; d < 3; d++) df[d] = FF * A[i + d * npt];
      hits++;
    }
    for (int d = 0; d < 3; d++) {
      double *p = F + i + d * npt;
      *p = mode == 0 ? df[d] : -*p + df[d];
    }
  }
  if (count && hits) atomicAdd(count, hits);
}

// per cell: the mean of its nodes and the largest distance of a node from it
__global__ void __launch_bounds__(REP_T) k_rep_sphere(int nrbc, const CellRec *__restrict__ cells,
                                                      int np, const double *__restrict__ x,
                                                      double *sph) {
  __shared__ double red[4][REP_T / 64];
  const int c = blockIdx.x;
  if (c >= nrbc) return;
  const CellRec cr = cells[c];
  const int npt = cr.nlat * cr.nlon, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double s[3] = {0., 0., 0.};
  for (int q = threadIdx.x; q < npt; q += REP_T)
    for (int d = 0; d < 3; d++) s[d] += x[cr.off + q + (size_t)np * d];
  for (int d = 0; d < 3; d++) {
    for (int o = 32; o > 0; o >>= 1) s[d] += __shfl_xor(s[d], o, 64);
    if (lane == 0) red[d][w] = s[d];
  }
  __syncthreads();
  double ctr[3];
  for (int d = 0; d < 3; d++) {
    double t = 0;
    for (int k = 0; k < REP_T / 64; k++) t += red[d][k];
    ctr[d] = t / npt;
  }
  double r2 = 0;
  for (int q = threadIdx.x; q < npt; q += REP_T) {
    double t = 0;
    for (int d = 0; d < 3; d++) {
      const double u = x[cr.off + q + (size_t)np * d] - ctr[d];
      t += u * u;
    }
    r2 = fmax(r2, t);
  }
  for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
  __syncthreads();
  if (lane == 0) red[3][w] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int k = 0; k < REP_T / 64; k++) m = fmax(m, red[3][k]);
    for (int d = 0; d < 3; d++) sph[4 * c + d] = ctr[d];
    sph[4 * c + 3] = sqrt(m);
  }
}

__device__ inline double min_image_norm(const double a[3], const double b[3], const double Lb[3],
                                        const double iLb[3]) {
  double s = 0;
  for (int d = 0; d < 3; d++) {
    double t = a[d] - b[d];
    t -= round(t * iLb[d]) * Lb[d];
    s += t * t;
  }
  return sqrt(s);
}

struct RepArgs {
  double Lb[3], iLb[3];
  double eps;
  int nrbc, np, nele, nvert;
  int cap_c, cap_e;  // candidate list capacities per cell
  const CellRec *cells;
  const int *node_cell;
  const double *x, *sph, *spl;
  const double *const *th, *const *phi, *const *proj;
  const double *wx, *wcen;
  const int *e2v;
  double wrad;  // largest triangle centroid-to-vertex distance
  int *cand_c, *cnt_c, *cand_e, *cnt_e, *overflow;
  double *dx;
  int *moved;
  unsigned long long *dmin;  // bits of the smallest separation found (positive doubles order as integers)
};

// candidate neighbours per cell: cells whose bounding spheres come within 2 epsDist, wall
// triangles whose centroid sphere comes within epsDist of the cell's sphere (with a
// relative margin of 1e-12 for the rounding of the sphere radii)
__global__ void __launch_bounds__(REP_T) k_rep_candidates(RepArgs a) {
  __shared__ int nc, ne;
  const int c = blockIdx.x;
  if (c >= a.nrbc) return;
  if (threadIdx.x == 0) nc = ne = 0;
  __syncthreads();
  const double *sc = a.sph + 4 * c;
  for (int j = threadIdx.x; j < a.nrbc; j += REP_T) {
    if (j == c) continue;
    const double *sj = a.sph + 4 * j;
    if (min_image_norm(sc, sj, a.Lb, a.iLb) <= (sc[3] + sj[3] + 2 * a.eps) * (1 + 1e-12)) {
      const int k = atomicAdd(&nc, 1);
      if (k < a.cap_c) a.cand_c[(size_t)c * a.cap_c + k] = j;
    }
  }
  for (int e = threadIdx.x; e < a.nele; e += REP_T) {
    const double ce[3] = {a.wcen[e], a.wcen[e + a.nele], a.wcen[e + 2 * (size_t)a.nele]};
    if (min_image_norm(sc, ce, a.Lb, a.iLb) <= (sc[3] + a.wrad + a.eps) * (1 + 1e-12)) {
      const int k = atomicAdd(&ne, 1);
      if (k < a.cap_e) a.cand_e[(size_t)c * a.cap_e + k] = e;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.cnt_c[c] = nc;
    a.cnt_e[c] = ne;
    if (nc > a.cap_c || ne > a.cap_e) atomicMax(a.overflow, max(nc, ne));
  }
}

// InterCellRepulsion's displacement of every node: Closest_Neighbor_Cell (nearest node of a
// candidate cell, then Spline_FindProjection on that cell's x spline) and
// Closest_Neighbor_Wall (MinDistToTri over candidate triangles); the closer of the two, if
// nearer than epsDist, gives dx = 0.5 xx (epsDist - |xx|)/|xx|, xx the minimum image of
// xi - xj.  One lane per node.
__global__ void __launch_bounds__(REP_T) k_inter_repulsion(RepArgs a) {
  const int p = blockIdx.x * REP_T + threadIdx.x;
  if (p >= a.np) return;
  const size_t np = a.np;
  const int c = a.node_cell[p];
  const double xi[3] = {a.x[p], a.x[p + np], a.x[p + 2 * np]};
  // nearest node of the candidate cells
  double best = HUGE_VAL;
  int bj = -1, bq = -1;
  const int ncand = min(a.cnt_c[c], a.cap_c);
  for (int k = 0; k < ncand; k++) {
    const int j = a.cand_c[(size_t)c * a.cap_c + k];
    const CellRec cj = a.cells[j];
    const int npt = cj.nlat * cj.nlon;
    for (int q = 0; q < npt; q++) {
      const size_t o = cj.off + q;
      double s = 0;
      {
#pragma clang fp contract(off)
        double xx[3];
        for (int d = 0; d < 3; d++) {
          xx[d] = xi[d] - a.x[o + np * d];
          xx[d] = xx[d] - round(xx[d] * a.iLb[d]) * a.Lb[d];
        }
        s = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      }
      if (s < best || (s == best && (j < bj || (j == bj && q < bq)))) {
        best = s;
        bj = j;
        bq = q;
      }
    }
  }
  double dist1 = best, x1[3] = {0., 0., 0.};
  if (bj >= 0 && best <= 2 * a.eps) {
    const CellRec cj = a.cells[bj];
    const int gt = cj.gtype, m = 2 * cj.nlat, n = cj.nlon;
    double th0 = a.th[gt][bq % cj.nlat], phi0 = a.phi[gt][bq / cj.nlat], x0[3], xt[3];
    for (int d = 0; d < 3; d++) x0[d] = a.x[cj.off + bq + np * d];
    {
#pragma clang fp contract(off)
      for (int d = 0; d < 3; d++) {
        double xx = xi[d] - x0[d];
        xx = xx - round(xx * a.iLb[d]) * a.Lb[d];
        xt[d] = x0[d] + xx;
      }
    }
    const double *sxp = a.spl + cj.spl + (size_t)m * n * 4 * SPL_X;
    spline_find_projection(sxp, m, n, a.proj[gt], xt, th0, phi0, x0);
    {
#pragma clang fp contract(off)
      double xx[3];
      for (int d = 0; d < 3; d++) xx[d] = xt[d] - x0[d];
      dist1 = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    }
    for (int d = 0; d < 3; d++) x1[d] = x0[d];
  }
  // nearest point of the candidate wall triangles
  double dist2 = HUGE_VAL, x2[3] = {0., 0., 0.};
  int be = -1;
  const int nce = min(a.cnt_e[c], a.cap_e);
  for (int k = 0; k < nce; k++) {
    const int e = a.cand_e[(size_t)c * a.cap_e + k];
    double xe[3][3], xt[3], s0, t0;
    for (int l = 0; l < 3; l++) {
      const int v = a.e2v[e + (size_t)a.nele * l] - 1;
      for (int d = 0; d < 3; d++) xe[l][d] = a.wx[v + (size_t)a.nvert * d];
    }
    {
#pragma clang fp contract(off)
      for (int d = 0; d < 3; d++) {
        double xx = xi[d] - xe[0][d];
        xx = xx - round(xx * a.iLb[d]) * a.Lb[d];
        xt[d] = xe[0][d] + xx;
      }
    }
    const double rr = min_dist_to_tri(xt, xe, s0, t0);
    if (rr < dist2 || (rr == dist2 && e < be)) {
#pragma clang fp contract(off)
      dist2 = rr;
      be = e;
      for (int d = 0; d < 3; d++) x2[d] = (1 - s0 - t0) * xe[0][d] + s0 * xe[1][d] + t0 * xe[2][d];
    }
  }
  const bool cell = dist1 < dist2;
  const double dist = cell ? dist1 : dist2;
  const double *xj = cell ? x1 : x2;
  double dx[3] = {0., 0., 0.};
  if (dist < a.eps) {
#pragma clang fp contract(off)
    double xx[3];
    for (int d = 0; d < 3; d++) {
      xx[d] = xi[d] - xj[d];
      xx[d] = xx[d] - round(xx[d] * a.iLb[d]) * a.Lb[d];
    }
    const double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    for (int d = 0; d < 3; d++) dx[d] = 0.5 * xx[d] * (a.eps - rr) / rr;
    atomicAdd(a.moved, 1);
  }
  if (dist <= 2 * a.eps) atomicMin(a.dmin, (unsigned long long)__double_as_longlong(dist));
  for (int d = 0; d < 3; d++) a.dx[p + np * d] = dx[d];
}

// apply the displacements (common/ModRepulsion.F90:367-396): cells with |viscRat| below
// viscRatThresh move node by node; with rigidsep the others move rigidly by the mean
// displacement of their displaced nodes
__global__ void __launch_bounds__(REP_T) k_apply_repulsion(int nrbc, const CellRec *__restrict__ cells,
                                                           int np, const double *__restrict__ visc,
                                                           double thresh, int rigidsep,
                                                           const double *__restrict__ dx, double *x) {
  __shared__ double red[4][REP_T / 64];
  const int c = blockIdx.x;
  if (c >= nrbc) return;
  const CellRec cr = cells[c];
  const int npt = cr.nlat * cr.nlon;
  if (fabs(visc[cr.celltype - 1]) < thresh) {
    for (int q = threadIdx.x; q < npt; q += REP_T)
      for (int d = 0; d < 3; d++) {
        const size_t o = cr.off + q + (size_t)np * d;
        x[o] = x[o] + dx[o];
      }
    return;
  }
  if (!rigidsep) return;
  double s[4] = {0., 0., 0., 0.};
  for (int q = threadIdx.x; q < npt; q += REP_T) {
    const double u0 = dx[cr.off + q], u1 = dx[cr.off + q + (size_t)np], u2 = dx[cr.off + q + 2 * (size_t)np];
    if (u0 * u0 + u1 * u1 + u2 * u2 > 0.) {
      s[0] += u0;
      s[1] += u1;
      s[2] += u2;
      s[3] += 1.;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < 4; k++) {
    for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
    if (lane == 0) red[k][w] = s[k];
  }
  __syncthreads();
  double t[4];
  for (int k = 0; k < 4; k++) {
    t[k] = 0;
    for (int v = 0; v < REP_T / 64; v++) t[k] += red[k][v];
  }
  if (t[3] == 0.) return;
  for (int q = threadIdx.x; q < npt; q += REP_T)
    for (int d = 0; d < 3; d++) {
      const size_t o = cr.off + q + (size_t)np * d;
      x[o] = x[o] + t[d] / t[3];
    }
}

static int require_cells(rbc3d_ctx *c) {
  if (!c) return RBC3D_EINVAL;
  if (c->nrbc == 0) {
    rbc_set_error("no cells set");
    return RBC3D_ESTATE;
  }
  return RBC3D_OK;
}

// the intra-cell repulsion on the block geometry of the last rbc_cells_geometry: mode 0
// writes df into out (blocks), mode 1 turns out = f (blocks) into -f + df
int rbc_intra_repulsion(rbc3d_ctx *c, int mode, double *out, int *count) {
  TypePtrs tp;
  RBC_TRY(rbc_type_ptrs(c, &tp));
  k_intra_repulsion<<<c->nrbc, REP_T, 0, c->stream>>>(c->nrbc, c->d_cells.p, tp.th, tp.phi,
                                                      c->d_mx.p, c->d_ma3.p, c->d_marea.p,
                                                      c->force_coef, mode, out, count);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_set_repulsion(rbc3d_ctx *c, int on, double force_coef, double visc_rat_thresh,
                                   int rigidsep) {
  if (!c || !std::isfinite(force_coef) || !(visc_rat_thresh >= 0)) return RBC3D_EINVAL;
  c->rep_on = on != 0;
  c->force_coef = force_coef;
  c->visc_rat_thresh = visc_rat_thresh;
  c->rigidsep = rigidsep != 0;
  return RBC3D_OK;
}

extern "C" int rbc3d_intra_cell_repulsion(rbc3d_ctx *c, double *df, int *npoints) {
  RBC_TRY(require_cells(c));
  RBC_TRY(rbc_cells_geometry(c, false));
  DBuf<double> d_df;
  DBuf<int> d_cnt;
  RBC_TRY(d_df.ensure(3 * (size_t)c->np));
  RBC_TRY(d_cnt.ensure(1));
  HIP_TRY(hipMemsetAsync(d_cnt.p, 0, sizeof(int), c->stream));
  RBC_TRY(rbc_intra_repulsion(c, 0, d_df.p, d_cnt.p));
  if (df)
    HIP_TRY(hipMemcpyAsync(df, d_df.p, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, c->stream));
  int n = 0;
  HIP_TRY(hipMemcpyAsync(&n, d_cnt.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (npoints) *npoints = n;
  return RBC3D_OK;
}

extern "C" int rbc3d_inter_cell_repulsion(rbc3d_ctx *c, double eps_dist, int *npoints,
                                          double *dist_min) {
  RBC_TRY(require_cells(c));
  if (!(eps_dist > 0) || !std::isfinite(eps_dist)) return RBC3D_EINVAL;
  hipStream_t st = c->stream;
  const int nrbc = c->nrbc, np = c->np;
  RBC_TRY(c->d_rep_sph.ensure(4 * (size_t)nrbc));
  RBC_TRY(c->d_rep_dx.ensure(3 * (size_t)np));
  RBC_TRY(c->d_rep_stat.ensure(2));
  RBC_TRY(c->d_rep_cnt.ensure(2 * (size_t)nrbc + 2));
  TypePtrs tp;
  RBC_TRY(rbc_type_ptrs(c, &tp));
  k_rep_sphere<<<nrbc, REP_T, 0, st>>>(nrbc, c->d_cells.p, np, c->d_x.p, c->d_rep_sph.p);
  HIP_TRY(hipGetLastError());
  RepArgs a;
  for (int d = 0; d < 3; d++) {
    a.Lb[d] = c->Lb[d];
    a.iLb[d] = c->iLb[d];
  }
  a.eps = eps_dist;
  a.nrbc = nrbc;
  a.np = np;
  WallData &w = c->wall;
  a.nele = w.nvert ? w.nele : 0;
  a.nvert = w.nvert;
  a.cells = c->d_cells.p;
  a.node_cell = c->d_node_cell.p;
  a.x = c->d_x.p;
  a.sph = c->d_rep_sph.p;
  a.spl = c->d_spl.p;
  a.th = tp.th;
  a.phi = tp.phi;
  a.proj = tp.proj;
  a.wx = w.d_x.p;
  a.wcen = w.d_cen.p;
  a.e2v = w.d_e2v.p;
  a.wrad = w.rmax;
  a.cnt_c = c->d_rep_cnt.p;
  a.cnt_e = c->d_rep_cnt.p + nrbc;
  a.overflow = c->d_rep_cnt.p + 2 * nrbc;
  a.dx = c->d_rep_dx.p;
  a.moved = (int *)c->d_rep_stat.p;
  a.dmin = (unsigned long long *)(c->d_rep_stat.p + 1);
  // candidate lists, grown until no cell overflows its capacity
  for (;;) {
    a.cap_c = c->rep_cap_c;
    a.cap_e = c->rep_cap_e;
    RBC_TRY(c->d_rep_cand_c.ensure((size_t)nrbc * a.cap_c));
    RBC_TRY(c->d_rep_cand_e.ensure((size_t)nrbc * a.cap_e + 1));
    a.cand_c = c->d_rep_cand_c.p;
    a.cand_e = c->d_rep_cand_e.p;
    HIP_TRY(hipMemsetAsync(a.overflow, 0, sizeof(int), st));
    k_rep_candidates<<<nrbc, REP_T, 0, st>>>(a);
    HIP_TRY(hipGetLastError());
    int ovf = 0;
    HIP_TRY(hipMemcpyAsync(&ovf, a.overflow, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (!ovf) break;
    c->rep_cap_c = std::max(c->rep_cap_c, std::min(ovf, nrbc));
    c->rep_cap_e = std::max(c->rep_cap_e, ovf);
  }
  const unsigned long long inf_bits = 0x7ff0000000000000ull;
  HIP_TRY(hipMemsetAsync(a.moved, 0, sizeof(int), st));
  HIP_TRY(hipMemcpyAsync(a.dmin, &inf_bits, sizeof(inf_bits), hipMemcpyHostToDevice, st));
  k_inter_repulsion<<<(np + REP_T - 1) / REP_T, REP_T, 0, st>>>(a);
  HIP_TRY(hipGetLastError());
  RBC_TRY(c->d_visc.ensure(8));
  HIP_TRY(hipMemcpyAsync(c->d_visc.p, c->conf.viscRat, sizeof(double) * c->conf.ntypes,
                         hipMemcpyHostToDevice, st));
  k_apply_repulsion<<<nrbc, REP_T, 0, st>>>(nrbc, c->d_cells.p, np, c->d_visc.p, c->visc_rat_thresh,
                                            c->rigidsep, c->d_rep_dx.p, c->d_x.p);
  HIP_TRY(hipGetLastError());
  int moved = 0;
  unsigned long long dm = 0;
  HIP_TRY(hipMemcpyAsync(&moved, a.moved, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(&dm, a.dmin, sizeof(dm), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (npoints) *npoints = moved;
  if (dist_min) {
    double v;
    memcpy(&v, &dm, sizeof(v));
    *dist_min = v;
  }
  c->rbc_geom_ready = false;
  c->x_moved = true;
  return RBC3D_OK;
}
