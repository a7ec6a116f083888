// Time stepping of the cells on the device (common/ModTimeInt.F90): Compute_Rbc_Vel,
// the cell update x += Ts v, ReboxRbcs, VolConstrainRbcs, FilterRbcs and the Euler loop
// TimeInt_Euler that strings them together with NoSlipWall.  Positions, geometry, forces
// and velocities stay in HBM; the host sees the per-cell areas and volumes only.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

// d_x += dt * d_vel (rbc%x = rbc%x + Ts*rbc%v, common/ModTimeInt.F90:131-136)
__global__ void k_move(size_t n, double dt, const double *__restrict__ v, double *x) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + dt * v[i];
}

// x = x + Ts (b1 k1 + b2 v): the update of TimeInt_RK2 (common/ModTimeInt.F90:246-250)
__global__ void k_rk2_update(size_t n, double Ts, double b1, double b2, const double *__restrict__ k1,
                             const double *__restrict__ v, double *x) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + Ts * (b1 * k1[i] + b2 * v[i]);
}

// x = x + Ts (3 v - v_1)/2: the update of TimeInt_AB2 (common/ModTimeInt.F90:314-317)
__global__ void k_ab2_update(size_t n, double Ts, const double *__restrict__ v,
                             const double *__restrict__ v1, double *x) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + Ts * (3 * v[i] - v1[i]) / 2;
}

// ReboxRbcs (common/ModTimeInt.F90:1006-1020): per cell and axis the centre of the
// coordinate range, and the cell shifted by whole boxes so that centre lies in [0, Lb)
__global__ void __launch_bounds__(256) k_rebox(int nrbc, const CellRec *__restrict__ cells, int np,
                                               double L0, double L1, double L2, double iL0,
                                               double iL1, double iL2, double *x) {
  __shared__ double red[2][4];
  const int c = blockIdx.x;
  if (c >= nrbc) return;
  const CellRec cr = cells[c];
  const int npt = cr.nlat * cr.nlon;
  const double L[3] = {L0, L1, L2}, iL[3] = {iL0, iL1, iL2};
  for (int d = 0; d < 3; d++) {
    double *X = x + cr.off + (size_t)np * d;
    double mx = -INFINITY, mn = INFINITY;
    for (int q = threadIdx.x; q < npt; q += 256) {
      mx = fmax(mx, X[q]);
      mn = fmin(mn, X[q]);
    }
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmax(mx, __shfl_xor(mx, o, 64));
      mn = fmin(mn, __shfl_xor(mn, o, 64));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = mx;
      red[1][threadIdx.x >> 6] = mn;
    }
    __syncthreads();
    mx = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    mn = fmin(fmin(red[1][0], red[1][1]), fmin(red[1][2], red[1][3]));
    const double xc = 0.5 * (mx + mn);
    const double sh = floor(xc * iL[d]) * L[d];
    __syncthreads();
    for (int q = threadIdx.x; q < npt; q += 256) X[q] = X[q] - sh;
  }
}

// VolConstrainRbcs (common/ModTimeInt.F90:938-963): x += fac a3 with
// fac = (vol_ref - vol) / area limited to epsDist/20 in magnitude.  The reference declares
// fac a default REAL and is built with -freal-4-real-8: it is double precision.
__global__ void k_vol_constrain(int np, const int *__restrict__ node_cell,
                                const double *__restrict__ fac, const double *__restrict__ a3,
                                double *x) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double f = fac[node_cell[p]];
  for (int d = 0; d < 3; d++) x[p + (size_t)np * d] = x[p + (size_t)np * d] + f * a3[p + (size_t)np * d];
}

static int require_cells(rbc3d_ctx *c) {
  if (!c) return RBC3D_EINVAL;
  if (c->nrbc == 0) {
    rbc_set_error("no cells set");
    return RBC3D_ESTATE;
  }
  return RBC3D_OK;
}

// Compute_Rbc_Vel (common/ModTimeInt.F90:716-808): the cells' geometry, -Shell_ResForce
// plus the intra-cell repulsion (when on) projected as the force density, Solve_RBC_Vel;
// the node velocities stay on the device for rbc3d_move_rbcs (and go to v when given)
extern "C" int rbc3d_compute_rbc_vel(rbc3d_ctx *c, double *v, int *niter, double *rnorm) {
  RBC_TRY(require_cells(c));
  if (!niter || !rnorm) return RBC3D_EINVAL;
  RBC_TRY(rbc_cells_geometry(c, true));
  RBC_TRY(rbc_cells_force(c));
  const int r = rbc3d_solve_rbc_vel(c, nullptr, niter, rnorm);
  if (r != RBC3D_OK && r != RBC3D_ENOCONV) return r;
  // the velocities: what Solve_RBC_Vel synthesised (d_shbuf)
  RBC_TRY(c->d_vel.ensure((size_t)3 * c->np + 1));
  HIP_TRY(hipMemcpyAsync(c->d_vel.p, c->d_shbuf.p, sizeof(double) * 3 * c->np,
                         hipMemcpyDeviceToDevice, c->stream));
  c->vel_ready = true;
  if (v)
    HIP_TRY(hipMemcpyAsync(v, c->d_vel.p, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return r;
}

extern "C" int rbc3d_move_rbcs(rbc3d_ctx *c, double dt) {
  RBC_TRY(require_cells(c));
  if (!c->vel_ready) {
    rbc_set_error("rbc3d_move_rbcs before rbc3d_compute_rbc_vel");
    return RBC3D_ESTATE;
  }
  const size_t n = (size_t)3 * c->np;
  k_move<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(n, dt, c->d_vel.p, c->d_x.p);
  HIP_TRY(hipGetLastError());
  c->rbc_geom_ready = false;
  c->x_moved = true;
  return RBC3D_OK;
}

extern "C" int rbc3d_rebox_rbcs(rbc3d_ctx *c) {
  RBC_TRY(require_cells(c));
  k_rebox<<<c->nrbc, 256, 0, c->stream>>>(c->nrbc, c->d_cells.p, c->np, c->Lb[0], c->Lb[1], c->Lb[2],
                                         c->iLb[0], c->iLb[1], c->iLb[2], c->d_x.p);
  HIP_TRY(hipGetLastError());
  c->rbc_geom_ready = false;
  c->x_moved = true;
  return RBC3D_OK;
}

extern "C" int rbc3d_vol_constrain_rbcs(rbc3d_ctx *c, double eps_dist) {
  RBC_TRY(require_cells(c));
  if (c->nref == 0) {
    rbc_set_error("rbc3d_vol_constrain_rbcs: no reference cells (rbc3d_set_ref_cells)");
    return RBC3D_ESTATE;
  }
  RBC_TRY(rbc_cells_geometry(c, false));
  std::vector<double> vol(c->nrbc), area(c->nrbc);
  HIP_TRY(hipMemcpy(vol.data(), c->d_mvol.p, sizeof(double) * c->nrbc, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(area.data(), c->d_marea.p, sizeof(double) * c->nrbc, hipMemcpyDeviceToHost));
  std::vector<double> fac(c->nrbc);
  for (int i = 0; i < c->nrbc; i++) {
    const int k = c->cells[i].celltype - 1;
    if (k >= c->nref) {
      rbc_set_error("rbc3d_vol_constrain_rbcs: no reference cell for cell type " + std::to_string(k + 1));
      return RBC3D_EINVAL;
    }
    const double f = (c->ref_vol[k] - vol[i]) / area[i];
    fac[i] = std::copysign(std::fmin(std::fabs(f), eps_dist / 20.), f);
  }
  DBuf<double> d_fac;
  RBC_TRY(d_fac.ensure(c->nrbc));
  HIP_TRY(hipMemcpy(d_fac.p, fac.data(), sizeof(double) * c->nrbc, hipMemcpyHostToDevice));
  k_vol_constrain<<<(c->np + 255) / 256, 256, 0, c->stream>>>(c->np, c->d_node_cell.p, d_fac.p,
                                                               c->d_a3.p, c->d_x.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->rbc_geom_ready = false;
  c->x_moved = true;
  return RBC3D_OK;
}

extern "C" int rbc3d_filter_rbcs(rbc3d_ctx *c) {
  RBC_TRY(require_cells(c));
  RBC_TRY(rbc_cells_project_x(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_get_rbc_coords(rbc3d_ctx *c, double *x) {
  RBC_TRY(require_cells(c));
  if (!x) return RBC3D_EINVAL;
  const size_t np = c->np;
  std::vector<double> h(3 * np);
  HIP_TRY(hipMemcpyAsync(h.data(), c->d_x.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (const CellRec &r : c->cells) {
    const int npt = r.nlat * r.nlon;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++) x[3 * (size_t)r.off + q + (size_t)npt * d] = h[r.off + q + np * d];
  }
  return RBC3D_OK;
}

// TimeInt_Euler (common/ModTimeInt.F90:109-178) for nsteps steps: Compute_Rbc_Vel,
// NoSlipWall (when a wall is set), x += Ts v, ReboxRbcs, then four rounds of
// VolConstrainRbcs, InterCellRepulsion (when the repulsion is on) and FilterRbcs.  time
// advances by Ts per step; per step the cell and wall GMRES iteration counts go to
// its_cell / its_wall when given (length nsteps).  A solve that reaches its maxit does
// not stop the run (its count is maxit), as the reference prints the residual and goes on.
extern "C" int rbc3d_timeint_euler(rbc3d_ctx *c, double Ts, int nsteps, double eps_dist,
                                   double noslip_rtol, int noslip_maxit, double *time,
                                   int *its_cell, int *its_wall) {
  RBC_TRY(require_cells(c));
  if (nsteps < 0 || !(Ts > 0) || !time) return RBC3D_EINVAL;
  for (int lt = 0; lt < nsteps; lt++) {
    int ni = 0, nw = 0, conv = 0;
    double rn = 0, vres = 0;
    const int rv = rbc3d_compute_rbc_vel(c, nullptr, &ni, &rn);
    if (rv != RBC3D_OK && rv != RBC3D_ENOCONV) return rv;
    if (c->wall.nvert) {
      const int r = rbc3d_noslip_wall(c, noslip_rtol, noslip_maxit, nullptr, &nw, &rn, &conv, &vres);
      if (r != RBC3D_OK && r != RBC3D_ENOCONV) return r;
    }
    RBC_TRY(rbc3d_move_rbcs(c, Ts));
    RBC_TRY(rbc3d_rebox_rbcs(c));
    for (int k = 0; k < 4; k++) {
      RBC_TRY(rbc3d_vol_constrain_rbcs(c, eps_dist));
      if (c->rep_on) RBC_TRY(rbc3d_inter_cell_repulsion(c, eps_dist, nullptr, nullptr));
      RBC_TRY(rbc3d_filter_rbcs(c));
    }
    *time += Ts;
    if (its_cell) its_cell[lt] = ni;
    if (its_wall) its_wall[lt] = nw;
  }
  return RBC3D_OK;
}

// Compute_Rbc_Vel, then NoSlipWall when a wall is set: the velocity evaluation of the
// integrators below; iteration counts to *ni, *nw
static int velocity_and_wall(rbc3d_ctx *c, double noslip_rtol, int noslip_maxit, int *ni, int *nw) {
  double rn = 0, vres = 0;
  int conv = 0;
  *ni = *nw = 0;
  const int rv = rbc3d_compute_rbc_vel(c, nullptr, ni, &rn);
  if (rv != RBC3D_OK && rv != RBC3D_ENOCONV) return rv;
  if (c->wall.nvert) {
    const int r = rbc3d_noslip_wall(c, noslip_rtol, noslip_maxit, nullptr, nw, &rn, &conv, &vres);
    if (r != RBC3D_OK && r != RBC3D_ENOCONV) return r;
  }
  return RBC3D_OK;
}

// TimeInt_RK2(alpha) (common/ModTimeInt.F90:183-271): per step k1 = Compute_Rbc_Vel +
// NoSlipWall at x, the cells moved to x + alpha Ts k1 for the second evaluation k2 and moved
// back by the same amount, then x + Ts (b1 k1 + b2 k2) with b2 = 1/(2 alpha), b1 = 1 - b2,
// and ReboxRbcs.  No volume constraint, repulsion or filter, as in the reference.
// its_cell / its_wall (optional) take 2 nsteps counts: the two evaluations of each step.
extern "C" int rbc3d_timeint_rk2(rbc3d_ctx *c, double alpha, double Ts, int nsteps,
                                 double noslip_rtol, int noslip_maxit, double *time, int *its_cell,
                                 int *its_wall) {
  RBC_TRY(require_cells(c));
  if (nsteps < 0 || !(Ts > 0) || !(alpha != 0) || !std::isfinite(alpha) || !time) return RBC3D_EINVAL;
  const double c2 = alpha, b2 = 1 / (2 * alpha), b1 = 1 - b2;
  const size_t n = (size_t)3 * c->np;
  const unsigned nb = (unsigned)((n + 255) / 256);
  RBC_TRY(c->d_vel_k1.ensure(n + 1));
  for (int lt = 0; lt < nsteps; lt++) {
    int ni, nw;
    RBC_TRY(velocity_and_wall(c, noslip_rtol, noslip_maxit, &ni, &nw));
    if (its_cell) its_cell[2 * lt] = ni;
    if (its_wall) its_wall[2 * lt] = nw;
    HIP_TRY(hipMemcpyAsync(c->d_vel_k1.p, c->d_vel.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           c->stream));
    k_move<<<nb, 256, 0, c->stream>>>(n, c2 * Ts, c->d_vel_k1.p, c->d_x.p);
    HIP_TRY(hipGetLastError());
    RBC_TRY(velocity_and_wall(c, noslip_rtol, noslip_maxit, &ni, &nw));
    if (its_cell) its_cell[2 * lt + 1] = ni;
    if (its_wall) its_wall[2 * lt + 1] = nw;
    k_move<<<nb, 256, 0, c->stream>>>(n, -(c2 * Ts), c->d_vel_k1.p, c->d_x.p);
    k_rk2_update<<<nb, 256, 0, c->stream>>>(n, Ts, b1, b2, c->d_vel_k1.p, c->d_vel.p, c->d_x.p);
    HIP_TRY(hipGetLastError());
    c->rbc_geom_ready = false;
    c->x_moved = true;
    RBC_TRY(rbc3d_rebox_rbcs(c));
    *time += Ts;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

// TimeInt_AB2 (common/ModTimeInt.F90:275-341): one Euler step x + Ts v from Compute_Rbc_Vel
// alone (no NoSlipWall, no rebox, time not advanced, as in the reference), then nsteps - 1
// steps of Compute_Rbc_Vel + NoSlipWall, x + Ts (3 v - v_1)/2 with v_1 the previous
// velocity, ReboxRbcs, time + Ts.  its_cell / its_wall (optional): nsteps counts, the
// first from the Euler step.
extern "C" int rbc3d_timeint_ab2(rbc3d_ctx *c, double Ts, int nsteps, double noslip_rtol,
                                 int noslip_maxit, double *time, int *its_cell, int *its_wall) {
  RBC_TRY(require_cells(c));
  if (nsteps < 1 || !(Ts > 0) || !time) return RBC3D_EINVAL;
  const size_t n = (size_t)3 * c->np;
  const unsigned nb = (unsigned)((n + 255) / 256);
  RBC_TRY(c->d_vel_k1.ensure(n + 1));
  int ni = 0;
  double rn = 0;
  const int rv = rbc3d_compute_rbc_vel(c, nullptr, &ni, &rn);
  if (rv != RBC3D_OK && rv != RBC3D_ENOCONV) return rv;
  if (its_cell) its_cell[0] = ni;
  if (its_wall) its_wall[0] = 0;
  RBC_TRY(rbc3d_move_rbcs(c, Ts));
  for (int lt = 1; lt < nsteps; lt++) {
    HIP_TRY(hipMemcpyAsync(c->d_vel_k1.p, c->d_vel.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                           c->stream));
    int nw;
    RBC_TRY(velocity_and_wall(c, noslip_rtol, noslip_maxit, &ni, &nw));
    if (its_cell) its_cell[lt] = ni;
    if (its_wall) its_wall[lt] = nw;
    k_ab2_update<<<nb, 256, 0, c->stream>>>(n, Ts, c->d_vel.p, c->d_vel_k1.p, c->d_x.p);
    HIP_TRY(hipGetLastError());
    c->rbc_geom_ready = false;
    c->x_moved = true;
    RBC_TRY(rbc3d_rebox_rbcs(c));
    *time += Ts;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

int rbc_refresh_moved(rbc3d_ctx *c) {
  if (!c->x_moved || c->nrbc == 0) return RBC3D_OK;
  RBC_TRY(rbc_cells_geometry(c, true));
  RBC_TRY(rbc_update_density(c, 1));
  RBC_TRY(rbc_update_density(c, 2));
  return RBC3D_OK;
}
