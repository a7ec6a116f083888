// Per-cell surface geometry and membrane mechanics on the GPU, the steps that come
// before every cell-BIE matvec of a timestep (Compute_Rbc_Vel, common/ModTimeInt.F90:
// 716-808): RBC_ComputeGeometry (common/ModRbc.F90:404-509), Shell_ResForce with
// Shell_ElasStrs and Shell_Bend (:815-969) and its covariant gradients (:527-654), and
// RBC_SphProject / FilterRbcs (:659-686; common/ModTimeInt.F90:908-934).
//
// One workgroup per cell (a persistent loop over the cells, cells of different grids
// side by side).  Every surface derivative is ShAnalGau followed by ShGradGau of one
// scalar field, run out of LDS with the per-grid tables (the building blocks of the
// transform kernels, sh_core.h); intermediate fields (tangent-basis tensors, their
// derivatives) go to a per-cell scratch block in HBM.  Pointwise steps follow the
// reference's arithmetic statement by statement, without FMA contraction, and the
// per-cell sums (area, centroid, volume, mean normal force) run in the reference's loop
// order, so the results match the CPU restatement to rounding.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "internal.h"
#include "sh_core.h"

#define MECH_T 256
#define MECH_MAXGRID 8
// scratch of Shell_ResForce per node: tau_t 4, bnd 4, dtau 8, dbnd 8, tau_n 2, dtau_n 4,
// tc 9, tc1 9, tc2 9
#define MECH_SCR 57

// tables of one Gauss grid nlat x 2 nlat
struct MechGrid {
  int nl, nn, mmax;
  const double *pw, *pt, *pd, *cs, *sth, *w;
};
struct MechGrids {
  int n;
  MechGrid g[MECH_MAXGRID];
};

struct MechTables {
  DBuf<double> pw, pt, pd, cs, sth, w;
};

static std::mutex g_mech_mu;
static std::map<std::pair<int, int>, MechTables *> g_mech_tabs;  // (device, nlat) -> tables

// Analysis table Pbar(m,n,th_i) w_i (the 2/nlon factor is applied in the sum), synthesis
// tables Pbar and dPbar/dth with the 1/2 of m = 0, cos/sin(m phi_j), sin(th_i) and the
// cell quadrature weights w_i 2 pi / nlon (RBC_Create, common/ModRbc.F90:93-94).
static int mech_grid(int nlat, MechGrid *out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mech_mu);
  MechTables *&T = g_mech_tabs[{dev, nlat}];
  const int nl = nlat, nn = 2 * nlat, mmax = std::min(nl, nn / 2 + 1);
  if (!T) {
    std::vector<double> th, wg, P;
    host_gauss(nl, th, wg);
    host_pbar(nl, mmax, th, P);
    std::vector<double> pw(P.size()), pt(P.size()), pd(P.size());
    for (int m = 0; m < mmax; m++)
      for (int n = 0; n < nl; n++)
        for (int i = 0; i < nl; i++) {
          const size_t o = ((size_t)m * nl + n) * nl + i;
          const double h = m == 0 ? 0.5 : 1.0;
          pw[o] = P[o] * wg[i];
          pt[o] = h * P[o];
          double d = (double)n * cos(th[i]) * P[o];
          if (n > m) {
            const double dn = n, dm = m;
            d -= sqrt((2 * dn + 1) * (dn * dn - dm * dm) / (2 * dn - 1)) * P[o - nl];
          }
          pd[o] = h * (d / sin(th[i]));
        }
    std::vector<double> cs((size_t)2 * mmax * nn), sth(nl), w(nl);
    for (int m = 0; m < mmax; m++)
      for (int j = 0; j < nn; j++) {
        const double ph = RBC_TWO_PI * j / nn;
        cs[(size_t)m * nn + j] = cos(m * ph);
        cs[(size_t)mmax * nn + m * nn + j] = sin(m * ph);
      }
    for (int i = 0; i < nl; i++) {
      sth[i] = sin(th[i]);
      w[i] = (RBC_TWO_PI / nn) * wg[i];
    }
    MechTables *t = new MechTables;
    int r = RBC3D_OK;
    auto up = [&r](DBuf<double> &b, const std::vector<double> &h) {
      if (r) return;
      r = b.ensure(h.size());
      if (!r && hipMemcpy(b.p, h.data(), h.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        r = RBC3D_EHIP;
    };
    up(t->pw, pw);
    up(t->pt, pt);
    up(t->pd, pd);
    up(t->cs, cs);
    up(t->sth, sth);
    up(t->w, w);
    if (r) {
      delete t;
      rbc_set_error("membrane mechanics: table upload failed");
      return r;
    }
    T = t;
  }
  out->nl = nl;
  out->nn = nn;
  out->mmax = mmax;
  out->pw = T->pw.p;
  out->pt = T->pt.p;
  out->pd = T->pd.p;
  out->cs = T->cs.p;
  out->sth = T->sth.p;
  out->w = T->w.p;
  return RBC3D_OK;
}

static size_t mech_lds(int nl) { return sizeof(double) * 10 * (size_t)nl * nl; }

__device__ __forceinline__ const MechGrid &grid_of(const MechGrids &G, int nl) {
  int k = 0;
  while (k + 1 < G.n && G.g[k].nl != nl) k++;
  return G.g[k];
}

// v = dg/dth, w = dg/dphi of one scalar field g (nl x nn, i + nl j; global memory) by
// ShAnalGau and ShGradGau (the operations of k_anal_batch and k_grad_batch)
__device__ void field_grad(const MechGrid &T, const double *g, double *v, double *w, double *sm) {
  const int nl = T.nl, nn = T.nn, mmax = T.mmax, np = nl * nn;
  double *f = sm, *C = f + np, *ab = C + 2 * mmax * nl, *G = ab + 2 * mmax * nl, *gg = G + 2 * mmax * nl;
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) f[t] = g[t];
  __syncthreads();
  lon_dft(f, nl, nl, nn, mmax, T.cs, mmax, C, C + mmax * nl);
  __syncthreads();
  leg_anal(C, C + mmax * nl, T.pw, nl, mmax, nl, ab, ab + mmax * nl, mmax, 2.0 / nn);
  __syncthreads();
  leg_synth(ab, ab + mmax * nl, mmax, T.pd, nl, nl, mmax, nl, G, G + mmax * nl);
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, T.cs, mmax, gg, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) v[t] = gg[t];
  leg_synth(ab, ab + mmax * nl, mmax, T.pt, nl, nl, mmax, nl, G, G + mmax * nl);
  __syncthreads();
  // d/dphi of (Ga cos m phi - Gb sin m phi) is (-m Gb) cos m phi - (m Ga) sin m phi
  for (int t = threadIdx.x; t < mmax * nl; t += blockDim.x) {
    const double mm = (double)(t / nl), ga = G[t], gb = G[mmax * nl + t];
    G[t] = -mm * gb;
    G[mmax * nl + t] = mm * ga;
  }
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, T.cs, mmax, gg, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) w[t] = gg[t];
  __syncthreads();
}

// RBC_SphProject of one field: ShAnalGau, ShFilter (n >= nlat0 dropped), ShSynthGau
__device__ void field_project(const MechGrid &T, int nlat0, double *g, double *sm) {
  const int nl = T.nl, nn = T.nn, mmax = T.mmax, np = nl * nn;
  double *f = sm, *C = f + np, *ab = C + 2 * mmax * nl, *G = ab + 2 * mmax * nl;
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) f[t] = g[t];
  __syncthreads();
  lon_dft(f, nl, nl, nn, mmax, T.cs, mmax, C, C + mmax * nl);
  __syncthreads();
  leg_anal(C, C + mmax * nl, T.pw, nl, mmax, nl, ab, ab + mmax * nl, mmax, 2.0 / nn);
  __syncthreads();
  for (int t = threadIdx.x; t < mmax * nl; t += blockDim.x)
    if (t / mmax >= nlat0) {
      ab[t] = 0.;
      ab[mmax * nl + t] = 0.;
    }
  __syncthreads();
  leg_synth(ab, ab + mmax * nl, mmax, T.pt, nl, nl, mmax, nl, G, G + mmax * nl);
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, T.cs, mmax, f, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) g[t] = f[t];
  __syncthreads();
}

struct GeoOut {
  const double *x;
  double *a1, *a2, *a3, *a, *ar, *b, *detj, *xc, *area, *vol, *mesh;
  double *scr;  // 6 np per cell
};

// RBC_ComputeGeometry (common/ModRbc.F90:404-509), one workgroup per cell
__global__ void __launch_bounds__(MECH_T) k_rbc_geometry(int ncell, const int *__restrict__ off,
                                                         const int *__restrict__ nlat_c,
                                                         MechGrids G, GeoOut o) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  for (int c = blockIdx.x; c < ncell; c += gridDim.x) {
    const MechGrid &T = grid_of(G, nlat_c[c]);
    const int nl = T.nl, nn = T.nn;
    const size_t np = (size_t)nl * nn, q0 = off[c];
    const double *X = o.x + 3 * q0;
    double *A1 = o.a1 + 3 * q0, *A2 = o.a2 + 3 * q0, *A3 = o.a3 + 3 * q0;
    double *Am = o.a + 4 * q0, *Ar = o.ar + 4 * q0, *B = o.b + 4 * q0, *J = o.detj + q0;
    double *S1 = o.scr + 6 * q0, *S2 = S1 + 3 * np;
    for (int d = 0; d < 3; d++) field_grad(T, X + d * np, A1 + d * np, A2 + d * np, sm);
    // surface tangents and normals (:423-455)
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      const int i = (int)(t % nl);
      double t1[3], t2[3];
      for (int d = 0; d < 3; d++) {
        t1[d] = A1[t + d * np];
        t2[d] = A2[t + d * np];
      }
      const double m11 = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
      const double m12 = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
      const double m22 = t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2];
      const double detA = m11 * m22 - m12 * m12;
      const double idetA = 1. / detA;
      Am[t] = m11;
      Am[t + 2 * np] = m12;
      Am[t + np] = m12;
      Am[t + 3 * np] = m22;
      Ar[t] = idetA * m22;
      Ar[t + 2 * np] = -idetA * m12;
      Ar[t + np] = -idetA * m12;
      Ar[t + 3 * np] = idetA * m11;
      J[t] = sqrt(detA) / T.sth[i];
      const double n0 = t1[1] * t2[2] - t1[2] * t2[1], n1 = t1[2] * t2[0] - t1[0] * t2[2],
                   n2 = t1[0] * t2[1] - t1[1] * t2[0];
      const double nn2 = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      A3[t] = n0 / nn2;
      A3[t + np] = n1 / nn2;
      A3[t + 2 * np] = n2 / nn2;
    }
    // curvature tensor (:458-476)
    for (int d = 0; d < 3; d++) field_grad(T, A3 + d * np, S1 + d * np, S2 + d * np, sm);
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      double s11 = 0, s12 = 0, s21 = 0, s22 = 0;
      for (int d = 0; d < 3; d++) {
        s11 += A1[t + d * np] * S1[t + d * np];
        s12 += A1[t + d * np] * S2[t + d * np];
        s21 += A2[t + d * np] * S1[t + d * np];
        s22 += A2[t + d * np] * S2[t + d * np];
      }
      B[t] = -s11;
      B[t + 2 * np] = -s12;
      B[t + np] = -s21;
      B[t + 3 * np] = -s22;
    }
    __syncthreads();
    // centroid and area (:478-491), volume (:493-501), mesh size (:504): one thread, the
    // reference's loop order
    if (threadIdx.x == 0) {
      double ar = 0, c3[3] = {0, 0, 0};
      for (int i = 0; i < nl; i++)
        for (int j = 0; j < nn; j++) {
          const double ds = J[i + (size_t)nl * j] * T.w[i];
          ar += ds;
          for (int d = 0; d < 3; d++) c3[d] += ds * X[i + (size_t)nl * j + d * np];
        }
      for (int d = 0; d < 3; d++) c3[d] = c3[d] / ar;
      double v = 0;
      for (int j = 0; j < nn; j++)
        for (int i = 0; i < nl; i++) {
          const size_t t = i + (size_t)nl * j;
          const double ds = J[t] * T.w[i];
          double s = 0;
          for (int d = 0; d < 3; d++) s += A3[t + d * np] * (X[t + d * np] - c3[d]);
          v += ds * s;
        }
      for (int d = 0; d < 3; d++) o.xc[3 * c + d] = c3[d];
      o.area[c] = ar;
      o.vol[c] = (1.0 / 3) * v;
      o.mesh[c] = sqrt(ar / ((double)nl * nl));
    }
    __syncthreads();
  }
}

// a_p^rcp at point t: a1_rcp = a_rcp(1,1) a1 + a_rcp(1,2) a2 (:437-438)
__device__ __forceinline__ void rcp_vec(const double *A1, const double *A2, const double *Ar,
                                        size_t t, size_t np, double r[2][3]) {
#pragma clang fp contract(off)
  for (int d = 0; d < 3; d++) {
    r[0][d] = Ar[t] * A1[t + d * np] + Ar[t + 2 * np] * A2[t + d * np];
    r[1][d] = Ar[t + np] * A1[t + d * np] + Ar[t + 3 * np] * A2[t + d * np];
  }
}

// RBC_CovarGrad_Tensor (common/ModRbc.F90:586-654): t (4 np, (p,q) at p + 2q) -> dt (8 np,
// (p,q,k) at p + 2(q + 2k)); tc, tc1, tc2 are 9 np of scratch each
__device__ void covar_grad_tensor(const MechGrid &T, const double *A1, const double *A2,
                                  const double *Ar, const double *t4, double *dt, double *tc,
                                  double *tc1, double *tc2, double *sm) {
#pragma clang fp contract(off)
  const size_t np = (size_t)T.nl * T.nn;
  for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
    const double t11 = t4[t], t21 = t4[t + np], t12 = t4[t + 2 * np], t22 = t4[t + 3 * np];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        const double x1i = A1[t + i * np], x2i = A2[t + i * np];
        const double x1j = A1[t + j * np], x2j = A2[t + j * np];
        tc[t + (i + 3 * j) * np] = t11 * x1i * x1j + t12 * x1i * x2j + t21 * x2i * x1j + t22 * x2i * x2j;
      }
  }
  for (int k = 0; k < 9; k++) field_grad(T, tc + k * np, tc1 + k * np, tc2 + k * np, sm);
  for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
    double rc[2][3];
    rcp_vec(A1, A2, Ar, t, np, rc);
    for (int k = 0; k < 2; k++) {
      const double *Tk = k == 0 ? tc1 : tc2;
      for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2; q++) {
          double s = 0;
          for (int i = 0; i < 3; i++) {
            double y = 0;
            for (int j = 0; j < 3; j++) y += Tk[t + (i + 3 * j) * np] * rc[q][j];
            s += rc[p][i] * y;
          }
          dt[t + (p + 2 * (q + 2 * k)) * np] = s;
        }
    }
  }
  __syncthreads();
}

struct ForceArgs {
  const int *off, *nlat, *ref_of;  // per cell: node offset, nlat, reference-cell offset
  const double *a1, *a2, *a3, *a, *ar, *b, *detj, *area;
  const double *ar_ref, *b_ref;    // reference cells' a_rcp, b (concatenated)
  const double *ES, *ED, *EB;      // per cell
  double *f;
  double *scr;  // MECH_SCR np per cell
};

// 2 x 2 column-major product c = a b of point t (arrays of 4 np)
__device__ __forceinline__ void mm2p(const double a[4], const double b[4], double c[4]) {
#pragma clang fp contract(off)
  c[0] = a[0] * b[0] + a[2] * b[1];
  c[1] = a[1] * b[0] + a[3] * b[1];
  c[2] = a[0] * b[2] + a[2] * b[3];
  c[3] = a[1] * b[2] + a[3] * b[3];
}

// Shell_ResForce (common/ModRbc.F90:815-900), one workgroup per cell
__global__ void __launch_bounds__(MECH_T) k_shell_res_force(int ncell, MechGrids G, ForceArgs o) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  __shared__ double fn_mean;
  for (int c = blockIdx.x; c < ncell; c += gridDim.x) {
    const MechGrid &T = grid_of(G, o.nlat[c]);
    const int nl = T.nl, nn = T.nn;
    const size_t np = (size_t)nl * nn, q0 = o.off[c], r0 = o.ref_of[c];
    const double *A1 = o.a1 + 3 * q0, *A2 = o.a2 + 3 * q0, *A3 = o.a3 + 3 * q0;
    const double *Am = o.a + 4 * q0, *Ar = o.ar + 4 * q0, *B = o.b + 4 * q0, *J = o.detj + q0;
    const double *Arf = o.ar_ref + 4 * r0, *Bf = o.b_ref + 4 * r0;
    double *F = o.f + 3 * q0;
    double *tau_t = o.scr + MECH_SCR * q0, *bnd = tau_t + 4 * np, *dtau = bnd + 4 * np, *dbnd = dtau + 8 * np;
    double *tau_n = dbnd + 8 * np, *dtau_n = tau_n + 2 * np, *tc = dtau_n + 4 * np;
    double *tc1 = tc + 9 * np, *tc2 = tc1 + 9 * np;
    const double ES = o.ES[c], ED = o.ED[c], EB = o.EB[c];
    // Shell_ElasStrs (:908-944) and Shell_Bend (:951-969)
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      double A[4], R[4], Rf[4], V2[4], Bc[4], Bfr[4], bm[4], bmr[4], d[4], tt[4];
      for (int k = 0; k < 4; k++) {
        A[k] = Am[t + k * np];
        R[k] = Ar[t + k * np];
        Rf[k] = Arf[t + k * np];
        Bc[k] = B[t + k * np];
        Bfr[k] = Bf[t + k * np];
      }
      mm2p(A, Rf, V2);
      const double lbd1 = V2[0] + V2[3] - 2.0;
      const double lbd2 = V2[0] * V2[3] - V2[2] * V2[1] - 1.0;
      const double JS = sqrt(V2[0] * V2[3] - V2[2] * V2[1]);
      for (int k = 0; k < 4; k++)
        tau_t[t + k * np] = 0.5 * ES / JS * (lbd1 + 1.0) * Rf[k] + 0.5 * JS * (ED * lbd2 - ES) * R[k];
      mm2p(R, Bc, bm);
      mm2p(Rf, Bfr, bmr);
      for (int k = 0; k < 4; k++) d[k] = bm[k] - bmr[k];
      mm2p(d, R, tt);
      for (int k = 0; k < 4; k++) bnd[t + k * np] = -EB * tt[k];
    }
    // normal stress from the bending moment (:845-851)
    covar_grad_tensor(T, A1, A2, Ar, bnd, dbnd, tc, tc1, tc2, sm);
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      tau_n[t] = dbnd[t + 0 * np] + dbnd[t + 5 * np];       // (1,1,1) + (2,1,2)
      tau_n[t + np] = dbnd[t + 2 * np] + dbnd[t + 7 * np];  // (1,2,1) + (2,2,2)
    }
    // tangential residual force (:853-866)
    covar_grad_tensor(T, A1, A2, Ar, tau_t, dtau, tc, tc1, tc2, sm);
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      double f1 = -dtau[t + 0 * np] - dtau[t + 5 * np];
      double f2 = -dtau[t + 2 * np] - dtau[t + 7 * np];
      const double tn0 = tau_n[t], tn1 = tau_n[t + np];
      const double y0 = B[t] * tn0 + B[t + 2 * np] * tn1;
      const double y1 = B[t + np] * tn0 + B[t + 3 * np] * tn1;
      f1 = f1 + (Ar[t] * y0 + Ar[t + 2 * np] * y1);
      f2 = f2 + (Ar[t + np] * y0 + Ar[t + 3 * np] * y1);
      F[t] = f1;
      F[t + np] = f2;
    }
    // normal residual force (:868-876): RBC_CovarGrad_Vec (:527-575) of tau_n
    for (size_t t = threadIdx.x; t < np; t += blockDim.x)
      for (int d = 0; d < 3; d++)
        tc[t + d * np] = tau_n[t] * A1[t + d * np] + tau_n[t + np] * A2[t + d * np];
    for (int d = 0; d < 3; d++) field_grad(T, tc + d * np, tc1 + d * np, tc2 + d * np, sm);
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      double rc[2][3];
      rcp_vec(A1, A2, Ar, t, np, rc);
      for (int p = 0; p < 2; p++)
        for (int k = 0; k < 2; k++) {
          const double *V = k == 0 ? tc1 : tc2;
          double s = 0;
          for (int d = 0; d < 3; d++) s += rc[p][d] * V[t + d * np];
          dtau_n[t + (p + 2 * k) * np] = s;
        }
      double s = 0;  // sum(tau_t * b) in array element order
      for (int k = 0; k < 4; k++) s += tau_t[t + k * np] * B[t + k * np];
      F[t + 2 * np] = -dtau_n[t] - dtau_n[t + 3 * np] - s;
    }
    __syncthreads();
    // the average normal force (:878-885), the reference's loop order
    if (threadIdx.x == 0) {
      double fnTot = 0.;
      for (int j = 0; j < nn; j++)
        for (int i = 0; i < nl; i++) {
          const size_t t = i + (size_t)nl * j;
          fnTot += F[t + 2 * np] * J[t] * T.w[i];
        }
      fn_mean = fnTot / o.area[c];
    }
    __syncthreads();
    // Cartesian components (:887-894)
    for (size_t t = threadIdx.x; t < np; t += blockDim.x) {
      const double f1 = F[t], f2 = F[t + np], f3 = F[t + 2 * np] - fn_mean;
      for (int d = 0; d < 3; d++)
        F[t + d * np] = f1 * A1[t + d * np] + f2 * A2[t + d * np] + f3 * A3[t + d * np];
    }
    __syncthreads();
  }
}

// RBC_SphProject of nvar fields per cell
__global__ void __launch_bounds__(MECH_T) k_rbc_sph_project(int ncell, const int *__restrict__ off,
                                                            const int *__restrict__ nlat_c,
                                                            const int *__restrict__ nlat0_c,
                                                            int nvar, MechGrids G, double *u) {
  extern __shared__ double sm[];
  for (int c = blockIdx.x; c < ncell; c += gridDim.x) {
    const MechGrid &T = grid_of(G, nlat_c[c]);
    const size_t np = (size_t)T.nl * T.nn;
    double *U = u + (size_t)nvar * off[c];
    for (int k = 0; k < nvar; k++) field_project(T, nlat0_c[c], U + k * np, sm);
  }
}

// ---------------------------------------------------------------------------
// host side: per-cell offsets, grid tables, transfers

struct CellBatch {
  std::vector<int> off, nl;
  size_t np = 0;
  MechGrids G{};
  size_t lds = 0;
};

static int cell_batch(int ncell, const int *nlat, CellBatch *cb) {
  cb->off.resize(ncell);
  cb->nl.assign(nlat, nlat + ncell);
  cb->np = 0;
  cb->G.n = 0;
  cb->lds = 0;
  for (int c = 0; c < ncell; c++) {
    if (nlat[c] < 2 || nlat[c] > 28) {
      rbc_set_error("membrane mechanics: nlat must lie in 2..28 (LDS of one cell)");
      return RBC3D_EINVAL;
    }
    cb->off[c] = (int)cb->np;
    cb->np += 2 * (size_t)nlat[c] * nlat[c];
    bool have = false;
    for (int k = 0; k < cb->G.n; k++) have = have || cb->G.g[k].nl == nlat[c];
    if (!have) {
      if (cb->G.n == MECH_MAXGRID) {
        rbc_set_error("membrane mechanics: more than 8 distinct grids in one call");
        return RBC3D_EINVAL;
      }
      RBC_TRY(mech_grid(nlat[c], &cb->G.g[cb->G.n]));
      cb->G.n++;
      cb->lds = std::max(cb->lds, mech_lds(nlat[c]));
    }
  }
  return RBC3D_OK;
}

static unsigned mech_blocks(int ncell) { return (unsigned)std::max(1, std::min(ncell, 2048)); }

// device-pointer core of rbc3d_rbc_compute_geometry
int rbc_rbc_geometry_dev(hipStream_t st, int ncell, const int *d_off, const int *d_nlat,
                         const MechGrids &G, size_t lds, const GeoOut &o) {
  hipLaunchKernelGGL(k_rbc_geometry, dim3(mech_blocks(ncell)), dim3(MECH_T), lds, st, ncell, d_off,
                     d_nlat, G, o);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

namespace {
struct Dev {
  std::vector<void *> ps;
  ~Dev() {
    for (void *p : ps) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T **p, size_t n) {
    *p = nullptr;
    if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      rbc_set_error("membrane mechanics: device allocation failed");
      return RBC3D_EHIP;
    }
    ps.push_back(*p);
    return RBC3D_OK;
  }
  template <typename T>
  int up(T **p, const T *h, size_t n) {
    RBC_TRY(alloc(p, n));
    HIP_TRY(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return RBC3D_OK;
  }
};
}  // namespace

static bool all_finite(const double *x, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(x[i])) return false;
  return true;
}

extern "C" int rbc3d_rbc_compute_geometry(int ncell, const int *nlat, const double *x, double *a1,
                                          double *a2, double *a3, double *a, double *a_rcp,
                                          double *b, double *detj, double *xc, double *area,
                                          double *vol, double *mesh_size) {
  if (ncell < 0 || (ncell > 0 && (!nlat || !x || !a1 || !a2 || !a3 || !a || !a_rcp || !b || !detj ||
                                  !xc || !area || !vol || !mesh_size)))
    return RBC3D_EINVAL;
  if (ncell == 0) return RBC3D_OK;
  CellBatch cb;
  RBC_TRY(cell_batch(ncell, nlat, &cb));
  const size_t np = cb.np;
  if (!all_finite(x, 3 * np)) {
    rbc_set_error("rbc3d_rbc_compute_geometry: non-finite coordinates");
    return RBC3D_ENONFINITE;
  }
  Dev D;
  int *d_off, *d_nl;
  GeoOut o;
  double *d_x;
  RBC_TRY(D.up(&d_off, cb.off.data(), ncell));
  RBC_TRY(D.up(&d_nl, cb.nl.data(), ncell));
  RBC_TRY(D.up(&d_x, x, 3 * np));
  o.x = d_x;
  RBC_TRY(D.alloc(&o.a1, 3 * np));
  RBC_TRY(D.alloc(&o.a2, 3 * np));
  RBC_TRY(D.alloc(&o.a3, 3 * np));
  RBC_TRY(D.alloc(&o.a, 4 * np));
  RBC_TRY(D.alloc(&o.ar, 4 * np));
  RBC_TRY(D.alloc(&o.b, 4 * np));
  RBC_TRY(D.alloc(&o.detj, np));
  RBC_TRY(D.alloc(&o.xc, 3 * (size_t)ncell));
  RBC_TRY(D.alloc(&o.area, ncell));
  RBC_TRY(D.alloc(&o.vol, ncell));
  RBC_TRY(D.alloc(&o.mesh, ncell));
  RBC_TRY(D.alloc(&o.scr, 6 * np));
  RBC_TRY(rbc_rbc_geometry_dev(nullptr, ncell, d_off, d_nl, cb.G, cb.lds, o));
  HIP_TRY(hipDeviceSynchronize());
  auto down = [](double *h, const double *d, size_t n) {
    return hipMemcpy(h, d, n * 8, hipMemcpyDeviceToHost) == hipSuccess;
  };
  bool ok = down(a1, o.a1, 3 * np) && down(a2, o.a2, 3 * np) && down(a3, o.a3, 3 * np) &&
            down(a, o.a, 4 * np) && down(a_rcp, o.ar, 4 * np) && down(b, o.b, 4 * np) &&
            down(detj, o.detj, np) && down(xc, o.xc, 3 * (size_t)ncell) && down(area, o.area, ncell) &&
            down(vol, o.vol, ncell) && down(mesh_size, o.mesh, ncell);
  if (!ok) {
    rbc_set_error("rbc3d_rbc_compute_geometry: copy to the host failed");
    return RBC3D_EHIP;
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_shell_res_force(int ncell, const int *nlat, const int *celltype,
                                     const double *a1, const double *a2, const double *a3,
                                     const double *a, const double *a_rcp, const double *b,
                                     const double *detj, const double *area, int ntypes,
                                     const int *nlat_ref, const double *a_rcp_ref,
                                     const double *b_ref, const double *ES, const double *ED,
                                     const double *EB, double *f) {
  if (ncell < 0 || ntypes < 1 || !nlat_ref || !a_rcp_ref || !b_ref || !ES || !ED || !EB ||
      (ncell > 0 && (!nlat || !celltype || !a1 || !a2 || !a3 || !a || !a_rcp || !b || !detj ||
                     !area || !f)))
    return RBC3D_EINVAL;
  if (ncell == 0) return RBC3D_OK;
  CellBatch cb;
  RBC_TRY(cell_batch(ncell, nlat, &cb));
  const size_t np = cb.np;
  // reference cells: one per cell type, concatenated in type order
  std::vector<int> roff(ntypes);
  size_t nref = 0;
  for (int k = 0; k < ntypes; k++) {
    roff[k] = (int)nref;
    nref += 2 * (size_t)nlat_ref[k] * nlat_ref[k];
  }
  std::vector<int> ref_of(ncell);
  std::vector<double> es(ncell), ed(ncell), eb(ncell);
  for (int c = 0; c < ncell; c++) {
    const int k = celltype[c] - 1;
    if (k < 0 || k >= ntypes || nlat_ref[k] != nlat[c]) {
      rbc_set_error("rbc3d_shell_res_force: cell " + std::to_string(c) +
                    " has no reference cell of its type and grid");
      return RBC3D_EINVAL;
    }
    ref_of[c] = roff[k];
    es[c] = ES[k];
    ed[c] = ED[k];
    eb[c] = EB[k];
  }
  if (!all_finite(a1, 3 * np) || !all_finite(a2, 3 * np) || !all_finite(a_rcp, 4 * np) ||
      !all_finite(b, 4 * np)) {
    rbc_set_error("rbc3d_shell_res_force: non-finite geometry");
    return RBC3D_ENONFINITE;
  }
  Dev D;
  ForceArgs o;
  int *d_off, *d_nl, *d_ref;
  double *p;
  RBC_TRY(D.up(&d_off, cb.off.data(), ncell));
  RBC_TRY(D.up(&d_nl, cb.nl.data(), ncell));
  RBC_TRY(D.up(&d_ref, ref_of.data(), ncell));
  o.off = d_off;
  o.nlat = d_nl;
  o.ref_of = d_ref;
  RBC_TRY(D.up(&p, a1, 3 * np));
  o.a1 = p;
  RBC_TRY(D.up(&p, a2, 3 * np));
  o.a2 = p;
  RBC_TRY(D.up(&p, a3, 3 * np));
  o.a3 = p;
  RBC_TRY(D.up(&p, a, 4 * np));
  o.a = p;
  RBC_TRY(D.up(&p, a_rcp, 4 * np));
  o.ar = p;
  RBC_TRY(D.up(&p, b, 4 * np));
  o.b = p;
  RBC_TRY(D.up(&p, detj, np));
  o.detj = p;
  RBC_TRY(D.up(&p, area, ncell));
  o.area = p;
  RBC_TRY(D.up(&p, a_rcp_ref, 4 * nref));
  o.ar_ref = p;
  RBC_TRY(D.up(&p, b_ref, 4 * nref));
  o.b_ref = p;
  RBC_TRY(D.up(&p, es.data(), ncell));
  o.ES = p;
  RBC_TRY(D.up(&p, ed.data(), ncell));
  o.ED = p;
  RBC_TRY(D.up(&p, eb.data(), ncell));
  o.EB = p;
  RBC_TRY(D.alloc(&o.f, 3 * np));
  RBC_TRY(D.alloc(&o.scr, MECH_SCR * np));
  hipLaunchKernelGGL(k_shell_res_force, dim3(mech_blocks(ncell)), dim3(MECH_T), cb.lds, nullptr,
                     ncell, cb.G, o);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(f, o.f, 3 * np * 8, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

extern "C" int rbc3d_rbc_sph_project(int ncell, const int *nlat, const int *nlat0, int nvar,
                                     double *u) {
  if (ncell < 0 || nvar < 1 || (ncell > 0 && (!nlat || !nlat0 || !u))) return RBC3D_EINVAL;
  if (ncell == 0) return RBC3D_OK;
  CellBatch cb;
  RBC_TRY(cell_batch(ncell, nlat, &cb));
  for (int c = 0; c < ncell; c++)
    if (nlat0[c] < 1 || nlat0[c] > nlat[c]) {
      rbc_set_error("rbc3d_rbc_sph_project: nlat0 must lie in 1..nlat");
      return RBC3D_EINVAL;
    }
  const size_t n = (size_t)nvar * cb.np;
  Dev D;
  int *d_off, *d_nl, *d_nl0;
  double *d_u;
  RBC_TRY(D.up(&d_off, cb.off.data(), ncell));
  RBC_TRY(D.up(&d_nl, cb.nl.data(), ncell));
  RBC_TRY(D.up(&d_nl0, nlat0, ncell));
  RBC_TRY(D.up(&d_u, u, n));
  hipLaunchKernelGGL(k_rbc_sph_project, dim3(mech_blocks(ncell)), dim3(MECH_T), cb.lds, nullptr,
                     ncell, d_off, d_nl, d_nl0, nvar, cb.G, d_u);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(u, d_u, n * 8, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// the context's cells: geometry refresh, membrane force and coordinate filter on device

// [3][NP] node arrays (the context's layout) <-> per-cell (nlat, nlon, 3) blocks
__global__ void k_nodes_to_blocks(int np, const int *__restrict__ node_cell,
                                  const CellRec *__restrict__ cells, const double *__restrict__ src,
                                  int ncomp, double *dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const CellRec cr = cells[node_cell[p]];
  const int npt = cr.nlat * cr.nlon, q = p - cr.off;
  for (int d = 0; d < ncomp; d++) dst[(size_t)ncomp * cr.off + q + (size_t)npt * d] = src[p + (size_t)np * d];
}
__global__ void k_blocks_to_nodes(int np, const int *__restrict__ node_cell,
                                  const CellRec *__restrict__ cells, const double *__restrict__ src,
                                  int ncomp, double scale, double *dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const CellRec cr = cells[node_cell[p]];
  const int npt = cr.nlat * cr.nlon, q = p - cr.off;
  for (int d = 0; d < ncomp; d++)
    dst[p + (size_t)np * d] = scale * src[(size_t)ncomp * cr.off + q + (size_t)npt * d];
}
// dS = detJ w at every node (the set_cells product) and the cell records' area and meshSize
__global__ void k_ds_nodes(int np, const int *__restrict__ node_cell, const int *__restrict__ node_ll,
                           const CellRec *__restrict__ cells, MechGrids G,
                           const double *__restrict__ detj, double *ds) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const CellRec cr = cells[node_cell[p]];
  const MechGrid &T = grid_of(G, cr.nlat);
  ds[p] = detj[p] * T.w[node_ll[p] & 0xffff];
}
__global__ void k_cellrec_area(int nrbc, CellRec *cells, const double *__restrict__ area,
                               const double *__restrict__ mesh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrbc) return;
  cells[i].area = area[i];
  cells[i].meshSize = mesh[i];
}

// per-cell tables of the context's cells (offsets, grids, reference cells), rebuilt when
// the cells are set again
static int mech_tables(rbc3d_ctx *c, CellBatch *cb) {
  std::vector<int> nl(c->nrbc), nl0(c->nrbc), ref(c->nrbc, 0);
  for (int i = 0; i < c->nrbc; i++) {
    nl[i] = c->cells[i].nlat;
    nl0[i] = c->cells[i].nlat0;
  }
  RBC_TRY(cell_batch(c->nrbc, nl.data(), cb));
  if (c->mech_epoch == c->geom_epoch) return RBC3D_OK;
  for (int i = 0; i < c->nrbc; i++)
    if (cb->off[i] != c->cells[i].off) {
      rbc_set_error("cell mechanics: node offsets differ from the cell records");
      return RBC3D_ESTATE;
    }
  hipStream_t st = c->stream;
  const size_t np = cb->np;
  RBC_TRY(c->d_moff.ensure(c->nrbc));
  RBC_TRY(c->d_mnlat.ensure(c->nrbc));
  RBC_TRY(c->d_mnlat0.ensure(c->nrbc));
  HIP_TRY(hipMemcpyAsync(c->d_moff.p, cb->off.data(), sizeof(int) * c->nrbc, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_mnlat.p, nl.data(), sizeof(int) * c->nrbc, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_mnlat0.p, nl0.data(), sizeof(int) * c->nrbc, hipMemcpyHostToDevice, st));
  for (DBuf<double> *b : {&c->d_mx, &c->d_ma1, &c->d_ma2, &c->d_ma3, &c->d_mf})
    RBC_TRY(b->ensure(3 * np + 1));
  for (DBuf<double> *b : {&c->d_ma, &c->d_mar, &c->d_mb}) RBC_TRY(b->ensure(4 * np + 1));
  RBC_TRY(c->d_mdetj.ensure(np + 1));
  RBC_TRY(c->d_mscr.ensure(MECH_SCR * np + 1));
  RBC_TRY(c->d_mxc.ensure(3 * (size_t)c->nrbc + 1));
  for (DBuf<double> *b : {&c->d_marea, &c->d_mvol, &c->d_mmesh}) RBC_TRY(b->ensure(c->nrbc + 1));
  RBC_TRY(c->d_vel.ensure(3 * np + 1));
  HIP_TRY(hipStreamSynchronize(st));
  c->mech_epoch = c->geom_epoch;
  return RBC3D_OK;
}

// RBC_ComputeGeometry of every cell from d_x: the block geometry, a3, detJ, dS and the
// cell records' area and meshSize.  With sources, also what Compute_Rbc_Vel does before
// its forces (common/ModTimeInt.F90:737-745): Rbc_BuildSurfaceSource(xFlag),
// SourceList_UpdateCoord, TargetList_Update -- the x/a3/detJ splines and the source and
// cell lists, on the device, under a new geometry epoch.
int rbc_cells_geometry(rbc3d_ctx *c, bool sources) {
  if (!c->nrbc) return RBC3D_OK;
  CellBatch cb;
  RBC_TRY(mech_tables(c, &cb));
  hipStream_t st = c->stream;
  const int np = c->np, nb = (np + 255) / 256;
  k_nodes_to_blocks<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_x.p, 3, c->d_mx.p);
  GeoOut o;
  o.x = c->d_mx.p;
  o.a1 = c->d_ma1.p;
  o.a2 = c->d_ma2.p;
  o.a3 = c->d_ma3.p;
  o.a = c->d_ma.p;
  o.ar = c->d_mar.p;
  o.b = c->d_mb.p;
  o.detj = c->d_mdetj.p;
  o.xc = c->d_mxc.p;
  o.area = c->d_marea.p;
  o.vol = c->d_mvol.p;
  o.mesh = c->d_mmesh.p;
  o.scr = c->d_mscr.p;
  RBC_TRY(rbc_rbc_geometry_dev(st, c->nrbc, c->d_moff.p, c->d_mnlat.p, cb.G, cb.lds, o));
  k_blocks_to_nodes<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_ma3.p, 3, 1.0, c->d_a3.p);
  k_blocks_to_nodes<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_mdetj.p, 1, 1.0,
                                        c->d_detj.p);
  k_ds_nodes<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_node_ll.p, c->d_cells.p, cb.G, c->d_detj.p,
                                 c->d_ds.p);
  k_cellrec_area<<<(c->nrbc + 255) / 256, 256, 0, st>>>(c->nrbc, c->d_cells.p, c->d_marea.p,
                                                        c->d_mmesh.p);
  HIP_TRY(hipGetLastError());
  std::vector<double> ar(c->nrbc), me(c->nrbc);
  HIP_TRY(hipMemcpyAsync(ar.data(), c->d_marea.p, sizeof(double) * c->nrbc, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(me.data(), c->d_mmesh.p, sizeof(double) * c->nrbc, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < c->nrbc; i++) {
    if (!std::isfinite(ar[i])) {
      rbc_set_error("cell geometry: non-finite area (the cell surface degenerated)");
      return RBC3D_ENONFINITE;
    }
    c->cells[i].area = ar[i];
    c->cells[i].meshSize = me[i];
  }
  if (!sources) {
    c->rbc_geom_ready = false;  // the splines and source list are not those of d_x
    return RBC3D_OK;
  }
  RBC_TRY(rbc_sh_build_splines(c, 0));
  RBC_TRY(rbc_build_source_list(c));
  c->geom_epoch++;
  c->mech_epoch = c->geom_epoch;  // same cells: the tables stay valid
  c->rbc_geom_ready = true;
  c->x_moved = false;
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

// rbc%f = -Shell_ResForce(rbc, rbcRef), AddIntraCellRepulsionForce (when the repulsion is
// on), RBC_SphProject(rbc, 3, rbc%f) and the f spline and source densities
// (common/ModTimeInt.F90:747-769)
int rbc_cells_force(rbc3d_ctx *c) {
  if (!c->nrbc) return RBC3D_OK;
  if (!c->rbc_geom_ready) {
    rbc_set_error("cell force before the cell geometry");
    return RBC3D_ESTATE;
  }
  if (c->nref == 0) {
    rbc_set_error("cell force: no reference cells (rbc3d_set_ref_cells)");
    return RBC3D_ESTATE;
  }
  CellBatch cb;
  RBC_TRY(mech_tables(c, &cb));
  hipStream_t st = c->stream;
  const int np = c->np;
  std::vector<int> ref(c->nrbc);
  std::vector<double> mod(3 * (size_t)c->nrbc);
  for (int i = 0; i < c->nrbc; i++) {
    const int k = c->cells[i].celltype - 1;
    if (k >= c->nref || c->ref_nlat[k] != c->cells[i].nlat) {
      rbc_set_error("cell force: cell " + std::to_string(i) + " has no reference cell of its grid");
      return RBC3D_EINVAL;
    }
    ref[i] = c->ref_off[k];
    mod[i] = c->ES[k];
    mod[c->nrbc + i] = c->ED[k];
    mod[2 * c->nrbc + i] = c->EB[k];
  }
  RBC_TRY(c->d_mref.ensure(c->nrbc));
  RBC_TRY(c->d_mmod.ensure(3 * (size_t)c->nrbc));
  HIP_TRY(hipMemcpyAsync(c->d_mref.p, ref.data(), sizeof(int) * c->nrbc, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_mmod.p, mod.data(), sizeof(double) * mod.size(), hipMemcpyHostToDevice, st));
  ForceArgs o;
  o.off = c->d_moff.p;
  o.nlat = c->d_mnlat.p;
  o.ref_of = c->d_mref.p;
  o.a1 = c->d_ma1.p;
  o.a2 = c->d_ma2.p;
  o.a3 = c->d_ma3.p;
  o.a = c->d_ma.p;
  o.ar = c->d_mar.p;
  o.b = c->d_mb.p;
  o.detj = c->d_mdetj.p;
  o.area = c->d_marea.p;
  o.ar_ref = c->d_ref_ar.p;
  o.b_ref = c->d_ref_b.p;
  o.ES = c->d_mmod.p;
  o.ED = c->d_mmod.p + c->nrbc;
  o.EB = c->d_mmod.p + 2 * c->nrbc;
  o.f = c->d_mf.p;
  o.scr = c->d_mscr.p;
  hipLaunchKernelGGL(k_shell_res_force, dim3(mech_blocks(c->nrbc)), dim3(MECH_T), cb.lds, st, c->nrbc,
                     cb.G, o);
  // rbc%f = -rbc%f, AddIntraCellRepulsionForce, then the projection; without the
  // repulsion force the sign is applied after the (linear) projection instead
  const bool intra = c->rep_on && c->force_coef != 0;
  if (intra) RBC_TRY(rbc_intra_repulsion(c, 1, c->d_mf.p, nullptr));
  hipLaunchKernelGGL(k_rbc_sph_project, dim3(mech_blocks(c->nrbc)), dim3(MECH_T), cb.lds, st, c->nrbc,
                     c->d_moff.p, c->d_mnlat.p, c->d_mnlat0.p, 3, cb.G, c->d_mf.p);
  const int nb = (np + 255) / 256;
  k_blocks_to_nodes<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_mf.p, 3,
                                        intra ? 1.0 : -1.0, c->d_f.p);
  HIP_TRY(hipGetLastError());
  RBC_TRY(rbc_update_density(c, 1));
  return RBC3D_OK;
}

// FilterRbcs (common/ModTimeInt.F90:908-934): RBC_SphProject of every cell's coordinates
int rbc_cells_project_x(rbc3d_ctx *c) {
  if (!c->nrbc) return RBC3D_OK;
  CellBatch cb;
  RBC_TRY(mech_tables(c, &cb));
  hipStream_t st = c->stream;
  const int np = c->np, nb = (np + 255) / 256;
  k_nodes_to_blocks<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_x.p, 3, c->d_mx.p);
  hipLaunchKernelGGL(k_rbc_sph_project, dim3(mech_blocks(c->nrbc)), dim3(MECH_T), cb.lds, st, c->nrbc,
                     c->d_moff.p, c->d_mnlat.p, c->d_mnlat0.p, 3, cb.G, c->d_mx.p);
  k_blocks_to_nodes<<<nb, 256, 0, st>>>(np, c->d_node_cell.p, c->d_cells.p, c->d_mx.p, 3, 1.0, c->d_x.p);
  HIP_TRY(hipGetLastError());
  c->rbc_geom_ready = false;
  c->x_moved = true;
  return RBC3D_OK;
}

// rbcRefs (common/ModTimeInt.F90:63-76): one reference cell per cell type, its geometry
// from x_ref, and the membrane moduli of the type
extern "C" int rbc3d_set_ref_cells(rbc3d_ctx *c, int ntypes, const int *nlat, const double *x,
                                   const double *ES, const double *ED, const double *EB) {
  if (!c || ntypes < 1 || ntypes > 8 || !nlat || !x || !ES || !ED || !EB) return RBC3D_EINVAL;
  size_t np = 0;
  std::vector<int> off(ntypes);
  for (int k = 0; k < ntypes; k++) {
    off[k] = (int)np;
    np += 2 * (size_t)nlat[k] * nlat[k];
  }
  std::vector<double> a1(3 * np), a2(3 * np), a3(3 * np), a(4 * np), ar(4 * np), b(4 * np), dj(np),
      xc(3 * (size_t)ntypes), area(ntypes), vol(ntypes), mesh(ntypes);
  RBC_TRY(rbc3d_rbc_compute_geometry(ntypes, nlat, x, a1.data(), a2.data(), a3.data(), a.data(),
                                     ar.data(), b.data(), dj.data(), xc.data(), area.data(),
                                     vol.data(), mesh.data()));
  RBC_TRY(c->d_ref_ar.ensure(4 * np));
  RBC_TRY(c->d_ref_b.ensure(4 * np));
  HIP_TRY(hipMemcpy(c->d_ref_ar.p, ar.data(), ar.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_ref_b.p, b.data(), b.size() * 8, hipMemcpyHostToDevice));
  c->nref = ntypes;
  c->ref_nlat.assign(nlat, nlat + ntypes);
  c->ref_off = off;
  c->ref_vol = vol;
  c->ES.assign(ES, ES + ntypes);
  c->ED.assign(ED, ED + ntypes);
  c->EB.assign(EB, EB + ntypes);
  return RBC3D_OK;
}
