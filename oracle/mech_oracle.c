/*
 * Test infrastructure: the CPU restatement of the reference's per-cell surface geometry
 * and membrane mechanics, the parity oracle of the GPU kernels in
 * rbc3d_amd/csrc/mech.hip.  Plain C, FP64, no FMA contraction (see Makefile).  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU baseline load it.
 *
 * Restated (each function cites the reference routine it follows):
 *   RBC_ComputeGeometry   common/ModRbc.F90:404-509
 *   RBC_CovarGrad_Vec     common/ModRbc.F90:527-575
 *   RBC_CovarGrad_Tensor  common/ModRbc.F90:586-654
 *   RBC_SphProject        common/ModRbc.F90:659-686 (ShFilter, common/ModSphpk.F90:423-432)
 *   Shell_ResForce        common/ModRbc.F90:815-900
 *   Shell_ElasStrs        common/ModRbc.F90:908-944
 *   Shell_Bend            common/ModRbc.F90:951-969
 *
 * Arrays are the reference's Fortran arrays of one cell: x(nlat,nlon,3),
 * a(nlat,nlon,2,2) (column-major, 0-based here), nlon = 2 nlat.  The spherical-harmonic
 * steps call the oracle's ShAnalGau / ShGradGau / ShSynthGau (rbc3d_oracle.c).
 *
 * Pins (tests/test_mech_oracle.py): a sphere's exact geometry (normals, detJ = R^2,
 * curvature tensor b = -a / R, area, volume, centroid); an ellipsoid's exact volume and
 * its normals against the gradient of the implicit surface; covariant gradients of
 * fields with closed-form derivatives; a stress-free reference state giving zero force;
 * the Laplace law 2 tau / R of a uniformly stretched sphere under the Skalak law; zero
 * net force and torque on deformed shapes; and the projection's idempotence.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "rbc3d_oracle.h"

#define TWO_PI_M 6.28318530717958647692

#define IX(i, j, d) ((i) + nlat * ((j) + nlon * (d)))
#define IT(i, j, p, q) ((i) + nlat * ((j) + nlon * ((p) + 2 * (q))))

/* v = dg/dth, w = dg/dphi of nt fields g(nlat,nlon,nt) (ShAnalGau + ShGradGau with
 * coefficient arrays (nlon/2+1, nlat, nt), as RBC_ComputeGeometry and the covariant
 * gradients declare them) */
static int field_grad(int nlat, int nt, const double *g, double *v, double *w) {
  const int nlon = 2 * nlat, ida = nlon / 2 + 1, jda = nlat;
  double *a = calloc((size_t)ida * jda * nt, sizeof(double));
  double *b = calloc((size_t)ida * jda * nt, sizeof(double));
  int r = orc_sh_anal_gau(nlat, nlon, nt, g, nlat, nlon, a, b, ida, jda);
  if (!r) r = orc_sh_grad_gau(nlat, nlon, nt, v, w, nlat, nlon, a, b, ida, jda);
  free(a);
  free(b);
  return r;
}

/* Gauss weights times 2 pi / nlon (RBC_Create, common/ModRbc.F90:93-94) */
static void cell_weights(int nlat, double *th, double *w) {
  orc_gaqd(nlat, th, w);
  for (int i = 0; i < nlat; i++) w[i] = (TWO_PI_M / (2 * nlat)) * w[i];
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* RBC_ComputeGeometry (common/ModRbc.F90:404-509) */
int orc_rbc_geometry(int nlat, const double *x, double *a1, double *a2, double *a3, double *a,
                     double *a_rcp, double *b, double *detj, double *xc, double *area,
                     double *vol, double *mesh) {
  const int nlon = 2 * nlat;
  const size_t np = (size_t)nlat * nlon;
  double *th = malloc(sizeof(double) * nlat), *w = malloc(sizeof(double) * nlat);
  cell_weights(nlat, th, w);
  int r = field_grad(nlat, 3, x, a1, a2);
  /* surface tangents and normals (:423-455) */
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      double t1[3], t2[3], n[3];
      for (int d = 0; d < 3; d++) {
        t1[d] = a1[IX(i, j, d)];
        t2[d] = a2[IX(i, j, d)];
      }
      double m11 = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
      double m12 = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
      double m22 = t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2];
      double detA = m11 * m22 - m12 * m12;
      double idetA = 1. / detA;
      a[IT(i, j, 0, 0)] = m11;
      a[IT(i, j, 0, 1)] = m12;
      a[IT(i, j, 1, 0)] = m12;
      a[IT(i, j, 1, 1)] = m22;
      a_rcp[IT(i, j, 0, 0)] = idetA * m22;
      a_rcp[IT(i, j, 0, 1)] = -idetA * m12;
      a_rcp[IT(i, j, 1, 0)] = -idetA * m12;
      a_rcp[IT(i, j, 1, 1)] = idetA * m11;
      detj[i + (size_t)nlat * j] = sqrt(detA) / sin(th[i]);
      cross3(t1, t2, n);
      double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int d = 0; d < 3; d++) a3[IX(i, j, d)] = n[d] / nn;
    }
  /* curvature tensor (:458-476) */
  double *a31 = malloc(sizeof(double) * 3 * np), *a32 = malloc(sizeof(double) * 3 * np);
  if (!r) r = field_grad(nlat, 3, a3, a31, a32);
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      double s11 = 0, s12 = 0, s21 = 0, s22 = 0;
      for (int d = 0; d < 3; d++) {
        s11 += a1[IX(i, j, d)] * a31[IX(i, j, d)];
        s12 += a1[IX(i, j, d)] * a32[IX(i, j, d)];
        s21 += a2[IX(i, j, d)] * a31[IX(i, j, d)];
        s22 += a2[IX(i, j, d)] * a32[IX(i, j, d)];
      }
      b[IT(i, j, 0, 0)] = -s11;
      b[IT(i, j, 0, 1)] = -s12;
      b[IT(i, j, 1, 0)] = -s21;
      b[IT(i, j, 1, 1)] = -s22;
    }
  /* centroid and area (:478-491), volume (:493-501), mesh size (:504) */
  double ar = 0, c[3] = {0, 0, 0};
  for (int i = 0; i < nlat; i++)
    for (int j = 0; j < nlon; j++) {
      double ds = detj[i + (size_t)nlat * j] * w[i];
      ar += ds;
      for (int d = 0; d < 3; d++) c[d] += ds * x[IX(i, j, d)];
    }
  for (int d = 0; d < 3; d++) c[d] = c[d] / ar;
  double v = 0;
  for (int j = 0; j < nlon; j++)
    for (int i = 0; i < nlat; i++) {
      double ds = detj[i + (size_t)nlat * j] * w[i], s = 0;
      for (int d = 0; d < 3; d++) s += a3[IX(i, j, d)] * (x[IX(i, j, d)] - c[d]);
      v += ds * s;
    }
  if (xc)
    for (int d = 0; d < 3; d++) xc[d] = c[d];
  if (area) *area = ar;
  if (vol) *vol = (1.0 / 3) * v;  /* THRD * vol */
  if (mesh) *mesh = sqrt(ar / ((double)nlat * nlat));
  free(a31);
  free(a32);
  free(th);
  free(w);
  return r;
}

/* a_i^rcp of one point: a1_rcp = a_rcp(1,1) a1 + a_rcp(1,2) a2 (:437-438) */
static void rcp_vectors(int nlat, int i, int j, const double *a1, const double *a2,
                        const double *a_rcp, double *r1, double *r2) {
  const int nlon = 2 * nlat;
  for (int d = 0; d < 3; d++) {
    r1[d] = a_rcp[IT(i, j, 0, 0)] * a1[IX(i, j, d)] + a_rcp[IT(i, j, 0, 1)] * a2[IX(i, j, d)];
    r2[d] = a_rcp[IT(i, j, 1, 0)] * a1[IX(i, j, d)] + a_rcp[IT(i, j, 1, 1)] * a2[IX(i, j, d)];
  }
}

/* RBC_CovarGrad_Tensor (common/ModRbc.F90:586-654): t(nlat,nlon,2,2) -> dt(nlat,nlon,2,2,2),
 * dt(.,.,i,j,k) = t^{ij}_k */
static int covar_grad_tensor(int nlat, const double *a1, const double *a2, const double *a_rcp,
                             const double *t, double *dt) {
  const int nlon = 2 * nlat;
  const size_t np = (size_t)nlat * nlon;
  double *tc = calloc(9 * np, sizeof(double)), *tc1 = malloc(sizeof(double) * 9 * np),
         *tc2 = malloc(sizeof(double) * 9 * np);
  /* tc(i,j) = t11 a1_i a1_j + t12 a1_i a2_j + t21 a2_i a1_j + t22 a2_i a2_j, the field
   * tc(:,:,i,j) at component i + 3 j */
  for (int jl = 0; jl < nlon; jl++)
    for (int il = 0; il < nlat; il++) {
      const double t11 = t[IT(il, jl, 0, 0)], t12 = t[IT(il, jl, 0, 1)],
                   t21 = t[IT(il, jl, 1, 0)], t22 = t[IT(il, jl, 1, 1)];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          const double x1i = a1[IX(il, jl, i)], x2i = a2[IX(il, jl, i)];
          const double x1j = a1[IX(il, jl, j)], x2j = a2[IX(il, jl, j)];
          tc[il + nlat * (jl + (size_t)nlon * (i + 3 * j))] =
              t11 * x1i * x1j + t12 * x1i * x2j + t21 * x2i * x1j + t22 * x2i * x2j;
        }
    }
  int r = field_grad(nlat, 9, tc, tc1, tc2);
  for (int jl = 0; jl < nlon && !r; jl++)
    for (int il = 0; il < nlat; il++) {
      double rc[2][3];
      rcp_vectors(nlat, il, jl, a1, a2, a_rcp, rc[0], rc[1]);
      for (int k = 0; k < 2; k++) {
        const double *T = k == 0 ? tc1 : tc2;
        for (int p = 0; p < 2; p++)
          for (int q = 0; q < 2; q++) {
            /* dot_product(a_p^rcp, matmul(tc_k, a_q^rcp)) */
            double s = 0;
            for (int i = 0; i < 3; i++) {
              double y = 0;
              for (int j = 0; j < 3; j++)
                y += T[il + nlat * (jl + (size_t)nlon * (i + 3 * j))] * rc[q][j];
              s += rc[p][i] * y;
            }
            dt[il + nlat * (jl + (size_t)nlon * (p + 2 * (q + 2 * k)))] = s;
          }
      }
    }
  free(tc);
  free(tc1);
  free(tc2);
  return r;
}

/* RBC_CovarGrad_Vec (common/ModRbc.F90:527-575): v(nlat,nlon,2) -> dv(nlat,nlon,2,2) */
static int covar_grad_vec(int nlat, const double *a1, const double *a2, const double *a_rcp,
                          const double *v, double *dv) {
  const int nlon = 2 * nlat;
  const size_t np = (size_t)nlat * nlon;
  double *vc = malloc(sizeof(double) * 3 * np), *vc1 = malloc(sizeof(double) * 3 * np),
         *vc2 = malloc(sizeof(double) * 3 * np);
  for (int j = 0; j < nlon; j++)
    for (int i = 0; i < nlat; i++)
      for (int d = 0; d < 3; d++)
        vc[IX(i, j, d)] = v[i + nlat * (j + (size_t)nlon * 0)] * a1[IX(i, j, d)] +
                          v[i + nlat * (j + (size_t)nlon * 1)] * a2[IX(i, j, d)];
  int r = field_grad(nlat, 3, vc, vc1, vc2);
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      double rc[2][3];
      rcp_vectors(nlat, i, j, a1, a2, a_rcp, rc[0], rc[1]);
      for (int p = 0; p < 2; p++)
        for (int k = 0; k < 2; k++) {
          const double *V = k == 0 ? vc1 : vc2;
          double s = 0;
          for (int d = 0; d < 3; d++) s += rc[p][d] * V[IX(i, j, d)];
          dv[IT(i, j, p, k)] = s;
        }
    }
  free(vc);
  free(vc1);
  free(vc2);
  return r;
}

/* 2 x 2 column-major product c = a b */
static void mm2(const double *a, const double *b, double *c) {
  c[0] = a[0] * b[0] + a[2] * b[1];
  c[1] = a[1] * b[0] + a[3] * b[1];
  c[2] = a[0] * b[2] + a[2] * b[3];
  c[3] = a[1] * b[2] + a[3] * b[3];
}
static void get2(int nlat, const double *t, int i, int j, double *m) {
  const int nlon = 2 * nlat;
  m[0] = t[IT(i, j, 0, 0)];
  m[1] = t[IT(i, j, 1, 0)];
  m[2] = t[IT(i, j, 0, 1)];
  m[3] = t[IT(i, j, 1, 1)];
}
static void put2(int nlat, double *t, int i, int j, const double *m) {
  const int nlon = 2 * nlat;
  t[IT(i, j, 0, 0)] = m[0];
  t[IT(i, j, 1, 0)] = m[1];
  t[IT(i, j, 0, 1)] = m[2];
  t[IT(i, j, 1, 1)] = m[3];
}

/* Shell_ResForce (common/ModRbc.F90:815-900) with Shell_ElasStrs (:908-944) and
 * Shell_Bend (:951-969); f(nlat,nlon,3) in Cartesian components */
int orc_shell_res_force(int nlat, const double *a1, const double *a2, const double *a3,
                        const double *a, const double *a_rcp, const double *b, const double *detj,
                        double area, const double *a_rcp_ref, const double *b_ref, double ES,
                        double ED, double EB, double *f) {
  const int nlon = 2 * nlat;
  const size_t np = (size_t)nlat * nlon;
  double *tau_t = malloc(sizeof(double) * 4 * np), *bnd = malloc(sizeof(double) * 4 * np);
  double *dtau_t = malloc(sizeof(double) * 8 * np), *dbnd = malloc(sizeof(double) * 8 * np);
  double *tau_n = malloc(sizeof(double) * 2 * np), *dtau_n = malloc(sizeof(double) * 4 * np);
  double *th = malloc(sizeof(double) * nlat), *w = malloc(sizeof(double) * nlat);
  cell_weights(nlat, th, w);
  for (size_t k = 0; k < 3 * np; k++) f[k] = 0.;
  for (int i = 0; i < nlat; i++)
    for (int j = 0; j < nlon; j++) {
      double A[4], Ar[4], Aref[4], V2[4], B[4], Bref[4], bm[4], bmr[4], d[4], t[4];
      get2(nlat, a, i, j, A);
      get2(nlat, a_rcp, i, j, Ar);
      get2(nlat, a_rcp_ref, i, j, Aref);
      /* Shell_ElasStrs: V2 = a . a_rcp_ref, Skalak law */
      mm2(A, Aref, V2);
      const double lbd1 = V2[0] + V2[3] - 2.0;
      const double lbd2 = V2[0] * V2[3] - V2[2] * V2[1] - 1.0;
      const double JS = sqrt(V2[0] * V2[3] - V2[2] * V2[1]);
      for (int k = 0; k < 4; k++)
        t[k] = 0.5 * ES / JS * (lbd1 + 1.0) * Aref[k] + 0.5 * JS * (ED * lbd2 - ES) * Ar[k];
      put2(nlat, tau_t, i, j, t);
      /* Shell_Bend: bnd = -EB (a_rcp b - a_rcp_ref b_ref) a_rcp */
      get2(nlat, b, i, j, B);
      get2(nlat, b_ref, i, j, Bref);
      mm2(Ar, B, bm);
      mm2(Aref, Bref, bmr);
      for (int k = 0; k < 4; k++) d[k] = bm[k] - bmr[k];
      mm2(d, Ar, t);
      for (int k = 0; k < 4; k++) t[k] = -EB * t[k];
      put2(nlat, bnd, i, j, t);
    }
  /* normal stress from the bending moment (:845-851) */
  int r = covar_grad_tensor(nlat, a1, a2, a_rcp, bnd, dbnd);
#define D5(arr, i, j, p, q, k) arr[(i) + nlat * ((j) + (size_t)nlon * ((p) + 2 * ((q) + 2 * (k))))]
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      tau_n[i + nlat * (j + (size_t)nlon * 0)] = D5(dbnd, i, j, 0, 0, 0) + D5(dbnd, i, j, 1, 0, 1);
      tau_n[i + nlat * (j + (size_t)nlon * 1)] = D5(dbnd, i, j, 0, 1, 0) + D5(dbnd, i, j, 1, 1, 1);
    }
  /* tangential residual force (:853-866) */
  if (!r) r = covar_grad_tensor(nlat, a1, a2, a_rcp, tau_t, dtau_t);
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      double Ar[4], B[4], tn[2], y[2];
      double f1 = -D5(dtau_t, i, j, 0, 0, 0) - D5(dtau_t, i, j, 1, 0, 1);
      double f2 = -D5(dtau_t, i, j, 0, 1, 0) - D5(dtau_t, i, j, 1, 1, 1);
      get2(nlat, a_rcp, i, j, Ar);
      get2(nlat, b, i, j, B);
      tn[0] = tau_n[i + nlat * (j + (size_t)nlon * 0)];
      tn[1] = tau_n[i + nlat * (j + (size_t)nlon * 1)];
      /* matmul(a_rcp, matmul(b, tau_n)) */
      y[0] = B[0] * tn[0] + B[2] * tn[1];
      y[1] = B[1] * tn[0] + B[3] * tn[1];
      f1 = f1 + (Ar[0] * y[0] + Ar[2] * y[1]);
      f2 = f2 + (Ar[1] * y[0] + Ar[3] * y[1]);
      f[IX(i, j, 0)] = f1;
      f[IX(i, j, 1)] = f2;
    }
#undef D5
  /* normal residual force (:868-876) */
  if (!r) r = covar_grad_vec(nlat, a1, a2, a_rcp, tau_n, dtau_n);
  for (int j = 0; j < nlon && !r; j++)
    for (int i = 0; i < nlat; i++) {
      double s = 0; /* sum(tau_t * b) in array element order */
      for (int q = 0; q < 2; q++)
        for (int p = 0; p < 2; p++) s += tau_t[IT(i, j, p, q)] * b[IT(i, j, p, q)];
      f[IX(i, j, 2)] = -dtau_n[IT(i, j, 0, 0)] - dtau_n[IT(i, j, 1, 1)] - s;
    }
  /* subtract the average normal force (:878-885) */
  double fnTot = 0.;
  for (int j = 0; j < nlon; j++)
    for (int i = 0; i < nlat; i++) fnTot += f[IX(i, j, 2)] * detj[i + (size_t)nlat * j] * w[i];
  for (int j = 0; j < nlon; j++)
    for (int i = 0; i < nlat; i++) f[IX(i, j, 2)] = f[IX(i, j, 2)] - fnTot / area;
  /* Cartesian components (:887-894) */
  for (int j = 0; j < nlon; j++)
    for (int i = 0; i < nlat; i++) {
      const double f1 = f[IX(i, j, 0)], f2 = f[IX(i, j, 1)], f3 = f[IX(i, j, 2)];
      for (int d = 0; d < 3; d++)
        f[IX(i, j, d)] = f1 * a1[IX(i, j, d)] + f2 * a2[IX(i, j, d)] + f3 * a3[IX(i, j, d)];
    }
  free(tau_t);
  free(bnd);
  free(dtau_t);
  free(dbnd);
  free(tau_n);
  free(dtau_n);
  free(th);
  free(w);
  return r;
}

/* RBC_SphProject (common/ModRbc.F90:659-686): ShAnalGau into (nlat, nlat, nvar) arrays,
 * ShFilter (modes n >= nlat0 set to zero), ShSynthGau */
int orc_rbc_sph_project(int nlat, int nlat0, int nvar, double *u) {
  const int nlon = 2 * nlat;
  double *a = calloc((size_t)nlat * nlat * nvar, sizeof(double));
  double *b = calloc((size_t)nlat * nlat * nvar, sizeof(double));
  int r = orc_sh_anal_gau(nlat, nlon, nvar, u, nlat, nlon, a, b, nlat, nlat);
  for (int k = 0; k < nvar && !r; k++)
    for (int n = nlat0; n < nlat; n++)
      for (int m = 0; m < nlat; m++) {
        a[m + (size_t)nlat * (n + (size_t)nlat * k)] = 0.;
        b[m + (size_t)nlat * (n + (size_t)nlat * k)] = 0.;
      }
  if (!r) r = orc_sh_synth_gau(nlat, nlon, nvar, u, nlat, nlon, a, b, nlat, nlat);
  free(a);
  free(b);
  return r;
}
