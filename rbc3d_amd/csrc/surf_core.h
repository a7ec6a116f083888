// Surface-geometry device helpers shared by the near-singular corrections (pair.hip),
// the wall integrals (wall.hip) and the repulsion (repulsion.hip): the polar-patch map,
// the spline projection of a point onto a cell surface and the point-triangle distance.
#pragma once

#include "internal.h"

// PolarPatch_Map (common/ModPolarPatch.F90:218-242)
__device__ inline void patch_map(double th0, double phi0, double dth, double dphi, double &th, double &phi) {
  double ct0 = cos(th0), st0 = sin(th0), cp0 = cos(phi0), sp0 = sin(phi0);
  double s1[3] = {ct0 * cp0, ct0 * sp0, -st0};
  double s2[3] = {-sp0, cp0, 0.};
  double x0[3] = {st0 * cp0, st0 * sp0, ct0};
  double xl[3] = {sin(dth) * cos(dphi), sin(dth) * sin(dphi), cos(dth)};
  double x[3];
  for (int d = 0; d < 3; d++) x[d] = xl[0] * s1[d] + xl[1] * s2[d] + xl[2] * x0[d];
  double nrm = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int d = 0; d < 3; d++) x[d] = x[d] / nrm;
  x[2] = fmax(-1., fmin(1., x[2]));
  th = acos(x[2]);
  phi = atan2(x[1], x[0]);
  if (phi < 0) phi += RBC_TWO_PI;
}

// Spline_FindProjection (common/ModSpline.F90:203-269) on the x spline sxp (m x n points)
// of a cell, with its grid type's projection table pj (patch directions and least-squares
// weights): from the seed (th0, phi0), up to three quadratic-fit steps toward the point
// of the surface closest to xt.  On return (th0, phi0) and x0 are that point.
__device__ inline void spline_find_projection(const double *__restrict__ sxp, int m, int n,
                                              const double *__restrict__ pj, const double xt[3],
                                              double &th0, double &phi0, double x0[3]) {
  spline_interp(sxp, m, n, th0, phi0, 3, x0);
#pragma nounroll
  for (int iter = 0; iter < 3; iter++) {
    const double s0 = sin(th0), c0 = cos(th0);
    double xx[3];
    for (int d = 0; d < 3; d++) xx[d] = x0[d] - xt[d];
    const double d20 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
    const double *W = pj + RBC_PJ_W + iter * 17 * 6;
    double c1 = W[1] * d20, c2 = W[2] * d20, c3 = W[3] * d20, c4 = W[4] * d20, c5 = W[5] * d20;
#pragma nounroll
    for (int k = 0; k < 16; k++) {
      // PolarPatch_Build (common/ModPolarPatch.F90:99-148) for patch point k
      const double *dr = pj + RBC_PJ_DIR + (iter * 16 + k) * 3;
      double y0, y1, y2;
      {
#pragma clang fp contract(off)
        y0 = c0 * dr[0] + 0. * dr[1] + s0 * dr[2];
        y1 = 0. * dr[0] + 1. * dr[1] + 0. * dr[2];
        y2 = -s0 * dr[0] + 0. * dr[1] + c0 * dr[2];
      }
      y2 = fmax(-1.0, fmin(1.0, y2));
      double ph = atan2(y1, y0) + phi0;
      const double thP = acos(y2);
      const double phP = ph - floor(ph * (1. / RBC_TWO_PI)) * RBC_TWO_PI;
      spline_interp(sxp, m, n, thP, phP, 3, xx);
      for (int d = 0; d < 3; d++) xx[d] -= xt[d];
      const double fk = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
      const double *w = W + (k + 1) * 6;
      c1 += w[1] * fk;
      c2 += w[2] * fk;
      c3 += w[3] * fk;
      c4 += w[4] * fk;
      c5 += w[5] * fk;
    }
    // Min_Quad_2D (common/ModBasicMath.F90:296-328)
    double l11 = 2. * c3, l12 = c4, l21 = c4, l22 = 2. * c5;
    double det = l11 * l22 - l12 * l21, xm0 = 0., xm1 = 0.;
    if (det > 0) {
      double idet = 1. / det, r1 = -c1, r2 = -c2;
      xm0 = idet * (l22 * r1 - l12 * r2);
      xm1 = idet * (-l21 * r1 + l11 * r2);
    }
    double thMinL = sqrt(xm0 * xm0 + xm1 * xm1), phiMinL = atan2(xm1, xm0);
    double thMin, phiMin, xMin[3];
    patch_map(th0, phi0, thMinL, phiMinL, thMin, phiMin);
    spline_interp(sxp, m, n, thMin, phiMin, 3, xMin);
    for (int d = 0; d < 3; d++) xx[d] = xMin[d] - xt[d];
    double d2Min = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
    if (d2Min > d20) break;
    th0 = thMin;
    phi0 = phiMin;
    for (int d = 0; d < 3; d++) x0[d] = xMin[d];
  }
}

// MinDistToTri (common/ModIntOnWalls.F90:480-577); x[l][d].  Evaluated without FMA
// contraction, operation for operation as the reference, so the in-cutoff test (and with
// it the self-matrix sparsity pattern) is the reference's bit for bit.
__device__ inline double min_dist_to_tri(const double xt[3], const double x[3][3], double &s0, double &t0) {
#pragma clang fp contract(off)
  double x12[3], x13[3], x1T[3];
  for (int d = 0; d < 3; d++) {
    x12[d] = x[1][d] - x[0][d];
    x13[d] = x[2][d] - x[0][d];
    x1T[d] = x[0][d] - xt[d];
  }
  double a = x12[0] * x12[0] + x12[1] * x12[1] + x12[2] * x12[2];
  double b = x12[0] * x13[0] + x12[1] * x13[1] + x12[2] * x13[2];
  double c = x13[0] * x13[0] + x13[1] * x13[1] + x13[2] * x13[2];
  double det = a * c - b * b, invDet = 1. / det;
  double d = x12[0] * x1T[0] + x12[1] * x1T[1] + x12[2] * x1T[2];
  double e = x13[0] * x1T[0] + x13[1] * x1T[1] + x13[2] * x1T[2];
  double f = x1T[0] * x1T[0] + x1T[1] * x1T[1] + x1T[2] * x1T[2];
  double s = b * e - c * d, t = b * d - a * e;
  int reg;
  if (s + t <= det) {
    if (s < 0)
      reg = t < 0 ? 4 : 3;
    else if (t < 0)
      reg = 5;
    else
      reg = 0;
  } else {
    if (s < 0)
      reg = 2;
    else if (t < 0)
      reg = 6;
    else
      reg = 1;
  }
  if (reg == 2)
    reg = (-(c + e) < 0) ? 3 : 1;
  else if (reg == 4)
    reg = (d < 0) ? 5 : 3;
  else if (reg == 6)
    reg = (b + e - a - d < 0) ? 1 : 5;
  if (reg == 0) {
    s = invDet * s;
    t = invDet * t;
  } else if (reg == 1) {
    s = (c + e - b - d) / (a - 2 * b + c);
    s = fmin(1., fmax(0., s));
    t = 1. - s;
  } else if (reg == 3) {
    s = 0.;
    t = -e / c;
    t = fmin(1., fmax(0., t));
  } else if (reg == 5) {
    t = 0.;
    s = -d / a;
    s = fmin(1., fmax(0., s));
  }
  s0 = s;
  t0 = t;
  return sqrt(a * s * s + 2 * b * s * t + c * t * t + 2 * d * s + 2 * e * t + f);
}
