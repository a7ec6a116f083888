/*
 * rbc3d_oracle.c -- CPU FP64 restatement of RBC3D's Ewald-split boundary-integral
 * matvec.  TEST INFRASTRUCTURE: the parity checker and the host-CPU baseline.
 * Nothing in the product path (rbc3d_amd/) links or calls this file.
 *
 * Each routine cites the reference routine it restates (paths relative to the
 * RBC3D source tree).  Third-party arithmetic the reference delegates to
 * SPHEREPACK 3.2 (shags/shsgs/shses/gaqd), FFTW 3.3.10 (PFFTW, FFT_Diff) and
 * LAPACK (dposv) is restated from those libraries' published definitions:
 *   - SPHEREPACK: normalised associated Legendre functions Pbar(m,n) with
 *     int_0^pi Pbar^2 sin = 1, no Condon-Shortley phase; synthesis
 *     g = sum_n .5 a(0,n) Pbar(0,n) + sum_{m>=1} sum_n Pbar(m,n)(a cos m phi - b sin m phi);
 *     Gauss analysis is the Gauss-weighted discrete projection.
 *   - FFTW: unnormalised transforms, sign as requested by the plan; c2r keeps
 *     only the real part of the zero and Nyquist frequencies of the halved axis.
 *   - dposv('U'): Cholesky A = U^T U and two triangular solves.
 *
 * Parity status of every function is listed in the test suite (tests/) and in
 * the oracle section of the design notes.
 */
#include "rbc3d_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_ 3.14159265358979323846
#define TWO_PI (2 * PI_)
#define I_2PI (1. / TWO_PI)
#define NTAB ORC_NTAB
#define NBR_MAX 32

static inline int imod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}
/* Fortran NINT: round half away from zero */
static inline double fnint(double x) { return round(x); }

/* ------------------------------------------------------------------------ */
typedef struct {
  int m, n, nvar;
  double hx, hy, ihx, ihy;
  double *u, *u1, *u2, *u12; /* (0:m-1, 0:n-1, nvar) */
} orc_spline;

typedef struct {
  int nlat, nlon;
  double radius;
  int nrad, nazm;
  double *thL, *phiL, *w;
  double *thG, *phiG; /* (nrad, nazm, nlat, nlon) */
} orc_patch;

typedef struct {
  int id, nlat0, nlat, nlon, celltype;
  double *th, *phi, *w;
  double *x, *a3, *detj, *f, *g;
  double area, meshSize;
  orc_spline sx, sa3, sdetj, sF, sG;
  orc_patch *patch;
  int off; /* node offset in the concatenated node arrays */
} orc_rbc;

typedef struct {
  int nPoint, cap;
  double *x, *f, *g, *a3, *Bcoef;
  int *indx;
  int Nc[3];
  double iLbNc[3];
  int *hoc, *next;
} orc_slist;

typedef struct {
  int id, nvert, nele;
  double *x, *f;
  int *e2v, *v2v, *indxVertGlb;
  double *a3, *area, *epsDist;
  int nnz;
  int *rowptr, *col;
  double *val;
  int ndof;
} orc_wall;

struct orc_sys {
  orc_conf conf;
  double Lb[3], iLb[3];
  double Acoef[ORC_MAXTYPES], Bcoef[ORC_MAXTYPES];
  double sl1[NTAB + 1], sl2[NTAB + 1], dl1[NTAB + 1], r_eps;
  double mtab[NTAB + 1];
  double gq7_rst[7][2], gq7_w[7];
  double duffy_r[4], duffy_w[4];
  int nrbc;
  orc_rbc *rbcs;
  int npatch;
  orc_patch patches[8];
  int nwall;
  orc_wall wall;
  orc_slist slr, slw;
  /* PME */
  int flag_sl, flag_dl;
  double *ff, *tt, *vv;
  double complex *ffC, *ttC, *vvC;
  double *bb, *qq;
  unsigned long long last_checksum;
  int gmres_restart; /* KSPGMRESSetRestart; 30 is PETSc's default */
  int terms;         /* AddIntOnRbcs parts evaluated (ORC_TERM_*), all by default */
  int deflate;       /* the reference's Deflate switch (common/ModVelSolver.F90:103-106) */
  double *sbq;       /* solid-body modes per node: sbq[(m * 3 + d) * nP + node], m < 6 */
};

/* ======================================================================== */
/* Ewald coefficients: common/ModEwaldFunc.F90:86-178 (tables built eagerly,  */
/* from the explicit alpha/rc, instead of lazily on first call :97-107)       */
static void build_ewald_tables(orc_sys *s) {
  double alpha = s->conf.alpha, rc = s->conf.rc;
  for (int i = 0; i <= NTAB; i++) {
    double r_t = sqrt(PI_ / alpha) * ((double)i * rc / NTAB);
    s->sl1[i] = erfc(r_t);
    s->sl2[i] = 2 / sqrt(alpha) * exp(-(r_t * r_t));
    s->dl1[i] = -8 / sqrt(PI_) *
                (exp(-(r_t * r_t)) * (1.5 * r_t + r_t * r_t * r_t) + 0.75 * sqrt(PI_) * erfc(r_t));
  }
  s->r_eps = 1.e-3 * sqrt(alpha / PI_);
}

void orc_ewald_sl(const orc_sys *s, double r, double *A, double *B) {
  /* ModEwaldFunc.F90:109-129 */
  if (r < s->r_eps) {
    *A = 0;
    *B = 0;
    return;
  }
  double sv = NTAB * r / s->conf.rc;
  double fi = floor(sv);
  if (fi >= NTAB) {
    *A = 0;
    *B = 0;
    return;
  }
  int i = (int)fi;
  double c1 = s->sl1[i] * (i + 1 - sv) + s->sl1[i + 1] * (sv - i);
  double c2 = s->sl2[i] * (i + 1 - sv) + s->sl2[i + 1] * (sv - i);
  double ir = 1. / r, ir2 = ir * ir;
  *A = c1 * ir * ir2 + c2 * ir2;
  *B = c1 * ir - c2;
}

void orc_ewald_dl(const orc_sys *s, double r, double *A) {
  /* ModEwaldFunc.F90:164-176 */
  if (r < s->r_eps) {
    *A = 0;
    return;
  }
  double sv = NTAB * r / s->conf.rc;
  double fi = floor(sv);
  if (fi >= NTAB) {
    *A = 0;
    return;
  }
  int i = (int)fi;
  double c1 = s->dl1[i] * (i + 1 - sv) + s->dl1[i + 1] * (sv - i);
  double r2 = r * r;
  *A = c1 / (r2 * r2 * r);
}

/* ======================================================================== */
/* MaskFunc_Exact / MaskFunc: common/ModBasicMath.F90:332-379.  The table is  */
/* built once (the reference's table lacks SAVE, :355; this is the intended   */
/* behaviour).                                                                */
static double mask_exact(double x) {
  double t = fabs(x);
  if (t < 0.01) return 1.;
  if (t > 0.99) return 0.;
  return exp(2 * (exp(-1. / t)) / (t - 1));
}
static void build_mask_table(orc_sys *s) {
  for (int i = 0; i <= NTAB; i++) s->mtab[i] = mask_exact((double)i / NTAB);
}
double orc_mask(const orc_sys *s, double x) {
  double sv = fabs(x) * NTAB;
  double fi = floor(sv);
  if (fi >= NTAB) return 0.;
  int i = (int)fi;
  return s->mtab[i] * (i + 1 - sv) + s->mtab[i + 1] * (sv - i);
}

/* ======================================================================== */
/* BsplineFunc: common/ModBasicMath.F90:392-417                              */
void orc_bspline(double xc, int P, int *imin, double *w) {
  double u[16];
  *imin = (int)floor(xc) - (P - 1);
  u[0] = *imin - (xc - P);
  for (int j = 1; j < P; j++) u[j] = u[j - 1] + 1;
  w[0] = 1.;
  for (int j = 1; j < P; j++) w[j] = 0.;
  for (int pp = 2; pp <= P; pp++) {
    for (int j = pp; j >= 2; j--)
      w[j - 1] = u[j - 1] / (pp - 1.) * w[j - 1] + (pp - u[j - 1]) / (pp - 1.) * w[j - 2];
    w[0] = u[0] / (pp - 1.) * w[0];
  }
}

/* ======================================================================== */
/* GauLeg: common/ModQuadRule.F90:135-208 (Newton iteration per root, each   */
/* root frozen once converged, as the WHERE-masked loop does).               */
void orc_gauleg(double x1, double x2, int n, double *x, double *w) {
  const double EPS = 3.e-14;
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  double *zz = malloc(sizeof(double) * m), *ppv = malloc(sizeof(double) * m);
  for (int j = 0; j < m; j++) {
    double z = cos(PI_ * ((j + 1) - 0.25) / (n + 0.5)), z1, p1 = 0, p2, p3, pp = 0;
    int unfinished = 1;
    for (int its = 1; its <= 10 && unfinished; its++) {
      p1 = 1.0;
      p2 = 0.0;
      for (int jj = 1; jj <= n; jj++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * jj - 1.0) * z * p2 - (jj - 1.0) * p3) / jj;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
      unfinished = fabs(z - z1) > EPS;
    }
    zz[j] = z;
    ppv[j] = pp;
  }
  for (int j = 0; j < m; j++) x[j] = xm - xl * zz[j];
  for (int j = 0; j < m; j++) x[n - 1 - j] = xm + xl * zz[j];
  for (int j = 0; j < m; j++) w[j] = 2.0 * xl / ((1.0 - zz[j] * zz[j]) * ppv[j] * ppv[j]);
  for (int j = 0; j < m; j++) w[n - 1 - j] = w[j];
  free(zz);
  free(ppv);
}

/* GauLeg_Sinh: common/ModQuadRule.F90:223-263 */
void orc_gauleg_sinh(double xmin, double xmax, double a, double b, int n, double *x, double *w) {
  double xm = 0.5 * (xmin + xmax), xl = 0.5 * (xmax - xmin);
  double a0 = (a - xm) / xl, b0 = b * xl;
  double f1 = (1 + a0) / b0, f2 = (1 - a0) / b0;
  double u1 = log(f1 + sqrt(1 + f1 * f1)), u2 = log(f2 + sqrt(1 + f2 * f2));
  double mu = 0.5 * (u1 + u2), eta = 0.5 * (u1 - u2);
  double sg[64];
  orc_gauleg(-1., 1., n, sg, w);
  for (int i = 0; i < n; i++) {
    x[i] = a0 + b0 * sinh(mu * sg[i] - eta);
    w[i] = w[i] * b0 * mu * cosh(mu * sg[i] - eta);
  }
  for (int i = 0; i < n; i++) {
    w[i] = xl * w[i];
    x[i] = xm + xl * x[i];
  }
}

/* ======================================================================== */
/* SPHEREPACK gaqd restated: Gauss-Legendre colatitudes theta_i ascending in */
/* (0,pi) and weights (sum 2) for int f(cos th) sin th dth.  Newton in theta. */
static double legendre_p(int n, double x, double *dp) {
  double p0 = 1, p1 = x;
  if (n == 0) {
    *dp = 0;
    return 1;
  }
  for (int k = 2; k <= n; k++) {
    double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *dp = n * (x * p1 - p0) / (x * x - 1);
  return p1;
}
void orc_gaqd(int nlat, double *th, double *wts) {
  int n = nlat;
  for (int k = 0; k < n; k++) {
    double t = PI_ * (k + 0.75) / (n + 0.5);
    for (int it = 0; it < 100; it++) {
      double dp, p = legendre_p(n, cos(t), &dp);
      double dt = p / (-sin(t) * dp);
      t -= dt;
      if (fabs(dt) < 1e-17) break;
    }
    /* one last Newton step for full accuracy */
    double dp, p = legendre_p(n, cos(t), &dp);
    t -= p / (-sin(t) * dp);
    p = legendre_p(n, cos(t), &dp);
    double st = sin(t);
    th[k] = t;
    wts[k] = 2. / (st * st * dp * dp);
  }
}

/* Normalised associated Legendre functions Pbar(m,n,theta), SPHEREPACK's   */
/* convention.  pbar[i + nlat*(n + nlat*m)] for m<mmax, n<nlat.              */
void orc_pbar_table(int nlat, const double *th, int mmax, double *pbar) {
  memset(pbar, 0, sizeof(double) * nlat * nlat * mmax);
  for (int i = 0; i < nlat; i++) {
    double c = cos(th[i]), s = sin(th[i]);
    double pmm = sqrt(0.5);
    for (int m = 0; m < mmax; m++) {
      if (m > 0) pmm = sqrt((2.0 * m + 1) / (2.0 * m)) * s * pmm;
      if (m >= nlat) break;
      pbar[i + nlat * (m + nlat * m)] = pmm;
      if (m + 1 < nlat) {
        double pm1 = sqrt(2.0 * m + 3) * c * pmm;
        pbar[i + nlat * ((m + 1) + nlat * m)] = pm1;
        double pa = pmm, pb = pm1;
        for (int n = m + 2; n < nlat; n++) {
          double nn = n, mm = m;
          double an = sqrt((4 * nn * nn - 1) / (nn * nn - mm * mm));
          double bn = sqrt(((2 * nn + 1) * ((nn - 1) * (nn - 1) - mm * mm)) /
                           ((2 * nn - 3) * (nn * nn - mm * mm)));
          double pn = an * c * pb - bn * pa;
          pbar[i + nlat * (n + nlat * m)] = pn;
          pa = pb;
          pb = pn;
        }
      }
    }
  }
}

/* shags (SPHEREPACK) via ShAnalGau common/ModSphpk.F90:74-105.
 * g(idg,jdg,nt), a/b(ida,jda,nt); a(m+1,n+1) for m<mmax=min(nlat,nlon/2+1), n<nlat. */
int orc_sh_anal_gau(int nlat, int nlon, int nt, const double *g, int idg, int jdg, double *a,
                    double *b, int ida, int jda) {
  int mmax = nlat < nlon / 2 + 1 ? nlat : nlon / 2 + 1;
  if (ida < mmax || jda < nlat || idg < nlat || jdg < nlon) return 1;
  double *th = malloc(sizeof(double) * nlat), *wt = malloc(sizeof(double) * nlat);
  orc_gaqd(nlat, th, wt);
  double *pb = malloc(sizeof(double) * nlat * nlat * mmax);
  orc_pbar_table(nlat, th, mmax, pb);
  double *C = malloc(sizeof(double) * nlat), *S = malloc(sizeof(double) * nlat);
  for (int k = 0; k < nt; k++) {
    const double *gk = g + (size_t)k * idg * jdg;
    double *ak = a + (size_t)k * ida * jda, *bk = b + (size_t)k * ida * jda;
    for (int m = 0; m < mmax; m++) {
      for (int i = 0; i < nlat; i++) {
        double cs = 0, sn = 0;
        for (int j = 0; j < nlon; j++) {
          double ph = TWO_PI * j / nlon;
          cs += gk[i + idg * j] * cos(m * ph);
          sn += gk[i + idg * j] * sin(m * ph);
        }
        C[i] = cs;
        S[i] = sn;
      }
      for (int n = m; n < nlat; n++) {
        double sa = 0, sb = 0;
        for (int i = 0; i < nlat; i++) {
          double p = pb[i + nlat * (n + nlat * m)] * wt[i];
          sa += p * C[i];
          sb += p * S[i];
        }
        ak[m + ida * n] = (2.0 / nlon) * sa;
        bk[m + ida * n] = -(2.0 / nlon) * sb;
      }
    }
  }
  free(th);
  free(wt);
  free(pb);
  free(C);
  free(S);
  return 0;
}

static int sh_synth(int nlat, int nlon, const double *th, int nt, double *g, int idg, int jdg,
                    const double *a, const double *b, int ida, int jda) {
  int mmax = nlat < nlon / 2 + 1 ? nlat : nlon / 2 + 1;
  if (mmax > ida) mmax = ida;
  int nmax = nlat < jda ? nlat : jda;
  double *pb = malloc(sizeof(double) * nlat * nlat * (mmax > 0 ? mmax : 1));
  orc_pbar_table(nlat, th, mmax, pb);
  double *Ga = malloc(sizeof(double) * nlat * mmax), *Gb = malloc(sizeof(double) * nlat * mmax);
  for (int k = 0; k < nt; k++) {
    double *gk = g + (size_t)k * idg * jdg;
    const double *ak = a + (size_t)k * ida * jda, *bk = b + (size_t)k * ida * jda;
    for (int m = 0; m < mmax; m++)
      for (int i = 0; i < nlat; i++) {
        double sa = 0, sb = 0;
        for (int n = m; n < nmax; n++) {
          double p = pb[i + nlat * (n + nlat * m)];
          sa += p * ak[m + ida * n];
          sb += p * bk[m + ida * n];
        }
        Ga[i + nlat * m] = (m == 0 ? 0.5 : 1.0) * sa;
        Gb[i + nlat * m] = sb;
      }
    for (int i = 0; i < nlat; i++)
      for (int j = 0; j < nlon; j++) {
        double ph = TWO_PI * j / nlon, v = 0;
        for (int m = 0; m < mmax; m++) v += Ga[i + nlat * m] * cos(m * ph) - Gb[i + nlat * m] * sin(m * ph);
        gk[i + idg * j] = v;
      }
  }
  free(pb);
  free(Ga);
  free(Gb);
  return 0;
}

/* shsgs (SPHEREPACK) via ShSynthGau common/ModSphpk.F90:240-270 */
int orc_sh_synth_gau(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                     const double *b, int ida, int jda) {
  if (idg < nlat || jdg < nlon) return 1;
  double *th = malloc(sizeof(double) * nlat), *wt = malloc(sizeof(double) * nlat);
  orc_gaqd(nlat, th, wt);
  int r = sh_synth(nlat, nlon, th, nt, g, idg, jdg, a, b, ida, jda);
  free(th);
  free(wt);
  return r;
}

/* shses (SPHEREPACK) via ShSynthEqu common/ModSphpk.F90:308-338: colatitudes */
/* th_i = i*pi/(nlat-1), i=0..nlat-1, poles included.                        */
int orc_sh_synth_equ(int nlat, int nlon, int nt, double *g, int idg, int jdg, const double *a,
                     const double *b, int ida, int jda) {
  if (idg < nlat || jdg < nlon) return 1;
  double *th = malloc(sizeof(double) * nlat);
  for (int i = 0; i < nlat; i++) th[i] = PI_ * i / (nlat - 1);
  int r = sh_synth(nlat, nlon, th, nt, g, idg, jdg, a, b, ida, jda);
  free(th);
  return r;
}

/* ShGradGau common/ModSphpk.F90:377-417 (gradgs, then w times sin th): v = dg/dth,
 * w = dg/dphi on the Gauss grid, from the shags coefficients of g.  dPbar/dth from
 * (1 - x^2) dP_n^m/dx = (n + m) P_{n-1}^m - n x P_n^m in the normalised functions. */
int orc_sh_grad_gau(int nlat, int nlon, int nt, double *v, double *w, int idv, int jdv,
                    const double *a, const double *b, int ida, int jda) {
  int mmax = nlat < nlon / 2 + 1 ? nlat : nlon / 2 + 1;
  if (mmax > ida) mmax = ida;
  int nmax = nlat < jda ? nlat : jda;
  if (idv < nlat || jdv < nlon) return 1;
  double *th = malloc(sizeof(double) * nlat), *wt = malloc(sizeof(double) * nlat);
  orc_gaqd(nlat, th, wt);
  double *pb = malloc(sizeof(double) * nlat * nlat * mmax);
  orc_pbar_table(nlat, th, mmax, pb);
  for (int k = 0; k < nt; k++) {
    double *vk = v + (size_t)k * idv * jdv, *wk = w + (size_t)k * idv * jdv;
    const double *ak = a + (size_t)k * ida * jda, *bk = b + (size_t)k * ida * jda;
    for (int i = 0; i < nlat; i++) {
      double ct = cos(th[i]), st = sin(th[i]);
      for (int j = 0; j < nlon; j++) {
        double ph = TWO_PI * j / nlon, sv = 0, sw = 0;
        for (int m = 0; m < mmax; m++) {
          double da = 0, db = 0, pa = 0, pbb = 0, h = m == 0 ? 0.5 : 1.0;
          for (int n = m; n < nmax; n++) {
            double p = pb[i + nlat * (n + nlat * m)];
            double dp = n * ct * p;
            if (n > m)
              dp -= sqrt((2.0 * n + 1) * ((double)n * n - (double)m * m) / (2.0 * n - 1)) *
                    pb[i + nlat * ((n - 1) + nlat * m)];
            dp = dp / st;
            da += dp * ak[m + ida * n];
            db += dp * bk[m + ida * n];
            pa += p * ak[m + ida * n];
            pbb += p * bk[m + ida * n];
          }
          sv += h * (da * cos(m * ph) - db * sin(m * ph));
          sw += h * m * (-pa * sin(m * ph) - pbb * cos(m * ph));
        }
        vk[i + idv * j] = sv;
        wk[i + idv * j] = sw;
      }
    }
  }
  free(th);
  free(wt);
  free(pb);
  return 0;
}

/* ======================================================================== */
/* FFT_Diff: common/ModFFT.F90:25-93.  r2c -> multiply by i*k/m (k<m/2,      */
/* Nyquist zeroed) -> c2r, unnormalised FFTW.  Written as the equivalent real */
/* operator u1(p) = -(2/m) sum_j u(j) sum_{k=1}^{m/2-1} k sin(2 pi k (p-j)/m). */
static void diff_matrix(int m, double *D) {
  for (int p = 0; p < m; p++)
    for (int j = 0; j < m; j++) {
      double s = 0;
      for (int k = 1; k < m / 2; k++) s += k * sin(TWO_PI * (double)((long)k * imod(p - j, m)) / m);
      D[p + m * j] = -(2.0 / m) * s;
    }
}
void orc_fft_diff(int m, int n, int nvar, const double *u, double *u1, double *u2) {
  if (u1) {
    double *D = malloc(sizeof(double) * m * m);
    diff_matrix(m, D);
    for (int l = 0; l < nvar; l++)
      for (int j = 0; j < n; j++)
        for (int p = 0; p < m; p++) {
          double s = 0;
          for (int q = 0; q < m; q++) s += D[p + m * q] * u[q + m * (j + n * l)];
          u1[p + m * (j + n * l)] = s;
        }
    free(D);
  }
  if (u2) {
    double *D = malloc(sizeof(double) * n * n);
    diff_matrix(n, D);
    for (int l = 0; l < nvar; l++)
      for (int i = 0; i < m; i++)
        for (int p = 0; p < n; p++) {
          double s = 0;
          for (int q = 0; q < n; q++) s += D[p + n * q] * u[i + m * (q + n * l)];
          u2[i + m * (p + n * l)] = s;
        }
    free(D);
  }
}

/* ======================================================================== */
/* Splines: common/ModSpline.F90                                             */
static void spline_create(orc_spline *sp, int m, int n, int nvar) {
  /* Spline_Create :22-47 */
  sp->m = m;
  sp->n = n;
  sp->nvar = nvar;
  size_t sz = sizeof(double) * m * n * nvar;
  sp->u = calloc(1, sz);
  sp->u1 = calloc(1, sz);
  sp->u2 = calloc(1, sz);
  sp->u12 = calloc(1, sz);
  sp->hx = TWO_PI / m;
  sp->hy = TWO_PI / n;
  sp->ihx = 1. / sp->hx;
  sp->ihy = 1. / sp->hy;
}
static void spline_free(orc_spline *sp) {
  free(sp->u);
  free(sp->u1);
  free(sp->u2);
  free(sp->u12);
  memset(sp, 0, sizeof(*sp));
}
/* Spline_Build_on_Sphere :121-142.  f(0:nlat, 0:nlon-1, nvar) equi-spaced */
static void spline_build_on_sphere(orc_spline *sp, const double *f) {
  int nlat = sp->m / 2, nlon = sp->n, m = sp->m;
  for (int l = 0; l < sp->nvar; l++)
    for (int j = 0; j < nlon; j++)
      for (int i = 0; i < nlat; i++) {
        sp->u[i + m * (j + nlon * l)] = f[i + (nlat + 1) * (j + nlon * l)];
        sp->u[nlat + i + m * (j + nlon * l)] = f[(nlat - i) + (nlat + 1) * (((j + nlon / 2) % nlon) + nlon * l)];
      }
  orc_fft_diff(sp->m, sp->n, sp->nvar, sp->u, sp->u1, sp->u2);
  orc_fft_diff(sp->m, sp->n, sp->nvar, sp->u1, NULL, sp->u12);
}
/* Spline_Interp :150-191 (kx = ky = 0 for sphere splines) */
static void spline_interp(const orc_spline *sp, double x, double y, double *f) {
  double fi1 = floor(x * sp->ihx), fj1 = floor(y * sp->ihy);
  double s = x * sp->ihx - fi1, t = y * sp->ihy - fj1;
  int i1 = imod((int)fi1, sp->m), j1 = imod((int)fj1, sp->n);
  int i2 = imod(i1 + 1, sp->m), j2 = imod(j1 + 1, sp->n);
  double cx[4] = {1 + s * s * (-3. + 2. * s), s * s * (3. - 2. * s), sp->hx * s * (1. + s * (-2. + s)),
                  sp->hx * s * s * (-1. + s)};
  double cy[4] = {1 + t * t * (-3. + 2. * t), t * t * (3. - 2. * t), sp->hy * t * (1. + t * (-2. + t)),
                  sp->hy * t * t * (-1. + t)};
  int m = sp->m, n = sp->n;
  for (int l = 0; l < sp->nvar; l++) {
    size_t o = (size_t)m * n * l;
#define U(a, i, j) a[o + (i) + m * (j)]
    double u[4][4] = {
        {U(sp->u, i1, j1), U(sp->u, i1, j2), U(sp->u2, i1, j1), U(sp->u2, i1, j2)},
        {U(sp->u, i2, j1), U(sp->u, i2, j2), U(sp->u2, i2, j1), U(sp->u2, i2, j2)},
        {U(sp->u1, i1, j1), U(sp->u1, i1, j2), U(sp->u12, i1, j1), U(sp->u12, i1, j2)},
        {U(sp->u1, i2, j1), U(sp->u1, i2, j2), U(sp->u12, i2, j1), U(sp->u12, i2, j2)}};
#undef U
    double v = 0;
    for (int a = 0; a < 4; a++) {
      double r = 0;
      for (int b = 0; b < 4; b++) r += u[a][b] * cy[b];
      v += cx[a] * r;
    }
    f[l] = v;
  }
}

/* ======================================================================== */
/* Polar patches: common/ModPolarPatch.F90                                   */
/* PolarPatch_Build :99-148 */
static void polar_patch_build(double th0, double phi0, int nth, const double *thL, int nphi,
                              const double *phiL, double *thG, double *phiG, int ld) {
  double s0 = sin(th0), c0 = cos(th0);
  for (int i = 0; i < nth; i++) {
    double st = sin(thL[i]), ct = cos(thL[i]);
    for (int j = 0; j < nphi; j++) {
      double sp = sin(phiL[j]), cp = cos(phiL[j]);
      double x0 = st * cp, x1 = st * sp, x2 = ct;
      double y0 = c0 * x0 + 0. * x1 + s0 * x2;
      double y1 = 0. * x0 + 1. * x1 + 0. * x2;
      double y2 = -s0 * x0 + 0. * x1 + c0 * x2;
      y2 = fmax(-1.0, fmin(1.0, y2));
      double ph = atan2(y1, y0) + phi0;
      thG[i + ld * j] = acos(y2);
      phiG[i + ld * j] = ph - floor(ph * I_2PI) * TWO_PI;
    }
  }
}
/* PolarPatch_Map :218-242 */
static void polar_patch_map(double th0, double phi0, double dth, double dphi, double *th, double *phi) {
  double s1[3] = {cos(th0) * cos(phi0), cos(th0) * sin(phi0), -sin(th0)};
  double s2[3] = {-sin(phi0), cos(phi0), 0.};
  double x0[3] = {sin(th0) * cos(phi0), sin(th0) * sin(phi0), cos(th0)};
  double xl[3] = {sin(dth) * cos(dphi), sin(dth) * sin(dphi), cos(dth)};
  double x[3];
  for (int d = 0; d < 3; d++) x[d] = xl[0] * s1[d] + xl[1] * s2[d] + xl[2] * x0[d];
  double nrm = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int d = 0; d < 3; d++) x[d] = x[d] / nrm;
  x[2] = fmax(-1., fmin(1., x[2]));
  *th = acos(x[2]);
  *phi = atan2(x[1], x[0]);
  if (*phi < 0) *phi += TWO_PI;
}
/* DistOnSphere :249-257 */
static double dist_on_sphere(double th0, double phi0, double th1, double phi1) {
  double d = cos(th0 - th1) - sin(th0) * sin(th1) * (1. - cos(phi0 - phi1));
  d = fmin(1., fmax(-1., d));
  return acos(d);
}
/* PolarPatch_FindPoints :161-209 */
static int polar_patch_find_points(double th0, double phi0, double r0, int nth, const double *ths,
                                   int nphi, const double *phis, int *ijs1, int *ijs2) {
  const double eps = 1.e-10;
  double h_phi = TWO_PI / nphi, ih_phi = 1.0 / h_phi;
  double cosr0 = cos(r0);
  int n = 0;
  for (int i = 0; i < nth; i++) {
    if (ths[i] <= th0 - r0) continue;
    if (ths[i] >= th0 + r0) break;
    double dth = ths[i] - th0;
    double dphi = 1. - (cos(dth) - cosr0) / ((sin(ths[i]) + eps) * (sin(th0) + eps));
    dphi = fmax(-1., fmin(1., dphi));
    dphi = acos(dphi);
    int jmin, jmax;
    if (dphi > PI_ - eps) {
      jmin = 1;
      jmax = nphi;
    } else {
      jmin = (int)ceil((phi0 - dphi - phis[0]) * ih_phi) + 1;
      jmax = (int)floor((phi0 + dphi - phis[0]) * ih_phi) + 1;
    }
    for (int j = jmin; j <= jmax; j++) {
      ijs1[n] = i + 1;
      ijs2[n] = imod(j - 1, nphi) + 1;
      n++;
    }
  }
  return n;
}
/* RbcPolarPatch_Create :27-75 */
static void patch_create(orc_sys *s, orc_patch *pt, const orc_rbc *rbc) {
  int nlat = rbc->nlat, nlon = rbc->nlon;
  pt->nlat = nlat;
  pt->nlon = nlon;
  double radius = PI_ / sqrt((double)nlat), h = PI_ / nlat;
  int nrad = 2 * (int)fnint(radius / h), nazm = 2 * nrad;
  pt->radius = radius;
  pt->nrad = nrad;
  pt->nazm = nazm;
  pt->thL = malloc(sizeof(double) * nrad);
  pt->w = malloc(sizeof(double) * nrad);
  pt->phiL = malloc(sizeof(double) * nazm);
  orc_gauleg(0., radius, nrad, pt->thL, pt->w);
  for (int i = 0; i < nrad; i++) {
    pt->w[i] = pt->w[i] * sin(pt->thL[i]) * (TWO_PI / nazm);
    pt->w[i] = pt->w[i] * orc_mask(s, pt->thL[i] / radius);
  }
  for (int j = 0; j < nazm; j++) pt->phiL[j] = j * TWO_PI / nazm;
  size_t per = (size_t)nrad * nazm;
  pt->thG = malloc(sizeof(double) * per * nlat * nlon);
  pt->phiG = malloc(sizeof(double) * per * nlat * nlon);
  for (int ilon = 0; ilon < nlon; ilon++)
    for (int ilat = 0; ilat < nlat; ilat++) {
      size_t o = per * (ilat + (size_t)nlat * ilon);
      polar_patch_build(rbc->th[ilat], rbc->phi[ilon], nrad, pt->thL, nazm, pt->phiL, pt->thG + o,
                        pt->phiG + o, nrad);
    }
}

/* ======================================================================== */
/* QuadFit_2D / Min_Quad_2D: common/ModBasicMath.F90:234-328, with LAPACK   */
/* dposv('U') restated as a Cholesky factorisation + two triangular solves.  */
static void dposv_u(int n, double *A /* n x n col-major */, double *b) {
  /* A = U^T U, upper triangle */
  for (int j = 0; j < n; j++) {
    double s = A[j + n * j];
    for (int k = 0; k < j; k++) s -= A[k + n * j] * A[k + n * j];
    double ujj = sqrt(s);
    A[j + n * j] = ujj;
    for (int i = j + 1; i < n; i++) {
      double t = A[j + n * i];
      for (int k = 0; k < j; k++) t -= A[k + n * j] * A[k + n * i];
      A[j + n * i] = t / ujj;
    }
  }
  for (int i = 0; i < n; i++) { /* U^T y = b */
    double t = b[i];
    for (int k = 0; k < i; k++) t -= A[k + n * i] * b[k];
    b[i] = t / A[i + n * i];
  }
  for (int i = n - 1; i >= 0; i--) { /* U x = y */
    double t = b[i];
    for (int k = i + 1; k < n; k++) t -= A[i + n * k] * b[k];
    b[i] = t / A[i + n * i];
  }
}
static void quad_fit_2d(int np, const double *xy /* (np,2) */, const double *f, double *c) {
  double lhs[36] = {0}, rhs[6] = {0};
  for (int i = 0; i < np; i++) {
    double xi = xy[i], yi = xy[i + np];
    double u[6] = {1., xi, yi, xi * xi, xi * yi, yi * yi};
    for (int a = 0; a < 6; a++)
      for (int b = a; b < 6; b++) lhs[a + 6 * b] += u[a] * u[b];
    for (int a = 0; a < 6; a++) rhs[a] += u[a] * f[i];
  }
  for (int a = 1; a < 6; a++)
    for (int b = 0; b < a; b++) lhs[a + 6 * b] = lhs[b + 6 * a];
  dposv_u(6, lhs, rhs);
  for (int a = 0; a < 6; a++) c[a] = rhs[a];
}
static void min_quad_2d(const double *c, double *xmin) {
  double l11 = 2. * c[3], l12 = c[4], l21 = c[4], l22 = 2. * c[5];
  double det = l11 * l22 - l12 * l21;
  double r1 = -c[1], r2 = -c[2];
  if (det > 0) {
    double idet = 1. / det;
    xmin[0] = idet * (l22 * r1 - l12 * r2);
    xmin[1] = idet * (-l21 * r1 + l11 * r2);
  } else {
    xmin[0] = 0.;
    xmin[1] = 0.;
  }
}

/* Spline_FindProjection: common/ModSpline.F90:203-269 */
static void spline_find_projection(const orc_spline *sp, const double *xtar, double *th0,
                                   double *phi0, double *x0) {
  enum { nth = 2, nphi = 8, NQ = nth * nphi + 1 };
  double thPatL[nth], phiPatL[nphi], thPat[nth * nphi], phiPat[nth * nphi];
  double xy[NQ * 2], d2[NQ], xx[3], c[6], xyMin[2], xMin[3];
  spline_interp(sp, *th0, *phi0, x0);
  double h = fmax(sp->hx, sp->hy);
  for (int iter = 1; iter <= 3; iter++) {
    for (int i = 0; i < nth; i++) thPatL[i] = (i + 1) * h / nth;
    for (int j = 0; j < nphi; j++) phiPatL[j] = j * TWO_PI / nphi;
    polar_patch_build(*th0, *phi0, nth, thPatL, nphi, phiPatL, thPat, phiPat, nth);
    xy[0] = 0.;
    xy[NQ] = 0.;
    spline_interp(sp, *th0, *phi0, xx);
    for (int d = 0; d < 3; d++) xx[d] -= xtar[d];
    d2[0] = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
    for (int ip = 0; ip < nphi; ip++)
      for (int it = 0; it < nth; it++) {
        int i = (it + 1) + nth * ip;
        xy[i] = thPatL[it] * cos(phiPatL[ip]);
        xy[i + NQ] = thPatL[it] * sin(phiPatL[ip]);
        spline_interp(sp, thPat[it + nth * ip], phiPat[it + nth * ip], xx);
        for (int d = 0; d < 3; d++) xx[d] -= xtar[d];
        d2[i] = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
      }
    quad_fit_2d(NQ, xy, d2, c);
    min_quad_2d(c, xyMin);
    double thMinL = sqrt(xyMin[0] * xyMin[0] + xyMin[1] * xyMin[1]);
    double phiMinL = atan2(xyMin[1], xyMin[0]);
    double thMin, phiMin;
    polar_patch_map(*th0, *phi0, thMinL, phiMinL, &thMin, &phiMin);
    spline_interp(sp, thMin, phiMin, xMin);
    for (int d = 0; d < 3; d++) xx[d] = xMin[d] - xtar[d];
    double d2Min = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
    if (d2Min > d2[0]) break;
    *th0 = thMin;
    *phi0 = phiMin;
    for (int d = 0; d < 3; d++) x0[d] = xMin[d];
    h = 0.5 * h;
  }
}

/* ======================================================================== */
/* Hash table: common/ModHashTable.F90                                       */
static void hash_index(const int Nc[3], const double iLbNc[3], const double *x, int *i1, int *i2,
                       int *i3) {
  /* HashTable_Index :67-91, SINGLE_NODE branch */
  *i1 = imod((int)floor(x[0] * iLbNc[0]), Nc[0]) + 1;
  *i2 = imod((int)floor(x[1] * iLbNc[1]), Nc[1]) + 1;
  *i3 = imod((int)floor(x[2] * iLbNc[2]), Nc[2]) + 1;
}
#define HOC(hoc, Nc, a, b, c) hoc[(a) + ((Nc)[0] + 2) * ((b) + ((Nc)[1] + 2) * (c))]
int orc_hashtable_build(const int Nc[3], const double iLbNc[3], int nPoint, const double *x,
                        int ldx, int *hoc, int *next) {
  /* HashTable_Build :23-58 (SINGLE_NODE ghost layers) */
  int n1 = Nc[0] + 2, n2 = Nc[1] + 2, n3 = Nc[2] + 2;
  for (int i = 0; i < n1 * n2 * n3; i++) hoc[i] = -1;
  for (int i = 0; i < nPoint; i++) next[i] = -1;
  for (int i = 0; i < nPoint; i++) {
    double xi[3] = {x[i], x[i + ldx], x[i + 2 * (size_t)ldx]};
    int i1, i2, i3;
    hash_index(Nc, iLbNc, xi, &i1, &i2, &i3);
    if (i3 >= 0 && i3 <= Nc[2] + 1) {
      next[i] = HOC(hoc, Nc, i1, i2, i3);
      HOC(hoc, Nc, i1, i2, i3) = i + 1;
    }
  }
  for (int b = 0; b < n2; b++)
    for (int c = 0; c < n3; c++) {
      HOC(hoc, Nc, 0, b, c) = HOC(hoc, Nc, Nc[0], b, c);
      HOC(hoc, Nc, Nc[0] + 1, b, c) = HOC(hoc, Nc, 1, b, c);
    }
  for (int a = 0; a < n1; a++)
    for (int c = 0; c < n3; c++) {
      HOC(hoc, Nc, a, 0, c) = HOC(hoc, Nc, a, Nc[1], c);
      HOC(hoc, Nc, a, Nc[1] + 1, c) = HOC(hoc, Nc, a, 1, c);
    }
  for (int a = 0; a < n1; a++)
    for (int b = 0; b < n2; b++) {
      HOC(hoc, Nc, a, b, 0) = HOC(hoc, Nc, a, b, Nc[2]);
      HOC(hoc, Nc, a, b, Nc[2] + 1) = HOC(hoc, Nc, a, b, 1);
    }
  return 0;
}

/* HashTable_ComputeNumBlocks :99-126 (one rank: nodeZmin=0, nodeZmax=Lb(3)) */
static void hash_num_blocks(const orc_sys *s, int Nc[3], double iLbNc[3]) {
  for (int d = 0; d < 3; d++) {
    Nc[d] = (int)floor(s->Lb[d] / s->conf.rc);
    if (Nc[d] < 3) Nc[d] = 3;
  }
  iLbNc[0] = s->iLb[0] * Nc[0];
  iLbNc[1] = s->iLb[1] * Nc[1];
  iLbNc[2] = Nc[2] / (s->Lb[2] - 0.);
}

/* ======================================================================== */
orc_sys *orc_create(const orc_conf *conf) {
  orc_sys *s = calloc(1, sizeof(orc_sys));
  s->conf = *conf;
  s->gmres_restart = 30;
  s->terms = ORC_TERM_ALL;
  for (int d = 0; d < 3; d++) {
    s->Lb[d] = conf->Lb[d];
    s->iLb[d] = 1. / conf->Lb[d];
  }
  /* initCOEFs common/ModConf.F90:327-344 */
  for (int l = 0; l < conf->ntypes; l++) {
    if (conf->viscRat[l] > 0) {
      s->Acoef[l] = 1. + conf->viscRat[l];
      s->Bcoef[l] = 1. - conf->viscRat[l];
    } else {
      s->Acoef[l] = 1.;
      s->Bcoef[l] = -1.;
    }
  }
  build_ewald_tables(s);
  build_mask_table(s);
  /* gqTri7: common/ModQuadRule.F90:70-94 */
  {
    double r = (6. - sqrt(15.)) / 21., sv = r, t = 1 - r - sv;
    double rr[3] = {r, sv, t}, ss[3] = {sv, t, r};
    for (int k = 0; k < 3; k++) {
      s->gq7_rst[k][0] = rr[k];
      s->gq7_rst[k][1] = ss[k];
      s->gq7_w[k] = (155. - sqrt(15.)) / 2400.;
    }
    r = (6. + sqrt(15.)) / 21.;
    sv = r;
    t = 1 - r - sv;
    double r2[3] = {r, sv, t}, s2[3] = {sv, t, r};
    for (int k = 0; k < 3; k++) {
      s->gq7_rst[3 + k][0] = r2[k];
      s->gq7_rst[3 + k][1] = s2[k];
      s->gq7_w[3 + k] = (155. + sqrt(15.)) / 2400.;
    }
    s->gq7_rst[6][0] = 1. / 3.;
    s->gq7_rst[6][1] = 1. / 3.;
    s->gq7_w[6] = 9. / 80.;
  }
  orc_gauleg(0., 1., 4, s->duffy_r, s->duffy_w); /* Tri_Int_Duffy :398-404 */
  return s;
}

static void slist_free(orc_slist *l) {
  free(l->x);
  free(l->f);
  free(l->g);
  free(l->a3);
  free(l->Bcoef);
  free(l->indx);
  free(l->hoc);
  free(l->next);
  memset(l, 0, sizeof(*l));
}

static void free_cells(orc_sys *s) {
  for (int i = 0; i < s->nrbc; i++) {
    orc_rbc *r = &s->rbcs[i];
    free(r->th);
    free(r->phi);
    free(r->w);
    free(r->x);
    free(r->a3);
    free(r->detj);
    free(r->f);
    free(r->g);
    spline_free(&r->sx);
    spline_free(&r->sa3);
    spline_free(&r->sdetj);
    spline_free(&r->sF);
    spline_free(&r->sG);
  }
  free(s->rbcs);
  s->rbcs = NULL;
  for (int p = 0; p < s->npatch; p++) {
    free(s->patches[p].thL);
    free(s->patches[p].phiL);
    free(s->patches[p].w);
    free(s->patches[p].thG);
    free(s->patches[p].phiG);
  }
  s->npatch = 0;
  s->nrbc = 0;
  slist_free(&s->slr);
}

static void free_wall(orc_sys *s) {
  orc_wall *w = &s->wall;
  free(w->x);
  free(w->f);
  free(w->e2v);
  free(w->v2v);
  free(w->indxVertGlb);
  free(w->a3);
  free(w->area);
  free(w->epsDist);
  free(w->rowptr);
  free(w->col);
  free(w->val);
  memset(w, 0, sizeof(*w));
  s->nwall = 0;
  slist_free(&s->slw);
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->sbq);
  free_cells(s);
  free_wall(s);
  free(s->ff);
  free(s->tt);
  free(s->vv);
  free(s->ffC);
  free(s->ttC);
  free(s->vvC);
  free(s->bb);
  free(s->qq);
  free(s);
}

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ======================================================================== */
/* Rbc_BuildSurfaceSource: common/ModRbc.F90:716-807.  Gauss analysis ->     */
/* ShFilter (n >= nlat0 zeroed) -> equi-spaced synthesis on nlat+1 rows ->    */
/* Spline_Build_on_Sphere.                                                    */
static void build_surface_spline(const orc_rbc *r, int nvar, const double *field, orc_spline *sp) {
  int nlat = r->nlat, nlon = r->nlon, ld = nlat + 1;
  double *va = calloc((size_t)ld * ld * nvar, sizeof(double));
  double *vb = calloc((size_t)ld * ld * nvar, sizeof(double));
  double *v = calloc((size_t)(nlat + 1) * nlon * nvar, sizeof(double));
  orc_sh_anal_gau(nlat, nlon, nvar, field, nlat, nlon, va, vb, ld, ld);
  for (int k = 0; k < nvar; k++) /* ShFilter common/ModSphpk.F90:423-432 */
    for (int n = r->nlat0; n < ld; n++)
      for (int m = 0; m < ld; m++) {
        va[m + ld * (n + ld * k)] = 0.;
        vb[m + ld * (n + ld * k)] = 0.;
      }
  orc_sh_synth_equ(nlat + 1, nlon, nvar, v, nlat + 1, nlon, va, vb, ld, ld);
  spline_build_on_sphere(sp, v);
  free(va);
  free(vb);
  free(v);
}

static orc_patch *find_or_make_patch(orc_sys *s, const orc_rbc *r) {
  for (int p = 0; p < s->npatch; p++)
    if (s->patches[p].nlat == r->nlat && s->patches[p].nlon == r->nlon) return &s->patches[p];
  orc_patch *pt = &s->patches[s->npatch++];
  patch_create(s, pt, r);
  return pt;
}

static void slist_alloc(orc_slist *l, int cap, int Nc[3], double iLbNc[3]) {
  l->cap = cap;
  l->x = calloc((size_t)3 * cap, sizeof(double));
  l->f = calloc((size_t)3 * cap, sizeof(double));
  l->g = calloc((size_t)3 * cap, sizeof(double));
  l->a3 = calloc((size_t)3 * cap, sizeof(double));
  l->Bcoef = calloc(cap, sizeof(double));
  l->indx = calloc((size_t)3 * cap, sizeof(int));
  l->next = calloc(cap, sizeof(int));
  memcpy(l->Nc, Nc, sizeof(int) * 3);
  memcpy(l->iLbNc, iLbNc, sizeof(double) * 3);
  l->hoc = calloc((size_t)(Nc[0] + 2) * (Nc[1] + 2) * (Nc[2] + 2), sizeof(int));
}

int orc_set_cells(orc_sys *s, int nrbc, const int *nlat0, const int *nlat, const int *celltype,
                  const double *x, const double *a3, const double *detj) {
  free_cells(s);
  s->nrbc = nrbc;
  /* surface ids: cells 1..nrbc, then the wall (GlobData_Init common/ModData.F90:40-49),
   * whichever was set first */
  if (s->nwall) {
    orc_wall *w = &s->wall;
    w->id = nrbc + 1;
    for (int e = 0; e < s->slw.nPoint; e++) s->slw.indx[e] = w->id;
  }
  s->rbcs = calloc(nrbc, sizeof(orc_rbc));
  size_t off = 0;
  int np = 0;
  for (int i = 0; i < nrbc; i++) {
    orc_rbc *r = &s->rbcs[i];
    r->id = i + 1; /* GlobData_Init common/ModData.F90:40-44 */
    r->nlat0 = nlat0[i];
    r->nlat = nlat[i];
    r->nlon = 2 * nlat[i];
    r->celltype = celltype[i];
    int nl = r->nlat, nn = r->nlon, npt = nl * nn;
    r->off = np;
    np += npt;
    /* RBC_Create grid common/ModRbc.F90:92-97 */
    r->th = malloc(sizeof(double) * nl);
    r->w = malloc(sizeof(double) * nl);
    r->phi = malloc(sizeof(double) * nn);
    orc_gaqd(nl, r->th, r->w);
    for (int k = 0; k < nl; k++) r->w[k] = (TWO_PI / nn) * r->w[k];
    for (int j = 0; j < nn; j++) r->phi[j] = j * TWO_PI / nn;
    r->x = malloc(sizeof(double) * 3 * npt);
    r->a3 = malloc(sizeof(double) * 3 * npt);
    r->detj = malloc(sizeof(double) * npt);
    r->f = calloc(3 * npt, sizeof(double));
    r->g = calloc(3 * npt, sizeof(double));
    memcpy(r->x, x + 3 * off, sizeof(double) * 3 * npt);
    memcpy(r->a3, a3 + 3 * off, sizeof(double) * 3 * npt);
    memcpy(r->detj, detj + off, sizeof(double) * npt);
    off += npt;
    /* area and mesh size: RBC_ComputeGeometry common/ModRbc.F90:482-505 */
    r->area = 0;
    for (int ilat = 0; ilat < nl; ilat++)
      for (int ilon = 0; ilon < nn; ilon++) r->area += r->detj[ilat + nl * ilon] * r->w[ilat];
    r->meshSize = sqrt(r->area / (nl * nl));
    spline_create(&r->sx, 2 * nl, nn, 3);
    spline_create(&r->sa3, 2 * nl, nn, 3);
    spline_create(&r->sdetj, 2 * nl, nn, 1);
    spline_create(&r->sF, 2 * nl, nn, 3);
    spline_create(&r->sG, 2 * nl, nn, 3);
    r->patch = find_or_make_patch(s, r);
  }
  /* splines of x, a3, detJ: Rbc_BuildSurfaceSource(xFlag), independent per cell */
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nrbc; i++) {
    orc_rbc *r = &s->rbcs[i];
    build_surface_spline(r, 3, r->x, &r->sx);
    build_surface_spline(r, 3, r->a3, &r->sa3);
    build_surface_spline(r, 1, r->detj, &r->sdetj);
  }
  /* SourceList_Create + SourceList_UpdateCoord common/ModSourceList.F90:27-151 */
  int Nc[3];
  double iLbNc[3];
  hash_num_blocks(s, Nc, iLbNc);
  orc_slist *l = &s->slr;
  slist_alloc(l, np, Nc, iLbNc);
  int p = 0;
  for (int i = 0; i < nrbc; i++) {
    orc_rbc *r = &s->rbcs[i];
    int nl = r->nlat, nn = r->nlon;
    for (int ilon = 0; ilon < nn; ilon++)
      for (int ilat = 0; ilat < nl; ilat++) {
        for (int d = 0; d < 3; d++) {
          l->x[p + (size_t)np * d] = r->x[ilat + nl * (ilon + nn * d)];
          l->a3[p + (size_t)np * d] = r->a3[ilat + nl * (ilon + nn * d)];
        }
        l->indx[p] = r->id;
        l->indx[p + np] = ilat + 1;
        l->indx[p + 2 * (size_t)np] = ilon + 1;
        l->Bcoef[p] = s->Bcoef[r->celltype - 1];
        p++;
      }
  }
  l->nPoint = p;
  orc_hashtable_build(l->Nc, l->iLbNc, l->nPoint, l->x, np, l->hoc, l->next);
  return 0;
}

int orc_set_cell_density(orc_sys *s, int which, const double *dens) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < s->nrbc; i++) {
    orc_rbc *r = &s->rbcs[i];
    int nl = r->nlat, nn = r->nlon, npt = nl * nn;
    double *dst = which == 1 ? r->f : r->g;
    memcpy(dst, dens + 3 * (size_t)r->off, sizeof(double) * 3 * npt);
    double *fd = malloc(sizeof(double) * 3 * npt);
    for (int k = 0; k < 3; k++)
      for (int q = 0; q < npt; q++) fd[q + npt * k] = dst[q + npt * k] * r->detj[q];
    build_surface_spline(r, 3, fd, which == 1 ? &r->sF : &r->sG);
    free(fd);
  }
  /* SourceList_UpdateDensity common/ModSourceList.F90:160-187 */
  orc_slist *l = &s->slr;
  int np = l->nPoint;
  for (int p = 0; p < np; p++) {
    orc_rbc *r = &s->rbcs[l->indx[p] - 1];
    int ilat = l->indx[p + np] - 1, ilon = l->indx[p + 2 * (size_t)np] - 1;
    int nl = r->nlat, nn = r->nlon;
    double dS = r->detj[ilat + nl * ilon] * r->w[ilat];
    double *src = which == 1 ? r->f : r->g, *dst = which == 1 ? l->f : l->g;
    for (int d = 0; d < 3; d++) dst[p + (size_t)np * d] = src[ilat + nl * (ilon + nn * d)] * dS;
  }
  return 0;
}

int orc_cell_spline(const orc_sys *s, int irbc, int which, double *u, double *u1, double *u2,
                    double *u12) {
  const orc_rbc *r = &s->rbcs[irbc];
  const orc_spline *sp = which == 0 ? &r->sx : which == 1 ? &r->sa3 : which == 2 ? &r->sdetj
                                      : which == 3 ? &r->sF : &r->sG;
  size_t n = (size_t)sp->m * sp->n * sp->nvar;
  memcpy(u, sp->u, n * sizeof(double));
  memcpy(u1, sp->u1, n * sizeof(double));
  memcpy(u2, sp->u2, n * sizeof(double));
  memcpy(u12, sp->u12, n * sizeof(double));
  return (int)n;
}

/* Test hooks: the pieces of the path on their own, so each can be pinned. */
int orc_set_terms(orc_sys *s, int terms) {
  if (terms & ~ORC_TERM_ALL) return 1;
  s->terms = terms;
  return 0;
}
static const orc_spline *cell_spline_of(const orc_rbc *r, int which) {
  return which == 0 ? &r->sx : which == 1 ? &r->sa3 : which == 2 ? &r->sdetj
         : which == 3 ? &r->sF : &r->sG;
}
static void spline_interp(const orc_spline *sp, double x, double y, double *f);
static void spline_find_projection(const orc_spline *sp, const double *xtar, double *th0,
                                   double *phi0, double *x0);
int orc_spline_eval(const orc_sys *s, int irbc, int which, double th, double phi, double *out) {
  if (irbc < 0 || irbc >= s->nrbc || which < 0 || which > 4) return 1;
  spline_interp(cell_spline_of(&s->rbcs[irbc], which), th, phi, out);
  return 0;
}
int orc_find_projection(const orc_sys *s, int irbc, const double *xtar, double *th0,
                        double *phi0, double *x0) {
  if (irbc < 0 || irbc >= s->nrbc) return 1;
  spline_find_projection(&s->rbcs[irbc].sx, xtar, th0, phi0, x0);
  return 0;
}
void orc_spline_interp_raw(int m, int n, int nvar, const double *u, const double *u1,
                           const double *u2, const double *u12, double x, double y, double *f) {
  orc_spline sp;
  sp.m = m;
  sp.n = n;
  sp.nvar = nvar;
  sp.hx = TWO_PI / m;
  sp.hy = TWO_PI / n;
  sp.ihx = 1. / sp.hx;
  sp.ihy = 1. / sp.hy;
  sp.u = (double *)u;
  sp.u1 = (double *)u1;
  sp.u2 = (double *)u2;
  sp.u12 = (double *)u12;
  spline_interp(&sp, x, y, f);
}
double orc_cell_mesh_size(const orc_sys *s, int irbc) {
  return (irbc < 0 || irbc >= s->nrbc) ? -1. : s->rbcs[irbc].meshSize;
}

/* ======================================================================== */
/* Target lists: TargetList_Create/Update common/ModTargetList.F90:26-135     */
int orc_ntargets_rbc(const orc_sys *s) { return s->slr.nPoint; }
int orc_ntargets_wall(const orc_sys *s) { return s->nwall ? s->wall.nvert : 0; }
int orc_tlist_rbc(const orc_sys *s, double *tx, int *tindx, double *Acoef) {
  int nT = s->slr.nPoint, p = 0;
  for (int i = 0; i < s->nrbc; i++) {
    const orc_rbc *r = &s->rbcs[i];
    int nl = r->nlat, nn = r->nlon;
    for (int ilon = 0; ilon < nn; ilon++)
      for (int ilat = 0; ilat < nl; ilat++) {
        for (int d = 0; d < 3; d++) tx[p + (size_t)nT * d] = r->x[ilat + nl * (ilon + nn * d)];
        tindx[p] = r->id;
        tindx[p + nT] = ilat + 1;
        tindx[p + 2 * (size_t)nT] = ilon + 1;
        Acoef[p] = s->Acoef[r->celltype - 1];
        p++;
      }
  }
  return p;
}
int orc_tlist_wall(const orc_sys *s, double *tx, int *tindx, double *Acoef) {
  if (!s->nwall) return 0;
  const orc_wall *w = &s->wall;
  int nT = w->nvert;
  for (int i = 0; i < nT; i++) {
    for (int d = 0; d < 3; d++) tx[i + (size_t)nT * d] = w->x[i + (size_t)nT * d];
    tindx[i] = w->id;
    tindx[i + nT] = i + 1;
    tindx[i + 2 * (size_t)nT] = -1;
    Acoef[i] = 2.;
  }
  return nT;
}

/* ======================================================================== */
/* RBC_SingInt: common/ModRbcSingInt.F90:29-90                               */
static void rbc_sing_int(const orc_sys *s, double c1, double c2, const orc_rbc *r, int ilat0,
                         int ilon0, double *dv) {
  const orc_patch *pt = r->patch;
  double xi[3], xj[3], fj[3], gj[3], a3j[3], xx[3], EA, EB;
  dv[0] = dv[1] = dv[2] = 0;
  spline_interp(&r->sx, r->th[ilat0], r->phi[ilon0], xi);
  size_t o = (size_t)pt->nrad * pt->nazm * (ilat0 + (size_t)pt->nlat * ilon0);
  for (int irad = 0; irad < pt->nrad; irad++)
    for (int iazm = 0; iazm < pt->nazm; iazm++) {
      double thj = pt->thG[o + irad + pt->nrad * iazm], phj = pt->phiG[o + irad + pt->nrad * iazm];
      spline_interp(&r->sx, thj, phj, xj);
      for (int d = 0; d < 3; d++) xx[d] = xj[d] - xi[d];
      double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      if (rr >= s->conf.rc) continue;
      if (c1 != 0) {
        spline_interp(&r->sF, thj, phj, fj);
        for (int d = 0; d < 3; d++) fj[d] = pt->w[irad] * fj[d];
        orc_ewald_sl(s, rr, &EA, &EB);
        double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
        for (int d = 0; d < 3; d++) dv[d] += c1 * (EA * xx[d] * xf + EB * fj[d]);
      }
      if (c2 != 0) {
        spline_interp(&r->sG, thj, phj, gj);
        for (int d = 0; d < 3; d++) gj[d] = pt->w[irad] * gj[d];
        spline_interp(&r->sa3, thj, phj, a3j);
        orc_ewald_dl(s, rr, &EA);
        double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
        double xn = xx[0] * a3j[0] + xx[1] * a3j[1] + xx[2] * a3j[2];
        for (int d = 0; d < 3; d++) dv[d] += c2 * (EA * xx[d] * xg * xn);
      }
    }
}

/* RBC_NearSingInt_Subtract: common/ModRbcSingInt.F90:175-226 */
static void near_sing_subtract(const orc_sys *s, double c1, double c2, const orc_rbc *r,
                               const double *xi, double th0, double phi0, double radPat, double *dv) {
  int nl = r->nlat, nn = r->nlon;
  int *ij1 = malloc(sizeof(int) * nl * nn), *ij2 = malloc(sizeof(int) * nl * nn);
  int npt = polar_patch_find_points(th0, phi0, radPat, nl, r->th, nn, r->phi, ij1, ij2);
  double xx[3], fj[3], gj[3], a3j[3], EA, EB;
  dv[0] = dv[1] = dv[2] = 0;
  for (int p = 0; p < npt; p++) {
    int ilat = ij1[p] - 1, ilon = ij2[p] - 1;
    for (int d = 0; d < 3; d++) xx[d] = r->x[ilat + nl * (ilon + nn * d)] - xi[d];
    double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    if (rr > s->conf.rc) continue;
    double dth = dist_on_sphere(th0, phi0, r->th[ilat], r->phi[ilon]);
    double mask = orc_mask(s, dth / radPat);
    double ds = r->detj[ilat + nl * ilon] * r->w[ilat];
    if (c1 != 0) {
      for (int d = 0; d < 3; d++) fj[d] = r->f[ilat + nl * (ilon + nn * d)] * ds;
      orc_ewald_sl(s, rr, &EA, &EB);
      double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
      for (int d = 0; d < 3; d++) dv[d] -= c1 * mask * (EA * xx[d] * xf + EB * fj[d]);
    }
    if (c2 != 0) {
      for (int d = 0; d < 3; d++) {
        gj[d] = r->g[ilat + nl * (ilon + nn * d)] * ds;
        a3j[d] = r->a3[ilat + nl * (ilon + nn * d)];
      }
      orc_ewald_dl(s, rr, &EA);
      double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
      double xn = xx[0] * a3j[0] + xx[1] * a3j[1] + xx[2] * a3j[2];
      for (int d = 0; d < 3; d++) dv[d] -= c2 * mask * (EA * xx[d] * xg * xn);
    }
  }
  free(ij1);
  free(ij2);
}

/* RBC_NearSingInt_ReAdd: common/ModRbcSingInt.F90:232-310 */
static void near_sing_readd(const orc_sys *s, double c1, double c2, const orc_rbc *r,
                            const double *xi, const double *x0, double th0, double phi0,
                            double radPat, double *dv) {
  enum { nrad = 16, nazm = 32 };
  double thPat[nrad], wtPat[nrad], phiPat[nazm], thG[nrad * nazm], phiG[nrad * nazm];
  double xj[3], fj[3], gj[3], a3j[3], xx[3], EA, EB;
  double d0 = xi[0] - x0[0], d1 = xi[1] - x0[1], d2 = xi[2] - x0[2];
  double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double sizePat = radPat * sqrt(r->area / (4 * PI_));
  dv[0] = dv[1] = dv[2] = 0;
  if (dist > 2.2250738585072014e-308) {
    dist = dist * (radPat / sizePat);
    orc_gauleg_sinh(0., radPat, 0., dist, nrad, thPat, wtPat);
  } else {
    orc_gauleg(0., radPat, nrad, thPat, wtPat);
  }
  for (int i = 0; i < nrad; i++) {
    wtPat[i] = wtPat[i] * sin(thPat[i]) * (TWO_PI / nazm);
    wtPat[i] = wtPat[i] * orc_mask(s, thPat[i] / radPat);
  }
  for (int j = 0; j < nazm; j++) phiPat[j] = j * TWO_PI / nazm;
  polar_patch_build(th0, phi0, nrad, thPat, nazm, phiPat, thG, phiG, nrad);
  for (int irad = 0; irad < nrad; irad++)
    for (int iazm = 0; iazm < nazm; iazm++) {
      double thj = thG[irad + nrad * iazm], phj = phiG[irad + nrad * iazm];
      spline_interp(&r->sx, thj, phj, xj);
      for (int d = 0; d < 3; d++) xx[d] = xj[d] - xi[d];
      double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      if (rr > s->conf.rc) continue;
      if (c1 != 0) {
        spline_interp(&r->sF, thj, phj, fj);
        for (int d = 0; d < 3; d++) fj[d] = wtPat[irad] * fj[d];
        orc_ewald_sl(s, rr, &EA, &EB);
        double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
        for (int d = 0; d < 3; d++) dv[d] += c1 * (EA * xx[d] * xf + EB * fj[d]);
      }
      if (c2 != 0) {
        spline_interp(&r->sa3, thj, phj, a3j);
        spline_interp(&r->sG, thj, phj, gj);
        for (int d = 0; d < 3; d++) gj[d] = wtPat[irad] * gj[d];
        orc_ewald_dl(s, rr, &EA);
        double xg = xx[0] * gj[0] + xx[1] * gj[1] + xx[2] * gj[2];
        double xn = xx[0] * a3j[0] + xx[1] * a3j[1] + xx[2] * a3j[2];
        for (int d = 0; d < 3; d++) dv[d] += c2 * (EA * xx[d] * xg * xn);
      }
    }
}

/* RBC_NearSingInt: common/ModRbcSingInt.F90:103-166 */
static void near_sing_int(const orc_sys *s, double c1, double c2, const orc_rbc *r,
                          const double *xi, const double *x0, double th0, double phi0, double *dv) {
  const orc_patch *pt = r->patch;
  double a30[3], dvt[3];
  spline_interp(&r->sa3, th0, phi0, a30);
  double dist = a30[0] * (xi[0] - x0[0]) + a30[1] * (xi[1] - x0[1]) + a30[2] * (xi[2] - x0[2]);
  dv[0] = dv[1] = dv[2] = 0;
  if (dist > 2 * r->meshSize) return;
  double sizePat = pt->radius * sqrt(r->area / (4 * PI_));
  double dist1 = copysign(0.01 * sizePat, dist);
  near_sing_subtract(s, c1, c2, r, xi, th0, phi0, pt->radius, dvt);
  for (int d = 0; d < 3; d++) dv[d] += dvt[d];
  if (fabs(dist) >= fabs(dist1)) {
    near_sing_readd(s, c1, c2, r, xi, x0, th0, phi0, pt->radius, dvt);
    for (int d = 0; d < 3; d++) dv[d] += dvt[d];
  } else {
    double xi1[3], dv1[3], dv0[3];
    for (int d = 0; d < 3; d++) xi1[d] = x0[d] + dist1 * a30[d];
    near_sing_readd(s, c1, c2, r, xi1, x0, th0, phi0, pt->radius, dv1);
    near_sing_readd(s, c1, c2, r, x0, x0, th0, phi0, pt->radius, dv0);
    if (c2 != 0) {
      double detJ0, g0[3];
      spline_interp(&r->sdetj, th0, phi0, &detJ0);
      spline_interp(&r->sG, th0, phi0, g0);
      for (int d = 0; d < 3; d++) g0[d] = g0[d] / detJ0;
      for (int d = 0; d < 3; d++) dv0[d] += (dist > 0 ? 1 : -1) * c2 * 4 * PI_ * g0[d];
    }
    for (int d = 0; d < 3; d++) dv[d] += dv0[d] + dist / dist1 * (dv1[d] - dv0[d]);
  }
}

/* ======================================================================== */
/* AddIntOnRbcs: common/ModIntOnRbcs.F90:25-158, with AddLinearInt :162-201  */
typedef struct {
  int n;
  int id[NBR_MAX], ilat[NBR_MAX], ilon[NBR_MAX];
  double dist[NBR_MAX];
} nbr_list;

/* NbrRbcList_Insert common/ModRbcSingInt.F90:341-364; returns 1 on overflow */
static int nbr_insert(nbr_list *L, int id, int ilat, int ilon, double dist) {
  for (int i = 0; i < L->n; i++)
    if (L->id[i] == id) {
      if (dist < L->dist[i]) {
        L->ilat[i] = ilat;
        L->ilon[i] = ilon;
        L->dist[i] = dist;
      }
      return 0;
    }
  if (L->n >= NBR_MAX) return 1;
  L->id[L->n] = id;
  L->ilat[L->n] = ilat;
  L->ilon[L->n] = ilon;
  L->dist[L->n] = dist;
  L->n++;
  return 0;
}

/* pair loop over the 27 bins for target i; fills the neighbour-cell list */
static void pair_loop(const orc_sys *s, double c1, double c2, const double *xi, int surfId_i,
                      double th_i, double phi_i, double Acoef_i, double *v, nbr_list *L,
                      long long *npairs, unsigned long long *csum, int ti, int *ovf) {
  const orc_slist *sl = &s->slr;
  int np = sl->nPoint;
  int irbc_i = surfId_i - 1;
  const orc_rbc *rbc_i = (irbc_i >= 0 && irbc_i < s->nrbc) ? &s->rbcs[irbc_i] : NULL;
  int i1, i2, i3;
  hash_index(sl->Nc, sl->iLbNc, xi, &i1, &i2, &i3);
  double EA, EB, xx[3];
  for (int j1 = (i1 - 1 > 0 ? i1 - 1 : 0); j1 <= (i1 + 1 < sl->Nc[0] + 1 ? i1 + 1 : sl->Nc[0] + 1); j1++)
    for (int j2 = (i2 - 1 > 0 ? i2 - 1 : 0); j2 <= (i2 + 1 < sl->Nc[1] + 1 ? i2 + 1 : sl->Nc[1] + 1); j2++)
      for (int j3 = (i3 - 1 > 0 ? i3 - 1 : 0); j3 <= (i3 + 1 < sl->Nc[2] + 1 ? i3 + 1 : sl->Nc[2] + 1); j3++) {
        int j = HOC(sl->hoc, sl->Nc, j1, j2, j3);
        while (j > 0) {
          int q = j - 1;
          for (int d = 0; d < 3; d++) {
            xx[d] = sl->x[q + (size_t)np * d] - xi[d];
            xx[d] = xx[d] - fnint(xx[d] * s->iLb[d]) * s->Lb[d];
          }
          double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
          if (!(rr > s->conf.rc)) {
            int surfId_j = sl->indx[q];
            const orc_rbc *rbc_j = &s->rbcs[surfId_j - 1];
            int jl = sl->indx[q + np] - 1, jn = sl->indx[q + 2 * (size_t)np] - 1;
            double mask;
            if (npairs) (*npairs)++;
            if (csum) *csum += (unsigned long long)(ti + 1) * (unsigned long long)(q + 1);
            if (surfId_i == surfId_j) {
              double dth = dist_on_sphere(th_i, phi_i, rbc_j->th[jl], rbc_j->phi[jn]);
              mask = orc_mask(s, dth / rbc_i->patch->radius);
            } else {
              mask = 0.;
              if (nbr_insert(L, surfId_j, jl + 1, jn + 1, rr)) *ovf = 1;
            }
            if (c1 != 0) {
              orc_ewald_sl(s, rr, &EA, &EB);
              const double *f = sl->f;
              double fj[3] = {f[q], f[q + np], f[q + 2 * (size_t)np]};
              double xf = xx[0] * fj[0] + xx[1] * fj[1] + xx[2] * fj[2];
              double cc = (1. - mask) * c1 / Acoef_i;
              for (int d = 0; d < 3; d++) v[d] += cc * (EA * xx[d] * xf + EB * fj[d]);
            }
            if (c2 != 0) {
              orc_ewald_dl(s, rr, &EA);
              const double *g = sl->g, *a3 = sl->a3;
              double xg = xx[0] * g[q] + xx[1] * g[q + np] + xx[2] * g[q + 2 * (size_t)np];
              double xn = xx[0] * a3[q] + xx[1] * a3[q + np] + xx[2] * a3[q + 2 * (size_t)np];
              double cc = (1. - mask) * c2 * sl->Bcoef[q] / Acoef_i;
              for (int d = 0; d < 3; d++) v[d] += cc * (EA * xx[d] * xg * xn);
            }
          }
          j = sl->next[q];
        }
      }
}

int orc_add_int_on_rbcs(orc_sys *s, double c1, double c2, int nT, const double *tx,
                        const int *tindx, const int *active, const double *Acoef, double *v,
                        long long *npairs_out) {
  if (s->nrbc == 0) return 0;
  long long npairs = 0;
  unsigned long long csum = 0;
  int ovf = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : npairs, csum) reduction(| : ovf)
  for (int i = 0; i < nT; i++) {
    if (!active[i]) continue;
    double xi[3] = {tx[i], tx[i + nT], tx[i + 2 * (size_t)nT]};
    int surfId_i = tindx[i];
    int irbc_i = surfId_i - 1;
    const orc_rbc *rbc_i = (irbc_i >= 0 && irbc_i < s->nrbc) ? &s->rbcs[irbc_i] : NULL;
    double acc[3] = {0, 0, 0}, th_i = 0, phi_i = 0;
    if (rbc_i) {
      th_i = rbc_i->th[tindx[i + nT] - 1];
      phi_i = rbc_i->phi[tindx[i + 2 * (size_t)nT] - 1];
    }
    nbr_list L;
    L.n = 0;
    long long np_i = 0;
    unsigned long long cs_i = 0;
    pair_loop(s, c1, c2, xi, surfId_i, th_i, phi_i, Acoef[i], acc, &L, &np_i, &cs_i, i, &ovf);
    if (!(s->terms & ORC_TERM_PAIR)) acc[0] = acc[1] = acc[2] = 0.;
    npairs += np_i;
    csum += cs_i;
    double dv[3];
    /* singular integration on the target's own surface */
    if (rbc_i && (s->terms & ORC_TERM_SELF)) {
      double c2Mod = c2 * s->Bcoef[rbc_i->celltype - 1];
      rbc_sing_int(s, c1, c2Mod, rbc_i, tindx[i + nT] - 1, tindx[i + 2 * (size_t)nT] - 1, dv);
      for (int d = 0; d < 3; d++) acc[d] += dv[d] / Acoef[i];
    }
    /* near-singular integration on neighbouring cells */
    for (int p = 0; p < ((s->terms & ORC_TERM_NEAR) ? L.n : 0); p++) {
      const orc_rbc *rj = &s->rbcs[L.id[p] - 1];
      int ilat0 = L.ilat[p] - 1, ilon0 = L.ilon[p] - 1;
      double th0 = rj->th[ilat0], phi0 = rj->phi[ilon0], x0[3], xt[3];
      int nl = rj->nlat, nn = rj->nlon;
      for (int d = 0; d < 3; d++) x0[d] = rj->x[ilat0 + nl * (ilon0 + nn * d)];
      for (int d = 0; d < 3; d++) {
        double xx = xi[d] - x0[d];
        xx = xx - fnint(xx * s->iLb[d]) * s->Lb[d];
        xt[d] = x0[d] + xx;
      }
      spline_find_projection(&rj->sx, xt, &th0, &phi0, x0);
      double c2Mod = c2 * s->Bcoef[rj->celltype - 1];
      near_sing_int(s, c1, c2Mod, rj, xt, x0, th0, phi0, dv);
      for (int d = 0; d < 3; d++) acc[d] += dv[d] / Acoef[i];
    }
    for (int d = 0; d < 3; d++) v[i + (size_t)nT * d] += acc[d];
  }
  if (npairs_out) *npairs_out = npairs;
  s->last_checksum = csum;
  if (ovf) return 2;
  /* AddLinearInt :162-201 */
  if (c2 != 0 && (s->terms & ORC_TERM_LINEAR)) {
    double xv[3] = {0, 0, 0};
    for (int ir = 0; ir < s->nrbc; ir++) {
      const orc_rbc *r = &s->rbcs[ir];
      int nl = r->nlat, nn = r->nlon;
      double B = s->Bcoef[r->celltype - 1];
      for (int ilon = 0; ilon < nn; ilon++)
        for (int ilat = 0; ilat < nl; ilat++) {
          double ds = r->w[ilat] * r->detj[ilat + nl * ilon], vn = 0;
          for (int d = 0; d < 3; d++)
            vn += r->g[ilat + nl * (ilon + nn * d)] * r->a3[ilat + nl * (ilon + nn * d)];
          for (int d = 0; d < 3; d++) xv[d] += B * r->x[ilat + nl * (ilon + nn * d)] * vn * ds;
        }
    }
    double pl = s->iLb[0] * s->iLb[1] * s->iLb[2];
    for (int d = 0; d < 3; d++) xv[d] = -8 * PI_ * pl * xv[d];
    for (int i = 0; i < nT; i++)
      if (active[i])
        for (int d = 0; d < 3; d++) v[i + (size_t)nT * d] += c2 * xv[d] / Acoef[i];
  }
  return 0;
}

unsigned long long orc_pair_checksum(const orc_sys *s) { return s->last_checksum; }

int orc_near_list(orc_sys *s, int nT, const double *tx, const int *tindx, int i, int *cells,
                  int *ilat, int *ilon, double *dist, int cap) {
  double xi[3] = {tx[i], tx[i + nT], tx[i + 2 * (size_t)nT]};
  double acc[3] = {0, 0, 0}, th_i = 0, phi_i = 0;
  int irbc_i = tindx[i] - 1;
  if (irbc_i >= 0 && irbc_i < s->nrbc) {
    th_i = s->rbcs[irbc_i].th[tindx[i + nT] - 1];
    phi_i = s->rbcs[irbc_i].phi[tindx[i + 2 * (size_t)nT] - 1];
  }
  nbr_list L;
  L.n = 0;
  int ovf = 0;
  pair_loop(s, 0., 0., xi, tindx[i], th_i, phi_i, 1., acc, &L, NULL, NULL, i, &ovf);
  for (int p = 0; p < L.n && p < cap; p++) {
    cells[p] = L.id[p];
    ilat[p] = L.ilat[p];
    ilon[p] = L.ilon[p];
    dist[p] = L.dist[p];
  }
  return ovf ? -1 : L.n;
}

/* ======================================================================== */
/* Walls: common/ModIntOnWalls.F90, common/ModWall.F90                       */
double orc_min_dist_to_tri(const double *xTar, const double *x, double *s0, double *t0) {
  /* MinDistToTri :480-577; x(3,3) col-major: x[l + 3*d] = vertex l, comp d */
  double x12[3], x13[3], x1T[3];
  for (int d = 0; d < 3; d++) {
    x12[d] = x[1 + 3 * d] - x[0 + 3 * d];
    x13[d] = x[2 + 3 * d] - x[0 + 3 * d];
    x1T[d] = x[0 + 3 * d] - xTar[d];
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
  switch (reg) {
  case 0:
    s = invDet * s;
    t = invDet * t;
    break;
  case 1:
    s = (c + e - b - d) / (a - 2 * b + c);
    s = fmin(1., fmax(0., s));
    t = 1. - s;
    break;
  case 3:
    s = 0.;
    t = -e / c;
    t = fmin(1., fmax(0., t));
    break;
  case 5:
    t = 0.;
    s = -d / a;
    s = fmin(1., fmax(0., s));
    break;
  }
  if (s0) *s0 = s;
  if (t0) *t0 = t;
  return sqrt(a * s * s + 2 * b * s * t + c * t * t + 2 * d * s + 2 * e * t + f);
}

static double tri_area(const double *x) {
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; d++) {
    a[d] = x[1 + 3 * d] - x[0 + 3 * d];
    b[d] = x[2 + 3 * d] - x[0 + 3 * d];
  }
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  return 0.5 * sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

/* Tri_Int_Regular :319-363; lhs(3,3,3) optional: lhs[l + 3*(i + 3*j)] */
static void tri_int_regular_full(const orc_sys *s, const double *x, const double *f,
                                 const double *xtar, double *rhs, double *lhs) {
  double detJ = 2 * tri_area(x), EA, EB;
  rhs[0] = rhs[1] = rhs[2] = 0;
  if (lhs) memset(lhs, 0, sizeof(double) * 27);
  for (int q = 0; q < 7; q++) {
    double sv = s->gq7_rst[q][0], t = s->gq7_rst[q][1];
    double xg[3], fg[3], xx[3];
    for (int d = 0; d < 3; d++) {
      xg[d] = (1. - sv - t) * x[0 + 3 * d] + sv * x[1 + 3 * d] + t * x[2 + 3 * d];
      fg[d] = (1. - sv - t) * f[0 + 3 * d] + sv * f[1 + 3 * d] + t * f[2 + 3 * d];
    }
    double ds = s->gq7_w[q] * detJ;
    for (int d = 0; d < 3; d++) {
      fg[d] = ds * fg[d];
      xx[d] = xtar[d] - xg[d];
    }
    double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
    orc_ewald_sl(s, rr, &EA, &EB);
    double xf = xx[0] * fg[0] + xx[1] * fg[1] + xx[2] * fg[2];
    for (int d = 0; d < 3; d++) rhs[d] += EA * xx[d] * xf + EB * fg[d];
    if (lhs) {
      double L[9];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) L[a + 3 * b] = ds * (EA * xx[a] * xx[b] + (a == b ? EB : 0.));
      double wv[3] = {1 - sv - t, sv, t};
      for (int l = 0; l < 3; l++)
        for (int k = 0; k < 9; k++) lhs[l + 3 * k] += wv[l] * L[k];
    }
  }
}
void orc_tri_int_regular(const orc_sys *s, const double *x, const double *f, const double *xtar,
                         double *rhs) {
  tri_int_regular_full(s, x, f, xtar, rhs, NULL);
}

/* Tri_Int_Duffy :373-465 */
static void tri_int_duffy_full(const orc_sys *s, const double *x, const double *f,
                               const double *xtar, double s0, double t0, double *rhs, double *lhs) {
  double x0[3], f0[3], EA, EB;
  for (int d = 0; d < 3; d++) {
    x0[d] = (1 - s0 - t0) * x[0 + 3 * d] + s0 * x[1 + 3 * d] + t0 * x[2 + 3 * d];
    f0[d] = (1 - s0 - t0) * f[0 + 3 * d] + s0 * f[1 + 3 * d] + t0 * f[2 + 3 * d];
  }
  rhs[0] = rhs[1] = rhs[2] = 0;
  if (lhs) memset(lhs, 0, sizeof(double) * 27);
  static const double st1[3][2] = {{0, 0}, {1, 0}, {0, 1}}, st2[3][2] = {{1, 0}, {0, 1}, {0, 0}};
  for (int n = 0; n < 3; n++) {
    int n2 = (n + 1) % 3;
    double x1[3], x2[3], f1[3], f2[3], a[3], b[3], nr[3];
    for (int d = 0; d < 3; d++) {
      x1[d] = x[n + 3 * d];
      x2[d] = x[n2 + 3 * d];
      f1[d] = f[n + 3 * d];
      f2[d] = f[n2 + 3 * d];
      a[d] = x1[d] - x0[d];
      b[d] = x2[d] - x0[d];
    }
    nr[0] = a[1] * b[2] - a[2] * b[1];
    nr[1] = a[2] * b[0] - a[0] * b[2];
    nr[2] = a[0] * b[1] - a[1] * b[0];
    double detJ = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double sv = s->duffy_r[i], t = sv * s->duffy_r[j];
        double xg[3], fg[3], xx[3];
        for (int d = 0; d < 3; d++) {
          xg[d] = (1. - sv) * x0[d] + (sv - t) * x1[d] + t * x2[d];
          fg[d] = (1. - sv) * f0[d] + (sv - t) * f1[d] + t * f2[d];
        }
        double ds = s->duffy_w[i] * s->duffy_w[j] * detJ * sv;
        for (int d = 0; d < 3; d++) {
          fg[d] = ds * fg[d];
          xx[d] = xtar[d] - xg[d];
        }
        double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
        orc_ewald_sl(s, rr, &EA, &EB);
        double xf = xx[0] * fg[0] + xx[1] * fg[1] + xx[2] * fg[2];
        for (int d = 0; d < 3; d++) rhs[d] += EA * xx[d] * xf + EB * fg[d];
        if (lhs) {
          double L[9];
          for (int aa = 0; aa < 3; aa++)
            for (int bb = 0; bb < 3; bb++)
              L[aa + 3 * bb] = ds * (EA * xx[aa] * xx[bb] + (aa == bb ? EB : 0.));
          double sG = (1. - sv) * s0 + (sv - t) * st1[n][0] + t * st2[n][0];
          double tG = (1. - sv) * t0 + (sv - t) * st1[n][1] + t * st2[n][1];
          double wv[3] = {1. - sG - tG, sG, tG};
          for (int l = 0; l < 3; l++)
            for (int k = 0; k < 9; k++) lhs[l + 3 * k] += wv[l] * L[k];
        }
      }
  }
}
void orc_tri_int_duffy(const orc_sys *s, const double *x, const double *f, const double *xtar,
                       double s0, double t0, double *rhs) {
  tri_int_duffy_full(s, x, f, xtar, s0, t0, rhs, NULL);
}

int orc_set_wall(orc_sys *s, int nvert, int nele, const double *x, const int *e2v) {
  free_wall(s);
  orc_wall *w = &s->wall;
  s->nwall = 1;
  w->id = s->nrbc + 1; /* GlobData_Init common/ModData.F90:46-49 */
  w->nvert = nvert;
  w->nele = nele;
  w->x = malloc(sizeof(double) * 3 * nvert);
  memcpy(w->x, x, sizeof(double) * 3 * nvert);
  w->f = calloc(3 * (size_t)nvert, sizeof(double));
  w->e2v = malloc(sizeof(int) * 3 * nele);
  memcpy(w->e2v, e2v, sizeof(int) * 3 * nele);
  /* Wall_Build_V2V common/ModWall.F90:66-114 */
  w->v2v = calloc(nvert, sizeof(int));
  {
    double eps = fmin(s->Lb[0], fmin(s->Lb[1], s->Lb[2])) * 1.e-5;
    double xmin[3], xmax[3];
    for (int d = 0; d < 3; d++) {
      xmin[d] = xmax[d] = x[(size_t)nvert * d];
      for (int i = 0; i < nvert; i++) {
        xmin[d] = fmin(xmin[d], x[i + (size_t)nvert * d]);
        xmax[d] = fmax(xmax[d], x[i + (size_t)nvert * d]);
      }
    }
    int *bd = malloc(sizeof(int) * nvert), nbd = 0;
    for (int i = 0; i < nvert; i++) {
      double m1 = 1e300, m2 = 1e300;
      for (int d = 0; d < 3; d++) {
        m1 = fmin(m1, fabs(x[i + (size_t)nvert * d] - xmin[d]));
        m2 = fmin(m2, fabs(x[i + (size_t)nvert * d] - xmax[d]));
      }
      if (m1 < eps || m2 < eps) bd[nbd++] = i;
    }
    for (int p1 = 0; p1 < nbd - 1; p1++) {
      int v1 = bd[p1];
      if (w->v2v[v1] > 0) continue;
      for (int p2 = p1 + 1; p2 < nbd; p2++) {
        int v2 = bd[p2];
        double mx = 0;
        for (int d = 0; d < 3; d++) {
          double xx = x[v2 + (size_t)nvert * d] - x[v1 + (size_t)nvert * d];
          xx = xx - fnint(xx * s->iLb[d]) * s->Lb[d];
          mx = fmax(mx, fabs(xx));
        }
        if (mx < eps) w->v2v[v2] = v1 + 1;
      }
    }
    free(bd);
  }
  /* indxVertGlb common/ModData.F90:50-63 */
  w->indxVertGlb = malloc(sizeof(int) * nvert);
  int p = 0;
  for (int i = 0; i < nvert; i++) {
    if (w->v2v[i] == 0)
      w->indxVertGlb[i] = ++p;
    else
      w->indxVertGlb[i] = w->indxVertGlb[w->v2v[i] - 1];
  }
  w->ndof = 3 * p;
  /* Wall_ComputeGeometry common/ModWall.F90:118-145 */
  w->a3 = malloc(sizeof(double) * 3 * nele);
  w->area = malloc(sizeof(double) * nele);
  w->epsDist = malloc(sizeof(double) * nele);
  for (int e = 0; e < nele; e++) {
    double xe[9];
    for (int l = 0; l < 3; l++)
      for (int d = 0; d < 3; d++) xe[l + 3 * d] = x[(e2v[e + (size_t)nele * l] - 1) + (size_t)nvert * d];
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; d++) {
      a[d] = xe[1 + 3 * d] - xe[3 * d];
      b[d] = xe[2 + 3 * d] - xe[3 * d];
    }
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    double n = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int d = 0; d < 3; d++) w->a3[e + (size_t)nele * d] = c[d] / n;
    w->area[e] = 0.5 * n;
    w->epsDist[e] = sqrt(w->area[e]);
  }
  /* SourceList for walls: centroids, SourceList_UpdateCoord :130-147 */
  int Nc[3];
  double iLbNc[3];
  hash_num_blocks(s, Nc, iLbNc);
  orc_slist *l = &s->slw;
  slist_alloc(l, nele, Nc, iLbNc);
  for (int e = 0; e < nele; e++) {
    for (int d = 0; d < 3; d++) {
      double sm = 0;
      for (int k = 0; k < 3; k++) sm += x[(e2v[e + (size_t)nele * k] - 1) + (size_t)nvert * d];
      l->x[e + (size_t)nele * d] = (1.0 / 3) * sm;
    }
    l->indx[e] = w->id;
    l->indx[e + nele] = e + 1;
    l->indx[e + 2 * (size_t)nele] = -1;
    l->Bcoef[e] = 0.;
  }
  l->nPoint = nele;
  orc_hashtable_build(l->Nc, l->iLbNc, nele, l->x, nele, l->hoc, l->next);
  return 0;
}

/* GMRES restart length of both solves (PETSc -ksp_gmres_restart) */
int orc_set_gmres_restart(orc_sys *s, int m) {
  if (m < 1 || m > 255) return 1;
  s->gmres_restart = m;
  return 0;
}

int orc_set_wall_density(orc_sys *s, const double *f) {
  if (!s->nwall) return 1;
  memcpy(s->wall.f, f, sizeof(double) * 3 * s->wall.nvert);
  return 0;
}

static void tri_gather(const orc_wall *w, int e, double *xe, double *fe) {
  for (int l = 0; l < 3; l++) {
    int iv = w->e2v[e + (size_t)w->nele * l] - 1;
    for (int d = 0; d < 3; d++) {
      xe[l + 3 * d] = w->x[iv + (size_t)w->nvert * d];
      if (fe) fe[l + 3 * d] = w->f[iv + (size_t)w->nvert * d];
    }
  }
}

/* PrepareSingIntOnWall :181-308, assembled as CSR (rows i + k*nvert, k=0..2).
 * Duplicate (row, col) entries from different triangles are summed, as
 * ADD_VALUES does; columns sorted ascending within each row. */
typedef struct {
  int col;
  double val;
} ent_t;
static int ent_cmp(const void *a, const void *b) {
  int ca = ((const ent_t *)a)->col, cb = ((const ent_t *)b)->col;
  return (ca > cb) - (ca < cb);
}
int orc_prepare_sing_int_on_wall(orc_sys *s) {
  orc_wall *w = &s->wall;
  const orc_slist *sl = &s->slw;
  int nv = w->nvert, nrow = 3 * nv;
  /* per vertex row-block: collect (col, 3x3) contributions */
  ent_t **rows = calloc(nrow, sizeof(ent_t *));
  int *cnt = calloc(nrow, sizeof(int)), *cap = calloc(nrow, sizeof(int));
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < nv; i++) {
    double xi[3] = {w->x[i], w->x[i + nv], w->x[i + 2 * (size_t)nv]};
    int i1, i2, i3;
    hash_index(sl->Nc, sl->iLbNc, xi, &i1, &i2, &i3);
    for (int j1 = (i1 - 1 > 0 ? i1 - 1 : 0); j1 <= (i1 + 1 < sl->Nc[0] + 1 ? i1 + 1 : sl->Nc[0] + 1); j1++)
      for (int j2 = (i2 - 1 > 0 ? i2 - 1 : 0); j2 <= (i2 + 1 < sl->Nc[1] + 1 ? i2 + 1 : sl->Nc[1] + 1); j2++)
        for (int j3 = (i3 - 1 > 0 ? i3 - 1 : 0); j3 <= (i3 + 1 < sl->Nc[2] + 1 ? i3 + 1 : sl->Nc[2] + 1); j3++) {
          int j = HOC(sl->hoc, sl->Nc, j1, j2, j3);
          while (j > 0) {
            int e = sl->indx[j - 1 + sl->nPoint] - 1;
            double xe[9], fe[9] = {0}, rhs[3], lhs[27], s0, t0;
            tri_gather(w, e, xe, NULL);
            for (int d = 0; d < 3; d++) {
              double sh = fnint((xi[d] - xe[3 * d]) * s->iLb[d]) * s->Lb[d];
              for (int l = 0; l < 3; l++) xe[l + 3 * d] += sh;
            }
            double rr = orc_min_dist_to_tri(xi, xe, &s0, &t0);
            if (!(rr > s->conf.rc)) {
              if (rr > w->epsDist[e])
                tri_int_regular_full(s, xe, fe, xi, rhs, lhs);
              else
                tri_int_duffy_full(s, xe, fe, xi, s0, t0, rhs, lhs);
              for (int l = 0; l < 3; l++) {
                int iv = w->e2v[e + (size_t)w->nele * l] - 1;
                for (int a = 0; a < 3; a++) {
                  int r = i + a * nv;
                  if (cnt[r] + 3 > cap[r]) {
                    cap[r] = cap[r] ? 2 * cap[r] : 96;
                    rows[r] = realloc(rows[r], sizeof(ent_t) * cap[r]);
                  }
                  for (int b = 0; b < 3; b++) {
                    rows[r][cnt[r]].col = iv + b * nv;
                    rows[r][cnt[r]].val = lhs[l + 3 * (a + 3 * b)];
                    cnt[r]++;
                  }
                }
              }
            }
            j = sl->next[j - 1];
          }
        }
  }
  /* compress */
  int nnz = 0;
  for (int r = 0; r < nrow; r++) {
    if (cnt[r]) qsort(rows[r], cnt[r], sizeof(ent_t), ent_cmp);
    int k = 0;
    for (int q = 0; q < cnt[r]; q++) {
      if (k > 0 && rows[r][k - 1].col == rows[r][q].col)
        rows[r][k - 1].val += rows[r][q].val;
      else
        rows[r][k++] = rows[r][q];
    }
    cnt[r] = k;
    nnz += k;
  }
  free(w->rowptr);
  free(w->col);
  free(w->val);
  w->rowptr = malloc(sizeof(int) * (nrow + 1));
  w->col = malloc(sizeof(int) * (size_t)nnz);
  w->val = malloc(sizeof(double) * (size_t)nnz);
  w->rowptr[0] = 0;
  for (int r = 0; r < nrow; r++) {
    w->rowptr[r + 1] = w->rowptr[r] + cnt[r];
    for (int q = 0; q < cnt[r]; q++) {
      w->col[w->rowptr[r] + q] = rows[r][q].col;
      w->val[w->rowptr[r] + q] = rows[r][q].val;
    }
    free(rows[r]);
  }
  w->nnz = nnz;
  free(rows);
  free(cnt);
  free(cap);
  return 0;
}
int orc_wall_self_nnz(const orc_sys *s) { return s->wall.nnz; }
int orc_wall_self_csr(const orc_sys *s, int *rowptr, int *col, double *val) {
  const orc_wall *w = &s->wall;
  memcpy(rowptr, w->rowptr, sizeof(int) * (3 * w->nvert + 1));
  memcpy(col, w->col, sizeof(int) * w->nnz);
  memcpy(val, w->val, sizeof(double) * w->nnz);
  return w->nnz;
}

/* AddIntOnWalls :37-128 */
int orc_add_int_on_walls(orc_sys *s, double c1, int nT, const double *tx, const int *tindx,
                         const int *active, const double *Acoef, double *v) {
  if (!s->nwall || nT == 0) return 0;
  orc_wall *w = &s->wall;
  int nv = w->nvert;
  if (tindx[0] == w->id) {
    /* SingIntOnWall :136-172: v += c1*(lhs*f) / Acoef */
    if (!w->rowptr) return 3;
    for (int r = 0; r < 3 * nv; r++) {
      double sum = 0;
      for (int q = w->rowptr[r]; q < w->rowptr[r + 1]; q++) {
        int cidx = w->col[q];
        sum += w->val[q] * w->f[(cidx % nv) + (size_t)nv * (cidx / nv)];
      }
      int i = r % nv, d = r / nv;
      v[i + (size_t)nT * d] += c1 * sum / Acoef[i];
    }
    return 0; /* one wall: all wall-target interactions are self-interactions */
  }
  const orc_slist *sl = &s->slw;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nT; i++) {
    if (!active[i]) continue;
    double xi[3] = {tx[i], tx[i + nT], tx[i + 2 * (size_t)nT]}, acc[3] = {0, 0, 0};
    int i1, i2, i3;
    hash_index(sl->Nc, sl->iLbNc, xi, &i1, &i2, &i3);
    for (int j1 = (i1 - 1 > 0 ? i1 - 1 : 0); j1 <= (i1 + 1 < sl->Nc[0] + 1 ? i1 + 1 : sl->Nc[0] + 1); j1++)
      for (int j2 = (i2 - 1 > 0 ? i2 - 1 : 0); j2 <= (i2 + 1 < sl->Nc[1] + 1 ? i2 + 1 : sl->Nc[1] + 1); j2++)
        for (int j3 = (i3 - 1 > 0 ? i3 - 1 : 0); j3 <= (i3 + 1 < sl->Nc[2] + 1 ? i3 + 1 : sl->Nc[2] + 1); j3++) {
          int j = HOC(sl->hoc, sl->Nc, j1, j2, j3);
          while (j > 0) {
            if (tindx[i] != sl->indx[j - 1]) {
              int e = sl->indx[j - 1 + sl->nPoint] - 1;
              double xe[9], fe[9], dv[3], s0, t0;
              tri_gather(w, e, xe, fe);
              for (int d = 0; d < 3; d++) {
                double sh = fnint((xi[d] - xe[3 * d]) * s->iLb[d]) * s->Lb[d];
                for (int l = 0; l < 3; l++) xe[l + 3 * d] += sh;
              }
              double rr = orc_min_dist_to_tri(xi, xe, &s0, &t0);
              if (!(rr > s->conf.rc)) {
                if (rr < w->epsDist[e])
                  tri_int_duffy_full(s, xe, fe, xi, s0, t0, dv, NULL);
                else
                  tri_int_regular_full(s, xe, fe, xi, dv, NULL);
                for (int d = 0; d < 3; d++) acc[d] += c1 * dv[d] / Acoef[i];
              }
            }
            j = sl->next[j - 1];
          }
        }
    for (int d = 0; d < 3; d++) v[i + (size_t)nT * d] += acc[d];
  }
  return 0;
}

/* ======================================================================== */
/* FFT used by the PME restatement (unnormalised, e^{sign 2 pi i jk/N}).     */
/* Mixed-radix decimation-in-time DFT of one line (smallest factor first, direct DFT for
 * a prime length), unnormalised, exp(sign 2 pi i j k / n).  The twiddles of every length
 * in the factor chain are tabulated once per plan, each entry the same expression the
 * transform would otherwise evaluate per term. */
static int fft_factor(int n) {
  for (int q = 2; q * q <= n; q++)
    if (n % q == 0) return q;
  return 0;
}
typedef struct {
  int nlev;
  int n[40];
  double complex *tw[40];
} fft_plan;
static void fft_plan_make(fft_plan *p, int n, int sign) {
  p->nlev = 0;
  int m = n;
  for (;;) {
    int li = p->nlev++;
    p->n[li] = m;
    p->tw[li] = malloc(sizeof(double complex) * (m > 0 ? m : 1));
    for (int a = 0; a < m; a++) p->tw[li][a] = cexp(sign * 2 * PI_ * I * (double)a / m);
    int r = m > 1 ? fft_factor(m) : 0;
    if (!r) break;
    m /= r;
  }
}
static void fft_plan_free(fft_plan *p) {
  for (int i = 0; i < p->nlev; i++) free(p->tw[i]);
  p->nlev = 0;
}
static void fft1(double complex *a, const fft_plan *p, int li, double complex *tmp) {
  int n = p->n[li];
  if (n == 1) return;
  const double complex *tw = p->tw[li];
  int r = fft_factor(n);
  if (!r) { /* prime length: direct DFT */
    for (int k = 0; k < n; k++) {
      double complex sm = 0;
      for (int j = 0; j < n; j++) sm += a[j] * tw[((long)j * k) % n];
      tmp[k] = sm;
    }
    memcpy(a, tmp, sizeof(double complex) * n);
    return;
  }
  int m = n / r;
  /* decimation in time: sub-sequences a[q + r*j] */
  double complex *sub = malloc(sizeof(double complex) * n);
  for (int q = 0; q < r; q++)
    for (int j = 0; j < m; j++) sub[q * m + j] = a[q + r * j];
  for (int q = 0; q < r; q++) fft1(sub + q * m, p, li + 1, tmp);
  for (int k = 0; k < n; k++) {
    double complex sm = 0;
    for (int q = 0; q < r; q++) sm += sub[q * m + (k % m)] * tw[((long)q * k) % n];
    a[k] = sm;
  }
  free(sub);
}

/* Forward 3-D transform in the reference's PFFTW convention (Get_That,
 * common/ModPFFTW.F90:148-167): x r2c (sign -1), y c2c (sign -1), z c2c (sign +1).
 * out_cplx: interleaved complex (0:N1/2, 0:N2-1, 0:N3-1), x fastest.  Lines are
 * independent and run on the OpenMP threads. */
void orc_fft3d_forward(int N1, int N2, int N3, const double *in, double *out_cplx) {
  int H = N1 / 2 + 1, L = N1 > N2 ? (N1 > N3 ? N1 : N3) : (N2 > N3 ? N2 : N3);
  double complex *out = (double complex *)out_cplx;
  fft_plan p1, p2, p3;
  fft_plan_make(&p1, N1, -1);
  fft_plan_make(&p2, N2, -1);
  fft_plan_make(&p3, N3, +1);
#pragma omp parallel
  {
    double complex *line = malloc(sizeof(double complex) * L), *tmp = malloc(sizeof(double complex) * L);
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < N3; k++)
      for (int j = 0; j < N2; j++) {
        for (int i = 0; i < N1; i++) line[i] = in[i + (size_t)N1 * (j + (size_t)N2 * k)];
        fft1(line, &p1, 0, tmp);
        for (int i = 0; i < H; i++) out[i + (size_t)H * (j + (size_t)N2 * k)] = line[i];
      }
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < N3; k++)
      for (int i = 0; i < H; i++) {
        for (int j = 0; j < N2; j++) line[j] = out[i + (size_t)H * (j + (size_t)N2 * k)];
        fft1(line, &p2, 0, tmp);
        for (int j = 0; j < N2; j++) out[i + (size_t)H * (j + (size_t)N2 * k)] = line[j];
      }
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i < H; i++) {
        for (int k = 0; k < N3; k++) line[k] = out[i + (size_t)H * (j + (size_t)N2 * k)];
        fft1(line, &p3, 0, tmp);
        for (int k = 0; k < N3; k++) out[i + (size_t)H * (j + (size_t)N2 * k)] = line[k];
      }
    free(line);
    free(tmp);
  }
  fft_plan_free(&p1);
  fft_plan_free(&p2);
  fft_plan_free(&p3);
}

/* Inverse (Get_R, common/ModPFFTW.F90:170-185): z c2c (sign -1), y c2c (+1),
 * x c2r (+1), unnormalised; c2r keeps Re of the kx = 0 and kx = N1/2 terms. */
void orc_fft3d_inverse(int N1, int N2, int N3, const double *in_cplx, double *out) {
  int H = N1 / 2 + 1, L = N1 > N2 ? (N1 > N3 ? N1 : N3) : (N2 > N3 ? N2 : N3);
  double complex *w = malloc(sizeof(double complex) * (size_t)H * N2 * N3);
  memcpy(w, in_cplx, sizeof(double complex) * (size_t)H * N2 * N3);
  fft_plan p1, p2, p3;
  fft_plan_make(&p1, N1, +1);
  fft_plan_make(&p2, N2, +1);
  fft_plan_make(&p3, N3, -1);
#pragma omp parallel
  {
    double complex *line = malloc(sizeof(double complex) * L), *tmp = malloc(sizeof(double complex) * L);
#pragma omp for collapse(2) schedule(static)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i < H; i++) {
        for (int k = 0; k < N3; k++) line[k] = w[i + (size_t)H * (j + (size_t)N2 * k)];
        fft1(line, &p3, 0, tmp);
        for (int k = 0; k < N3; k++) w[i + (size_t)H * (j + (size_t)N2 * k)] = line[k];
      }
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < N3; k++)
      for (int i = 0; i < H; i++) {
        for (int j = 0; j < N2; j++) line[j] = w[i + (size_t)H * (j + (size_t)N2 * k)];
        fft1(line, &p2, 0, tmp);
        for (int j = 0; j < N2; j++) w[i + (size_t)H * (j + (size_t)N2 * k)] = line[j];
      }
#pragma omp for collapse(2) schedule(static)
    for (int k = 0; k < N3; k++)
      for (int j = 0; j < N2; j++) {
        for (int i = 0; i < H; i++) line[i] = w[i + (size_t)H * (j + (size_t)N2 * k)];
        line[0] = creal(line[0]);
        if (N1 % 2 == 0) line[N1 / 2] = creal(line[N1 / 2]);
        for (int i = H; i < N1; i++) line[i] = conj(line[N1 - i]);
        fft1(line, &p1, 0, tmp);
        for (int i = 0; i < N1; i++) out[i + (size_t)N1 * (j + (size_t)N2 * k)] = creal(line[i]);
      }
    free(line);
    free(tmp);
  }
  free(w);
  fft_plan_free(&p1);
  fft_plan_free(&p2);
  fft_plan_free(&p3);
}

/* ======================================================================== */
/* PME: common/ModPME.F90                                                    */
static void pme_init(orc_sys *s) {
  /* PME_Init :252-338 */
  int N1 = s->conf.Nb[0], N2 = s->conf.Nb[1], N3 = s->conf.Nb[2], P = s->conf.P, H = N1 / 2 + 1;
  size_t ng = (size_t)N1 * N2 * N3, nc = (size_t)H * N2 * N3;
  s->ff = calloc(3 * ng, sizeof(double));
  s->tt = calloc(9 * ng, sizeof(double));
  s->vv = calloc(3 * ng, sizeof(double));
  s->ffC = calloc(3 * nc, sizeof(double complex));
  s->ttC = calloc(9 * nc, sizeof(double complex));
  s->vvC = calloc(3 * nc, sizeof(double complex));
  s->bb = malloc(sizeof(double) * nc);
  s->qq = malloc(sizeof(double) * 3 * nc);
  /* wave numbers, with the z sign flip :283-306 (stored here (kx,ky,kz), kx fastest) */
  for (int k = 0; k < N3; k++)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i < H; i++) {
        size_t o = i + (size_t)H * (j + (size_t)N2 * k);
        double q1 = i * s->iLb[0];
        double q2 = (j < N2 / 2) ? j * s->iLb[1] : (j - N2) * s->iLb[1];
        double q3 = (k < N3 / 2) ? k * s->iLb[2] : (k - N3) * s->iLb[2];
        s->qq[o] = q1;
        s->qq[o + nc] = q2;
        s->qq[o + 2 * nc] = -q3;
        s->bb[o] = 1.;
      }
  s->bb[0] = 0.;
  double MP[16];
  int imin;
  orc_bspline(P + 2.220446049250313e-16, P, &imin, MP);
  int Ns[3] = {N1, N2, N3}, cnt[3] = {H, N2, N3};
  double *fac[3];
  for (int ii = 0; ii < 3; ii++) {
    fac[ii] = malloc(sizeof(double) * cnt[ii]);
    for (int k = 0; k < cnt[ii]; k++) {
      double complex b = 0;
      for (int m = 0; m <= P - 2; m++) b += MP[m] * cexp(TWO_PI * I * k * m / (double)Ns[ii]);
      b = cexp(TWO_PI * I * k * (P - 1.) / (double)Ns[ii]) / b;
      double bv = cabs(b);
      fac[ii][k] = bv * bv;
    }
  }
  for (int k = 0; k < N3; k++)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i < H; i++) {
        size_t o = i + (size_t)H * (j + (size_t)N2 * k);
        s->bb[o] = ((s->bb[o] * fac[0][i]) * fac[1][j]) * fac[2][k];
      }
  for (int ii = 0; ii < 3; ii++) free(fac[ii]);
}

/* Distrib_Source :405-443 */
static void distrib_source(orc_sys *s, const double *x, double c1, double c2, const double *f,
                           const double *t, int klo, int khi) {
  int N1 = s->conf.Nb[0], N2 = s->conf.Nb[1], N3 = s->conf.Nb[2], P = s->conf.P;
  size_t ng = (size_t)N1 * N2 * N3;
  double wx[16], wy[16], wz[16];
  int imin, jmin, kmin;
  orc_bspline(x[0] * (N1 / s->Lb[0]), P, &imin, wx);
  orc_bspline(x[1] * (N2 / s->Lb[1]), P, &jmin, wy);
  orc_bspline(x[2] * (N3 / s->Lb[2]), P, &kmin, wz);
  for (int k0 = 0; k0 < P; k0++) {
    int k = imod(kmin + k0, N3);
    if (k < klo || k >= khi) continue;
    for (int j0 = 0; j0 < P; j0++)
      for (int i0 = 0; i0 < P; i0++) {
        int j = imod(jmin + j0, N2), i = imod(imin + i0, N1);
        double wxyz = wx[i0] * wy[j0] * wz[k0];
        size_t o = i + (size_t)N1 * (j + (size_t)N2 * k);
        if (s->flag_sl)
          for (int c = 0; c < 3; c++) s->ff[o + ng * c] += (c1 * wxyz) * f[c];
        if (s->flag_dl)
          for (int c = 0; c < 9; c++) s->tt[o + ng * c] += (c2 * wxyz) * t[c];
      }
  }
}

/* PME_Distrib_Source :58-133 */
int orc_pme_distrib_source(orc_sys *s, double c1, double c2, int cells, int walls) {
  if (!s->ff) pme_init(s);
  int N1 = s->conf.Nb[0], N2 = s->conf.Nb[1], N3 = s->conf.Nb[2];
  size_t ng = (size_t)N1 * N2 * N3;
  s->flag_sl = fabs(c1) > 1.e-10;
  s->flag_dl = fabs(c2) > 1.e-10;
  memset(s->ff, 0, sizeof(double) * 3 * ng);
  memset(s->tt, 0, sizeof(double) * 9 * ng);
  /* each thread owns a slab of z planes and adds, in source order, every contribution
   * that lands there: the same sums in the same order as one thread */
#pragma omp parallel
  {
  int nth = omp_get_num_threads(), ith = omp_get_thread_num();
  int klo = (int)((long)N3 * ith / nth), khi = (int)((long)N3 * (ith + 1) / nth);
  if (cells && s->nrbc) {
    const orc_slist *l = &s->slr;
    int np = l->nPoint;
    for (int p = 0; p < np; p++) {
      double x[3], f[3] = {0, 0, 0}, t[9] = {0};
      for (int d = 0; d < 3; d++) x[d] = l->x[p + (size_t)np * d];
      if (s->flag_sl)
        for (int d = 0; d < 3; d++) f[d] = l->f[p + (size_t)np * d];
      if (s->flag_dl) {
        double g[3], a3[3];
        for (int d = 0; d < 3; d++) {
          g[d] = l->g[p + (size_t)np * d];
          a3[d] = l->a3[p + (size_t)np * d] * l->Bcoef[p];
        }
        /* tt(...,ii,jj) = g(ii)*a3(jj); stored [ii + 3*jj] */
        for (int ii = 0; ii < 3; ii++)
          for (int jj = 0; jj < 3; jj++) t[ii + 3 * jj] = g[ii] * a3[jj];
      }
      distrib_source(s, x, c1, c2, f, t, klo, khi);
    }
  }
  if (walls && s->nwall) {
    const orc_wall *w = &s->wall;
    double t[9] = {0};
    for (int e = 0; e < w->nele; e++) {
      double xe[9], fe[9], x[3], f[3] = {0, 0, 0};
      tri_gather(w, e, xe, fe);
      for (int d = 0; d < 3; d++) x[d] = (1.0 / 3) * (xe[3 * d] + xe[1 + 3 * d] + xe[2 + 3 * d]);
      if (s->flag_sl)
        for (int d = 0; d < 3; d++)
          f[d] = (1.0 / 3) * (fe[3 * d] + fe[1 + 3 * d] + fe[2 + 3 * d]) * w->area[e];
      distrib_source(s, x, c1, c2, f, t, klo, khi);
    }
  }
  }
  return 0;
}

/* PME_Transform :137-222 */
int orc_pme_transform(orc_sys *s) {
  int N1 = s->conf.Nb[0], N2 = s->conf.Nb[1], N3 = s->conf.Nb[2], H = N1 / 2 + 1;
  size_t ng = (size_t)N1 * N2 * N3, nc = (size_t)H * N2 * N3;
  double alpha = s->conf.alpha;
  if (s->flag_sl)
    for (int c = 0; c < 3; c++) orc_fft3d_forward(N1, N2, N3, s->ff + ng * c, (double *)(s->ffC + nc * c));
  if (s->flag_dl)
    for (int c = 0; c < 9; c++) orc_fft3d_forward(N1, N2, N3, s->tt + ng * c, (double *)(s->ttC + nc * c));
  double vol = s->Lb[0] * s->Lb[1] * s->Lb[2];
#pragma omp parallel for
  for (size_t o = 0; o < nc; o++) {
    if (o == 0) {
      for (int c = 0; c < 3; c++) s->vvC[o + nc * c] = 0;
      continue;
    }
    double q[3] = {s->qq[o], s->qq[o + nc], s->qq[o + 2 * nc]}, qt[3];
    for (int d = 0; d < 3; d++) qt[d] = sqrt(PI_ * alpha) * q[d];
    double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    double q2t = PI_ * alpha * q2, e = exp(-q2t);
    double phi0 = e / q2t, phi1 = (e + phi0) / q2t;
    double complex vc[3] = {0, 0, 0};
    if (s->flag_sl) {
      double complex f[3] = {s->ffC[o], s->ffC[o + nc], s->ffC[o + 2 * nc]};
      double complex qf = qt[0] * f[0] + qt[1] * f[1] + qt[2] * f[2];
      for (int d = 0; d < 3; d++) vc[d] += 2 * alpha / vol * phi1 * (q2t * f[d] - qt[d] * qf);
    }
    if (s->flag_dl) {
      double complex T[3][3];
      for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 3; jj++) T[ii][jj] = s->ttC[o + nc * (ii + 3 * jj)];
      double complex tr = T[0][0] + T[1][1] + T[2][2], qT[3], Tq[3], dvc[3];
      for (int jj = 0; jj < 3; jj++) qT[jj] = q[0] * T[0][jj] + q[1] * T[1][jj] + q[2] * T[2][jj];
      for (int ii = 0; ii < 3; ii++) Tq[ii] = T[ii][0] * q[0] + T[ii][1] * q[1] + T[ii][2] * q[2];
      double complex qTq = q[0] * Tq[0] + q[1] * Tq[1] + q[2] * Tq[2];
      for (int d = 0; d < 3; d++) {
        dvc[d] = I * 4 * PI_ * alpha / vol * phi0 * (q[d] * tr + qT[d] + Tq[d]);
        dvc[d] -= I * 8 * PI_ * PI_ * alpha * alpha / vol * phi1 * qTq * q[d];
        vc[d] -= dvc[d];
      }
    }
    for (int d = 0; d < 3; d++) s->vvC[o + nc * d] = s->bb[o] * vc[d];
  }
  for (int c = 0; c < 3; c++) orc_fft3d_inverse(N1, N2, N3, (double *)(s->vvC + nc * c), s->vv + ng * c);
  return 0;
}

/* Interp_Vel :450-489 (single rank: the halo planes equal the periodic images) */
static void interp_vel(const orc_sys *s, const double *x, double *v) {
  int N1 = s->conf.Nb[0], N2 = s->conf.Nb[1], N3 = s->conf.Nb[2], P = s->conf.P;
  size_t ng = (size_t)N1 * N2 * N3;
  double wx[16], wy[16], wz[16];
  int imin, jmin, kmin;
  orc_bspline(x[2] * (N3 / s->Lb[2]), P, &kmin, wz);
  orc_bspline(x[1] * (N2 / s->Lb[1]), P, &jmin, wy);
  orc_bspline(x[0] * (N1 / s->Lb[0]), P, &imin, wx);
  v[0] = v[1] = v[2] = 0;
  for (int k0 = 0; k0 < P; k0++) {
    int k = imod(kmin + k0, N3);
    for (int j0 = 0; j0 < P; j0++)
      for (int i0 = 0; i0 < P; i0++) {
        int j = imod(jmin + j0, N2), i = imod(imin + i0, N1);
        double wxyz = wx[i0] * wy[j0] * wz[k0];
        size_t o = i + (size_t)N1 * (j + (size_t)N2 * k);
        for (int c = 0; c < 3; c++) v[c] += wxyz * s->vv[o + ng * c];
      }
  }
}

/* PME_Add_Interp_Vel :228-248 */
int orc_pme_add_interp_vel(orc_sys *s, int nT, const double *tx, const int *active,
                           const double *Acoef, double *v) {
#pragma omp parallel for
  for (int i = 0; i < nT; i++) {
    if (!active[i]) continue;
    double x[3] = {tx[i], tx[i + nT], tx[i + 2 * (size_t)nT]}, dv[3];
    interp_vel(s, x, dv);
    for (int d = 0; d < 3; d++) v[i + (size_t)nT * d] += dv[d] / Acoef[i];
  }
  return 0;
}

int orc_pme_grid(const orc_sys *s, int which, double *out) {
  size_t ng = (size_t)s->conf.Nb[0] * s->conf.Nb[1] * s->conf.Nb[2];
  const double *src = which == 0 ? s->ff : which == 1 ? s->tt : s->vv;
  size_t n = which == 1 ? 9 * ng : 3 * ng;
  memcpy(out, src, n * sizeof(double));
  return (int)n;
}

/* ======================================================================== */
/* CalcVelocityField: common/ModPostProcess.F90:29-61                        */
int orc_calc_velocity_field(orc_sys *s, int nT, const double *tx, const int *tindx,
                            const int *active, const double *Acoef, double *v) {
  memset(v, 0, sizeof(double) * 3 * nT);
  double c1 = 1.0 / (4. * PI_), c2 = 1.0 / (4. * PI_);
  int rc = orc_add_int_on_rbcs(s, c1, c2, nT, tx, tindx, active, Acoef, v, NULL);
  if (rc) return rc;
  orc_add_int_on_walls(s, c1, nT, tx, tindx, active, Acoef, v);
  orc_pme_distrib_source(s, c1, c2, 1, 1);
  orc_pme_transform(s);
  orc_pme_add_interp_vel(s, nT, tx, active, Acoef, v);
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < nT; i++) v[i + (size_t)nT * d] += s->conf.vBkg[d];
  return 0;
}

/* ======================================================================== */
/* Glob_Sph_Trans: common/ModVelSolver.F90:641-719                           */
int orc_cell_dof(const orc_sys *s) {
  int dof = 0;
  for (int i = 0; i < s->nrbc; i++) dof += 3 * s->rbcs[i].nlat0 * s->rbcs[i].nlat0;
  return dof;
}
int orc_glob_sph_trans(orc_sys *s, double *v, double *c, int direction) {
  int nP = s->slr.nPoint;
  int pv = 0, pc = 0;
  for (int ir = 0; ir < s->nrbc; ir++) {
    orc_rbc *r = &s->rbcs[ir];
    int nl = r->nlat, nn = r->nlon, n0 = r->nlat0, nvert = nl * nn;
    double *vt = malloc(sizeof(double) * 3 * nvert);
    double *va = calloc((size_t)nl * nl * 3, sizeof(double)), *vb = calloc((size_t)nl * nl * 3, sizeof(double));
    if (direction == 1) {
      for (int d = 0; d < 3; d++)
        for (int q = 0; q < nvert; q++) vt[q + (size_t)nvert * d] = v[pv + q + (size_t)nP * d];
      pv += nvert;
      orc_sh_anal_gau(nl, nn, 3, vt, nl, nn, va, vb, nl, nl);
      for (int m = 0; m < n0; m++)
        for (int n = m; n < n0; n++) {
          for (int d = 0; d < 3; d++) c[pc + d] = va[m + nl * (n + nl * d)];
          pc += 3;
        }
      for (int m = 1; m < n0; m++)
        for (int n = m; n < n0; n++) {
          for (int d = 0; d < 3; d++) c[pc + d] = vb[m + nl * (n + nl * d)];
          pc += 3;
        }
    } else {
      for (int m = 0; m < n0; m++)
        for (int n = m; n < n0; n++) {
          for (int d = 0; d < 3; d++) va[m + nl * (n + nl * d)] = c[pc + d];
          pc += 3;
        }
      for (int m = 1; m < n0; m++)
        for (int n = m; n < n0; n++) {
          for (int d = 0; d < 3; d++) vb[m + nl * (n + nl * d)] = c[pc + d];
          pc += 3;
        }
      orc_sh_synth_gau(nl, nn, 3, vt, nl, nn, va, vb, nl, nl);
      for (int d = 0; d < 3; d++)
        for (int q = 0; q < nvert; q++) v[pv + q + (size_t)nP * d] = vt[q + (size_t)nvert * d];
      pv += nvert;
    }
    free(vt);
    free(va);
    free(vb);
  }
  return 0;
}

/* MyMatMult: common/ModVelSolver.F90:523-601 (no deflation).  cell_active (or NULL for
 * every cell) limits the velocity evaluation to the nodes of the listed cells, so b is
 * the reference's only on those cells' coefficient blocks (a bounded CPU sample of a
 * large system; the sources, the PME grids and the linear term are always complete). */
static int cell_matmult_masked(orc_sys *s, const double *u, double *b, const int *cell_active);
static void deflate_matmult(orc_sys *s, double *v);
int orc_cell_matmult(orc_sys *s, const double *u, double *b) {
  return cell_matmult_masked(s, u, b, NULL);
}
int orc_cell_matmult_sample(orc_sys *s, const double *u, double *b, const int *cell_active) {
  return cell_matmult_masked(s, u, b, cell_active);
}
static int cell_matmult_masked(orc_sys *s, const double *u, double *b, const int *cell_active) {
  int nP = s->slr.nPoint;
  double *g = calloc((size_t)3 * nP, sizeof(double)), *v = calloc((size_t)3 * nP, sizeof(double));
  double *uc = malloc(sizeof(double) * orc_cell_dof(s));
  memcpy(uc, u, sizeof(double) * orc_cell_dof(s));
  orc_glob_sph_trans(s, g, uc, -1);
  /* rbc%g = reshape(g); Rbc_BuildSurfaceSource(gFlag); SourceList_UpdateDensity(G) */
  double *gc = malloc(sizeof(double) * 3 * nP);
  size_t off = 0;
  for (int ir = 0; ir < s->nrbc; ir++) {
    int nvert = s->rbcs[ir].nlat * s->rbcs[ir].nlon;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < nvert; q++) gc[3 * off + q + (size_t)nvert * d] = g[off + q + (size_t)nP * d];
    off += nvert;
  }
  orc_set_cell_density(s, 2, gc);
  free(gc);
  double c1 = 0., c2 = -1. / (4. * PI_);
  double *tx = malloc(sizeof(double) * 3 * nP), *Ac = malloc(sizeof(double) * nP);
  int *ti = malloc(sizeof(int) * 3 * nP), *act = malloc(sizeof(int) * nP);
  orc_tlist_rbc(s, tx, ti, Ac);
  for (int i = 0; i < nP; i++) act[i] = cell_active ? (cell_active[ti[i] - 1] != 0) : 1;
  int rc = orc_add_int_on_rbcs(s, c1, c2, nP, tx, ti, act, Ac, v, NULL);
  orc_pme_distrib_source(s, c1, c2, 1, 0);
  orc_pme_transform(s);
  orc_pme_add_interp_vel(s, nP, tx, act, Ac, v);
  for (size_t q = 0; q < 3 * (size_t)nP; q++) v[q] += g[q];
  if (s->deflate) deflate_matmult(s, v);
  orc_glob_sph_trans(s, v, b, 1);
  free(g);
  free(v);
  free(uc);
  free(tx);
  free(Ac);
  free(ti);
  free(act);
  return rc;
}

/* ======================================================================== */
/* Deflation of the cell BIE (Pozrikidis), common/ModVelSolver.F90:144-449.  Node
 * vectors v are (nP,3) component-major over the concatenated cells; the reference
 * addresses cell irbc at (irbc-1)*nlat*nlon with the first cell's grid, which is each
 * cell's own offset when every cell has the same grid (the only case it supports);
 * here each cell's own offset is used. */
int orc_set_deflation(orc_sys *s, int on) {
  s->deflate = on != 0;
  return 0;
}
int orc_sb_modes(const orc_sys *s, double *modes) {
  if (!s->sbq) return 1;
  memcpy(modes, s->sbq, sizeof(double) * 18 * (size_t)s->slr.nPoint);
  return 0;
}

/* surface integral of sum_d a_d b_d over cell r, the reference's loop order (outer index
 * first argument of the loop nest: ilat_outer = 1 for do ilat / do ilon) */
static double cell_dot(const orc_rbc *r, size_t nP, const double *a, size_t lda,
                       const double *b, size_t ldb, int ilat_outer) {
  int nl = r->nlat, nn = r->nlon;
  double fac = 0.;
  if (ilat_outer) {
    for (int ilat = 0; ilat < nl; ilat++)
      for (int ilon = 0; ilon < nn; ilon++) {
        const size_t q = (size_t)ilat + (size_t)nl * ilon;
        const double ds = r->detj[q] * r->w[ilat];
        double t = 0;
        for (int d = 0; d < 3; d++) t += a[q + lda * d] * b[q + ldb * d];
        fac = fac + t * ds;
      }
  } else {
    for (int ilon = 0; ilon < nn; ilon++)
      for (int ilat = 0; ilat < nl; ilat++) {
        const size_t q = (size_t)ilat + (size_t)nl * ilon;
        const double ds = r->detj[q] * r->w[ilat];
        double t = 0;
        for (int d = 0; d < 3; d++) t += a[q + lda * d] * b[q + ldb * d];
        fac = fac + t * ds;
      }
  }
  (void)nP;
  return fac;
}

/* DeflateGetSBModes :354-449: per cell the 6 rigid-body velocity fields (translations,
 * rotations about the centroid), Gram-Schmidt orthonormalised in the surface inner
 * product.  The centroid is RBC_ComputeGeometry's (common/ModRbc.F90:478-491). */
static void deflate_get_sb_modes(orc_sys *s) {
  const size_t nP = s->slr.nPoint;
  free(s->sbq);
  s->sbq = calloc(18 * nP, sizeof(double));
  for (int ir = 0; ir < s->nrbc; ir++) {
    const orc_rbc *r = &s->rbcs[ir];
    const int nl = r->nlat, nn = r->nlon, npt = nl * nn;
    double xc[3] = {0, 0, 0}, ar = 0;
    for (int ilat = 0; ilat < nl; ilat++)
      for (int ilon = 0; ilon < nn; ilon++) {
        const int q = ilat + nl * ilon;
        const double ds = r->detj[q] * r->w[ilat];
        ar += ds;
        for (int d = 0; d < 3; d++) xc[d] += ds * r->x[q + npt * d];
      }
    for (int d = 0; d < 3; d++) xc[d] = xc[d] / ar;
    double *pp = malloc(sizeof(double) * 18 * npt), *qq = malloc(sizeof(double) * 18 * npt);
    for (int q = 0; q < npt; q++) {
      double X[3];
      for (int d = 0; d < 3; d++) X[d] = r->x[q + npt * d] - xc[d];
      const double v[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, -X[2], X[1]},
                              {X[2], 0, -X[0]}, {-X[1], X[0], 0}};
      for (int m = 0; m < 6; m++)
        for (int d = 0; d < 3; d++) pp[q + (size_t)npt * (d + 3 * m)] = v[m][d];
    }
    for (int m = 0; m < 6; m++) {
      double *qm = qq + (size_t)npt * 3 * m;
      const double *pm = pp + (size_t)npt * 3 * m;
      memcpy(qm, pm, sizeof(double) * 3 * npt);
      for (int n = 1; n < m + 1; n++) {
        /* Proj(pp_m, qq_{n-1}) :603-621: fn / fd * qq_{n-1} */
        const double *qn = qq + (size_t)npt * 3 * (n - 1);
        const double fn = cell_dot(r, nP, pm, npt, qn, npt, 1);
        const double fd = cell_dot(r, nP, qn, npt, qn, npt, 1);
        for (int k = 0; k < 3 * npt; k++) qm[k] = qm[k] - fn / fd * qn[k];
      }
    }
    for (int m = 0; m < 6; m++) {
      const double *qm = qq + (size_t)npt * 3 * m;
      const double fac = cell_dot(r, nP, qm, npt, qm, npt, 1);
      for (int d = 0; d < 3; d++)
        for (int q = 0; q < npt; q++)
          s->sbq[(size_t)(m * 3 + d) * nP + r->off + q] = qm[q + (size_t)npt * d] / sqrt(fac);
    }
    free(pp);
    free(qq);
  }
}

/* DeflateRHS :219-274: FF += alpha/(1-alpha) a3 <FF, a3> / area, alpha = Bcoef/Acoef */
static void deflate_rhs(orc_sys *s, double *v) {
  const size_t nP = s->slr.nPoint;
  for (int ir = 0; ir < s->nrbc; ir++) {
    const orc_rbc *r = &s->rbcs[ir];
    const int npt = r->nlat * r->nlon;
    const double alpha = s->Bcoef[r->celltype - 1] / s->Acoef[r->celltype - 1];
    double *F = v + r->off;
    const double fac = cell_dot(r, nP, F, nP, r->a3, npt, 0);
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++)
        F[q + nP * d] = F[q + nP * d] + alpha / (1. - alpha) * r->a3[q + (size_t)npt * d] * fac / r->area;
  }
}

/* DeflateMatMult :281-349: vv += alpha a3 <g, a3> / area - alpha sum_m qq_m <g, qq_m> */
static void deflate_matmult(orc_sys *s, double *v) {
  const size_t nP = s->slr.nPoint;
  for (int ir = 0; ir < s->nrbc; ir++) {
    const orc_rbc *r = &s->rbcs[ir];
    const int npt = r->nlat * r->nlon;
    const double alpha = s->Bcoef[r->celltype - 1] / s->Acoef[r->celltype - 1];
    double *V = v + r->off;
    double fac = cell_dot(r, nP, r->g, npt, r->a3, npt, 0);
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++)
        V[q + nP * d] = V[q + nP * d] + alpha * r->a3[q + (size_t)npt * d] * fac / r->area;
    for (int m = 0; m < 6; m++) {
      const double *qm = s->sbq + (size_t)m * 3 * nP + r->off;
      fac = cell_dot(r, nP, r->g, npt, qm, nP, 0);
      for (int d = 0; d < 3; d++)
        for (int q = 0; q < npt; q++)
          V[q + nP * d] = V[q + nP * d] - alpha * qm[q + nP * d] * fac;
    }
  }
}

/* DeflateRecoverSol :144-214 on the node velocities v (nP,3); rigid cells (viscRat <= 0)
 * keep their solution as the double-layer potential g */
static void deflate_recover(orc_sys *s, double *v) {
  const size_t nP = s->slr.nPoint;
  for (int ir = 0; ir < s->nrbc; ir++) {
    orc_rbc *r = &s->rbcs[ir];
    const int npt = r->nlat * r->nlon;
    const double alpha = s->Bcoef[r->celltype - 1] / s->Acoef[r->celltype - 1];
    double *vv = malloc(sizeof(double) * 3 * npt), *vw = malloc(sizeof(double) * 3 * npt);
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++) vv[q + (size_t)npt * d] = v[r->off + q + nP * d];
    double afac;
    if (s->conf.viscRat[r->celltype - 1] > 0) {
      memcpy(vw, vv, sizeof(double) * 3 * npt);
      afac = alpha / (alpha + 1.);
    } else {
      memset(vw, 0, sizeof(double) * 3 * npt);
      memcpy(r->g, vv, sizeof(double) * 3 * npt);
      afac = -1. / 2.;
    }
    for (int m = 0; m < 6; m++) {
      const double *qm = s->sbq + (size_t)m * 3 * nP + r->off;
      const double fac = cell_dot(r, nP, vv, npt, qm, nP, 0);
      for (int d = 0; d < 3; d++)
        for (int q = 0; q < npt; q++)
          vw[q + (size_t)npt * d] = vw[q + (size_t)npt * d] - afac * qm[q + nP * d] * fac;
    }
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++) v[r->off + q + nP * d] = vw[q + (size_t)npt * d];
    free(vv);
    free(vw);
  }
}

/* Compute_Rhs: common/ModVelSolver.F90:455-515 (c1 = 1/4pi, c2 = 0) */
int orc_cell_rhs(orc_sys *s, double *rhs) {
  int nP = s->slr.nPoint;
  double *v = calloc((size_t)3 * nP, sizeof(double));
  double *tx = malloc(sizeof(double) * 3 * nP), *Ac = malloc(sizeof(double) * nP);
  int *ti = malloc(sizeof(int) * 3 * nP), *act = malloc(sizeof(int) * nP);
  orc_tlist_rbc(s, tx, ti, Ac);
  for (int i = 0; i < nP; i++) act[i] = 1;
  double c1 = 1. / (4. * PI_), c2 = 0.;
  int rc = orc_add_int_on_rbcs(s, c1, c2, nP, tx, ti, act, Ac, v, NULL);
  orc_add_int_on_walls(s, c1, nP, tx, ti, act, Ac, v);
  orc_pme_distrib_source(s, c1, c2, 1, 1);
  orc_pme_transform(s);
  orc_pme_add_interp_vel(s, nP, tx, act, Ac, v);
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < nP; i++) v[i + (size_t)nP * d] += 2. * s->conf.vBkg[d] / Ac[i];
  if (s->deflate) deflate_rhs(s, v);
  orc_glob_sph_trans(s, v, rhs, 1);
  free(v);
  free(tx);
  free(Ac);
  free(ti);
  free(act);
  return rc;
}

/* ======================================================================== */
/* Wall no-slip operator: common/ModNoSlip.F90                               */
int orc_wall_dof(const orc_sys *s) { return s->nwall ? s->wall.ndof : 0; }

/* AssembleArray :361-384 */
static void assemble(const orc_wall *w, double *u /* (nvert,3) */, double *u1d, int direction) {
  int nv = w->nvert;
  for (int i = 0; i < nv; i++) {
    int g = w->indxVertGlb[i];
    for (int d = 0; d < 3; d++) {
      if (direction == 1)
        u1d[3 * (g - 1) + d] = u[i + (size_t)nv * d];
      else
        u[i + (size_t)nv * d] = u1d[3 * (g - 1) + d];
    }
  }
}

/* MyMatMult :255-308 */
int orc_wall_matmult(orc_sys *s, const double *u, double *b) {
  orc_wall *w = &s->wall;
  int nv = w->nvert;
  double *f = calloc((size_t)3 * nv, sizeof(double)), *v = calloc((size_t)3 * nv, sizeof(double));
  assemble(w, f, (double *)u, -1);
  memcpy(w->f, f, sizeof(double) * 3 * nv);
  double *tx = malloc(sizeof(double) * 3 * nv), *Ac = malloc(sizeof(double) * nv);
  int *ti = malloc(sizeof(int) * 3 * nv), *act = malloc(sizeof(int) * nv);
  orc_tlist_wall(s, tx, ti, Ac);
  for (int i = 0; i < nv; i++) act[i] = 1;
  double c1 = 1. / (4. * PI_);
  orc_add_int_on_walls(s, c1, nv, tx, ti, act, Ac, v);
  orc_pme_distrib_source(s, c1, 0., 0, 1);
  orc_pme_transform(s);
  orc_pme_add_interp_vel(s, nv, tx, act, Ac, v);
  assemble(w, v, b, 1);
  free(f);
  free(v);
  free(tx);
  free(Ac);
  free(ti);
  free(act);
  return 0;
}

/* Compute_Wall_Residual_Vel :153-193 */
int orc_wall_residual_vel(orc_sys *s, double *v) {
  orc_wall *w = &s->wall;
  int nv = w->nvert;
  memset(v, 0, sizeof(double) * 3 * nv);
  double *tx = malloc(sizeof(double) * 3 * nv), *Ac = malloc(sizeof(double) * nv);
  int *ti = malloc(sizeof(int) * 3 * nv), *act = malloc(sizeof(int) * nv);
  orc_tlist_wall(s, tx, ti, Ac);
  for (int i = 0; i < nv; i++) act[i] = 1;
  double c1 = 1. / (4. * PI_), c2 = 1. / (4. * PI_);
  int rc = orc_add_int_on_rbcs(s, c1, c2, nv, tx, ti, act, Ac, v, NULL);
  orc_add_int_on_walls(s, c1, nv, tx, ti, act, Ac, v);
  orc_pme_distrib_source(s, c1, c2, 1, 1);
  orc_pme_transform(s);
  orc_pme_add_interp_vel(s, nv, tx, act, Ac, v);
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < nv; i++) v[i + (size_t)nv * d] += s->conf.vBkg[d];
  free(tx);
  free(Ac);
  free(ti);
  free(act);
  return rc;
}

/* ======================================================================== */
/* GMRES(m), PETSc KSPGMRES semantics as used by common/ModVelSolver.F90:74-90
 * and common/ModNoSlip.F90:70-87: classical Gram-Schmidt (no refinement), no
 * preconditioner, convergence when the Hessenberg residual estimate <=
 * max(rtol*||r0||, atol) (KSPConvergedDefault: ||r0|| is ||b|| when the initial
 * guess is nonzero), explicit residual at restarts, divergence at 1e5*||r0||.
 * m is PETSc's restart length (-ksp_gmres_restart), 30 unless set. */
typedef int (*orc_matvec_fn)(orc_sys *, const double *, double *);

static double dotp(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

static int orc_gmres(orc_sys *s, orc_matvec_fn mv, int n, const double *b, double *x, int x0nz,
                     double rtol, int maxit, int *niter, double *rnorm) {
  const int m = s->gmres_restart;
  double *V = malloc(sizeof(double) * (size_t)n * (m + 1));
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
  double *g = malloc(sizeof(double) * (m + 1)), *h = malloc(sizeof(double) * (m + 1));
  double *y = malloc(sizeof(double) * m);
  int its = 0, reason = 0, first = 1;
  double ttol = 0, rnorm0 = 0, res = 0;
  if (!x0nz) memset(x, 0, sizeof(double) * n);
  while (1) {
    if (x0nz || !first) {
      mv(s, x, V);
      for (int i = 0; i < n; i++) V[i] = b[i] - V[i];
    } else {
      memcpy(V, b, sizeof(double) * n);
    }
    res = sqrt(dotp(n, V, V));
    if (first) {
      rnorm0 = res;
      ttol = fmax(rtol * (x0nz ? sqrt(dotp(n, b, b)) : rnorm0), 1e-50);
    }
    first = 0;
    if (res == 0.0 || res <= ttol) {
      reason = 1;
      break;
    }
    for (int i = 0; i < n; i++) V[i] *= 1.0 / res;
    for (int i = 0; i <= m; i++) g[i] = 0;
    g[0] = res;
    int it = 0, happy = 0;
    while (it < m && its < maxit) {
      double *vj = V + (size_t)n * it, *vn = V + (size_t)n * (it + 1);
      mv(s, vj, vn);
      for (int i = 0; i <= it; i++) h[i] = dotp(n, V + (size_t)n * i, vn);
      for (int q = 0; q < n; q++) {
        double t = vn[q];
        for (int i = 0; i <= it; i++) t -= h[i] * V[(size_t)n * i + q];
        vn[q] = t;
      }
      double tt = sqrt(dotp(n, vn, vn));
      for (int i = 0; i <= it; i++) H[(size_t)i * m + it] = h[i];
      H[(size_t)(it + 1) * m + it] = tt;
      if (tt > fabs(g[it]) * 1.0e-30)
        for (int q = 0; q < n; q++) vn[q] *= 1.0 / tt;
      else
        happy = 1;
      for (int i = 0; i < it; i++) {
        double a = H[(size_t)i * m + it], bb = H[(size_t)(i + 1) * m + it];
        H[(size_t)i * m + it] = cs[i] * a + sn[i] * bb;
        H[(size_t)(i + 1) * m + it] = -sn[i] * a + cs[i] * bb;
      }
      double a = H[(size_t)it * m + it], bb = H[(size_t)(it + 1) * m + it];
      double tden = sqrt(a * a + bb * bb);
      if (tden != 0.0) {
        cs[it] = a / tden;
        sn[it] = bb / tden;
        H[(size_t)it * m + it] = tden;
        H[(size_t)(it + 1) * m + it] = 0.;
        g[it + 1] = -sn[it] * g[it];
        g[it] = cs[it] * g[it];
      } else {
        cs[it] = 1;
        sn[it] = 0;
        g[it + 1] = 0;
      }
      res = fabs(g[it + 1]);
      it++;
      its++;
      if (res <= ttol) {
        reason = 1;
        break;
      }
      if (res >= 1e5 * rnorm0) {
        reason = -1;
        break;
      }
      if (happy) {
        reason = 1;
        break;
      }
    }
    for (int i = it - 1; i >= 0; i--) {
      double t = g[i];
      for (int k = i + 1; k < it; k++) t -= H[(size_t)i * m + k] * y[k];
      y[i] = t / H[(size_t)i * m + i];
    }
    for (int q = 0; q < n; q++) {
      double t = x[q];
      for (int j = 0; j < it; j++) t += y[j] * V[(size_t)n * j + q];
      x[q] = t;
    }
    if (reason != 0 || its >= maxit) break;
  }
  free(V);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(h);
  free(y);
  *niter = its;
  *rnorm = res;
  return reason > 0;
}

/* NoSlipWall: common/ModNoSlip.F90:45-150 */
int orc_noslip_wall(orc_sys *s, double rtol, int maxit, const double *finc0, int *niter,
                    double *rnorm, double *vres_max) {
  orc_wall *w = &s->wall;
  int nv = w->nvert, nd = w->ndof;
  if (!w->rowptr) orc_prepare_sing_int_on_wall(s);
  double *f0 = malloc(sizeof(double) * 3 * nv), *v = malloc(sizeof(double) * 3 * nv);
  double *rhs = malloc(sizeof(double) * nd), *x = calloc(nd, sizeof(double));
  memcpy(f0, w->f, sizeof(double) * 3 * nv);
  orc_wall_residual_vel(s, v);
  for (int i = 0; i < 3 * nv; i++) v[i] = -v[i];
  assemble(w, v, rhs, 1);
  if (finc0) memcpy(x, finc0, sizeof(double) * nd);
  int conv = orc_gmres(s, orc_wall_matmult, nd, rhs, x, finc0 != NULL, rtol, maxit, niter, rnorm);
  double *finc = malloc(sizeof(double) * 3 * nv);
  assemble(w, finc, x, -1);
  for (int i = 0; i < 3 * nv; i++) w->f[i] = f0[i] + finc[i];
  if (vres_max) {
    orc_wall_residual_vel(s, v);
    double mx = 0;
    for (int i = 0; i < 3 * nv; i++) mx = fmax(mx, fabs(v[i]));
    *vres_max = mx;
  }
  free(f0);
  free(v);
  free(rhs);
  free(x);
  free(finc);
  return conv;
}

int orc_get_wall_density(const orc_sys *s, double *f) {
  memcpy(f, s->wall.f, sizeof(double) * 3 * s->wall.nvert);
  return 0;
}

/* Solve_RBC_Vel: common/ModVelSolver.F90:44-132 (+ g := v, common/ModTimeInt.F90:773-790);
 * sol0: previous solution (in/out, length cell_dof), the nonzero initial guess */
int orc_solve_rbc_vel(orc_sys *s, double *sol, double *v, int *niter, double *rnorm) {
  int dof = orc_cell_dof(s), nP = s->slr.nPoint;
  double *rhs = malloc(sizeof(double) * dof);
  orc_cell_rhs(s, rhs);
  if (s->deflate) deflate_get_sb_modes(s);
  int all_one = 1;
  for (int l = 0; l < s->conf.ntypes; l++) all_one = all_one && s->conf.viscRat[l] == 1.0;
  *niter = 0;
  *rnorm = 0;
  if (all_one)
    memcpy(sol, rhs, sizeof(double) * dof);
  else
    orc_gmres(s, orc_cell_matmult, dof, rhs, sol, 1, 1e-11, 200, niter, rnorm);
  double *c = malloc(sizeof(double) * dof);
  memcpy(c, sol, sizeof(double) * dof);
  orc_glob_sph_trans(s, v, c, -1);
  if (s->deflate) deflate_recover(s, v);
  /* g := v for cells of finite viscosity (common/ModTimeInt.F90:780-784); a rigid cell
   * keeps the potential the recovery stored */
  double *gc = malloc(sizeof(double) * 3 * nP);
  size_t off = 0;
  for (int ir = 0; ir < s->nrbc; ir++) {
    const orc_rbc *r = &s->rbcs[ir];
    int nvert = r->nlat * r->nlon;
    const int finite = s->conf.viscRat[r->celltype - 1] > 0;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < nvert; q++)
        gc[3 * off + q + (size_t)nvert * d] =
            finite ? v[off + q + (size_t)nP * d] : r->g[q + (size_t)nvert * d];
    off += nvert;
  }
  orc_set_cell_density(s, 2, gc);
  free(gc);
  free(c);
  free(rhs);
  return 0;
}
