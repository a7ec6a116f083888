// Context, host-side table construction and the C ABI (include/rbc3d.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "internal.h"

static thread_local std::string g_err;
void rbc_set_error(const std::string &msg) { g_err = msg; }
extern "C" const char *rbc3d_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Host tables

// Gauss-Legendre colatitudes, ascending, and weights summing to 2 (what
// SPHEREPACK's gaqd returns; RBC_Create, common/ModRbc.F90:95-97).
static void gauss_colat(int n, std::vector<double> &th, std::vector<double> &w) {
  th.resize(n);
  w.resize(n);
  auto leg = [n](double x, double &dp) {
    double p0 = 1, p1 = x;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1);
    return p1;
  };
  for (int k = 0; k < n; k++) {
    double t = RBC_PI * (k + 0.75) / (n + 0.5), dp, p;
    for (int it = 0; it < 100; it++) {
      p = leg(cos(t), dp);
      double dt = p / (-sin(t) * dp);
      t -= dt;
      if (fabs(dt) < 1e-17) break;
    }
    p = leg(cos(t), dp);
    t -= p / (-sin(t) * dp);
    p = leg(cos(t), dp);
    th[k] = t;
    double st = sin(t);
    w[k] = 2. / (st * st * dp * dp);
  }
}

// SPHEREPACK-normalised Pbar(m,n,theta): out[(m*nl + n)*nt + i]
static void pbar(int nl, int mmax, const std::vector<double> &th, std::vector<double> &out) {
  int nt = (int)th.size();
  out.assign((size_t)mmax * nl * nt, 0.);
  for (int i = 0; i < nt; i++) {
    double c = cos(th[i]), s = sin(th[i]), pmm = sqrt(0.5);
    for (int m = 0; m < mmax; m++) {
      if (m > 0) pmm = sqrt((2.0 * m + 1) / (2.0 * m)) * s * pmm;
      if (m >= nl) break;
      out[((size_t)m * nl + m) * nt + i] = pmm;
      if (m + 1 < nl) {
        double pa = pmm, pb = sqrt(2.0 * m + 3) * c * pmm;
        out[((size_t)m * nl + m + 1) * nt + i] = pb;
        for (int n = m + 2; n < nl; n++) {
          double nn = n, mm = m;
          double an = sqrt((4 * nn * nn - 1) / (nn * nn - mm * mm));
          double bn = sqrt(((2 * nn + 1) * ((nn - 1) * (nn - 1) - mm * mm)) /
                           ((2 * nn - 3) * (nn * nn - mm * mm)));
          double pn = an * c * pb - bn * pa;
          out[((size_t)m * nl + n) * nt + i] = pn;
          pa = pb;
          pb = pn;
        }
      }
    }
  }
}

// GauLeg (common/ModQuadRule.F90:135-208)
static void gauleg(double x1, double x2, int n, double *x, double *w) {
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  std::vector<double> zz(m), ppv(m);
  for (int j = 0; j < m; j++) {
    double z = cos(RBC_PI * ((j + 1) - 0.25) / (n + 0.5)), z1, p1 = 0, p2, p3, pp = 0;
    bool unfinished = true;
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
      unfinished = fabs(z - z1) > 3.e-14;
    }
    zz[j] = z;
    ppv[j] = pp;
  }
  for (int j = 0; j < m; j++) x[j] = xm - xl * zz[j];
  for (int j = 0; j < m; j++) x[n - 1 - j] = xm + xl * zz[j];
  for (int j = 0; j < m; j++) w[j] = 2.0 * xl / ((1.0 - zz[j] * zz[j]) * ppv[j] * ppv[j]);
  for (int j = 0; j < m; j++) w[n - 1 - j] = w[j];
}

static double mask_exact(double x) {  // MaskFunc_Exact, common/ModBasicMath.F90:332-347
  double t = fabs(x);
  if (t < 0.01) return 1.;
  if (t > 0.99) return 0.;
  return exp(2 * (exp(-1. / t)) / (t - 1));
}

static double mask_tab(const std::vector<double> &mt, double x) {
  double s = fabs(x) * RBC_NTAB, fi = floor(s);
  if (fi >= RBC_NTAB) return 0.;
  int i = (int)fi;
  return mt[i] * (i + 1 - s) + mt[i + 1] * (s - i);
}

// PolarPatch_Build (common/ModPolarPatch.F90:99-148) for one patch origin
static void patch_build(double th0, double phi0, int nth, const double *thL, int nphi,
                        const double *phiL, double *thG, double *phiG) {
  double s0 = sin(th0), c0 = cos(th0);
  for (int i = 0; i < nth; i++) {
    double st = sin(thL[i]), ct = cos(thL[i]);
    for (int j = 0; j < nphi; j++) {
      double sp = sin(phiL[j]), cp = cos(phiL[j]);
      double x0 = st * cp, x1 = st * sp, x2 = ct;
      double y0 = c0 * x0 + 0. * x1 + s0 * x2;
      double y1 = 0. * x0 + 1. * x1 + 0. * x2;
      double y2 = -s0 * x0 + 0. * x1 + c0 * x2;
      y2 = std::max(-1.0, std::min(1.0, y2));
      double ph = atan2(y1, y0) + phi0;
      thG[i + nth * j] = acos(y2);
      phiG[i + nth * j] = ph - floor(ph * (1. / RBC_TWO_PI)) * RBC_TWO_PI;
    }
  }
}

template <typename T>
static int upload(DBuf<T> &b, const std::vector<T> &h, hipStream_t st) {
  // synchronous: the host vectors are temporaries of the caller
  RBC_TRY(b.ensure(h.size()));
  HIP_TRY(hipStreamSynchronize(st));
  if (!h.empty()) HIP_TRY(hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return RBC3D_OK;
}

static int make_gtype(rbc3d_ctx *c, int nlat, GridType **out) {
  for (auto *g : c->gtypes)
    if (g->nlat == nlat) {
      *out = g;
      return RBC3D_OK;
    }
  GridType *g = new GridType();
  c->gtypes.push_back(g);
  g->nlat = nlat;
  g->nlon = 2 * nlat;
  int nl = nlat, nn = g->nlon;
  g->mmax = std::min(nl, nn / 2 + 1);
  std::vector<double> wg;
  gauss_colat(nl, g->th, wg);
  g->w.resize(nl);
  for (int i = 0; i < nl; i++) g->w[i] = (RBC_TWO_PI / nn) * wg[i];
  g->phi.resize(nn);
  for (int j = 0; j < nn; j++) g->phi[j] = j * RBC_TWO_PI / nn;
  hipStream_t st = c->stream;
  RBC_TRY(upload(g->d_th, g->th, st));
  RBC_TRY(upload(g->d_phi, g->phi, st));
  RBC_TRY(upload(g->d_w, g->w, st));
  // SH tables
  std::vector<double> P, pw((size_t)g->mmax * nl * nl), pg((size_t)g->mmax * nl * nl),
      pws((size_t)g->mmax * nl * nl);
  pbar(nl, g->mmax, g->th, P);
  for (int m = 0; m < g->mmax; m++)
    for (int n = 0; n < nl; n++)
      for (int i = 0; i < nl; i++) {
        size_t o = ((size_t)m * nl + n) * nl + i;
        pw[o] = (2.0 / nn) * wg[i] * P[o];
        pws[o] = P[o] * wg[i];
        pg[o] = (m == 0 ? 0.5 : 1.0) * P[o];
      }
  RBC_TRY(upload(g->d_pw, pw, st));
  RBC_TRY(upload(g->d_pws, pws, st));
  RBC_TRY(upload(g->d_pg, pg, st));
  int ne = nl + 1, mmax_e = std::min(ne, nn / 2 + 1);
  std::vector<double> the(ne), Pe, pe((size_t)mmax_e * ne * ne);
  for (int i = 0; i < ne; i++) the[i] = RBC_PI * i / (ne - 1);
  pbar(ne, mmax_e, the, Pe);
  for (int m = 0; m < mmax_e; m++)
    for (int n = 0; n < ne; n++)
      for (int i = 0; i < ne; i++) {
        size_t o = ((size_t)m * ne + n) * ne + i;
        pe[o] = (m == 0 ? 0.5 : 1.0) * Pe[o];
      }
  RBC_TRY(upload(g->d_pe, pe, st));
  std::vector<double> cs((size_t)2 * (nl + 1) * nn);
  for (int m = 0; m <= nl; m++)
    for (int j = 0; j < nn; j++) {
      double ph = RBC_TWO_PI * j / nn;
      cs[(size_t)m * nn + j] = cos(m * ph);
      cs[(size_t)(nl + 1) * nn + m * nn + j] = sin(m * ph);
    }
  RBC_TRY(upload(g->d_cs, cs, st));
  // FFT_Diff (common/ModFFT.F90:25-93) as a periodic derivative operator
  auto dmat = [](int m) {
    std::vector<double> D((size_t)m * m);
    for (int p = 0; p < m; p++)
      for (int q = 0; q < m; q++) {
        double s = 0;
        int dq = ((p - q) % m + m) % m;
        for (int k = 1; k < m / 2; k++) s += k * sin(RBC_TWO_PI * (double)((long)k * dq) / m);
        D[(size_t)p * m + q] = -(2.0 / m) * s;
      }
    return D;
  };
  RBC_TRY(upload(g->d_Dm, dmat(2 * nl), st));
  RBC_TRY(upload(g->d_Dn, dmat(nn), st));
  // polar patch (RbcPolarPatch_Create, common/ModPolarPatch.F90:27-75)
  std::vector<double> mt(RBC_NTAB + 1);
  for (int i = 0; i <= RBC_NTAB; i++) mt[i] = mask_exact((double)i / RBC_NTAB);
  g->radius = RBC_PI / sqrt((double)nl);
  double h = RBC_PI / nl;
  g->nrad = 2 * (int)std::round(g->radius / h);
  g->nazm = 2 * g->nrad;
  int nr = g->nrad, na = g->nazm;
  std::vector<double> thL(nr), wL(nr), phiL(na);
  gauleg(0., g->radius, nr, thL.data(), wL.data());
  for (int i = 0; i < nr; i++) {
    wL[i] = wL[i] * sin(thL[i]) * (RBC_TWO_PI / na);
    wL[i] = wL[i] * mask_tab(mt, thL[i] / g->radius);
  }
  for (int j = 0; j < na; j++) phiL[j] = j * RBC_TWO_PI / na;
  std::vector<double> thG((size_t)nl * nn * nr * na), phiG(thG.size());
  for (int ilon = 0; ilon < nn; ilon++)
    for (int ilat = 0; ilat < nl; ilat++) {
      size_t o = (size_t)nr * na * (ilat + (size_t)nl * ilon);
      patch_build(g->th[ilat], g->phi[ilon], nr, thL.data(), na, phiL.data(), &thG[o], &phiG[o]);
    }
  RBC_TRY(upload(g->d_thG, thG, st));
  RBC_TRY(upload(g->d_phiG, phiG, st));
  RBC_TRY(upload(g->d_pwt, wL, st));
  // same-surface mask MaskFunc(DistOnSphere/radius) by (ilat_i, ilat_j, ilon_j - ilon_i)
  std::vector<double> mk((size_t)nl * nl * nn);
  for (int a = 0; a < nl; a++)
    for (int b = 0; b < nl; b++)
      for (int d = 0; d < nn; d++) {
        double th0 = g->th[a], th1 = g->th[b], phi0 = g->phi[0], phi1 = g->phi[d];
        double dd = cos(th0 - th1) - sin(th0) * sin(th1) * (1. - cos(phi0 - phi1));
        dd = std::min(1., std::max(-1., dd));
        mk[((size_t)a * nl + b) * nn + d] = mask_tab(mt, acos(dd) / g->radius);
      }
  RBC_TRY(upload(g->d_mask, mk, st));
  std::vector<float> mkf(mk.begin(), mk.end());  // the split form's pair kernel
  RBC_TRY(upload(g->d_maskf, mkf, st));
  // Spline_FindProjection (common/ModSpline.F90:203-269): in each of its 3 iterations
  // the 17-point quadratic fit of QuadFit_2D has a design fixed by the patch size h
  // alone, so its least-squares solve is done here once: c = sum_k W_k d2_k with
  // W = A^-1 [u_0 .. u_16], A = sum_k u_k u_k^T, u = (1, x, y, x^2, xy, y^2).  The
  // local patch directions (sin th cos ph, sin th sin ph, cos th) are tabulated too.
  {
    std::vector<double> pj(RBC_PJ_SIZE, 0.);
    double hp = std::max(RBC_TWO_PI / (2 * nl), RBC_TWO_PI / nn);
    for (int it3 = 0; it3 < 3; it3++) {
      double xy[17][2] = {{0., 0.}};
      for (int ip = 0; ip < 8; ip++)
        for (int it = 0; it < 2; it++) {
          int i = (it + 1) + 2 * ip;
          double thL = (it + 1) * hp / 2, phL = ip * RBC_TWO_PI / 8;
          xy[i][0] = thL * cos(phL);
          xy[i][1] = thL * sin(phL);
          double *d = &pj[RBC_PJ_DIR + (size_t)(it3 * 16 + (i - 1)) * 3];
          d[0] = sin(thL) * cos(phL);
          d[1] = sin(thL) * sin(phL);
          d[2] = cos(thL);
        }
      long double A[6][12] = {};
      double U[17][6];
      for (int k = 0; k < 17; k++) {
        double x = xy[k][0], y = xy[k][1];
        double u[6] = {1., x, y, x * x, x * y, y * y};
        for (int a = 0; a < 6; a++) {
          U[k][a] = u[a];
          for (int b = 0; b < 6; b++) A[a][b] += (long double)u[a] * u[b];
        }
      }
      for (int a = 0; a < 6; a++) A[a][6 + a] = 1;
      for (int col = 0; col < 6; col++) {  // Gauss-Jordan, partial pivoting
        int pv = col;
        for (int r = col + 1; r < 6; r++)
          if (fabsl(A[r][col]) > fabsl(A[pv][col])) pv = r;
        for (int k = 0; k < 12; k++) std::swap(A[col][k], A[pv][k]);
        long double d = A[col][col];
        for (int k = 0; k < 12; k++) A[col][k] /= d;
        for (int r = 0; r < 6; r++)
          if (r != col) {
            long double f = A[r][col];
            for (int k = 0; k < 12; k++) A[r][k] -= f * A[col][k];
          }
      }
      for (int k = 0; k < 17; k++)
        for (int a = 0; a < 6; a++) {
          long double w = 0;
          for (int b = 0; b < 6; b++) w += A[a][6 + b] * U[k][b];
          pj[RBC_PJ_W + (size_t)(it3 * 17 + k) * 6 + a] = (double)w;
        }
      hp = 0.5 * hp;
    }
    RBC_TRY(upload(g->d_proj, pj, st));
  }
  *out = g;
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// stage timing
static hipEvent_t ev_get(StageTimer &t) {
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
unsigned long long g_dbuf_allocs = 0;

void rbc_stage_begin(rbc3d_ctx *c, int stage) {
  if (!c->timer.on || c->capturing) return;
  StageTimer::Rec r;
  r.stage = stage;
  r.a = ev_get(c->timer);
  r.b = ev_get(c->timer);
  (void)hipEventRecord(r.a, c->stream);
  c->timer.pending.push_back(r);
}
void rbc_stage_end(rbc3d_ctx *c, int stage) {
  if (!c->timer.on || c->capturing) return;
  for (auto it = c->timer.pending.rbegin(); it != c->timer.pending.rend(); ++it)
    if (it->stage == stage) {
      (void)hipEventRecord(it->b, c->stream);
      return;
    }
}
void rbc_stage_flush(rbc3d_ctx *c) {
  if (c->timer.pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->timer.pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->timer.ms[r.stage] += ms;
      c->timer.cnt[r.stage]++;
    }
    c->timer.pool.push_back(r.a);
    c->timer.pool.push_back(r.b);
  }
  c->timer.pending.clear();
}

static const char *kStageNames[ST_NSTAGE] = {
    "target_sort", "pair", "nbr_list", "self_sing", "near_proj", "near_sing", "linear_dl", "combine",
    "sh_glob", "sh_spline", "spread_prep", "spread", "fft", "green", "gather", "wall"};

extern "C" int rbc3d_timing(rbc3d_ctx *c, int enable) {
  if (!c) return RBC3D_EINVAL;
  rbc_stage_flush(c);
  c->timer.on = enable != 0;
  if (enable < 0) {
    for (int s = 0; s < ST_NSTAGE; s++) {
      c->timer.ms[s] = 0;
      c->timer.cnt[s] = 0;
    }
    c->timer.on = false;
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_stage_time(rbc3d_ctx *c, int stage, double *ms, long long *count,
                                const char **name) {
  if (!c || stage < 0 || stage >= ST_NSTAGE) return RBC3D_EINVAL;
  rbc_stage_flush(c);
  if (ms) *ms = c->timer.ms[stage];
  if (count) *count = c->timer.cnt[stage];
  if (name) *name = kStageNames[stage];
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" int rbc3d_create(const rbc3d_conf *conf, rbc3d_ctx **out) {
  if (!conf || !out) return RBC3D_EINVAL;
  if (conf->P < 2 || conf->P > 12 || conf->rc <= 0 || conf->alpha <= 0 || conf->ntypes < 1 ||
      conf->ntypes > 8) {
    rbc_set_error("rbc3d_create: invalid parameters");
    return RBC3D_EINVAL;
  }
  for (int d = 0; d < 3; d++)
    if (conf->Nb[d] < conf->P || conf->Lb[d] <= 0 || conf->rc >= 0.5 * conf->Lb[d]) {
      rbc_set_error("rbc3d_create: invalid box, cut-off (rc must be below Lb/2) or PME grid");
      return RBC3D_EINVAL;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    rbc_set_error("rbc3d_create: no HIP device");
    return RBC3D_EHIP;
  }
  rbc3d_ctx *c = new rbc3d_ctx();
  c->conf = *conf;
  for (int d = 0; d < 3; d++) {
    c->Lb[d] = conf->Lb[d];
    c->iLb[d] = 1. / conf->Lb[d];
  }
  for (int l = 0; l < conf->ntypes; l++) {  // initCOEFs, common/ModConf.F90:327-344
    if (conf->viscRat[l] > 0) {
      c->Acoef[l] = 1. + conf->viscRat[l];
      c->Bcoef[l] = 1. - conf->viscRat[l];
    } else {
      c->Acoef[l] = 1.;
      c->Bcoef[l] = -1.;
    }
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    rbc_set_error("rbc3d_create: stream");
    delete c;
    return RBC3D_EHIP;
  }
  // Ewald tables (common/ModEwaldFunc.F90:97-107,152-162) and MaskFunc table
  std::vector<double> t1(RBC_NTAB + 1), t2(RBC_NTAB + 1), t3(RBC_NTAB + 1), mt(RBC_NTAB + 1);
  double alpha = conf->alpha, rc = conf->rc;
  for (int i = 0; i <= RBC_NTAB; i++) {
    double r_t = sqrt(RBC_PI / alpha) * ((double)i * rc / RBC_NTAB);
    t1[i] = erfc(r_t);
    t2[i] = 2 / sqrt(alpha) * exp(-(r_t * r_t));
    t3[i] = -8 / sqrt(RBC_PI) *
            (exp(-(r_t * r_t)) * (1.5 * r_t + r_t * r_t * r_t) + 0.75 * sqrt(RBC_PI) * erfc(r_t));
    mt[i] = mask_exact((double)i / RBC_NTAB);
  }
  c->r_eps = 1.e-3 * sqrt(alpha / RBC_PI);
  int r = upload(c->d_sl1, t1, c->stream);
  if (!r) r = upload(c->d_sl2, t2, c->stream);
  if (!r) r = upload(c->d_dl1, t3, c->stream);
  if (!r) r = upload(c->d_mtab, mt, c->stream);
  {
    // the same tables in FP32 for the split form's pair kernel
    std::vector<float> tf(3 * (size_t)(RBC_NTAB + 1));
    for (int i = 0; i <= RBC_NTAB; i++) {
      tf[i] = (float)t1[i];
      tf[(RBC_NTAB + 1) + i] = (float)t2[i];
      tf[2 * (RBC_NTAB + 1) + i] = (float)t3[i];
    }
    if (!r) r = upload(c->d_tabf, tf, c->stream);
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        ncu > 0)
      c->ncu = ncu;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        lds >= 65536)
      c->lds_max = lds;
  }
  // HashTable_ComputeNumBlocks (common/ModHashTable.F90:99-126), one rank
  for (int d = 0; d < 3; d++) c->Nc[d] = std::max((int)floor(c->Lb[d] / rc), 3);
  if (c->Nc[0] > 1023 || c->Nc[1] > 1023 || c->Nc[2] > 1023) {
    rbc_set_error("rbc3d_create: more than 1023 cell-list bins on an axis");
    rbc3d_destroy(c);
    return RBC3D_EINVAL;
  }
  c->iLbNc[0] = c->iLb[0] * c->Nc[0];
  c->iLbNc[1] = c->iLb[1] * c->Nc[1];
  c->iLbNc[2] = c->Nc[2] / (c->Lb[2] - 0.);
  c->nbins = c->Nc[0] * c->Nc[1] * c->Nc[2];
  // pair-kernel grid: twice the reference's resolution, offsets within +-R bins
  // whose box distance to the target bin is within rc (a small margin keeps the
  // enumeration conservative); an axis with 2R+1 >= Nf enumerates every bin once
  {
    int R[3], lo[3], hi[3];
    double wmin[3];
    for (int d = 0; d < 3; d++) {
      c->Nf[d] = 2 * c->Nc[d];
      c->iLbNf[d] = c->iLb[d] * c->Nf[d];
      wmin[d] = c->Lb[d] / c->Nf[d];
      R[d] = (int)ceil(rc / wmin[d] * (1 + 1e-9));
      if (2 * R[d] + 1 >= c->Nf[d]) {
        lo[d] = 0;
        hi[d] = c->Nf[d] - 1;
      } else {
        lo[d] = -R[d];
        hi[d] = R[d];
      }
    }
    c->pair_img = 1;
    for (int d = 0; d < 3; d++)
      if (c->Nf[d] < 2 * R[d] + 3) c->pair_img = 0;
    // kept offsets grouped into x-rows: for each (b, q) the admissible a form one
    // contiguous range (the box distance grows with |a|), and the source bins of a
    // row are contiguous in the bin-sorted source order
    std::vector<int> offs;
    for (int b = lo[1]; b <= hi[1]; b++)
      for (int q = lo[2]; q <= hi[2]; q++) {
        int amin = 1 << 20, amax = -(1 << 20);
        for (int a = lo[0]; a <= hi[0]; a++) {
          int o[3] = {a, b, q};
          double g2 = 0;
          for (int d = 0; d < 3; d++) {
            bool all = (lo[d] == 0 && hi[d] == c->Nf[d] - 1);
            int k = all ? 0 : std::max(std::abs(o[d]) - 1, 0);
            g2 += (k * wmin[d]) * (k * wmin[d]);
          }
          if (g2 > rc * rc * (1 + 1e-9)) continue;
          amin = std::min(amin, a);
          amax = std::max(amax, a);
        }
        if (amin > amax) continue;
        unsigned code = (unsigned)(amin + 128) | ((unsigned)(amax + 128) << 8) |
                        ((unsigned)(b + 128) << 16) | ((unsigned)(q + 128) << 24);
        offs.push_back((int)code);
      }
    c->nfbins = c->Nf[0] * c->Nf[1] * c->Nf[2];
    c->noffs = offs.size();
    if (!r) r = upload(c->d_offs, offs, c->stream);
  }
  if (!r) r = hipStreamSynchronize(c->stream) == hipSuccess ? 0 : RBC3D_EHIP;
  if (r) {
    rbc3d_destroy(c);
    return r;
  }
  *out = c;
  return RBC3D_OK;
}

template <typename T>
static void rel(DBuf<T> &b) {
  b.release();
}

static void drop_graph(rbc3d_ctx *c);

extern "C" int rbc3d_destroy(rbc3d_ctx *c) {
  if (!c) return RBC3D_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  c->d_gin.release();
  c->d_gout.release();
  for (auto *g : c->gtypes) {
    for (DBuf<double> *b : {&g->d_th, &g->d_phi, &g->d_w, &g->d_pw, &g->d_pws, &g->d_pg, &g->d_pe, &g->d_cs,
                            &g->d_Dm, &g->d_Dn, &g->d_thG, &g->d_phiG, &g->d_pwt, &g->d_mask,
                            &g->d_proj})
      b->release();
    delete g;
  }
  for (DBuf<double> *b : {&c->d_sl1, &c->d_sl2, &c->d_dl1, &c->d_mtab, &c->d_x, &c->d_a3, &c->d_detj,
                          &c->d_ds, &c->d_f, &c->d_g, &c->d_fs, &c->d_gs, &c->d_spl, &c->d_src_rec,
                          &c->d_tx, &c->d_tA, &c->d_v, &c->d_vacc, &c->d_nbr_dist, &c->d_dvitem,
                          &c->d_scratch, &c->d_cell_sph, &c->d_chunk_sph, &c->d_grid, &c->d_cgrid, &c->d_vgrid,
                          &c->d_bb, &c->d_tw, &c->d_pme_src, &c->d_shbuf, &c->d_shbuf2,
                          &c->wall.d_x, &c->wall.d_f, &c->wall.d_area, &c->wall.d_eps,
                          &c->wall.d_cen, &c->wall.d_val})
    b->release();
  for (DBuf<int> *b : {&c->d_node_cell, &c->d_node_ll, &c->d_src_bin, &c->d_src_order,
                       &c->d_bin_start, &c->d_tmp_i, &c->d_tindx, &c->d_tact, &c->d_tbin,
                       &c->d_torder, &c->d_tbin_start, &c->d_chunk_off, &c->d_nbr_cnt,
                       &c->d_nbr_cell, &c->d_nbr_node, &c->d_cell_cand, &c->d_work, &c->d_work_cnt, &c->d_flag,
                       &c->d_cellbin_start, &c->d_cellbin_order, &c->d_pme_key, &c->d_pme_order,
                       &c->d_pme_start, &c->wall.d_e2v, &c->wall.d_glb, &c->wall.d_bin_start,
                       &c->wall.d_order, &c->wall.d_rowptr, &c->wall.d_col})
    b->release();
  c->d_cells.release();
  c->d_cub_tmp.release();
  c->d_tp_ptr.release();
  c->d_tp_int.release();
  c->d_offs.release();
  c->d_cell_sol.release();
  c->d_tmp_f.release();
  c->d_fbin_start.release();
  c->d_worklist.release();
  c->d_ikey.release();
  c->d_ival.release();
  c->d_islot.release();
  c->d_linpart.release();
  rbc_stage_flush(c);
  for (auto e : c->timer.pool) (void)hipEventDestroy(e);
  for (auto e : c->gmres.ev)
    if (e) (void)hipEventDestroy(e);
  if (c->gmres.hpin) (void)hipHostFree(c->gmres.hpin);
  if (c->gmres.hcoh) (void)hipHostFree(c->gmres.hcoh);
  c->d_pairstat.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return RBC3D_OK;
}

extern "C" int rbc3d_set_cells(rbc3d_ctx *c, int nrbc, const int *nlat0, const int *nlat,
                               const int *celltype, const double *x, const double *a3,
                               const double *detj) {
  if (!c || nrbc < 0 || (nrbc > 0 && (!nlat0 || !nlat || !celltype || !x || !a3 || !detj)))
    return RBC3D_EINVAL;
  int prev_dof = 0;
  for (const CellRec &r : c->cells) prev_dof += 3 * r.nlat0 * r.nlat0;
  const int prev_nrbc = c->nrbc;
  c->nrbc = nrbc;
  c->cells.resize(nrbc);
  int np = 0, coff = 0, choff = 0;
  long long spl = 0;
  for (int i = 0; i < nrbc; i++) {
    if (nlat[i] < 2 || nlat[i] > 64 || nlat0[i] < 1 || nlat0[i] > nlat[i] || celltype[i] < 1 ||
        celltype[i] > c->conf.ntypes) {
      rbc_set_error("rbc3d_set_cells: invalid cell " + std::to_string(i));
      return RBC3D_EINVAL;
    }
    GridType *g;
    RBC_TRY(make_gtype(c, nlat[i], &g));
    CellRec &r = c->cells[i];
    r.off = np;
    r.gtype = (int)(std::find(c->gtypes.begin(), c->gtypes.end(), g) - c->gtypes.begin());
    r.nlat = g->nlat;
    r.nlon = g->nlon;
    r.nlat0 = nlat0[i];
    r.celltype = celltype[i];
    r.Acoef = c->Acoef[celltype[i] - 1];
    r.Bcoef = c->Bcoef[celltype[i] - 1];
    r.radius = g->radius;
    r.spl = spl;
    r.coff = coff;
    coff += 3 * r.nlat0 * r.nlat0;
    r.choff = choff;
    choff += (r.nlat * r.nlon + 63) / 64;
    int npt = g->nlat * g->nlon;
    // area and mesh size (RBC_ComputeGeometry, common/ModRbc.F90:482-505)
    double area = 0;
    for (int ilat = 0; ilat < g->nlat; ilat++)
      for (int ilon = 0; ilon < g->nlon; ilon++)
        area += detj[(size_t)np + ilat + (size_t)g->nlat * ilon] * g->w[ilat];
    r.area = area;
    r.meshSize = sqrt(area / (g->nlat * g->nlat));
    np += npt;
    spl += (long long)SPL_NVAR * 2 * g->nlat * g->nlon * 4;
  }
  c->np = np;
  c->sbq_ready = false;
  c->x_moved = false;
  c->nchunks = choff;
  c->geom_epoch++;
  // the previous cell-BIE solution is the next solve's initial guess only for the same
  // set of unknowns
  if (nrbc != prev_nrbc || coff != prev_dof) c->cell_sol_init = false;
  // surface ids follow the cells: the wall is nrbc + 1 (GlobData_Init,
  // common/ModData.F90:40-49), whichever of cells and wall was set first
  c->wall.id = nrbc + 1;
  c->spl_total = spl;
  hipStream_t st = c->stream;
  RBC_TRY(upload(c->d_cells, c->cells, st));
  for (DBuf<double> *b : {&c->d_x, &c->d_a3, &c->d_f, &c->d_g, &c->d_fs, &c->d_gs})
    RBC_TRY(b->ensure((size_t)3 * np + 1));
  RBC_TRY(c->d_detj.ensure(np + 1));
  RBC_TRY(c->d_ds.ensure(np + 1));
  RBC_TRY(c->d_spl.ensure(spl + 1));
  // node arrays: [d][np] from the per-cell (nlat,nlon,3) blocks
  std::vector<double> hx((size_t)3 * np), ha((size_t)3 * np), hd(np), hds(np);
  std::vector<int> hcell(np), hll(np);
  size_t off = 0;
  for (int i = 0; i < nrbc; i++) {
    const CellRec &r = c->cells[i];
    GridType *g = c->gtypes[r.gtype];
    int npt = r.nlat * r.nlon;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++) {
        hx[r.off + q + (size_t)np * d] = x[3 * off + q + (size_t)npt * d];
        ha[r.off + q + (size_t)np * d] = a3[3 * off + q + (size_t)npt * d];
      }
    for (int q = 0; q < npt; q++) {
      int ilat = q % r.nlat, ilon = q / r.nlat;
      hd[r.off + q] = detj[off + q];
      hds[r.off + q] = detj[off + q] * g->w[ilat];
      hcell[r.off + q] = i;
      hll[r.off + q] = ilat | (ilon << 16);
    }
    off += npt;
  }
  for (double v : hx)
    if (!std::isfinite(v)) {
      rbc_set_error("rbc3d_set_cells: non-finite coordinate");
      return RBC3D_ENONFINITE;
    }
  RBC_TRY(upload(c->d_node_cell, hcell, st));
  RBC_TRY(upload(c->d_node_ll, hll, st));
  HIP_TRY(hipMemcpyAsync(c->d_x.p, hx.data(), hx.size() * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_a3.p, ha.data(), ha.size() * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_detj.p, hd.data(), hd.size() * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_ds.p, hds.data(), hds.size() * 8, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemsetAsync(c->d_f.p, 0, sizeof(double) * 3 * np, st));
  HIP_TRY(hipMemsetAsync(c->d_g.p, 0, sizeof(double) * 3 * np, st));
  HIP_TRY(hipMemsetAsync(c->d_fs.p, 0, sizeof(double) * 3 * np, st));
  HIP_TRY(hipMemsetAsync(c->d_gs.p, 0, sizeof(double) * 3 * np, st));
  HIP_TRY(hipMemsetAsync(c->d_spl.p, 0, sizeof(double) * (spl + 1), st));
  RBC_TRY(rbc_sh_build_splines(c, 0));  // x, a3, detJ splines
  RBC_TRY(rbc_build_source_list(c));
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

extern "C" int rbc3d_set_cell_density(rbc3d_ctx *c, int which, const double *dens) {
  if (!c || (which != 1 && which != 2) || !dens) return RBC3D_EINVAL;
  if (c->nrbc == 0) return RBC3D_ESTATE;
  int np = c->np;
  std::vector<double> h((size_t)3 * np);
  size_t off = 0;
  for (int i = 0; i < c->nrbc; i++) {
    const CellRec &r = c->cells[i];
    int npt = r.nlat * r.nlon;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < npt; q++) h[r.off + q + (size_t)np * d] = dens[3 * off + q + (size_t)npt * d];
    off += npt;
  }
  DBuf<double> &dst = which == 1 ? c->d_f : c->d_g;
  HIP_TRY(hipMemcpyAsync(dst.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
  RBC_TRY(rbc_update_density(c, which));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_cell_spline(rbc3d_ctx *c, int irbc, int which, double *u, double *u1,
                                 double *u2, double *u12) {
  if (!c || irbc < 0 || irbc >= c->nrbc || which < 0 || which > 4) return RBC3D_EINVAL;
  const CellRec &r = c->cells[irbc];
  int m = 2 * r.nlat, n = r.nlon;
  int v0 = which == 0 ? SPL_X : which == 1 ? SPL_A3 : which == 2 ? SPL_DETJ : which == 3 ? SPL_F : SPL_G;
  int nvar = which == 2 ? 1 : 3;
  size_t per = (size_t)m * n * 4;
  std::vector<double> h(per * nvar);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(h.data(), c->d_spl.p + r.spl + per * v0, per * nvar * 8, hipMemcpyDeviceToHost));
  for (int l = 0; l < nvar; l++)
    for (int q = 0; q < m * n; q++) {
      size_t o = per * l + 4 * (size_t)q;
      u[q + (size_t)m * n * l] = h[o];
      u1[q + (size_t)m * n * l] = h[o + 1];
      u2[q + (size_t)m * n * l] = h[o + 2];
      u12[q + (size_t)m * n * l] = h[o + 3];
    }
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// target lists: host -> device scratch
static int upload_targets(rbc3d_ctx *c, int nT, const double *tx, const int *tindx,
                          const int *active, const double *A, const double *v) {
  hipStream_t st = c->stream;
  RBC_TRY(c->d_tx.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tA.ensure(nT + 1));
  RBC_TRY(c->d_v.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tindx.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tact.ensure(nT + 1));
  if (nT == 0) return RBC3D_OK;
  for (int i = 0; i < 3 * nT; i++)
    if (!std::isfinite(tx[i])) {
      rbc_set_error("non-finite target coordinate");
      return RBC3D_ENONFINITE;
    }
  HIP_TRY(hipMemcpyAsync(c->d_tx.p, tx, sizeof(double) * 3 * nT, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_tA.p, A, sizeof(double) * nT, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_v.p, v, sizeof(double) * 3 * nT, hipMemcpyHostToDevice, st));
  if (tindx) HIP_TRY(hipMemcpyAsync(c->d_tindx.p, tindx, sizeof(int) * 3 * nT, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(c->d_tact.p, active, sizeof(int) * nT, hipMemcpyHostToDevice, st));
  return RBC3D_OK;
}

static int download_v(rbc3d_ctx *c, int nT, double *v) {
  if (nT == 0) return RBC3D_OK;
  HIP_TRY(hipMemcpyAsync(v, c->d_v.p, sizeof(double) * 3 * nT, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_add_int_on_rbcs(rbc3d_ctx *c, double c1, double c2, int nT, const double *tx,
                                     const int *tindx, const int *active, const double *Acoef,
                                     double *v) {
  if (!c || nT < 0 || (nT > 0 && (!tx || !tindx || !active || !Acoef || !v))) return RBC3D_EINVAL;
  RBC_TRY(rbc_refresh_moved(c));
  if (c->nrbc == 0 || nT == 0) return RBC3D_OK;
  RBC_TRY(upload_targets(c, nT, tx, tindx, active, Acoef, v));
  RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, nT, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                  c->d_v.p, false));
  return download_v(c, nT, v);
}

extern "C" int rbc3d_pair_stats(rbc3d_ctx *c, long long *npairs, unsigned long long *checksum) {
  if (!c) return RBC3D_EINVAL;
  // the statistics of the last AddIntOnRbcs stay on the device until asked for
  if (c->d_pairstat.p) {
    unsigned long long hs[2];
    HIP_TRY(hipMemcpyAsync(hs, c->d_pairstat.p, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->last_npairs = (long long)hs[0];
    c->last_checksum = hs[1];
  }
  if (npairs) *npairs = c->last_npairs;
  if (checksum) *checksum = c->last_checksum;
  return RBC3D_OK;
}

extern "C" int rbc3d_set_precision(rbc3d_ctx *c, int mode) {
  if (!c || (mode != RBC3D_PREC_FP64 && mode != RBC3D_PREC_SPLIT)) return RBC3D_EINVAL;
  if (c->prec != mode) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->prec = mode;
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_set_split_pairs(rbc3d_ctx *c, int fp32) {
  if (!c) return RBC3D_EINVAL;
  if (c->split_pair32 != (fp32 != 0)) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->split_pair32 = fp32 != 0;
    c->mv_epoch = -1;  // a recorded matvec graph holds the other pair kernel
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_set_gmres_restart(rbc3d_ctx *c, int restart) {
  if (!c || restart < 1 || restart > RBC3D_GMRES_MAX_RESTART) {
    rbc_set_error("rbc3d_set_gmres_restart: restart must be in 1..255");
    return RBC3D_EINVAL;
  }
  c->gmres_restart = restart;
  return RBC3D_OK;
}

extern "C" int rbc3d_set_geometry_reuse(rbc3d_ctx *c, int enable) {
  if (!c) return RBC3D_EINVAL;
  c->geom_reuse = enable != 0;
  c->cache_epoch = -1;
  return RBC3D_OK;
}

extern "C" int rbc3d_near_stats(rbc3d_ctx *c, long long *nitems, long long *nwork) {
  if (!c) return RBC3D_EINVAL;
  if (nitems) *nitems = c->last_nitems;
  if (nwork) *nwork = c->last_nwork;
  return RBC3D_OK;
}

extern "C" int rbc3d_pme_distrib_source(rbc3d_ctx *c, double c1, double c2, int cells, int walls) {
  if (!c) return RBC3D_EINVAL;
  if (cells) RBC_TRY(rbc_refresh_moved(c));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, cells, walls));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_pme_transform(rbc3d_ctx *c) {
  if (!c) return RBC3D_EINVAL;
  if (!c->pme_ready) return RBC3D_ESTATE;
  RBC_TRY(rbc_pme_transform(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_pme_add_interp_vel(rbc3d_ctx *c, int nT, const double *tx, const int *active,
                                        const double *Acoef, double *v) {
  if (!c || nT < 0 || (nT > 0 && (!tx || !active || !Acoef || !v))) return RBC3D_EINVAL;
  if (!c->pme_ready) return RBC3D_ESTATE;
  if (nT == 0) return RBC3D_OK;
  RBC_TRY(upload_targets(c, nT, tx, nullptr, active, Acoef, v));
  RBC_TRY(rbc_pme_gather(c, nT, c->d_tx.p, c->d_tact.p, c->d_tA.p, c->d_v.p));
  return download_v(c, nT, v);
}

extern "C" int rbc3d_pme_grid(rbc3d_ctx *c, int which, double *out) {
  if (!c || !out || which < 0 || which > 2) return RBC3D_EINVAL;
  if (!c->pme_ready) return RBC3D_ESTATE;
  size_t ng = (size_t)c->conf.Nb[0] * c->conf.Nb[1] * c->conf.Nb[2];
  HIP_TRY(hipStreamSynchronize(c->stream));
  // the spread writes only the fields its coefficients make nonzero
  if (which == 0 && !c->flag_sl)
    std::fill(out, out + 3 * ng, 0.);
  else if (which == 1 && !c->flag_dl)
    std::fill(out, out + 9 * ng, 0.);
  else if (c->pme_grid_prec == RBC3D_PREC_SPLIT) {
    // FP32 grids of the split form, widened on the host
    const size_t off = which == 1 ? 3 * ng : 0, n = (which == 1 ? 9 : 3) * ng;
    const float *src = which == 2 ? (const float *)c->d_vgrid.p : (const float *)c->d_grid.p + off;
    std::vector<float> tmp(n);
    HIP_TRY(hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) out[i] = tmp[i];
  } else if (which == 0)
    HIP_TRY(hipMemcpy(out, c->d_grid.p, 3 * ng * 8, hipMemcpyDeviceToHost));
  else if (which == 1)
    HIP_TRY(hipMemcpy(out, c->d_grid.p + 3 * ng, 9 * ng * 8, hipMemcpyDeviceToHost));
  else
    HIP_TRY(hipMemcpy(out, c->d_vgrid.p, 3 * ng * 8, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

// standalone SH transforms on host arrays (device round trip)
static int sh_host(int kind, int nlat, int nlon, int nt, double *g, int idg, int jdg, double *a,
                   double *b, int ida, int jda) {
  if (nlat < 2 || nlon < 2 || nt < 1 || idg < nlat || jdg < nlon || !g || !a || !b) return RBC3D_EINVAL;
  int mmax = std::min(nlat, nlon / 2 + 1);
  if (ida < mmax || jda < nlat) return RBC3D_EINVAL;
  size_t ng = (size_t)idg * jdg * nt, na = (size_t)ida * jda * nt;
  DBuf<double> bg, ba, bb;  // released on every return path
  RBC_TRY(bg.ensure(ng));
  RBC_TRY(ba.ensure(na));
  RBC_TRY(bb.ensure(na));
  double *dg = bg.p, *da = ba.p, *db = bb.p;
  int r;
  if (kind == 0) {
    HIP_TRY(hipMemcpy(dg, g, ng * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(da, a, na * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db, b, na * 8, hipMemcpyHostToDevice));
    r = rbc_sh_anal_batch(nlat, nlon, nt, dg, idg, jdg, da, db, ida, jda, 0);
    if (!r) HIP_TRY(hipMemcpy(a, da, na * 8, hipMemcpyDeviceToHost));
    if (!r) HIP_TRY(hipMemcpy(b, db, na * 8, hipMemcpyDeviceToHost));
  } else {
    HIP_TRY(hipMemcpy(dg, g, ng * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(da, a, na * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db, b, na * 8, hipMemcpyHostToDevice));
    r = rbc_sh_synth_batch(nlat, nlon, nt, dg, idg, jdg, da, db, ida, jda, kind == 2, 0);
    if (!r) HIP_TRY(hipMemcpy(g, dg, ng * 8, hipMemcpyDeviceToHost));
  }
  return r;
}
extern "C" int rbc3d_sh_anal_gau(int nlat, int nlon, int nt, const double *g, int idg, int jdg,
                                 double *a, double *b, int ida, int jda) {
  return sh_host(0, nlat, nlon, nt, (double *)g, idg, jdg, a, b, ida, jda);
}
extern "C" int rbc3d_sh_synth_gau(int nlat, int nlon, int nt, double *g, int idg, int jdg,
                                  const double *a, const double *b, int ida, int jda) {
  return sh_host(1, nlat, nlon, nt, g, idg, jdg, (double *)a, (double *)b, ida, jda);
}
extern "C" int rbc3d_sh_synth_equ(int nlat, int nlon, int nt, double *g, int idg, int jdg,
                                  const double *a, const double *b, int ida, int jda) {
  return sh_host(2, nlat, nlon, nt, g, idg, jdg, (double *)a, (double *)b, ida, jda);
}

int rbc_sh_grad_batch(int nlat, int nlon, int nt, double *d_v, double *d_w, int idv, int jdv,
                      const double *d_a, const double *d_b, int ida, int jda, hipStream_t st);
extern "C" int rbc3d_sh_grad_gau(int nlat, int nlon, int nt, double *v, double *w, int idv,
                                 int jdv, const double *a, const double *b, int ida, int jda) {
  if (nlat < 2 || nlon < 2 || nt < 1 || idv < nlat || jdv < nlon || !v || !w || !a || !b)
    return RBC3D_EINVAL;
  int mmax = std::min(nlat, nlon / 2 + 1);
  if (ida < mmax || jda < nlat) return RBC3D_EINVAL;
  size_t ng = (size_t)idv * jdv * nt, na = (size_t)ida * jda * nt;
  DBuf<double> bv, bw, ba, bb;  // released on every return path
  RBC_TRY(bv.ensure(ng));
  RBC_TRY(bw.ensure(ng));
  RBC_TRY(ba.ensure(na));
  RBC_TRY(bb.ensure(na));
  HIP_TRY(hipMemcpy(bv.p, v, ng * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(bw.p, w, ng * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ba.p, a, na * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(bb.p, b, na * 8, hipMemcpyHostToDevice));
  RBC_TRY(rbc_sh_grad_batch(nlat, nlon, nt, bv.p, bw.p, idv, jdv, ba.p, bb.p, ida, jda, 0));
  HIP_TRY(hipMemcpy(v, bv.p, ng * 8, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(w, bw.p, ng * 8, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

extern "C" int rbc3d_cell_dof(rbc3d_ctx *c, int *dof) {
  if (!c || !dof) return RBC3D_EINVAL;
  int s = 0;
  for (auto &r : c->cells) s += 3 * r.nlat0 * r.nlat0;
  *dof = s;
  return RBC3D_OK;
}

extern "C" int rbc3d_glob_sph_trans(rbc3d_ctx *c, double *v, double *cf, int direction) {
  if (!c || !v || !cf || (direction != 1 && direction != -1)) return RBC3D_EINVAL;
  if (c->nrbc == 0) return RBC3D_OK;
  int dof;
  rbc3d_cell_dof(c, &dof);
  size_t nv = (size_t)3 * c->np;
  RBC_TRY(c->d_shbuf.ensure(nv + 1));
  RBC_TRY(c->d_shbuf2.ensure(dof + 1));
  if (direction == 1) {
    HIP_TRY(hipMemcpyAsync(c->d_shbuf.p, v, nv * 8, hipMemcpyHostToDevice, c->stream));
    RBC_TRY(rbc_sh_glob(c, c->d_shbuf.p, c->d_shbuf2.p, 1));
    HIP_TRY(hipMemcpyAsync(cf, c->d_shbuf2.p, (size_t)dof * 8, hipMemcpyDeviceToHost, c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(c->d_shbuf2.p, cf, (size_t)dof * 8, hipMemcpyHostToDevice, c->stream));
    RBC_TRY(rbc_sh_glob(c, c->d_shbuf2.p, c->d_shbuf.p, -1));
    HIP_TRY(hipMemcpyAsync(v, c->d_shbuf.p, nv * 8, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// composite operators

__global__ void k_tlist_rbc(const CellRec *__restrict__ cells, const int *__restrict__ node_cell,
                            const int *__restrict__ node_ll, const double *__restrict__ x, int np,
                            double *tx, int *tindx, int *act, double *A) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  int ci = node_cell[p], ll = node_ll[p];
  for (int d = 0; d < 3; d++) tx[p + (size_t)np * d] = x[p + (size_t)np * d];
  tindx[p] = ci + 1;
  tindx[p + np] = (ll & 0xffff) + 1;
  tindx[p + 2 * (size_t)np] = (ll >> 16) + 1;
  act[p] = 1;
  A[p] = cells[ci].Acoef;
}

__global__ void k_axpy_bkg(double *v, int n, const double *__restrict__ A, double s0, double s1,
                           double s2, int divide) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double f = divide ? 1.0 / A[p] : 1.0;
  double s[3] = {s0, s1, s2};
  for (int d = 0; d < 3; d++) v[p + (size_t)n * d] += divide ? s[d] / A[p] : s[d] * f;
}

__global__ void k_add(double *v, const double *__restrict__ g, size_t n) {
  size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (p < n) v[p] += g[p];
}

// scatter node-ordered g (np,3) into the cell density and rebuild splines + list density
int rbc_set_g_nodes(rbc3d_ctx *c, const double *d_gnodes) {
  HIP_TRY(hipMemcpyAsync(c->d_g.p, d_gnodes, sizeof(double) * 3 * c->np, hipMemcpyDeviceToDevice,
                         c->stream));
  return rbc_update_density(c, 2);
}

static int ensure_rbc_targets(rbc3d_ctx *c) {
  int np = c->np;
  RBC_TRY(c->d_tx.ensure((size_t)3 * np + 1));
  RBC_TRY(c->d_tA.ensure(np + 1));
  RBC_TRY(c->d_v.ensure((size_t)3 * np + 1));
  RBC_TRY(c->d_tindx.ensure((size_t)3 * np + 1));
  RBC_TRY(c->d_tact.ensure(np + 1));
  k_tlist_rbc<<<(np + 255) / 256, 256, 0, c->stream>>>(c->d_cells.p, c->d_node_cell.p,
                                                         c->d_node_ll.p, c->d_x.p, np, c->d_tx.p,
                                                         c->d_tindx.p, c->d_tact.p, c->d_tA.p);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

__global__ void k_zero(double *v, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.;
}

// MyMatMult (common/ModVelSolver.F90:523-601) on device vectors; queues work only (the
// host reads nothing back), so a stream capture can record it
static int cell_matmult_body(rbc3d_ctx *c, const double *d_u, double *d_b) {
  int np = c->np;
  hipStream_t st = c->stream;
  RBC_TRY(c->d_shbuf.ensure((size_t)3 * np + 1));
  RBC_TRY(c->d_vacc.ensure((size_t)3 * np + 1));
  // g = synthesis(u); rbc%g <- g; Rbc_BuildSurfaceSource(gFlag); SourceList_UpdateDensity(G)
  RBC_TRY(rbc_sh_glob(c, d_u, c->d_shbuf.p, -1));
  RBC_TRY(rbc_set_g_nodes(c, c->d_shbuf.p));
  RBC_TRY(ensure_rbc_targets(c));
  k_zero<<<(unsigned)((3 * (size_t)np + 255) / 256), 256, 0, st>>>(c->d_v.p, 3 * (size_t)np);
  double c1 = 0., c2 = -1. / (4. * RBC_PI);
  RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, np, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                  c->d_v.p, true));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, 1, 0));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather(c, np, c->d_tx.p, c->d_tact.p, c->d_tA.p, c->d_v.p));
  size_t n3 = (size_t)3 * np;
  k_add<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(c->d_v.p, c->d_shbuf.p, n3);
  HIP_TRY(hipGetLastError());
  RBC_TRY(rbc_defl_matmult(c, c->d_shbuf.p, c->d_v.p));
  RBC_TRY(rbc_sh_glob(c, c->d_v.p, d_b, 1));
  return RBC3D_OK;
}

static void drop_graph(rbc3d_ctx *c) {
  if (c->mv_exec) (void)hipGraphExecDestroy(c->mv_exec);
  if (c->mv_graph) (void)hipGraphDestroy(c->mv_graph);
  c->mv_exec = nullptr;
  c->mv_graph = nullptr;
  c->mv_epoch = -1;
}

extern "C" int rbc3d_set_graph(rbc3d_ctx *c, int on) {
  if (!c) return RBC3D_EINVAL;
  c->use_graph = on != 0;
  if (!c->use_graph) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
  }
  return RBC3D_OK;
}

// The graph form of the cell matvec (geometry reuse on): the matvec at fixed geometry is
// a fixed sequence of launches on fixed buffers, recorded once into a hipGraph on the
// context's own input and output vectors and replayed per call (u copied in, b out).
// A matvec whose geometry epoch, precision form or buffer allocations differ from the
// recorded one runs eagerly (which brings every buffer to size and fills the geometry
// caches) and records a new graph.
static int cell_matmult_graph(rbc3d_ctx *c, const double *d_u, double *d_b) {
  hipStream_t st = c->stream;
  int dof;
  RBC_TRY(rbc3d_cell_dof(c, &dof));
  RBC_TRY(c->d_gin.ensure(dof + 1));
  RBC_TRY(c->d_gout.ensure(dof + 1));
  const size_t bytes = sizeof(double) * (size_t)dof;
  const bool valid = c->mv_exec && c->mv_epoch == c->geom_epoch && c->mv_prec == c->prec &&
                     c->mv_allocs == g_dbuf_allocs && c->cache_epoch == c->geom_epoch;
  HIP_TRY(hipMemcpyAsync(c->d_gin.p, d_u, bytes, hipMemcpyDeviceToDevice, st));
  if (valid) {
    HIP_TRY(hipGraphLaunch(c->mv_exec, st));
    HIP_TRY(hipMemcpyAsync(d_b, c->d_gout.p, bytes, hipMemcpyDeviceToDevice, st));
    return RBC3D_OK;
  }
  HIP_TRY(hipStreamSynchronize(st));
  drop_graph(c);
  RBC_TRY(cell_matmult_body(c, c->d_gin.p, c->d_gout.p));  // this call's result
  HIP_TRY(hipMemcpyAsync(d_b, c->d_gout.p, bytes, hipMemcpyDeviceToDevice, st));
  RBC_TRY(rbc3d_sync(c));
  rbc_stage_flush(c);
  if (c->cache_epoch != c->geom_epoch) return RBC3D_OK;  // nothing kept to replay against
  const unsigned long long allocs = g_dbuf_allocs;
  c->capturing = true;
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int r = cell_matmult_body(c, c->d_gin.p, c->d_gout.p);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  c->capturing = false;
  if (r != RBC3D_OK || e != hipSuccess || !g || g_dbuf_allocs != allocs) {
    if (g) (void)hipGraphDestroy(g);
    if (r != RBC3D_OK) return r;
    rbc_set_error(std::string("cell matvec graph capture failed: ") +
                  (e != hipSuccess ? hipGetErrorString(e) : "a buffer was allocated while recording"));
    return RBC3D_EHIP;
  }
  c->mv_graph = g;
  HIP_TRY(hipGraphInstantiate(&c->mv_exec, g, nullptr, nullptr, 0));
  c->mv_epoch = c->geom_epoch;
  c->mv_prec = c->prec;
  c->mv_allocs = allocs;
  return RBC3D_OK;
}

extern "C" int rbc3d_graph_ready(rbc3d_ctx *c, int *ready) {
  if (!c || !ready) return RBC3D_EINVAL;
  *ready = c->mv_exec && c->mv_epoch == c->geom_epoch && c->mv_prec == c->prec &&
           c->mv_allocs == g_dbuf_allocs;
  return RBC3D_OK;
}

extern "C" int rbc3d_cell_matmult_dev(rbc3d_ctx *c, const double *d_u, double *d_b) {
  if (!c || !d_u || !d_b) return RBC3D_EINVAL;
  if (c->nrbc == 0) return RBC3D_ESTATE;
  RBC_TRY(rbc_refresh_moved(c));
  if (c->use_graph && c->geom_reuse) return cell_matmult_graph(c, d_u, d_b);
  return cell_matmult_body(c, d_u, d_b);
}

extern "C" int rbc3d_cell_matmult(rbc3d_ctx *c, const double *u, double *b) {
  if (!c || !u || !b) return RBC3D_EINVAL;
  int dof;
  rbc3d_cell_dof(c, &dof);
  RBC_TRY(c->d_shbuf2.ensure(2 * (size_t)dof + 2));
  double *du = c->d_shbuf2.p, *db = c->d_shbuf2.p + dof + 1;
  HIP_TRY(hipMemcpyAsync(du, u, (size_t)dof * 8, hipMemcpyHostToDevice, c->stream));
  RBC_TRY(rbc3d_cell_matmult_dev(c, du, db));
  HIP_TRY(hipMemcpyAsync(b, db, (size_t)dof * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_cell_rhs_dev(rbc3d_ctx *c, double *d_rhs) {
  if (!c || !d_rhs) return RBC3D_EINVAL;
  if (c->nrbc == 0) return RBC3D_ESTATE;
  RBC_TRY(rbc_refresh_moved(c));
  int np = c->np, dof;
  rbc3d_cell_dof(c, &dof);
  hipStream_t st = c->stream;
  RBC_TRY(ensure_rbc_targets(c));
  HIP_TRY(hipMemsetAsync(c->d_v.p, 0, sizeof(double) * 3 * np, st));
  double c1 = 1. / (4. * RBC_PI), c2 = 0.;
  RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, np, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                  c->d_v.p, true));
  if (c->wall.nvert)
    RBC_TRY(rbc_add_int_on_walls_dev(c, c1, np, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                     c->d_v.p, false));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, 1, 1));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather(c, np, c->d_tx.p, c->d_tact.p, c->d_tA.p, c->d_v.p));
  k_axpy_bkg<<<(np + 255) / 256, 256, 0, st>>>(c->d_v.p, np, c->d_tA.p, 2. * c->conf.vBkg[0],
                                                2. * c->conf.vBkg[1], 2. * c->conf.vBkg[2], 1);
  RBC_TRY(rbc_defl_rhs(c, c->d_v.p));
  RBC_TRY(rbc_sh_glob(c, c->d_v.p, d_rhs, 1));
  return RBC3D_OK;
}

extern "C" int rbc3d_cell_rhs(rbc3d_ctx *c, double *rhs) {
  if (!c || !rhs) return RBC3D_EINVAL;
  int dof;
  rbc3d_cell_dof(c, &dof);
  RBC_TRY(c->d_shbuf2.ensure(dof + 1));
  RBC_TRY(rbc3d_cell_rhs_dev(c, c->d_shbuf2.p));
  HIP_TRY(hipMemcpyAsync(rhs, c->d_shbuf2.p, (size_t)dof * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_calc_velocity_field(rbc3d_ctx *c, int nT, const double *tx, const int *tindx,
                                         const int *active, const double *Acoef, double *v) {
  if (!c || nT < 0 || (nT > 0 && (!tx || !tindx || !active || !Acoef || !v))) return RBC3D_EINVAL;
  if (nT == 0) return RBC3D_OK;
  RBC_TRY(rbc_refresh_moved(c));
  for (int i = 0; i < 3 * nT; i++) v[i] = 0.;
  RBC_TRY(upload_targets(c, nT, tx, tindx, active, Acoef, v));
  double c1 = 1.0 / (4. * RBC_PI), c2 = 1.0 / (4. * RBC_PI);
  if (c->nrbc)
    RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, nT, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                    c->d_v.p, false));
  if (c->wall.nvert)
    RBC_TRY(rbc_add_int_on_walls_dev(c, c1, nT, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                     c->d_v.p, tindx[0] == c->wall.id));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, 1, 1));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather(c, nT, c->d_tx.p, c->d_tact.p, c->d_tA.p, c->d_v.p));
  k_axpy_bkg<<<(nT + 255) / 256, 256, 0, c->stream>>>(c->d_v.p, nT, c->d_tA.p, c->conf.vBkg[0],
                                                       c->conf.vBkg[1], c->conf.vBkg[2], 0);
  HIP_TRY(hipGetLastError());
  return download_v(c, nT, v);
}

extern "C" int rbc3d_stream(rbc3d_ctx *c, void **s) {
  if (!c || !s) return RBC3D_EINVAL;
  *s = (void *)c->stream;
  return RBC3D_OK;
}

extern "C" int rbc3d_sync(rbc3d_ctx *c) {
  if (!c) return RBC3D_EINVAL;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}
