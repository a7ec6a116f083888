// Wall (vessel) surfaces as Tri3 meshes (common/ModIntOnWalls.F90, ModWall.F90,
// ModNoSlip.F90):
//   * the wall source list (triangle centroids) on the reference's cell list, so the
//     visited (target, triangle) set is the reference's 27-bin set;
//   * AddIntOnWalls for targets off the wall: one lane per target, 7-point rule or the
//     Duffy rule near the triangle (Tri_Int_Regular/Tri_Int_Duffy, MinDistToTri);
//   * PrepareSingIntOnWall as a 3x3-block sparse matrix assembled on the GPU: count,
//     fill per-contribution blocks, segmented sort by column, sum duplicates;
//   * the wall no-slip operator (MyMatMult :255-308) = block SpMV + wall-only PME,
//     and Compute_Wall_Residual_Vel :153-193.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "internal.h"
#include "surf_core.h"

struct TriRules {
  double r7[7][2], w7[7];  // gqTri7 (common/ModQuadRule.F90:70-94)
  double rd[4], wd[4];     // GauLeg(0,1,4) for the Duffy rule (Tri_Int_Duffy :398-404)
};

static void gauleg_h(double x1, double x2, int n, double *x, double *w) {
  int m = (n + 1) / 2;
  double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
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
    x[j] = xm - xl * z;
    x[n - 1 - j] = xm + xl * z;
    w[j] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - 1 - j] = w[j];
  }
}

static TriRules make_rules() {
  TriRules t;
  double r = (6. - sqrt(15.)) / 21., s = r, u = 1 - r - s;
  double a1[3] = {r, s, u}, b1[3] = {s, u, r};
  for (int k = 0; k < 3; k++) {
    t.r7[k][0] = a1[k];
    t.r7[k][1] = b1[k];
    t.w7[k] = (155. - sqrt(15.)) / 2400.;
  }
  r = (6. + sqrt(15.)) / 21.;
  s = r;
  u = 1 - r - s;
  double a2[3] = {r, s, u}, b2[3] = {s, u, r};
  for (int k = 0; k < 3; k++) {
    t.r7[3 + k][0] = a2[k];
    t.r7[3 + k][1] = b2[k];
    t.w7[3 + k] = (155. + sqrt(15.)) / 2400.;
  }
  t.r7[6][0] = 1. / 3.;
  t.r7[6][1] = 1. / 3.;
  t.w7[6] = 9. / 80.;
  gauleg_h(0., 1., 4, t.rd, t.wd);
  return t;
}

struct EwT {
  const double *sl1, *sl2;
  double reps, rc;
};

// Tri_Int_Regular (:319-363) / Tri_Int_Duffy (:373-465).  With lhs != nullptr the
// influence blocks lhs[l][a][b] are accumulated (and f is ignored).
__device__ void tri_int(const TriRules &R, const EwT &ew, const double x[3][3], const double f[3][3],
                        const double xt[3], bool duffy, double s0, double t0, double rhs[3],
                        double (*lhs)[3][3]) {
  rhs[0] = rhs[1] = rhs[2] = 0;
  if (lhs)
    for (int l = 0; l < 3; l++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) lhs[l][a][b] = 0;
  if (!duffy) {
    double e1[3], e2[3], cr[3];
    for (int d = 0; d < 3; d++) {
      e1[d] = x[1][d] - x[0][d];
      e2[d] = x[2][d] - x[0][d];
    }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double detJ = 2 * (0.5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]));
    for (int q = 0; q < 7; q++) {
      double s = R.r7[q][0], t = R.r7[q][1];
      double xg[3], fg[3], xx[3], EA, EB;
      for (int d = 0; d < 3; d++) {
        xg[d] = (1. - s - t) * x[0][d] + s * x[1][d] + t * x[2][d];
        fg[d] = (1. - s - t) * f[0][d] + s * f[1][d] + t * f[2][d];
      }
      double ds = R.w7[q] * detJ;
      for (int d = 0; d < 3; d++) {
        fg[d] = ds * fg[d];
        xx[d] = xt[d] - xg[d];
      }
      double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
      ewald_sl(ew.sl1, ew.sl2, rr, ew.reps, ew.rc, EA, EB);
      double xf = xx[0] * fg[0] + xx[1] * fg[1] + xx[2] * fg[2];
      for (int d = 0; d < 3; d++) rhs[d] += EA * xx[d] * xf + EB * fg[d];
      if (lhs) {
        double wv[3] = {1 - s - t, s, t};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            double L = ds * (EA * xx[a] * xx[b] + (a == b ? EB : 0.));
            for (int l = 0; l < 3; l++) lhs[l][a][b] += wv[l] * L;
          }
      }
    }
    return;
  }
  double x0[3], f0[3];
  for (int d = 0; d < 3; d++) {
    x0[d] = (1 - s0 - t0) * x[0][d] + s0 * x[1][d] + t0 * x[2][d];
    f0[d] = (1 - s0 - t0) * f[0][d] + s0 * f[1][d] + t0 * f[2][d];
  }
  const double st1[3][2] = {{0, 0}, {1, 0}, {0, 1}}, st2[3][2] = {{1, 0}, {0, 1}, {0, 0}};
  for (int n = 0; n < 3; n++) {
    int n2 = n == 2 ? 0 : n + 1;
    double a[3], b[3], nr[3];
    for (int d = 0; d < 3; d++) {
      a[d] = x[n][d] - x0[d];
      b[d] = x[n2][d] - x0[d];
    }
    nr[0] = a[1] * b[2] - a[2] * b[1];
    nr[1] = a[2] * b[0] - a[0] * b[2];
    nr[2] = a[0] * b[1] - a[1] * b[0];
    double detJ = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        double s = R.rd[i], t = s * R.rd[j];
        double xg[3], fg[3], xx[3], EA, EB;
        for (int d = 0; d < 3; d++) {
          xg[d] = (1. - s) * x0[d] + (s - t) * x[n][d] + t * x[n2][d];
          fg[d] = (1. - s) * f0[d] + (s - t) * f[n][d] + t * f[n2][d];
        }
        double ds = R.wd[i] * R.wd[j] * detJ * s;
        for (int d = 0; d < 3; d++) {
          fg[d] = ds * fg[d];
          xx[d] = xt[d] - xg[d];
        }
        double rr = sqrt(xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2]);
        ewald_sl(ew.sl1, ew.sl2, rr, ew.reps, ew.rc, EA, EB);
        double xf = xx[0] * fg[0] + xx[1] * fg[1] + xx[2] * fg[2];
        for (int d = 0; d < 3; d++) rhs[d] += EA * xx[d] * xf + EB * fg[d];
        if (lhs) {
          double sG = (1. - s) * s0 + (s - t) * st1[n][0] + t * st2[n][0];
          double tG = (1. - s) * t0 + (s - t) * st1[n][1] + t * st2[n][1];
          double wv[3] = {1. - sG - tG, sG, tG};
          for (int aa = 0; aa < 3; aa++)
            for (int bb = 0; bb < 3; bb++) {
              double L = ds * (EA * xx[aa] * xx[bb] + (aa == bb ? EB : 0.));
              for (int l = 0; l < 3; l++) lhs[l][aa][bb] += wv[l] * L;
            }
        }
      }
  }
}

struct WallArgs {
  double Lb[3], iLb[3], iLbNc[3];
  int Nc[3], nvert, nele, wall_id;
  double rc;
  TriRules R;
  EwT ew;
  const double *wx, *wf, *eps;  // vertices (nvert,3), traction (nvert,3), epsDist
  const int *e2v;               // (nele,3) 1-based
  const int *bin_start, *order; // centroid cell list (coarse bins)
  const double *cen;            // triangle centroids (nele,3)
  double far2;                  // (rc + largest centroid-vertex distance)^2, padded
};

// A triangle whose centroid (minimum image) is farther from xi than rc plus the
// largest centroid-to-vertex distance of the mesh is farther than rc in every image:
// this skips it before its vertices are loaded, and skips nothing the exact distance
// test would keep.
__device__ __forceinline__ bool tri_far(const WallArgs &a, const double xi[3], int e) {
  double s2 = 0;
  for (int d = 0; d < 3; d++) {
    double t = a.cen[e + (size_t)a.nele * d] - xi[d];
    t -= rint(t * a.iLb[d]) * a.Lb[d];
    s2 += t * t;
  }
  return s2 > a.far2;
}

__device__ __forceinline__ void load_tri(const WallArgs &a, int e, double x[3][3], double f[3][3]) {
  for (int l = 0; l < 3; l++) {
    int iv = a.e2v[e + (size_t)a.nele * l] - 1;
    for (int d = 0; d < 3; d++) {
      x[l][d] = a.wx[iv + (size_t)a.nvert * d];
      if (f) f[l][d] = a.wf[iv + (size_t)a.nvert * d];
    }
  }
}

__device__ __forceinline__ void hash_bins(const WallArgs &a, const double xi[3], int &i1, int &i2, int &i3) {
  i1 = imod_d((int)floor(xi[0] * a.iLbNc[0]), a.Nc[0]);
  i2 = imod_d((int)floor(xi[1] * a.iLbNc[1]), a.Nc[1]);
  i3 = imod_d((int)floor(xi[2] * a.iLbNc[2]), a.Nc[2]);
}

// translate the triangle next to xi (the shift of vertex 1, :103-108)
__device__ __forceinline__ void shift_tri(const WallArgs &a, const double xi[3], double x[3][3]) {
#pragma clang fp contract(off)
  for (int d = 0; d < 3; d++) {
    double sh = round((xi[d] - x[0][d]) * a.iLb[d]) * a.Lb[d];
    for (int l = 0; l < 3; l++) x[l][d] += sh;
  }
}

// AddIntOnWalls for targets that are not wall vertices (common/ModIntOnWalls.F90:80-127)
__global__ void __launch_bounds__(128) k_wall_targets(WallArgs a, double c1, int nT,
                                                      const double *__restrict__ tx,
                                                      const int *__restrict__ tindx,
                                                      const int *__restrict__ tact,
                                                      const double *__restrict__ tA, double *v) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT || !tact[ti]) return;
  double xi[3] = {tx[ti], tx[ti + nT], tx[ti + 2 * (size_t)nT]};
  if (tindx[ti] == a.wall_id) return;
  int i1, i2, i3;
  hash_bins(a, xi, i1, i2, i3);
  double acc[3] = {0, 0, 0};
  for (int nb = 0; nb < 27; nb++) {
    int n1 = imod_d(i1 + nb / 9 - 1, a.Nc[0]), n2 = imod_d(i2 + (nb / 3) % 3 - 1, a.Nc[1]),
        n3 = imod_d(i3 + nb % 3 - 1, a.Nc[2]);
    int b = n1 + a.Nc[0] * (n2 + a.Nc[1] * n3);
    for (int k = a.bin_start[b]; k < a.bin_start[b + 1]; k++) {
      int e = a.order[k];
      if (tri_far(a, xi, e)) continue;
      double x[3][3], f[3][3], dv[3], s0, t0;
      load_tri(a, e, x, f);
      shift_tri(a, xi, x);
      double rr = min_dist_to_tri(xi, x, s0, t0);
      if (rr > a.rc) continue;
      tri_int(a.R, a.ew, x, f, xi, rr < a.eps[e], s0, t0, dv, nullptr);
      for (int d = 0; d < 3; d++) acc[d] += c1 * dv[d] / tA[ti];
    }
  }
  for (int d = 0; d < 3; d++) v[ti + (size_t)nT * d] += acc[d];
}

// The same sums with the targets sorted by coarse bin: a workgroup takes up to 64
// targets of one bin; its four wavefronts hold the same targets and take every fourth
// of the 27 neighbouring bins, streaming that bin's triangles through their own LDS
// tile, 64 at a time, in the order k_wall_targets visits them.  The four partial sums
// are added in wavefront order (deterministic; the order of the additions differs from
// k_wall_targets' single running sum by rounding only).
#define WT_T 64
#define WT_W 4
__global__ void k_wall_chunks(int nbins, const int *__restrict__ start, int *cnt) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbins) return;
  cnt[b] = b < nbins ? (start[b + 1] - start[b] + WT_T - 1) / WT_T : 0;
}

__global__ void __launch_bounds__(WT_T *WT_W) k_wall_targets_tiled(
    WallArgs a, double c1, int nT, const double *__restrict__ tx, const int *__restrict__ tindx,
    const int *__restrict__ tact, const double *__restrict__ tA, const int *__restrict__ torder,
    const int *__restrict__ tstart, const int *__restrict__ chunk_off, int nbins, double *v) {
  __shared__ double sx[WT_W][9][WT_T], sf[WT_W][9][WT_T], scen[WT_W][3][WT_T], seps[WT_W][WT_T];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, blk = blockIdx.x;
  if (blk >= chunk_off[nbins]) return;
  int lo = 0, hi = nbins;  // chunk_off[lo] <= blk < chunk_off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chunk_off[mid] <= blk)
      lo = mid;
    else
      hi = mid;
  }
  const int bin = lo;
  const int k = tstart[bin] + (blk - chunk_off[lo]) * WT_T + lane;
  const bool have = k < tstart[bin + 1];
  const int ti = have ? torder[k] : 0;
  const bool act = have && tact[ti] && tindx[ti] != a.wall_id;
  double xi[3] = {0, 0, 0};
  if (have) {
    xi[0] = tx[ti];
    xi[1] = tx[ti + nT];
    xi[2] = tx[ti + 2 * (size_t)nT];
  }
  const int i1 = bin % a.Nc[0], i2 = (bin / a.Nc[0]) % a.Nc[1], i3 = bin / (a.Nc[0] * a.Nc[1]);
  double acc[3] = {0, 0, 0};
  // each wavefront walks its own bins with its own tile: only wavefront-local
  // synchronisation (the LDS writes of one wavefront are read by the same wavefront)
  for (int nb = wv; nb < 27; nb += WT_W) {
    int n1 = imod_d(i1 + nb / 9 - 1, a.Nc[0]), n2 = imod_d(i2 + (nb / 3) % 3 - 1, a.Nc[1]),
        n3 = imod_d(i3 + nb % 3 - 1, a.Nc[2]);
    int b = n1 + a.Nc[0] * (n2 + a.Nc[1] * n3);
    const int s0b = a.bin_start[b], s1b = a.bin_start[b + 1];
    for (int base = s0b; base < s1b; base += WT_T) {
      const int cnt = min(WT_T, s1b - base);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane < cnt) {
        const int e = a.order[base + lane];
        for (int l = 0; l < 3; l++) {
          const int iv = a.e2v[e + (size_t)a.nele * l] - 1;
          for (int d = 0; d < 3; d++) {
            sx[wv][3 * l + d][lane] = a.wx[iv + (size_t)a.nvert * d];
            sf[wv][3 * l + d][lane] = a.wf[iv + (size_t)a.nvert * d];
          }
        }
        for (int d = 0; d < 3; d++) scen[wv][d][lane] = a.cen[e + (size_t)a.nele * d];
        seps[wv][lane] = a.eps[e];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (!act) continue;
      for (int j = 0; j < cnt; j++) {
        double s2 = 0;
        for (int d = 0; d < 3; d++) {
          double t = scen[wv][d][j] - xi[d];
          t -= rint(t * a.iLb[d]) * a.Lb[d];
          s2 += t * t;
        }
        if (s2 > a.far2) continue;
        double x[3][3], f[3][3], dv[3], s0, t0;
        for (int l = 0; l < 3; l++)
          for (int d = 0; d < 3; d++) {
            x[l][d] = sx[wv][3 * l + d][j];
            f[l][d] = sf[wv][3 * l + d][j];
          }
        shift_tri(a, xi, x);
        double rr = min_dist_to_tri(xi, x, s0, t0);
        if (rr > a.rc) continue;
        tri_int(a.R, a.ew, x, f, xi, rr < seps[wv][j], s0, t0, dv, nullptr);
        for (int d = 0; d < 3; d++) acc[d] += c1 * dv[d] / tA[ti];
      }
    }
  }
  // partial sums of wavefronts 1..3 through LDS (reusing wavefront 0's tile), added by
  // wavefront 0 in order
  __syncthreads();
  double(*red)[3][WT_T] = reinterpret_cast<double(*)[3][WT_T]>(&sx[0][0][0]);
  if (wv > 0)
    for (int d = 0; d < 3; d++) red[wv - 1][d][lane] = acc[d];
  __syncthreads();
  if (wv == 0 && act) {
    for (int w = 0; w < WT_W - 1; w++)
      for (int d = 0; d < 3; d++) acc[d] += red[w][d][lane];
    for (int d = 0; d < 3; d++) v[ti + (size_t)nT * d] += acc[d];
  }
}

// Self-matrix assembly (PrepareSingIntOnWall :181-308), pass 1: triangles per vertex
__global__ void __launch_bounds__(128) k_wall_count(WallArgs a, int *cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nvert) return;
  double xi[3] = {a.wx[i], a.wx[i + a.nvert], a.wx[i + 2 * (size_t)a.nvert]};
  int i1, i2, i3, n = 0;
  hash_bins(a, xi, i1, i2, i3);
  for (int nb = 0; nb < 27; nb++) {
    int n1 = imod_d(i1 + nb / 9 - 1, a.Nc[0]), n2 = imod_d(i2 + (nb / 3) % 3 - 1, a.Nc[1]),
        n3 = imod_d(i3 + nb % 3 - 1, a.Nc[2]);
    int b = n1 + a.Nc[0] * (n2 + a.Nc[1] * n3);
    for (int k = a.bin_start[b]; k < a.bin_start[b + 1]; k++) {
      int e = a.order[k];
      if (tri_far(a, xi, e)) continue;
      double x[3][3], s0, t0;
      load_tri(a, e, x, nullptr);
      shift_tri(a, xi, x);
      if (!(min_dist_to_tri(xi, x, s0, t0) > a.rc)) n++;
    }
  }
  cnt[i] = n;
}

// pass 2: one 3x3 block per (vertex row, triangle, triangle vertex)
__global__ void __launch_bounds__(128) k_wall_fill(WallArgs a, const int *__restrict__ off,
                                                   int *ekey, int *eidx, double *eval) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nvert) return;
  double xi[3] = {a.wx[i], a.wx[i + a.nvert], a.wx[i + 2 * (size_t)a.nvert]};
  int i1, i2, i3;
  hash_bins(a, xi, i1, i2, i3);
  long long p = 3LL * off[i];
  for (int nb = 0; nb < 27; nb++) {
    int n1 = imod_d(i1 + nb / 9 - 1, a.Nc[0]), n2 = imod_d(i2 + (nb / 3) % 3 - 1, a.Nc[1]),
        n3 = imod_d(i3 + nb % 3 - 1, a.Nc[2]);
    int b = n1 + a.Nc[0] * (n2 + a.Nc[1] * n3);
    for (int k = a.bin_start[b]; k < a.bin_start[b + 1]; k++) {
      int e = a.order[k];
      if (tri_far(a, xi, e)) continue;
      double x[3][3], f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, s0, t0, rhs[3], lhs[3][3][3];
      load_tri(a, e, x, nullptr);
      shift_tri(a, xi, x);
      double rr = min_dist_to_tri(xi, x, s0, t0);
      if (rr > a.rc) continue;
      tri_int(a.R, a.ew, x, f, xi, !(rr > a.eps[e]), s0, t0, rhs, lhs);
      for (int l = 0; l < 3; l++) {
        int iv = a.e2v[e + (size_t)a.nele * l] - 1;
        ekey[p] = iv;
        eidx[p] = (int)p;
        for (int q = 0; q < 9; q++) eval[9 * p + q] = lhs[l][q / 3][q % 3];
        p++;
      }
    }
  }
}

// pass 3 (after a segmented sort of each row by column): sum equal columns
__global__ void __launch_bounds__(128) k_wall_reduce(int nvert, const int *__restrict__ off,
                                                     const int *__restrict__ skey,
                                                     const int *__restrict__ sidx,
                                                     const double *__restrict__ eval, int *rowcnt,
                                                     int *bcol, double *bval, size_t nbt, int pass) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvert) return;
  long long s = 3LL * off[i], e = 3LL * off[i + 1];
  int n = 0;
  long long out = pass ? rowcnt[i] : 0;
  for (long long k = s; k < e; k++) {
    bool first = (k == s) || skey[k] != skey[k - 1];
    if (first) {
      if (pass) {
        bcol[out + n] = skey[k];
        for (int q = 0; q < 9; q++) bval[q * nbt + out + n] = 0.;
      }
      n++;
    }
    if (pass) {
      const double *src = eval + 9 * (size_t)sidx[k];
      for (int q = 0; q < 9; q++) bval[q * nbt + out + n - 1] += src[q];
    }
  }
  if (!pass) rowcnt[i] = n;
}

// The 3x3 blocks of the self-matrix are symmetric: each is a Stokeslet-type tensor
// EA x x^T + EB I integrated against a scalar hat function (their two off-diagonal
// copies differ only by rounding, 3e-16 relative).  After assembly every block is
// kept as its symmetric part in six component arrays (B00, B11, B22, B01, B02, B12):
// 6 values + 1 column index per block instead of 9 + 1.
__global__ void k_wall_sym6(size_t nb, const double *__restrict__ b9, double *b6) {
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k >= nb) return;
  b6[k] = b9[k];
  b6[nb + k] = b9[4 * nb + k];
  b6[2 * nb + k] = b9[8 * nb + k];
  b6[3 * nb + k] = 0.5 * (b9[nb + k] + b9[3 * nb + k]);
  b6[4 * nb + k] = 0.5 * (b9[2 * nb + k] + b9[6 * nb + k]);
  b6[5 * nb + k] = 0.5 * (b9[5 * nb + k] + b9[7 * nb + k]);
}

// v_i += c1 (L f)_i / Acoef_i over wall vertices (SingIntOnWall :136-172)
// v_i += c1/A_i * sum_k B_k f_(col k): one wavefront per vertex row, lanes over the
// row's blocks.  The six component arrays make each of a lane's value loads one
// coalesced 64-lane access; the 3 traction components of the block's column are
// gathered (the traction vector stays in L2).
// V: the block scalar (double, or float in the split form); sums are FP64 either way.
#ifndef RBC_SPMV_NT
#define RBC_SPMV_NT 1  // the streamed matrix read with nontemporal loads (measured faster)
#endif
// One wavefront per target row of the self-matrix: v(i) (+)= c1 / A_i sum_j B_ij f_j.
// FROM_U: the sources are read from the de-duplicated vector u (3 per unique vertex,
// col holding unique indices); else from the vertex array f(nvert, 3).  ASSIGN: v is
// written rather than accumulated (the same values: the sum starts from zero).
template <typename V, bool FROM_U = false, bool ASSIGN = false>
__global__ void __launch_bounds__(256) k_wall_spmv(int nvert, double c1, const int *__restrict__ rowptr,
                                                   const int *__restrict__ bcol,
                                                   const V *__restrict__ bval,
                                                   const double *__restrict__ f,
                                                   const double *__restrict__ tA, double *v) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (i >= nvert) return;
  const size_t nbt = rowptr[nvert];
  double s0 = 0, s1 = 0, s2 = 0;
  auto block = [&](int k) {
#if RBC_SPMV_NT
    const int j = __builtin_nontemporal_load(bcol + k);
    const double b00 = __builtin_nontemporal_load(bval + k), b11 = __builtin_nontemporal_load(bval + nbt + k),
                 b22 = __builtin_nontemporal_load(bval + 2 * nbt + k);
    const double b01 = __builtin_nontemporal_load(bval + 3 * nbt + k), b02 = __builtin_nontemporal_load(bval + 4 * nbt + k),
                 b12 = __builtin_nontemporal_load(bval + 5 * nbt + k);
#else
    const int j = bcol[k];
    const double b00 = bval[k], b11 = bval[nbt + k], b22 = bval[2 * nbt + k];
    const double b01 = bval[3 * nbt + k], b02 = bval[4 * nbt + k], b12 = bval[5 * nbt + k];
#endif
    double f0, f1, f2;
    if constexpr (FROM_U) {
      f0 = f[3 * j];
      f1 = f[3 * j + 1];
      f2 = f[3 * j + 2];
    } else {
      f0 = f[j];
      f1 = f[j + nvert];
      f2 = f[j + 2 * (size_t)nvert];
    }
    s0 += b00 * f0 + b01 * f1 + b02 * f2;
    s1 += b01 * f0 + b11 * f1 + b12 * f2;
    s2 += b02 * f0 + b12 * f1 + b22 * f2;
  };
  // four blocks per lane per step: their column, value and source loads are issued
  // together (the source gathers depend on the column loads)
  int k = rowptr[i] + lane;
  const int kend = rowptr[i + 1];
  for (; k + 192 < kend; k += 256) {
    block(k);
    block(k + 64);
    block(k + 128);
    block(k + 192);
  }
  for (; k < kend; k += 64) block(k);
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o, 64);
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if (lane == 0) {
    const double A = tA[i];
    if constexpr (ASSIGN) {
      v[i] = c1 * s0 / A;
      v[i + nvert] = c1 * s1 / A;
      v[i + 2 * (size_t)nvert] = c1 * s2 / A;
    } else {
      v[i] += c1 * s0 / A;
      v[i + nvert] += c1 * s1 / A;
      v[i + 2 * (size_t)nvert] += c1 * s2 / A;
    }
  }
}

__global__ void k_to_float(const double *__restrict__ x, size_t n, float *y) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

static WallArgs wall_args(rbc3d_ctx *c) {
  WallArgs a;
  WallData &w = c->wall;
  for (int d = 0; d < 3; d++) {
    a.Lb[d] = c->Lb[d];
    a.iLb[d] = c->iLb[d];
    a.iLbNc[d] = c->iLbNc[d];
    a.Nc[d] = c->Nc[d];
  }
  a.nvert = w.nvert;
  a.nele = w.nele;
  a.wall_id = w.id;
  a.rc = c->conf.rc;
  a.R = make_rules();
  a.ew.sl1 = c->d_sl1.p;
  a.ew.sl2 = c->d_sl2.p;
  a.ew.reps = c->r_eps;
  a.ew.rc = c->conf.rc;
  a.wx = w.d_x.p;
  a.wf = w.d_f.p;
  a.eps = w.d_eps.p;
  a.e2v = w.d_e2v.p;
  a.bin_start = w.d_bin_start.p;
  a.order = w.d_order.p;
  a.cen = w.d_cen.p;
  const double far = (c->conf.rc + w.rmax) * (1 + 1e-9) + 1e-12;
  a.far2 = far * far;
  return a;
}

// ---------------------------------------------------------------------------
extern "C" int rbc3d_set_wall(rbc3d_ctx *c, int nvert, int nele, const double *x, const int *e2v) {
  if (!c || nvert < 3 || nele < 1 || !x || !e2v) return RBC3D_EINVAL;
  for (int i = 0; i < 3 * nele; i++)
    if (e2v[i] < 1 || e2v[i] > nvert) {
      rbc_set_error("rbc3d_set_wall: e2v out of range");
      return RBC3D_EINVAL;
    }
  for (int i = 0; i < 3 * nvert; i++)
    if (!std::isfinite(x[i])) return RBC3D_ENONFINITE;
  WallData &w = c->wall;
  w.nvert = nvert;
  w.nele = nele;
  w.id = c->nrbc + 1;  // GlobData_Init (common/ModData.F90:40-49)
  w.x.assign(x, x + 3 * (size_t)nvert);
  w.e2v.assign(e2v, e2v + 3 * (size_t)nele);
  // Wall_Build_V2V (common/ModWall.F90:66-114)
  w.v2v.assign(nvert, 0);
  {
    double eps = std::min(c->Lb[0], std::min(c->Lb[1], c->Lb[2])) * 1.e-5;
    double xmin[3], xmax[3];
    for (int d = 0; d < 3; d++) {
      xmin[d] = xmax[d] = x[(size_t)nvert * d];
      for (int i = 0; i < nvert; i++) {
        xmin[d] = std::min(xmin[d], x[i + (size_t)nvert * d]);
        xmax[d] = std::max(xmax[d], x[i + (size_t)nvert * d]);
      }
    }
    std::vector<int> bd;
    for (int i = 0; i < nvert; i++) {
      double m1 = 1e300, m2 = 1e300;
      for (int d = 0; d < 3; d++) {
        m1 = std::min(m1, fabs(x[i + (size_t)nvert * d] - xmin[d]));
        m2 = std::min(m2, fabs(x[i + (size_t)nvert * d] - xmax[d]));
      }
      if (m1 < eps || m2 < eps) bd.push_back(i);
    }
    for (size_t p1 = 0; p1 + 1 < bd.size(); p1++) {
      int v1 = bd[p1];
      if (w.v2v[v1] > 0) continue;
      for (size_t p2 = p1 + 1; p2 < bd.size(); p2++) {
        int v2 = bd[p2];
        double mx = 0;
        for (int d = 0; d < 3; d++) {
          double xx = x[v2 + (size_t)nvert * d] - x[v1 + (size_t)nvert * d];
          xx = xx - round(xx * c->iLb[d]) * c->Lb[d];
          mx = std::max(mx, fabs(xx));
        }
        if (mx < eps) w.v2v[v2] = v1 + 1;
      }
    }
  }
  // global vertex numbering (common/ModData.F90:50-63)
  w.glb.assign(nvert, 0);
  int p = 0;
  for (int i = 0; i < nvert; i++) w.glb[i] = w.v2v[i] == 0 ? ++p : w.glb[w.v2v[i] - 1];
  w.ndof = 3 * p;
  // Wall_ComputeGeometry (common/ModWall.F90:118-145) and centroids
  std::vector<double> area(nele), eps(nele), cen((size_t)3 * nele), vn((size_t)3 * p, 0.);
  double rmax = 0;  // largest centroid-to-vertex distance
  for (int e = 0; e < nele; e++) {
    double xe[3][3];
    for (int l = 0; l < 3; l++)
      for (int d = 0; d < 3; d++) xe[l][d] = x[(e2v[e + (size_t)nele * l] - 1) + (size_t)nvert * d];
    double a[3], b[3], cr[3];
    for (int d = 0; d < 3; d++) {
      a[d] = xe[1][d] - xe[0][d];
      b[d] = xe[2][d] - xe[0][d];
    }
    cr[0] = a[1] * b[2] - a[2] * b[1];
    cr[1] = a[2] * b[0] - a[0] * b[2];
    cr[2] = a[0] * b[1] - a[1] * b[0];
    double n = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    area[e] = 0.5 * n;
    eps[e] = sqrt(area[e]);
    // area-weighted vertex normals, merged over the duplicated seam vertices
    for (int l = 0; l < 3; l++) {
      const int gi = w.glb[e2v[e + (size_t)nele * l] - 1] - 1;
      for (int d = 0; d < 3; d++) vn[3 * (size_t)gi + d] += cr[d];
    }
    for (int d = 0; d < 3; d++) cen[e + (size_t)nele * d] = (1.0 / 3) * ((xe[0][d] + xe[1][d]) + xe[2][d]);
    for (int l = 0; l < 3; l++) {
      double r2 = 0;
      for (int d = 0; d < 3; d++) {
        const double t = xe[l][d] - cen[e + (size_t)nele * d];
        r2 += t * t;
      }
      rmax = std::max(rmax, sqrt(r2));
    }
  }
  w.rmax = rmax;
  // dedup map for AssembleArray (common/ModNoSlip.F90:361-384): the last vertex with a
  // given global index supplies that unknown's value
  std::vector<int> last(p, -1);
  for (int i = 0; i < nvert; i++) last[w.glb[i] - 1] = i;
  w.vnorm.assign((size_t)3 * nvert, 0.);
  for (int i = 0; i < nvert; i++) {
    const double *v = &vn[3 * (size_t)(w.glb[i] - 1)];
    const double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int d = 0; d < 3; d++) w.vnorm[i + (size_t)nvert * d] = nn > 0 ? v[d] / nn : 0.;
  }
  hipStream_t st = c->stream;
  RBC_TRY(w.d_x.ensure((size_t)3 * nvert));
  RBC_TRY(w.d_f.ensure((size_t)3 * nvert));
  RBC_TRY(w.d_area.ensure(nele));
  RBC_TRY(w.d_eps.ensure(nele));
  RBC_TRY(w.d_cen.ensure((size_t)3 * nele));
  RBC_TRY(w.d_e2v.ensure((size_t)3 * nele));
  RBC_TRY(w.d_glb.ensure(nvert + p + 1));
  RBC_TRY(w.d_vnorm.ensure((size_t)3 * nvert));
  HIP_TRY(hipMemcpy(w.d_vnorm.p, w.vnorm.data(), sizeof(double) * 3 * nvert, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_x.p, x, sizeof(double) * 3 * nvert, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(w.d_f.p, 0, sizeof(double) * 3 * nvert));
  HIP_TRY(hipMemcpy(w.d_area.p, area.data(), sizeof(double) * nele, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_eps.p, eps.data(), sizeof(double) * nele, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_cen.p, cen.data(), sizeof(double) * 3 * nele, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_e2v.p, e2v, sizeof(int) * 3 * nele, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_glb.p, w.glb.data(), sizeof(int) * nvert, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_glb.p + nvert, last.data(), sizeof(int) * p, hipMemcpyHostToDevice));
  // wall source list on the reference's cell list
  DBuf<int> bins;
  RBC_TRY(bins.ensure(nele + 1));
  RBC_TRY(w.d_order.ensure(nele + 1));
  RBC_TRY(w.d_bin_start.ensure(c->nbins + 1));
  RBC_TRY(rbc_hash_sort(c, nele, w.d_cen.p, nele, bins.p, w.d_order.p, w.d_bin_start.p));
  bins.release();
  w.prepared = false;
  w.valf_ready = false;
  w.lmodes_ready = false;
  w.wsp_ready = false;
  w.wgt_ready = false;
  w.op_ready = false;
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

extern "C" int rbc3d_set_wall_density(rbc3d_ctx *c, const double *f) {
  if (!c || !f) return RBC3D_EINVAL;
  if (!c->wall.nvert) return RBC3D_ESTATE;
  HIP_TRY(hipMemcpy(c->wall.d_f.p, f, sizeof(double) * 3 * c->wall.nvert, hipMemcpyHostToDevice));
  return RBC3D_OK;
}

int rbc_wall_prepare(rbc3d_ctx *c) {
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  hipStream_t st = c->stream;
  WallArgs a = wall_args(c);
  int nv = w.nvert;
  DBuf<int> cnt, off;
  RBC_TRY(cnt.ensure(nv + 1));
  RBC_TRY(off.ensure(nv + 1));
  k_wall_count<<<(nv + 127) / 128, 128, 0, st>>>(a, cnt.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemsetAsync(cnt.p + nv, 0, sizeof(int), st));
  size_t tmp = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, cnt.p, off.p, nv + 1, st));
  RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->d_cub_tmp.p, tmp, cnt.p, off.p, nv + 1, st));
  int ntri = 0;
  HIP_TRY(hipMemcpyAsync(&ntri, off.p + nv, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  long long ne = 3LL * ntri;
  if (ne > 0x7fffffffLL) {
    rbc_set_error("wall self-matrix too large");
    return RBC3D_EINVAL;
  }
  DBuf<int> ekey, eidx, skey, sidx, segoff, rowcnt;
  DBuf<double> eval;
  RBC_TRY(ekey.ensure(ne + 1));
  RBC_TRY(eidx.ensure(ne + 1));
  RBC_TRY(skey.ensure(ne + 1));
  RBC_TRY(sidx.ensure(ne + 1));
  RBC_TRY(eval.ensure(9 * ne + 9));
  k_wall_fill<<<(nv + 127) / 128, 128, 0, st>>>(a, off.p, ekey.p, eidx.p, eval.p);
  HIP_TRY(hipGetLastError());
  // segment offsets in entries: 3*off
  RBC_TRY(segoff.ensure(nv + 1));
  {
    std::vector<int> h(nv + 1);
    HIP_TRY(hipMemcpy(h.data(), off.p, sizeof(int) * (nv + 1), hipMemcpyDeviceToHost));
    for (auto &v : h) v *= 3;
    HIP_TRY(hipMemcpy(segoff.p, h.data(), sizeof(int) * (nv + 1), hipMemcpyHostToDevice));
  }
  tmp = 0;
  int bits = 1;
  while ((1 << bits) < nv) bits++;
  HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp, ekey.p, skey.p, eidx.p, sidx.p,
                                                      (int)ne, nv, segoff.p, segoff.p + 1, 0, bits, st));
  RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
  HIP_TRY(hipcub::DeviceSegmentedRadixSort::SortPairs(c->d_cub_tmp.p, tmp, ekey.p, skey.p, eidx.p,
                                                      sidx.p, (int)ne, nv, segoff.p, segoff.p + 1,
                                                      0, bits, st));
  RBC_TRY(rowcnt.ensure(nv + 1));
  k_wall_reduce<<<(nv + 127) / 128, 128, 0, st>>>(nv, off.p, skey.p, sidx.p, eval.p, rowcnt.p,
                                                  nullptr, nullptr, 0, 0);
  HIP_TRY(hipMemsetAsync(rowcnt.p + nv, 0, sizeof(int), st));
  RBC_TRY(w.d_rowptr.ensure(nv + 1));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, rowcnt.p, w.d_rowptr.p, nv + 1, st));
  RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->d_cub_tmp.p, tmp, rowcnt.p, w.d_rowptr.p, nv + 1, st));
  int nb = 0;
  HIP_TRY(hipMemcpyAsync(&nb, w.d_rowptr.p + nv, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (9LL * nb > 0x7fffffffLL) {
    rbc_set_error("wall self-matrix too large");
    return RBC3D_EINVAL;
  }
  RBC_TRY(w.d_col.ensure(nb + 1));
  {
    DBuf<double> b9;
    RBC_TRY(b9.ensure(9 * (size_t)nb + 9));
    k_wall_reduce<<<(nv + 127) / 128, 128, 0, st>>>(nv, off.p, skey.p, sidx.p, eval.p, w.d_rowptr.p,
                                                    w.d_col.p, b9.p, (size_t)nb, 1);
    HIP_TRY(hipGetLastError());
    RBC_TRY(w.d_val.ensure(6 * (size_t)nb + 6));
    k_wall_sym6<<<(unsigned)(((size_t)nb + 255) / 256), 256, 0, st>>>((size_t)nb, b9.p, w.d_val.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
  }
  w.nnz = 9LL * nb;
  w.prepared = true;
  w.valf_ready = false;
  w.lmodes_ready = false;
  return RBC3D_OK;
}

extern "C" int rbc3d_prepare_sing_int_on_wall(rbc3d_ctx *c) {
  if (!c) return RBC3D_EINVAL;
  return rbc_wall_prepare(c);
}

extern "C" int rbc3d_wall_self_nnz(rbc3d_ctx *c, long long *nnz) {
  if (!c || !nnz) return RBC3D_EINVAL;
  *nnz = c->wall.nnz;
  return RBC3D_OK;
}

int rbc_add_int_on_walls_dev(rbc3d_ctx *c, double c1, int nT, const double *d_tx, const int *d_tindx,
                             const int *d_act, const double *d_A, double *d_v, bool wall_targets) {
  WallData &w = c->wall;
  if (!w.nvert || nT == 0) return RBC3D_OK;
  hipStream_t st = c->stream;
  rbc_stage_begin(c, ST_WALL);
  if (wall_targets) {
    // wall vertices: the precomputed self-interaction; with one wall that is all
    if (!w.prepared) RBC_TRY(rbc_wall_prepare(c));
    const unsigned nblk = (unsigned)(((size_t)w.nvert * 64 + 255) / 256);
    if (c->prec == RBC3D_PREC_SPLIT) {
      if (!w.valf_ready) {
        const size_t nv6 = (size_t)w.nnz / 9 * 6;  // 6 values per symmetric 3x3 block
        RBC_TRY(w.d_valf.ensure(nv6 + 1));
        k_to_float<<<(unsigned)((nv6 + 255) / 256), 256, 0, st>>>(w.d_val.p, nv6, w.d_valf.p);
        HIP_TRY(hipGetLastError());
        w.valf_ready = true;
      }
      k_wall_spmv<float><<<nblk, 256, 0, st>>>(w.nvert, c1, w.d_rowptr.p, w.d_col.p, w.d_valf.p,
                                               w.d_f.p, d_A, d_v);
    } else {
      k_wall_spmv<double><<<nblk, 256, 0, st>>>(w.nvert, c1, w.d_rowptr.p, w.d_col.p, w.d_val.p,
                                                w.d_f.p, d_A, d_v);
    }
  } else {
    // targets sorted by coarse bin, then one wavefront per 64 targets of a bin
    const int nbins = c->nbins;
    RBC_TRY(w.d_tg_bin.ensure(nT + 1));
    RBC_TRY(w.d_tg_order.ensure(nT + 1));
    RBC_TRY(w.d_tg_start.ensure(nbins + 1));
    RBC_TRY(w.d_tg_chunk.ensure(2 * (size_t)(nbins + 1)));
    RBC_TRY(rbc_hash_sort(c, nT, d_tx, nT, w.d_tg_bin.p, w.d_tg_order.p, w.d_tg_start.p));
    int *ccnt = w.d_tg_chunk.p, *coff = ccnt + (nbins + 1);
    k_wall_chunks<<<(nbins + 1 + 255) / 256, 256, 0, st>>>(nbins, w.d_tg_start.p, ccnt);
    HIP_TRY(hipGetLastError());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, ccnt, coff, nbins + 1, st));
    RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->d_cub_tmp.p, tmp, ccnt, coff, nbins + 1, st));
    const unsigned nblk = (unsigned)(nbins + (nT + WT_T - 1) / WT_T + 1);
    k_wall_targets_tiled<<<nblk, WT_T * WT_W, 0, st>>>(wall_args(c), c1, nT, d_tx, d_tindx, d_act, d_A,
                                                w.d_tg_order.p, w.d_tg_start.p, coff, nbins, d_v);
  }
  rbc_stage_end(c, ST_WALL);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_add_int_on_walls(rbc3d_ctx *c, double c1, int nT, const double *tx,
                                      const int *tindx, const int *active, const double *Acoef,
                                      double *v) {
  if (!c || nT < 0 || (nT > 0 && (!tx || !tindx || !active || !Acoef || !v))) return RBC3D_EINVAL;
  if (nT == 0 || !c->wall.nvert) return RBC3D_OK;
  hipStream_t st = c->stream;
  RBC_TRY(c->d_tx.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tA.ensure(nT + 1));
  RBC_TRY(c->d_v.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tindx.ensure((size_t)3 * nT + 1));
  RBC_TRY(c->d_tact.ensure(nT + 1));
  HIP_TRY(hipMemcpy(c->d_tx.p, tx, sizeof(double) * 3 * nT, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_tA.p, Acoef, sizeof(double) * nT, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_v.p, v, sizeof(double) * 3 * nT, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_tindx.p, tindx, sizeof(int) * 3 * nT, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_tact.p, active, sizeof(int) * nT, hipMemcpyHostToDevice));
  // a target list that starts with this wall's id is the wall's own vertex list
  // (TargetList_Create order: cells, then walls)
  RBC_TRY(rbc_add_int_on_walls_dev(c, c1, nT, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                   c->d_v.p, tindx[0] == c->wall.id));
  HIP_TRY(hipMemcpyAsync(v, c->d_v.p, sizeof(double) * 3 * nT, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// no-slip operator on the de-duplicated traction vector
__global__ void k_wall_tlist(int nvert, int id, const double *__restrict__ x, double *tx, int *tindx,
                             int *act, double *A) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvert) return;
  for (int d = 0; d < 3; d++) tx[i + (size_t)nvert * d] = x[i + (size_t)nvert * d];
  tindx[i] = id;
  tindx[i + nvert] = i + 1;
  tindx[i + 2 * (size_t)nvert] = -1;
  act[i] = 1;
  A[i] = 2.;
}

// AssembleArray(u, u1D, -1): u(i,:) = u1D(3 glb(i) - 2 : 3 glb(i))
__global__ void k_assemble_in(int nvert, const int *__restrict__ glb, const double *__restrict__ u1d,
                              double *u) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvert) return;
  int g = glb[i] - 1;
  for (int d = 0; d < 3; d++) u[i + (size_t)nvert * d] = u1d[3 * g + d];
}
// AssembleArray(u, u1D, +1): the last vertex of each global index supplies it
__global__ void k_assemble_out(int nuniq, int nvert, const int *__restrict__ last,
                               const double *__restrict__ u, double *u1d) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nuniq) return;
  int i = last[g];
  for (int d = 0; d < 3; d++) u1d[3 * g + d] = u[i + (size_t)nvert * d];
}

static int wall_targets_dev(rbc3d_ctx *c) {
  WallData &w = c->wall;
  int nv = w.nvert;
  RBC_TRY(c->d_tx.ensure((size_t)3 * nv + 1));
  RBC_TRY(c->d_tA.ensure(nv + 1));
  RBC_TRY(c->d_v.ensure((size_t)3 * nv + 1));
  RBC_TRY(c->d_tindx.ensure((size_t)3 * nv + 1));
  RBC_TRY(c->d_tact.ensure(nv + 1));
  k_wall_tlist<<<(nv + 255) / 256, 256, 0, c->stream>>>(nv, w.id, w.d_x.p, c->d_tx.p, c->d_tindx.p,
                                                        c->d_tact.p, c->d_tA.p);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

__global__ void k_col_unique(long long nb, const int *__restrict__ col, const int *__restrict__ glb,
                             int *colu) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k < nb) colu[k] = glb[col[k]] - 1;
}
__global__ void k_e2v_unique(int nele, const int *__restrict__ e2v, const int *__restrict__ glb,
                             int *e2u) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 3 * nele) e2u[k] = glb[e2v[k] - 1] - 1;
}
__global__ void k_fill2(int n, double *a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 2.;
}

// The operator on the de-duplicated vector without assembling it into the vertex array:
// the self-matrix columns and the triangles' vertices are mapped to unique indices once,
// the SpMV and the spread read u directly, and the gather writes b for the last vertex
// of each unique index (AssembleArray's choice) as v_self + v_pme / A, the same sums as
// assembling, accumulating and de-duplicating in turn.  (The SpMV on a second stream
// beside the PME chain measured slower than in turn: both are bound by memory.)
static int wall_matmult_fused(rbc3d_ctx *c, const double *d_u, double *d_b) {
  WallData &w = c->wall;
  hipStream_t st = c->stream;
  const int nv = w.nvert, nu = w.ndof / 3;
  if (!w.op_ready) {
    const long long nb = w.nnz / 9;
    RBC_TRY(w.d_colu.ensure(nb + 1));
    RBC_TRY(w.d_e2u.ensure(3 * (size_t)w.nele + 1));
    RBC_TRY(w.d_A2.ensure(nv + 1));
    k_col_unique<<<(unsigned)((nb + 255) / 256), 256, 0, st>>>(nb, w.d_col.p, w.d_glb.p, w.d_colu.p);
    k_e2v_unique<<<(3 * w.nele + 255) / 256, 256, 0, st>>>(w.nele, w.d_e2v.p, w.d_glb.p, w.d_e2u.p);
    k_fill2<<<(nv + 255) / 256, 256, 0, st>>>(nv, w.d_A2.p);
    HIP_TRY(hipGetLastError());
    w.op_ready = true;
  }
  RBC_TRY(c->d_v.ensure((size_t)3 * nv + 1));
  const double c1 = 1. / (4. * RBC_PI);
  RBC_TRY(rbc_pme_init(c));
  {
    hipStream_t ss = st;
    rbc_stage_begin(c, ST_WALL);
    const unsigned nblk = (unsigned)(((size_t)nv * 64 + 255) / 256);
    if (c->prec == RBC3D_PREC_SPLIT) {
      if (!w.valf_ready) {
        const size_t nv6 = (size_t)w.nnz / 9 * 6;
        RBC_TRY(w.d_valf.ensure(nv6 + 1));
        k_to_float<<<(unsigned)((nv6 + 255) / 256), 256, 0, ss>>>(w.d_val.p, nv6, w.d_valf.p);
        w.valf_ready = true;
      }
      k_wall_spmv<float, true, true><<<nblk, 256, 0, ss>>>(nv, c1, w.d_rowptr.p, w.d_colu.p,
                                                           w.d_valf.p, d_u, w.d_A2.p, c->d_v.p);
    } else {
      k_wall_spmv<double, true, true><<<nblk, 256, 0, ss>>>(nv, c1, w.d_rowptr.p, w.d_colu.p,
                                                            w.d_val.p, d_u, w.d_A2.p, c->d_v.p);
    }
    HIP_TRY(hipGetLastError());
    rbc_stage_end(c, ST_WALL);
  }
  RBC_TRY(rbc_pme_spread_wall_u(c, c1, d_u, w.d_e2u.p));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather_wall_out(c, w.d_A2.p, c->d_v.p, w.d_glb.p + nv, nu, d_b));
  return RBC3D_OK;
}

int rbc_wall_matmult_dev(rbc3d_ctx *c, const double *d_u, double *d_b) {
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  if (!w.prepared) RBC_TRY(rbc_wall_prepare(c));
  if (rbc_wall_spread_applies(c)) return wall_matmult_fused(c, d_u, d_b);
  hipStream_t st = c->stream;
  int nv = w.nvert, nu = w.ndof / 3;
  RBC_TRY(wall_targets_dev(c));
  k_assemble_in<<<(nv + 255) / 256, 256, 0, st>>>(nv, w.d_glb.p, d_u, w.d_f.p);
  HIP_TRY(hipMemsetAsync(c->d_v.p, 0, sizeof(double) * 3 * nv, st));
  double c1 = 1. / (4. * RBC_PI);
  RBC_TRY(rbc_add_int_on_walls_dev(c, c1, nv, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                   c->d_v.p, true));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, 0., 0, 1));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather_wall(c, c->d_tA.p, c->d_v.p));
  k_assemble_out<<<(nu + 255) / 256, 256, 0, st>>>(nu, nv, w.d_glb.p + nv, c->d_v.p, d_b);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_wall_dof(rbc3d_ctx *c, int *dof) {
  if (!c || !dof) return RBC3D_EINVAL;
  *dof = c->wall.ndof;
  return RBC3D_OK;
}

extern "C" int rbc3d_wall_matmult(rbc3d_ctx *c, const double *u, double *b) {
  if (!c || !u || !b) return RBC3D_EINVAL;
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  int nd = w.ndof;
  RBC_TRY(c->d_shbuf2.ensure(2 * (size_t)nd + 2));
  double *du = c->d_shbuf2.p, *db = du + nd + 1;
  HIP_TRY(hipMemcpy(du, u, sizeof(double) * nd, hipMemcpyHostToDevice));
  RBC_TRY(rbc_wall_matmult_dev(c, du, db));
  HIP_TRY(hipMemcpyAsync(b, db, sizeof(double) * nd, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

__global__ void k_add_const(double *v, int n, double s0, double s1, double s2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] += s0;
  v[i + n] += s1;
  v[i + 2 * (size_t)n] += s2;
}

// The cells' part of Compute_Wall_Residual_Vel (common/ModNoSlip.F90:153-193): the
// cell sources' single- and double-layer velocity (real space and PME) at the wall
// vertices, into d_out (nvert,3)
int rbc_wall_cell_vel_dev(rbc3d_ctx *c, double *d_out) {
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  hipStream_t st = c->stream;
  int nv = w.nvert;
  RBC_TRY(wall_targets_dev(c));
  HIP_TRY(hipMemsetAsync(d_out, 0, sizeof(double) * 3 * nv, st));
  if (!c->nrbc) return RBC3D_OK;
  double c1 = 1. / (4. * RBC_PI), c2 = 1. / (4. * RBC_PI);
  RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, nv, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                  d_out, false));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, 1, 0));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather_wall(c, c->d_tA.p, d_out));
  return RBC3D_OK;
}

extern "C" int rbc3d_wall_cell_velocity(rbc3d_ctx *c, double *v) {
  if (!c || !v) return RBC3D_EINVAL;
  RBC_TRY(rbc_refresh_moved(c));
  RBC_TRY(c->d_shbuf2.ensure((size_t)3 * c->wall.nvert + 1));
  RBC_TRY(rbc_wall_cell_vel_dev(c, c->d_shbuf2.p));
  HIP_TRY(hipMemcpyAsync(v, c->d_shbuf2.p, sizeof(double) * 3 * c->wall.nvert,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_wall_vertex_normals(rbc3d_ctx *c, double *n) {
  if (!c || !n) return RBC3D_EINVAL;
  if (!c->wall.nvert) return RBC3D_ESTATE;
  std::copy(c->wall.vnorm.begin(), c->wall.vnorm.end(), n);
  return RBC3D_OK;
}

// Compute_Wall_Residual_Vel (common/ModNoSlip.F90:153-193) into d_v (nvert,3)
int rbc_wall_residual_dev(rbc3d_ctx *c) {
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  if (!w.prepared) RBC_TRY(rbc_wall_prepare(c));
  hipStream_t st = c->stream;
  int nv = w.nvert;
  RBC_TRY(wall_targets_dev(c));
  HIP_TRY(hipMemsetAsync(c->d_v.p, 0, sizeof(double) * 3 * nv, st));
  double c1 = 1. / (4. * RBC_PI), c2 = 1. / (4. * RBC_PI);
  if (c->nrbc)
    RBC_TRY(rbc_add_int_on_rbcs_dev(c, c1, c2, nv, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                    c->d_v.p, false));
  RBC_TRY(rbc_add_int_on_walls_dev(c, c1, nv, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                   c->d_v.p, true));
  RBC_TRY(rbc_pme_init(c));
  RBC_TRY(rbc_pme_spread(c, c1, c2, 1, 1));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather_wall(c, c->d_tA.p, c->d_v.p));
  k_add_const<<<(nv + 255) / 256, 256, 0, st>>>(c->d_v.p, nv, c->conf.vBkg[0], c->conf.vBkg[1],
                                                c->conf.vBkg[2]);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_wall_residual_vel(rbc3d_ctx *c, double *v) {
  if (!c || !v) return RBC3D_EINVAL;
  RBC_TRY(rbc_refresh_moved(c));
  RBC_TRY(rbc_wall_residual_dev(c));
  HIP_TRY(hipMemcpyAsync(v, c->d_v.p, sizeof(double) * 3 * c->wall.nvert, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}
