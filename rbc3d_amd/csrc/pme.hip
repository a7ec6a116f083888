// Smooth particle-mesh Ewald on the GPU (common/ModPME.F90, FFT conventions of
// common/ModPFFTW.F90):
//   spread  -- owner-computes per grid brick: sources are sorted by the brick of
//              their B-spline stencil origin; a workgroup owns one brick, stages
//              the sources of the bricks whose stencils reach it through LDS and
//              each lane sums the contributions to its own grid points (no atomics);
//   FFT     -- in-LDS mixed-radix (2,3,4,5) Stockham FFTs over batches of lines:
//              x r2c, y c2c (sign -1), z c2c (sign +1); inverse z (-1), y (+1), x c2r
//              keeping the real part of the zero and Nyquist x-frequencies;
//   Green   -- elementwise k-space multiply with the B-factor;
//   gather  -- one lane per target, P^3 stencil.
// Every stage is templated on the grid scalar: double for the plain form, float for
// the split form (same kernels, half the bytes; the B-spline weights and the
// Green's-function factors are formed in FP64 and rounded).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <complex>

#include "internal.h"
#include "fft_fixed.h"

#define SPR_B 8        // grid brick edge
#define SPR_T 256
#define FFT_T 256
#ifndef RBC_FX_ON
#define RBC_FX_ON true  // the compile-time-size kernels where an instance exists
#endif
#ifndef RBC_XY_LDS
#define RBC_XY_LDS (80 * 1024)
#endif
#ifndef RBC_ZG_LDS
#define RBC_ZG_LDS (40 * 1024)
#endif
#define FFT_ELEMS 2048 // complex elements per FFT workgroup (lines x length)

template <typename T>
struct CxT;
template <>
struct CxT<double> {
  typedef double2 t;
};
template <>
struct CxT<float> {
  typedef float2 t;
};

// ---------------------------------------------------------------------------
// B-spline weights: BsplineFunc (common/ModBasicMath.F90:392-417), on the host (tables,
// the wall spread's fixed weights) and the device (sources that move)
__host__ __device__ inline int bspline_d(double xc, int P, double *w) {
  double u[12];
  int imin = (int)floor(xc) - (P - 1);
  u[0] = imin - (xc - P);
  for (int j = 1; j < P; j++) u[j] = u[j - 1] + 1;
  w[0] = 1.;
  for (int j = 1; j < P; j++) w[j] = 0.;
  for (int pp = 2; pp <= P; pp++) {
    for (int j = pp; j >= 2; j--)
      w[j - 1] = u[j - 1] / (pp - 1.) * w[j - 1] + (pp - u[j - 1]) / (pp - 1.) * w[j - 2];
    w[0] = u[0] / (pp - 1.) * w[0];
  }
  return imin;
}

static void factor_radices(int n, std::vector<int> &r) {
  r.clear();
  while (n % 4 == 0) {
    r.push_back(4);
    n /= 4;
  }
  while (n % 2 == 0) {
    r.push_back(2);
    n /= 2;
  }
  while (n % 3 == 0) {
    r.push_back(3);
    n /= 3;
  }
  while (n % 5 == 0) {
    r.push_back(5);
    n /= 5;
  }
  if (n != 1) r.push_back(-n);  // unsupported factor
}

// ---------------------------------------------------------------------------
int rbc_pme_init(rbc3d_ctx *c) {
  if (c->pme_ready) return RBC3D_OK;
  int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], P = c->conf.P;
  for (int ax = 0; ax < 3; ax++) {
    std::vector<int> r;
    factor_radices(c->conf.Nb[ax], r);
    if (!r.empty() && r.back() < 0) {
      rbc_set_error("PME grid sizes must factor into 2, 3 and 5");
      return RBC3D_EINVAL;
    }
    if (c->conf.Nb[ax] > FFT_ELEMS) {
      rbc_set_error("PME grid size too large");
      return RBC3D_EINVAL;
    }
  }
  int H = N1 / 2 + 1;
  size_t ng = (size_t)N1 * N2 * N3, nc = (size_t)H * N2 * N3;
  if (12 * ng >= (1ull << 31) || 15 * nc >= (1ull << 31)) {
    rbc_set_error("PME grid too large for the FFT kernels' 32-bit offsets");
    return RBC3D_EINVAL;
  }
  RBC_TRY(c->d_grid.ensure(12 * ng));
  RBC_TRY(c->d_cgrid.ensure(2 * 15 * nc));
  RBC_TRY(c->d_vgrid.ensure(3 * ng));
  // B-factor per axis (PME_Init :309-336); bb = ((1*bx)*by)*bz, bb(0,0,0) = 0
  double MP[16];
  bspline_d(P + 2.220446049250313e-16, P, MP);
  std::vector<double> fac((size_t)H + N2 + N3);
  int Ns[3] = {N1, N2, N3}, cnt[3] = {H, N2, N3}, off = 0;
  for (int ax = 0; ax < 3; ax++) {
    for (int k = 0; k < cnt[ax]; k++) {
      std::complex<double> b = 0, I(0, 1);
      for (int m = 0; m <= P - 2; m++) b += MP[m] * std::exp(RBC_TWO_PI * I * (double)k * (double)m / (double)Ns[ax]);
      b = std::exp(RBC_TWO_PI * I * (double)k * (P - 1.) / (double)Ns[ax]) / b;
      double bv = std::abs(b);
      fac[off + k] = bv * bv;
    }
    off += cnt[ax];
  }
  RBC_TRY(c->d_bb.ensure(fac.size()));
  HIP_TRY(hipMemcpy(c->d_bb.p, fac.data(), fac.size() * 8, hipMemcpyHostToDevice));
  // twiddles exp(-2 pi i m / N) for the three lengths
  std::vector<double> tw;
  for (int ax = 0; ax < 3; ax++) {
    c->tw_off[ax] = (int)(tw.size() / 2);
    int N = Ns[ax];
    for (int m = 0; m < N; m++) {
      long double ang = 2.0L * 3.141592653589793238462643383279502884L * m / N;
      tw.push_back((double)cosl(ang));
      tw.push_back((double)-sinl(ang));
    }
  }
  RBC_TRY(c->d_tw.ensure(tw.size()));
  HIP_TRY(hipMemcpy(c->d_tw.p, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
  std::vector<float> twf(tw.begin(), tw.end());
  RBC_TRY(c->d_twf.ensure(twf.size()));
  HIP_TRY(hipMemcpy(c->d_twf.p, twf.data(), twf.size() * 4, hipMemcpyHostToDevice));
  c->pme_ready = true;
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// spread
// record layout (doubles): [0..2] imin,jmin,kmin (as doubles) [3..3+3P) weights,
// then nf field values (c1*f for single layer, c2*g_i*(a3_j Bcoef) for double layer)
template <typename R>
__global__ void k_spread_prep(int ns, int ncell, const double *__restrict__ x, int ldx,
                              const double *__restrict__ fs, const double *__restrict__ gs,
                              const double *__restrict__ a3, const double *__restrict__ Bn,
                              const int *__restrict__ node_cell, const CellRec *__restrict__ cells,
                              int nele, const double *__restrict__ wx, const double *__restrict__ wf,
                              const int *__restrict__ e2v, int nvert, const double *__restrict__ warea,
                              double ih0, double ih1, double ih2, int N1, int N2, int N3, int P,
                              int flag_sl, int flag_dl, int nbx, int nby, int nbz, R *rec,
                              int reclen, int *key, int *idx) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ns) return;
  double xs[3], f[3] = {0, 0, 0}, t[9];
  for (int q = 0; q < 9; q++) t[q] = 0;
  if (p < ncell) {
    for (int d = 0; d < 3; d++) xs[d] = x[p + (size_t)ldx * d];
    if (flag_sl)
      for (int d = 0; d < 3; d++) f[d] = fs[p + (size_t)ldx * d];
    if (flag_dl) {
      double B = cells[node_cell[p]].Bcoef, g[3], n[3];
      for (int d = 0; d < 3; d++) {
        g[d] = gs[p + (size_t)ldx * d];
        n[d] = a3[p + (size_t)ldx * d] * B;
      }
      for (int ii = 0; ii < 3; ii++)
        for (int jj = 0; jj < 3; jj++) t[ii + 3 * jj] = g[ii] * n[jj];
    }
  } else {
    // wall triangle centroid, f = mean vertex traction * area (PME_Distrib_Source :108-130)
    int e = p - ncell;
    double xe[3][3], fe[3][3];
    for (int l = 0; l < 3; l++) {
      int iv = e2v[e + (size_t)nele * l] - 1;
      for (int d = 0; d < 3; d++) {
        xe[l][d] = wx[iv + (size_t)nvert * d];
        fe[l][d] = wf[iv + (size_t)nvert * d];
      }
    }
    for (int d = 0; d < 3; d++) xs[d] = (1.0 / 3) * (xe[0][d] + xe[1][d] + xe[2][d]);
    if (flag_sl)
      for (int d = 0; d < 3; d++) f[d] = (1.0 / 3) * (fe[0][d] + fe[1][d] + fe[2][d]) * warea[e];
  }
  R *r = rec + (size_t)reclen * p;
  double w[12];
  int imin = bspline_d(xs[0] * ih0, P, w);
  for (int q = 0; q < P; q++) r[3 + q] = w[q];
  int jmin = bspline_d(xs[1] * ih1, P, w);
  for (int q = 0; q < P; q++) r[3 + P + q] = w[q];
  int kmin = bspline_d(xs[2] * ih2, P, w);
  for (int q = 0; q < P; q++) r[3 + 2 * P + q] = w[q];
  int i0 = imod_d(imin, N1), j0 = imod_d(jmin, N2), k0 = imod_d(kmin, N3);
  r[0] = i0;
  r[1] = j0;
  r[2] = k0;
  int o = 3 + 3 * P;
  for (int d = 0; d < 3; d++) r[o + d] = f[d];
  for (int q = 0; q < 9; q++) r[o + 3 + q] = t[q];
  // brick of the stencil origin
  key[p] = (i0 / SPR_B) + nbx * ((j0 / SPR_B) + nby * (k0 / SPR_B));
  idx[p] = p;
  (void)nbz;
}

template <typename T>
struct SpreadArgs {
  int N1, N2, N3, P, nbx, nby, nbz, reclen, ns;
  FastDiv frl;  // LDS record length
  T c1, c2;
  int flag_sl, flag_dl;
  const T *rec;
  const int *order, *brick_start;
  T *grid;  // 12 fields [f][k][j][i]
};

// One workgroup per brick of the grid.  A source with stencil origin g covers
// g..g+P-1 (mod N), so brick [b0, b0+B) receives sources with origin in
// [b0-P+1, b0+B-1] (mod N): the bricks that overlap that range on each axis.
// SL / DL select the single-layer (3) and double-layer (9) fields; sources stream
// through LDS in tiles of SPR_TILE records holding only what the fields need.
#define SPR_TILE 64
template <bool SL, bool DL, typename T>
__global__ void __launch_bounds__(SPR_T) k_spread(SpreadArgs<T> a) {
  constexpr int NF = (SL ? 3 : 0) + (DL ? 9 : 0);
  constexpr int MAXP = 12;
  __shared__ T srec[SPR_TILE * (3 + 3 * MAXP + NF)];
  const int bid = blockIdx.x;
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  const int ib0 = bx * SPR_B, jb0 = by * SPR_B, kb0 = bz * SPR_B;
  const int P = a.P;
  const int rl = 3 + 3 * P + NF;  // LDS record: origin, weights, fields
  // lists of source bricks per axis (deduplicated, at most 3 + wrap)
  int lst[3][4], nl[3];
  int Ns[3] = {a.N1, a.N2, a.N3}, b0s[3] = {ib0, jb0, kb0};
  for (int ax = 0; ax < 3; ax++) {
    nl[ax] = 0;
    int lo = b0s[ax] - P + 1, hi = b0s[ax] + SPR_B - 1;
    for (int g = lo; g <= hi; g++) {
      int gg = imod_d(g, Ns[ax]);
      int br = gg / SPR_B;
      bool dup = false;
      for (int q = 0; q < nl[ax]; q++) dup |= lst[ax][q] == br;
      if (!dup && nl[ax] < 4) lst[ax][nl[ax]++] = br;
    }
  }
  // my grid points (two per lane for B = 8)
  const int npts = SPR_B * SPR_B * SPR_B;
  T acc[2][NF > 0 ? NF : 1];
  int gi[2], gj[2], gk[2];
  bool valid[2];
  for (int s = 0; s < 2; s++) {
    int pt = threadIdx.x + s * SPR_T;
    gi[s] = ib0 + pt % SPR_B;
    gj[s] = jb0 + (pt / SPR_B) % SPR_B;
    gk[s] = kb0 + pt / (SPR_B * SPR_B);
    valid[s] = pt < npts && gi[s] < a.N1 && gj[s] < a.N2 && gk[s] < a.N3;
    for (int f = 0; f < NF; f++) acc[s][f] = 0;
  }
  const int grl = a.reclen;  // global record: origin, weights, f(3), t(9)
  const int goff = SL ? 0 : 3;  // first global field slot used
  for (int qz = 0; qz < nl[2]; qz++)
    for (int qy = 0; qy < nl[1]; qy++)
      for (int qx = 0; qx < nl[0]; qx++) {
        int sb = lst[0][qx] + a.nbx * (lst[1][qy] + a.nby * lst[2][qz]);
        int s0 = a.brick_start[sb], s1 = a.brick_start[sb + 1];
        for (int base = s0; base < s1; base += SPR_TILE) {
          const int cnt = min(SPR_TILE, s1 - base);
          __syncthreads();
          for (int t = threadIdx.x; t < cnt * rl; t += SPR_T) {
            const int q = fd_div(t, a.frl), e = t - q * rl;
            const int ge = e < 3 + 3 * P ? e : e + goff;
            srec[t] = a.rec[(size_t)grl * a.order[base + q] + ge];
          }
          __syncthreads();
          for (int q = 0; q < cnt; q++) {
            const T *r = srec + q * rl;
            const int i0 = (int)r[0], j0 = (int)r[1], k0 = (int)r[2];
            for (int s = 0; s < 2; s++) {
              if (!valid[s]) continue;
              int di = gi[s] - i0, dj = gj[s] - j0, dk = gk[s] - k0;
              di += di < 0 ? a.N1 : 0;
              dj += dj < 0 ? a.N2 : 0;
              dk += dk < 0 ? a.N3 : 0;
              if (di >= P || dj >= P || dk >= P) continue;
              const T wxyz = r[3 + di] * r[3 + P + dj] * r[3 + 2 * P + dk];
              const T *fv = r + 3 + 3 * P;
              if (SL) {
                const T cw = a.c1 * wxyz;
                for (int f = 0; f < 3; f++) acc[s][f] += cw * fv[f];
              }
              if (DL) {
                const T cw = a.c2 * wxyz;
                constexpr int o = SL ? 3 : 0;
                for (int f = 0; f < 9; f++) acc[s][o + f] += cw * fv[o + f];
              }
            }
          }
        }
      }
  const size_t ng = (size_t)a.N1 * a.N2 * a.N3;
  for (int s = 0; s < 2; s++) {
    if (!valid[s]) continue;
    const size_t o = gi[s] + (size_t)a.N1 * (gj[s] + (size_t)a.N2 * gk[s]);
    for (int f = 0; f < NF; f++) a.grid[o + ng * (goff + f)] = acc[s][f];
  }
}

// Single-layer spread in two passes (used when only the three single-layer fields
// are spread, P <= SPT_PMAX and every grid size is at least 3 SPR_B: the wall-only
// spread of every no-slip GMRES iteration).  Pass 1: one workgroup per source
// brick accumulates its own sources into an LDS tile of (SPR_B + P - 1)^3 points,
// the tile of points their stencils reach, and stores it; every source is handled
// once.  Pass 2: every grid point adds the tiles that reach it, in increasing brick
// order.  Within pass 1 each wavefront owns a fixed set of tile z-planes and lanes
// take a plane's P x P stencil points, so each tile point is updated by one lane in
// source order: deterministic, no atomics.
#define SPT_PMAX 8
#define SPT_E (SPR_B + SPT_PMAX - 1)  // tile edge bound
template <typename T>
__global__ void __launch_bounds__(SPR_T) k_spread_tile(SpreadArgs<T> a, T *tiles) {
  const int P = a.P, E = SPR_B + P - 1, E3 = E * E * E;
  __shared__ T acc[3 * SPT_E * SPT_E * SPT_E];  // field f at f * E3
  __shared__ T srec[SPR_TILE * (3 + 3 * SPT_PMAX + 3)];
  const int bid = blockIdx.x;
  const int s0 = a.brick_start[bid], s1 = a.brick_start[bid + 1];
  if (s0 == s1) return;  // empty brick: no tile (pass 2 skips it)
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  const int ib0 = bx * SPR_B, jb0 = by * SPR_B, kb0 = bz * SPR_B;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nwv = SPR_T / 64;
  for (int t = threadIdx.x; t < 3 * E3; t += SPR_T) acc[t] = (T)0;
  const int rl = 3 + 3 * P + 3;
  for (int base = s0; base < s1; base += SPR_TILE) {
    const int cnt = min(SPR_TILE, s1 - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt * rl; t += SPR_T) {
      const int q = fd_div(t, a.frl), e = t - q * rl;
      srec[t] = a.rec[(size_t)a.reclen * a.order[base + q] + e];
    }
    __syncthreads();
    for (int q = 0; q < cnt; q++) {
      const T *r = srec + q * rl;
      // the source's origin relative to the brick (0..SPR_B-1 on each axis)
      const int oi = (int)r[0] - ib0, oj = (int)r[1] - jb0, ok = (int)r[2] - kb0;
      const T *fv = r + 3 + 3 * P;
      for (int tz = wv; tz < E; tz += nwv) {  // this wavefront's tile planes
        const int dk = tz - ok;
        if (dk < 0 || dk >= P) continue;
        const T wz = a.c1 * r[3 + 2 * P + dk];
        for (int t = lane; t < P * P; t += 64) {
          const int di = t % P, dj = t / P;
          const int pt = (oi + di) + E * ((oj + dj) + E * tz);
          const T w = r[3 + di] * r[3 + P + dj] * wz;
#pragma unroll
          for (int f = 0; f < 3; f++) acc[f * E3 + pt] += w * fv[f];
        }
      }
    }
  }
  __syncthreads();
  T *out = tiles + (size_t)bid * 3 * E3;
  for (int t = threadIdx.x; t < 3 * E3; t += SPR_T) out[t] = acc[t];
}

// pass 2: grid point g sums the tiles of the bricks b with (g - SPR_B b) mod N < E
template <typename T>
__global__ void k_spread_merge(SpreadArgs<T> a, const T *__restrict__ tiles) {
  const size_t ng = (size_t)a.N1 * a.N2 * a.N3;
  const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (o >= ng) return;
  const int P = a.P, E = SPR_B + P - 1, E3 = E * E * E;
  const int g[3] = {(int)(o % a.N1), (int)((o / a.N1) % a.N2), (int)(o / ((size_t)a.N1 * a.N2))};
  const int Ns[3] = {a.N1, a.N2, a.N3}, nbs[3] = {a.nbx, a.nby, a.nbz};
  // candidate bricks per axis in increasing brick index, with the tile coordinate
  int cb[3][3], cl[3][3], nc[3];
  for (int ax = 0; ax < 3; ax++) {
    nc[ax] = 0;
    int b0 = g[ax] / SPR_B;
    int cand[3] = {b0 - 2, b0 - 1, b0};
    for (int k = 0; k < 3; k++) {
      int b = cand[k];
      b += b < 0 ? nbs[ax] : 0;
      int l = g[ax] - SPR_B * b;
      l += l < 0 ? Ns[ax] : 0;
      if (l < E) {
        cb[ax][nc[ax]] = b;
        cl[ax][nc[ax]] = l;
        nc[ax]++;
      }
    }
    // keep increasing brick order (a wrapped candidate may come first)
    for (int i = 1; i < nc[ax]; i++)
      for (int j = i; j > 0 && cb[ax][j] < cb[ax][j - 1]; j--) {
        int tb = cb[ax][j], tl = cl[ax][j];
        cb[ax][j] = cb[ax][j - 1];
        cl[ax][j] = cl[ax][j - 1];
        cb[ax][j - 1] = tb;
        cl[ax][j - 1] = tl;
      }
  }
  T s0 = 0, s1 = 0, s2 = 0;
  for (int kz = 0; kz < nc[2]; kz++)
    for (int ky = 0; ky < nc[1]; ky++)
      for (int kx = 0; kx < nc[0]; kx++) {
        const int b = cb[0][kx] + a.nbx * (cb[1][ky] + a.nby * cb[2][kz]);
        if (a.brick_start[b] == a.brick_start[b + 1]) continue;
        const T *t = tiles + (size_t)b * 3 * E3 + cl[0][kx] + E * (cl[1][ky] + E * cl[2][kz]);
        s0 += t[0];
        s1 += t[E3];
        s2 += t[2 * E3];
      }
  a.grid[o] = s0;
  a.grid[o + ng] = s1;
  a.grid[o + 2 * ng] = s2;
}

// Wall-only single-layer spread (every no-slip GMRES iteration).  The sources are the
// wall triangles' centroids, which do not move: once per wall and grid, each source's
// stencil origin and B-spline weights are computed, and each 4^3 brick of the grid gets
// the list of sources whose stencil reaches it, in ascending source order, with the
// stencil origin relative to the brick.  A spread is one wavefront per brick, a lane
// per grid point.  Per round of 64 sources each lane stages one source's weights on
// the brick's 4 x 4 (x, y) points (wx wy, zero off the stencil), on its 4 z planes, and
// its value; then every lane walks the round's sources in order and adds
// c1 ((wx wy) wz) f to its point.  Every point adds its sources in ascending order with
// the reference's expressions (PME_Distrib_Source :108-130, Distrib_Source :405-443)
// and no FMA contraction: the grid is the CPU restatement's bit for bit (a point off a
// source's stencil adds an exact zero).  Bricks without sources write zeros.  The split
// form stages and sums in FP32, the scalar of its grid.
#define WSB 4          // brick edge in grid points (a brick is 64 points)
#define WSB_CH 64      // sources staged per round
#define WSB_PMAX 8
template <typename T>
struct WspArgs {
  int N1, N2, N3, P, nbx, nby, nbz, nele, nvert, nbr;
  double c1;
  const int *bstart, *bsrc, *brel;  // per brick: its sources, their origins (packed)
  const double *wts;                // per source 3 P weights
  const int *e2v;
  const double *wf, *area;
  const double *u;    // when set: tractions from the de-duplicated vector u(3, nunique)
  const int *e2u;     //   through the triangles' unique vertex indices e2u(nele, 3)
  double *fs;  // per-source values f(nele, 3)
  T *grid;
};

// per-triangle source values: mean vertex traction times area (PME_Distrib_Source
// :108-130), fs(nele, 3)
template <typename T>
__global__ void k_wsp_src(WspArgs<T> a) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nele) return;
  if (a.u) {
    const int g0 = a.e2u[e], g1 = a.e2u[e + a.nele], g2 = a.e2u[e + 2 * (size_t)a.nele];
    for (int d = 0; d < 3; d++)
      a.fs[e + (size_t)a.nele * d] =
          (1.0 / 3) * (a.u[3 * g0 + d] + a.u[3 * g1 + d] + a.u[3 * g2 + d]) * a.area[e];
    return;
  }
  const int v0 = a.e2v[e] - 1, v1 = a.e2v[e + a.nele] - 1, v2 = a.e2v[e + 2 * (size_t)a.nele] - 1;
  for (int d = 0; d < 3; d++) {
    const size_t o = (size_t)a.nvert * d;
    a.fs[e + (size_t)a.nele * d] = (1.0 / 3) * (a.wf[v0 + o] + a.wf[v1 + o] + a.wf[v2 + o]) * a.area[e];
  }
}

// one source staged for a brick: wx wy on the 16 (x, y) points, wz on the 4 z planes,
// the value; in the grid's scalar A (the split form stages and sums in FP32)
template <typename A>
struct WspStage {
  A xy[16], z[4], f[3];
};
template <typename T>
__device__ __forceinline__ void wsp_fetch(WspStage<T> &r, const WspArgs<T> &a, int k) {
#pragma clang fp contract(off)
  const int e = a.bsrc[k], rp = a.brel[k], P = a.P;
  const double *w = a.wts + (size_t)3 * P * e;
  double wd[3][WSB];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int rel = ((rp >> (8 * d)) & 255) - 16;  // brick-local stencil start
#pragma unroll
    for (int l = 0; l < WSB; l++) {
      const int q = l - rel;
      wd[d][l] = (q >= 0 && q < P) ? w[d * P + q] : 0.;
    }
  }
#pragma unroll
  for (int ly = 0; ly < WSB; ly++)
#pragma unroll
    for (int lx = 0; lx < WSB; lx++) r.xy[lx + WSB * ly] = (T)(wd[0][lx] * wd[1][ly]);
#pragma unroll
  for (int l = 0; l < WSB; l++) r.z[l] = (T)wd[2][l];
#pragma unroll
  for (int d = 0; d < 3; d++) r.f[d] = (T)a.fs[e + (size_t)a.nele * d];
}

template <typename T>
struct Vec4T;
template <>
struct Vec4T<double> {
  typedef double4 t;
};
template <>
struct Vec4T<float> {
  typedef float4 t;
};

template <typename T>
__global__ void __launch_bounds__(64) k_wsp_brick(WspArgs<T> a) {
#pragma clang fp contract(off)
  typedef typename Vec4T<T>::t V4;
  __shared__ T sxy[WSB_CH][16 + 16 / sizeof(T)];  // rows padded: conflict-free 16-byte stores
  __shared__ V4 sz[WSB_CH];
  __shared__ V4 sf[WSB_CH];
  // consecutive bricks of an x-row on one XCD: in FP32 a brick row is 16 bytes, half a
  // write sector, and the neighbouring brick's half then merges in the same L2
  const int b = xcd_block(a.nbr);
  if (b < 0) return;
  const int lane = threadIdx.x;
  const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
  const int lxy = lane & 15, lz = lane >> 4;
  const int gi = bx * WSB + (lane & 3), gj = by * WSB + ((lane >> 2) & 3), gk = bz * WSB + lz;
  T acc0 = 0., acc1 = 0., acc2 = 0.;
  const T c1 = (T)a.c1;
  const int s0 = a.bstart[b], s1 = a.bstart[b + 1];
  WspStage<T> r;
  if (s0 + lane < s1) wsp_fetch(r, a, s0 + lane);
  for (int base = s0; base < s1; base += WSB_CH) {
    const int cnt = min(WSB_CH, s1 - base);
    __syncthreads();
    if (lane < cnt) {
#pragma unroll
      for (int q = 0; q < 16; q++) sxy[lane][q] = r.xy[q];
      sz[lane] = V4{r.z[0], r.z[1], r.z[2], r.z[3]};
      sf[lane] = V4{r.f[0], r.f[1], r.f[2], (T)0};
    }
    __syncthreads();
    // the next round's sources load while this round is summed
    if (base + WSB_CH + lane < s1) wsp_fetch(r, a, base + WSB_CH + lane);
    const T *zrow = reinterpret_cast<const T *>(sz) + lz;
#pragma unroll 8
    for (int t = 0; t < cnt; t++) {
      const T wxyz = sxy[t][lxy] * zrow[4 * t];
      const T cw = c1 * wxyz;
      const V4 f = sf[t];
      acc0 = acc0 + cw * f.x;
      acc1 = acc1 + cw * f.y;
      acc2 = acc2 + cw * f.z;
    }
  }
  if (gi < a.N1 && gj < a.N2 && gk < a.N3) {
    const size_t ng = (size_t)a.N1 * a.N2 * a.N3;
    const size_t g = gi + (size_t)a.N1 * (gj + (size_t)a.N2 * gk);
    a.grid[g] = acc0;
    a.grid[g + ng] = acc1;
    a.grid[g + 2 * ng] = acc2;
  }
}

// whether the wall-only spread above applies: B-spline order <= 8 and every grid
// dimension at least P + 4 (a stencil then reaches a brick through one image only);
// else the general spread runs
static bool wall_spread_ok(const rbc3d_ctx *c) {
  const int P = c->conf.P;
  return P <= WSB_PMAX && c->conf.Nb[0] >= P + WSB && c->conf.Nb[1] >= P + WSB &&
         c->conf.Nb[2] >= P + WSB;
}

static int wall_spread_setup(rbc3d_ctx *c) {
  WallData &w = c->wall;
  if (w.wsp_ready) return RBC3D_OK;
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], P = c->conf.P;
  const int nele = w.nele;
  const int nbx = (N1 + WSB - 1) / WSB, nby = (N2 + WSB - 1) / WSB, nbz = (N3 + WSB - 1) / WSB;
  const double ih[3] = {N1 / c->Lb[0], N2 / c->Lb[1], N3 / c->Lb[2]};
  const int Nd[3] = {N1, N2, N3};
  std::vector<double> cen((size_t)3 * nele);
  HIP_TRY(hipMemcpy(cen.data(), w.d_cen.p, sizeof(double) * 3 * nele, hipMemcpyDeviceToHost));
  std::vector<double> wts((size_t)3 * P * nele);
  // per brick: (source, brick-local stencil origin + 16 per axis, packed in bytes)
  std::vector<std::vector<std::pair<int, int>>> lists((size_t)nbx * nby * nbz);
  for (int e = 0; e < nele; e++) {
    int bl[3][WSB_PMAX], rl[3][WSB_PMAX], nb[3];  // distinct bricks per axis, local origin
    for (int d = 0; d < 3; d++) {
      const int o = bspline_d(cen[e + (size_t)nele * d] * ih[d], P, &wts[(size_t)3 * P * e + d * P]);
      nb[d] = 0;
      for (int q = 0; q < P; q++) {
        const int br = (((o + q) % Nd[d]) + Nd[d]) % Nd[d] / WSB;
        int u = 0;
        while (u < nb[d] && bl[d][u] != br) u++;
        if (u == nb[d]) {
          // the stencil start relative to the brick, in [-(P - 1), WSB - 1]
          int rel = (((o - br * WSB) % Nd[d]) + Nd[d]) % Nd[d];
          if (rel > WSB - 1) rel -= Nd[d];
          bl[d][nb[d]] = br;
          rl[d][nb[d]++] = rel + 16;
        }
      }
    }
    for (int u = 0; u < nb[0]; u++)
      for (int v = 0; v < nb[1]; v++)
        for (int x = 0; x < nb[2]; x++)
          lists[bl[0][u] + (size_t)nbx * (bl[1][v] + (size_t)nby * bl[2][x])].push_back(
              {e, rl[0][u] | (rl[1][v] << 8) | (rl[2][x] << 16)});
  }
  std::vector<int> bstart(lists.size() + 1, 0), bsrc, brel;
  for (size_t b = 0; b < lists.size(); b++) {
    bstart[b + 1] = bstart[b] + (int)lists[b].size();
    for (auto &pr : lists[b]) {
      bsrc.push_back(pr.first);
      brel.push_back(pr.second);
    }
  }
  if (bsrc.empty()) {
    bsrc.push_back(0);
    brel.push_back(0);
  }
  RBC_TRY(w.d_wsb_start.ensure(bstart.size()));
  RBC_TRY(w.d_wsb_src.ensure(bsrc.size()));
  RBC_TRY(w.d_wsb_rel.ensure(brel.size()));
  RBC_TRY(w.d_wsb_w.ensure(wts.size()));
  RBC_TRY(w.d_wsb_f.ensure(3 * (size_t)nele));
  HIP_TRY(hipMemcpy(w.d_wsb_start.p, bstart.data(), bstart.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_wsb_src.p, bsrc.data(), bsrc.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_wsb_rel.p, brel.data(), brel.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_wsb_w.p, wts.data(), wts.size() * sizeof(double), hipMemcpyHostToDevice));
  w.wsp_ready = true;
  return RBC3D_OK;
}

template <typename T>
static int wall_spread(rbc3d_ctx *c, double c1, const double *d_u = nullptr,
                       const int *d_e2u = nullptr) {
  WallData &w = c->wall;
  RBC_TRY(wall_spread_setup(c));
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2];
  WspArgs<T> a;
  a.N1 = N1;
  a.N2 = N2;
  a.N3 = N3;
  a.P = c->conf.P;
  a.nbx = (N1 + WSB - 1) / WSB;
  a.nby = (N2 + WSB - 1) / WSB;
  a.nbz = (N3 + WSB - 1) / WSB;
  a.nbr = a.nbx * a.nby * a.nbz;
  a.nele = w.nele;
  a.nvert = w.nvert;
  a.c1 = c1;
  a.bstart = w.d_wsb_start.p;
  a.bsrc = w.d_wsb_src.p;
  a.brel = w.d_wsb_rel.p;
  a.wts = w.d_wsb_w.p;
  a.e2v = w.d_e2v.p;
  a.wf = w.d_f.p;
  a.area = w.d_area.p;
  a.u = d_u;
  a.e2u = d_e2u;
  a.fs = w.d_wsb_f.p;
  a.grid = (T *)c->d_grid.p;
  rbc_stage_begin(c, ST_SPREAD);
  k_wsp_src<T><<<(w.nele + 255) / 256, 256, 0, c->stream>>>(a);
  k_wsp_brick<T><<<xcd_grid(a.nbr), 64, 0, c->stream>>>(a);
  rbc_stage_end(c, ST_SPREAD);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// brick_start[b] = first sorted position with brick >= b, b = 0..nbr
__global__ void k_brick_start(int n, const int *__restrict__ skey, int nbr, int *start) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbr) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (skey[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  start[b] = lo;
}

template <typename T>
static int pme_spread_t(rbc3d_ctx *c, double c1, double c2, int cells, int walls) {
  hipStream_t st = c->stream;
  int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], P = c->conf.P;
  c->flag_sl = fabs(c1) > 1.e-10;
  c->flag_dl = fabs(c2) > 1.e-10;
  if (!(cells && c->nrbc) && walls && c->wall.nvert && c->flag_sl && !c->flag_dl &&
      wall_spread_ok(c))
    return wall_spread<T>(c, c1);
  int ncell = (cells && c->nrbc) ? c->np : 0;
  int nele = (walls && c->wall.nvert) ? c->wall.nele : 0;
  int ns = ncell + nele;
  int nbx = (N1 + SPR_B - 1) / SPR_B, nby = (N2 + SPR_B - 1) / SPR_B, nbz = (N3 + SPR_B - 1) / SPR_B;
  int nbr = nbx * nby * nbz;
  int reclen = 3 + 3 * P + 12;
  RBC_TRY(c->d_pme_src.ensure((size_t)reclen * ns + 1));
  RBC_TRY(c->d_pme_key.ensure(2 * (size_t)ns + 2));
  RBC_TRY(c->d_pme_order.ensure(2 * (size_t)ns + 2));
  RBC_TRY(c->d_pme_start.ensure(nbr + 1));
  int *key = c->d_pme_key.p, *key2 = c->d_pme_key.p + ns + 1;
  int *idx = c->d_pme_order.p + ns + 1, *order = c->d_pme_order.p;
  if (ns > 0) {
    rbc_stage_begin(c, ST_SPREAD_PREP);
    k_spread_prep<T><<<(ns + 255) / 256, 256, 0, st>>>(
        ns, ncell, c->d_x.p, c->np, c->d_fs.p, c->d_gs.p, c->d_a3.p, nullptr, c->d_node_cell.p,
        c->d_cells.p, nele, c->wall.d_x.p, c->wall.d_f.p, c->wall.d_e2v.p, c->wall.nvert,
        c->wall.d_area.p, N1 / c->Lb[0], N2 / c->Lb[1], N3 / c->Lb[2], N1, N2, N3, P, c->flag_sl,
        c->flag_dl, nbx, nby, nbz, (T *)c->d_pme_src.p, reclen, key, idx);
    rbc_stage_end(c, ST_SPREAD_PREP);
    HIP_TRY(hipGetLastError());
  }
  // The sources' order by brick depends on the positions only: with geometry reuse, the
  // cell sources' sort (order, brick starts) is kept for the geometry epoch it was made
  // at.  Every call that sorts replaces it, so the kept one is the last sort made.
  const bool keep = cells && !walls && ncell > 0 && c->geom_reuse;
  const bool sorted = keep && c->spread_sort_epoch == c->geom_epoch && c->spread_sort_ns == ns;
  if (!sorted) {
    if (ns > 0) {
      int bits = 1;
      while ((1 << bits) < nbr) bits++;
      size_t tmp = 0;
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key, key2, idx, order, ns, 0, bits, st));
      RBC_TRY(c->d_cub_tmp.ensure(tmp + 16));
      HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->d_cub_tmp.p, tmp, key, key2, idx, order, ns, 0,
                                                 bits, st));
    }
    k_brick_start<<<(nbr + 1 + 255) / 256, 256, 0, st>>>(ns, key2, nbr, c->d_pme_start.p);
    HIP_TRY(hipGetLastError());
  }
  c->spread_sort_epoch = keep ? c->geom_epoch : -1;
  c->spread_sort_ns = ns;
  SpreadArgs<T> a;
  a.N1 = N1;
  a.N2 = N2;
  a.N3 = N3;
  a.P = P;
  a.nbx = nbx;
  a.nby = nby;
  a.nbz = nbz;
  a.reclen = reclen;
  a.ns = ns;
  a.c1 = (T)c1;
  a.c2 = (T)c2;
  a.flag_sl = c->flag_sl;
  a.flag_dl = c->flag_dl;
  a.rec = (const T *)c->d_pme_src.p;
  a.order = order;
  a.brick_start = c->d_pme_start.p;
  a.grid = (T *)c->d_grid.p;
  if (P > 12) {
    rbc_set_error("B-spline order above 12");
    return RBC3D_EINVAL;
  }
  a.frl = make_fastdiv(3 + 3 * P + (c->flag_sl ? 3 : 0) + (c->flag_dl ? 9 : 0));
  rbc_stage_begin(c, ST_SPREAD);
  // (three or more bricks per axis keep pass 2's candidate bricks distinct)
  const bool tile_path = c->flag_sl && !c->flag_dl && P <= SPT_PMAX && N1 >= 3 * SPR_B &&
                         N2 >= 3 * SPR_B && N3 >= 3 * SPR_B;
  if (tile_path) {
    const size_t E = SPR_B + P - 1, tsz = 3 * E * E * E;
    RBC_TRY(c->d_spread_tiles.ensure((tsz * nbr * sizeof(T) + 7) / 8));
    T *tiles = (T *)c->d_spread_tiles.p;
    a.frl = make_fastdiv(3 + 3 * P + 3);
    k_spread_tile<T><<<nbr, SPR_T, 0, st>>>(a, tiles);
    const size_t ng = (size_t)N1 * N2 * N3;
    k_spread_merge<T><<<(unsigned)((ng + 255) / 256), 256, 0, st>>>(a, tiles);
  } else if (c->flag_sl && c->flag_dl)
    k_spread<true, true, T><<<nbr, SPR_T, 0, st>>>(a);
  else if (c->flag_sl)
    k_spread<true, false, T><<<nbr, SPR_T, 0, st>>>(a);
  else if (c->flag_dl)
    k_spread<false, true, T><<<nbr, SPR_T, 0, st>>>(a);
  rbc_stage_end(c, ST_SPREAD);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

bool rbc_wall_spread_applies(const rbc3d_ctx *c) { return c->wall.nvert > 0 && wall_spread_ok(c); }

int rbc_pme_spread_wall_u(rbc3d_ctx *c, double c1, const double *d_u, const int *d_e2u) {
  if (!rbc_wall_spread_applies(c)) return RBC3D_EINVAL;
  c->flag_sl = fabs(c1) > 1.e-10;
  c->flag_dl = false;
  return c->prec == RBC3D_PREC_SPLIT ? wall_spread<float>(c, c1, d_u, d_e2u)
                                     : wall_spread<double>(c, c1, d_u, d_e2u);
}

int rbc_pme_spread(rbc3d_ctx *c, double c1, double c2, int cells, int walls) {
  return c->prec == RBC3D_PREC_SPLIT ? pme_spread_t<float>(c, c1, c2, cells, walls)
                                     : pme_spread_t<double>(c, c1, c2, cells, walls);
}

// ---------------------------------------------------------------------------
// FFT: batches of lines, Stockham autosort in LDS.
template <typename T>
struct FftPlan {
  int N, nrad, rad[16];
  int sign;
  const typename CxT<T>::t *tw;  // exp(-2 pi i m/N)
  FastDiv fNR[16], fNs[16];  // N / rad[s] and the span Ns before pass s
};

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C>
__device__ __forceinline__ C twid(const C *tw, int m, int sign) {  // m < N
  C t = tw[m];
  if (sign > 0) t.y = -t.y;
  return t;
}
template <typename C>
__device__ __forceinline__ C cpx(decltype(C().x) x, decltype(C().x) y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}

// FFT of TL lines of length N stored line-major in buf (line l at l*N); uses tmp.
// Returns pointer to the buffer holding the result.
// One Stockham pass of radix R (compile-time, so the butterfly lives in registers).
template <int R, typename T>
__device__ __forceinline__ void fft_pass(const typename CxT<T>::t *in, typename CxT<T>::t *out,
                                         int TL, int N, int Ns, const FastDiv &fNR,
                                         const FastDiv &fNs, const FftPlan<T> &pl,
                                         const typename CxT<T>::t *tw) {
  typedef typename CxT<T>::t C;
  const int NR = fNR.d;
  const int tws = NR / Ns;  // twiddle stride N / (Ns R)
  for (int t = threadIdx.x; t < TL * NR; t += blockDim.x) {
    const int l = fd_div32(t, fNR), j = t - l * NR;
    const C *src = in + l * N;
    C *dst = out + l * N;
    const int jq = fd_div32(j, fNs), k = j - jq * Ns;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      C x = src[j + r * NR];
      if (r && k) x = cmul(x, twid(tw, r * k * tws, pl.sign));
      v[r] = x;
    }
    const int base = jq * Ns * R + k;
    if constexpr (R == 2) {
      dst[base] = cpx<C>(v[0].x + v[1].x, v[0].y + v[1].y);
      dst[base + Ns] = cpx<C>(v[0].x - v[1].x, v[0].y - v[1].y);
    } else if constexpr (R == 4) {
      C a0 = cpx<C>(v[0].x + v[2].x, v[0].y + v[2].y);
      C a1 = cpx<C>(v[0].x - v[2].x, v[0].y - v[2].y);
      C b0 = cpx<C>(v[1].x + v[3].x, v[1].y + v[3].y);
      C b1 = cpx<C>(v[1].x - v[3].x, v[1].y - v[3].y);
      // multiply b1 by (sign * i)
      C jb1 = pl.sign < 0 ? cpx<C>(b1.y, -b1.x) : cpx<C>(-b1.y, b1.x);
      dst[base] = cpx<C>(a0.x + b0.x, a0.y + b0.y);
      dst[base + Ns] = cpx<C>(a1.x + jb1.x, a1.y + jb1.y);
      dst[base + 2 * Ns] = cpx<C>(a0.x - b0.x, a0.y - b0.y);
      dst[base + 3 * Ns] = cpx<C>(a1.x - jb1.x, a1.y - jb1.y);
    } else if constexpr (R == 3) {
      // y_q = sum_r v_r w^(qr), w = exp(sign 2 pi i / 3): y_1,2 = m +- i sign (sqrt3/2) d
      typedef decltype(C().x) Re;
      const Re hs = (Re)pl.sign * (Re)0.86602540378443864676;
      const C t = cpx<C>(v[1].x + v[2].x, v[1].y + v[2].y);
      const C d = cpx<C>(v[1].x - v[2].x, v[1].y - v[2].y);
      const C m = cpx<C>(v[0].x - (Re)0.5 * t.x, v[0].y - (Re)0.5 * t.y);
      const Re ux = -hs * d.y, uy = hs * d.x;
      dst[base] = cpx<C>(v[0].x + t.x, v[0].y + t.y);
      dst[base + Ns] = cpx<C>(m.x + ux, m.y + uy);
      dst[base + 2 * Ns] = cpx<C>(m.x - ux, m.y - uy);
    } else if constexpr (R == 5) {
      // w = exp(sign 2 pi i / 5); with a_k = v_k + v_(5-k), b_k = v_k - v_(5-k):
      // y_1,4 = v_0 + c1 a_1 + c2 a_2 +- i sign (s1 b_1 + s2 b_2)
      // y_2,3 = v_0 + c2 a_1 + c1 a_2 +- i sign (s2 b_1 - s1 b_2)
      typedef decltype(C().x) Re;
      const Re c1 = (Re)0.30901699437494742410, c2 = (Re)-0.80901699437494742410;
      const Re sg = (Re)pl.sign;
      const Re s1 = sg * (Re)0.95105651629515357212, s2 = sg * (Re)0.58778525229247312917;
      const C a1 = cpx<C>(v[1].x + v[4].x, v[1].y + v[4].y), b1 = cpx<C>(v[1].x - v[4].x, v[1].y - v[4].y);
      const C a2 = cpx<C>(v[2].x + v[3].x, v[2].y + v[3].y), b2 = cpx<C>(v[2].x - v[3].x, v[2].y - v[3].y);
      const C m1 = cpx<C>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
      const C m2 = cpx<C>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
      // i (s b) = (-(s b).y, (s b).x)
      const Re u1x = -(s1 * b1.y + s2 * b2.y), u1y = s1 * b1.x + s2 * b2.x;
      const Re u2x = -(s2 * b1.y - s1 * b2.y), u2y = s2 * b1.x - s1 * b2.x;
      dst[base] = cpx<C>(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
      dst[base + Ns] = cpx<C>(m1.x + u1x, m1.y + u1y);
      dst[base + 4 * Ns] = cpx<C>(m1.x - u1x, m1.y - u1y);
      dst[base + 2 * Ns] = cpx<C>(m2.x + u2x, m2.y + u2y);
      dst[base + 3 * Ns] = cpx<C>(m2.x - u2x, m2.y - u2y);
    } else {
#pragma unroll
      for (int q = 0; q < R; q++) {
        C sum = v[0];
#pragma unroll
        for (int r = 1; r < R; r++) {
          C w = twid(tw, ((q * r) % R) * NR, pl.sign);
          C p = cmul(v[r], w);
          sum.x += p.x;
          sum.y += p.y;
        }
        dst[base + q * Ns] = sum;
      }
    }
  }
}

// FFT of TL lines of length N stored line-major in buf (line l at l*N); uses tmp.
// Returns pointer to the buffer holding the result.
template <typename T>
__device__ typename CxT<T>::t *fft_lds(typename CxT<T>::t *buf, typename CxT<T>::t *tmp, int TL,
                                       const FftPlan<T> &pl, const typename CxT<T>::t *tw) {
  typedef typename CxT<T>::t C;
  int N = pl.N, Ns = 1;
  C *in = buf, *out = tmp;
  for (int s = 0; s < pl.nrad; s++) {
    const int R = pl.rad[s];
    if (R == 2)
      fft_pass<2, T>(in, out, TL, N, Ns, pl.fNR[s], pl.fNs[s], pl, tw);
    else if (R == 3)
      fft_pass<3, T>(in, out, TL, N, Ns, pl.fNR[s], pl.fNs[s], pl, tw);
    else if (R == 4)
      fft_pass<4, T>(in, out, TL, N, Ns, pl.fNR[s], pl.fNs[s], pl, tw);
    else
      fft_pass<5, T>(in, out, TL, N, Ns, pl.fNR[s], pl.fNs[s], pl, tw);
    __syncthreads();
    C *t2 = in;
    in = out;
    out = t2;
    Ns *= R;
  }
  return in;
}

// complex lines: base(l) = (l % A)*sa + (l / A)*sb, element stride es (32-bit: the
// host checks that every offset fits)
struct LineMap {
  FastDiv A;  // lines per outer index
  int sa, sb, es;
  int nlines;
};

template <typename T>
__global__ void __launch_bounds__(FFT_T) k_fft_c2c(typename CxT<T>::t *data, FftPlan<T> pl,
                                                   LineMap lm, int TL, FastDiv fTL) {
  typedef typename CxT<T>::t C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  C *buf = lds, *tmp = lds + (size_t)TL * pl.N, *stw = tmp + (size_t)TL * pl.N;
  // the axis' twiddle table in LDS (read by every pass; the barrier after the load
  // phase below makes it visible)
  // consecutive blocks of lines share 128-byte lines of the grid: they run on one XCD
  // (workgroups go to the 8 XCDs round-robin; the grid is a multiple of 8 workgroups)
  const int blk = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (blk * TL >= lm.nlines) return;
  for (int m = threadIdx.x; m < pl.N; m += blockDim.x) stw[m] = pl.tw[m];
  const int l0 = blk * TL;
  const int N = pl.N;
  // load: lines fastest (adjacent lines are adjacent in memory when sa = 1)
  for (int t = threadIdx.x; t < TL * N; t += blockDim.x) {
    const int e = fd_div32(t, fTL), l = t - e * TL, L = l0 + l;
    if (L >= lm.nlines) continue;
    const int Lq = fd_div32(L, lm.A), Lr = L - Lq * (int)lm.A.d;
    const int o = Lr * lm.sa + Lq * lm.sb + e * lm.es;
    buf[l * N + e] = data[o];
  }
  __syncthreads();
  C *res = fft_lds<T>(buf, tmp, TL, pl, stw);
  for (int t = threadIdx.x; t < TL * N; t += blockDim.x) {
    const int e = fd_div32(t, fTL), l = t - e * TL, L = l0 + l;
    if (L >= lm.nlines) continue;
    const int Lq = fd_div32(L, lm.A), Lr = L - Lq * (int)lm.A.d;
    const int o = Lr * lm.sa + Lq * lm.sb + e * lm.es;
    data[o] = res[l * N + e];
  }
}

// x r2c: real rows of N1 -> complex rows of H = N1/2+1
template <typename T>
__global__ void __launch_bounds__(FFT_T) k_fft_r2c(const T *in, typename CxT<T>::t *out,
                                                   FftPlan<T> pl, int nrows, int TL, FastDiv fN,
                                                   FastDiv fH) {
  typedef typename CxT<T>::t C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  C *buf = lds, *tmp = lds + (size_t)TL * pl.N, *stw = tmp + (size_t)TL * pl.N;
  // the axis' twiddle table in LDS (read by every pass; the barrier after the load
  // phase below makes it visible)
  for (int m = threadIdx.x; m < pl.N; m += blockDim.x) stw[m] = pl.tw[m];
  const int N = pl.N, H = N / 2 + 1, r0 = blockIdx.x * TL;
  // the block's rows are contiguous in memory: element t of the block is row t / N
  for (int t = threadIdx.x; t < TL * N; t += blockDim.x) {
    if (r0 + (int)fd_div32(t, fN) >= nrows) continue;
    buf[t] = cpx<C>(in[r0 * N + t], 0.);
  }
  __syncthreads();
  C *res = fft_lds<T>(buf, tmp, TL, pl, stw);
  for (int t = threadIdx.x; t < TL * H; t += blockDim.x) {
    const int l = fd_div32(t, fH), e = t - l * H, R = r0 + l;
    if (R >= nrows) continue;
    out[R * H + e] = res[l * N + e];
  }
}

// x c2r: complex rows of H -> real rows of N1 (real part of the zero and Nyquist
// frequencies only, the halfcomplex convention of FFTW's c2r)
template <typename T>
__global__ void __launch_bounds__(FFT_T) k_fft_c2r(const typename CxT<T>::t *in, T *out,
                                                   FftPlan<T> pl, int nrows, int TL, FastDiv fN) {
  typedef typename CxT<T>::t C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  C *buf = lds, *tmp = lds + (size_t)TL * pl.N, *stw = tmp + (size_t)TL * pl.N;
  // the axis' twiddle table in LDS (read by every pass; the barrier after the load
  // phase below makes it visible)
  for (int m = threadIdx.x; m < pl.N; m += blockDim.x) stw[m] = pl.tw[m];
  const int N = pl.N, H = N / 2 + 1, r0 = blockIdx.x * TL;
  for (int t = threadIdx.x; t < TL * N; t += blockDim.x) {
    const int l = fd_div32(t, fN), e = t - l * N, R = r0 + l;
    if (R >= nrows) continue;
    C v;
    if (e == 0) {
      v = cpx<C>(in[R * H].x, 0.);
    } else if (2 * e == N) {
      v = cpx<C>(in[R * H + e].x, 0.);
    } else if (e < H) {
      v = in[R * H + e];
    } else {
      C w = in[R * H + (N - e)];
      v = cpx<C>(w.x, -w.y);
    }
    buf[t] = v;
  }
  __syncthreads();
  C *res = fft_lds<T>(buf, tmp, TL, pl, stw);
  for (int t = threadIdx.x; t < TL * N; t += blockDim.x) {
    if (r0 + (int)fd_div32(t, fN) >= nrows) continue;
    out[r0 * N + t] = res[t].x;
  }
}

template <typename T>
static FftPlan<T> make_plan(rbc3d_ctx *c, int ax, int sign) {
  FftPlan<T> p;
  p.N = c->conf.Nb[ax];
  std::vector<int> r;
  factor_radices(p.N, r);
  p.nrad = (int)r.size();
  for (int i = 0; i < p.nrad; i++) p.rad[i] = r[i];
  p.sign = sign;
  if (sizeof(T) == 8)
    p.tw = (const typename CxT<T>::t *)(c->d_tw.p + 2 * (size_t)c->tw_off[ax]);
  else
    p.tw = (const typename CxT<T>::t *)(c->d_twf.p + 2 * (size_t)c->tw_off[ax]);
  int Ns = 1;
  for (int i = 0; i < p.nrad; i++) {
    p.fNR[i] = make_fastdiv(p.N / p.rad[i]);
    p.fNs[i] = make_fastdiv(Ns);
    Ns *= p.rad[i];
  }
  return p;
}

// ~1024 complex elements per workgroup (2048 in FP32): 32 KiB of LDS for the two
// buffers, so five workgroups fit a CU.  (A 48 KiB budget -- longer contiguous runs
// in the strided passes, fewer workgroups -- measured slower on the tube's grid.)
template <typename T>
static int lines_per_block(int N) { return std::max(1, (int)(8192 / sizeof(T)) / N); }

template <typename T>
static int fft_c2c(rbc3d_ctx *c, typename CxT<T>::t *data, int ax, int sign, int nf) {
  int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], H = N1 / 2 + 1;
  FftPlan<T> pl = make_plan<T>(c, ax, sign);
  LineMap lm;
  if ((long long)nf * H * N2 * N3 >= (1LL << 31)) {
    rbc_set_error("PME grid too large for 32-bit FFT offsets");
    return RBC3D_EINVAL;
  }
  if (ax == 1) {  // lines over (kx, kz, f), stride H
    lm.A = make_fastdiv(H);
    lm.sa = 1;
    lm.sb = H * N2;
    lm.es = H;
    lm.nlines = H * N3 * nf;
  } else {  // ax == 2: lines over (kx, ky, f), stride H*N2
    lm.A = make_fastdiv(H * N2);
    lm.sa = 1;
    lm.sb = H * N2 * N3;
    lm.es = H * N2;
    lm.nlines = H * N2 * nf;
  }
  int TL = lines_per_block<T>(pl.N);
  size_t lds = sizeof(typename CxT<T>::t) * (2 * (size_t)TL * pl.N + pl.N);
  const unsigned nblk = (unsigned)(((lm.nlines + TL - 1) / TL + 7) / 8 * 8);
  k_fft_c2c<T><<<nblk, FFT_T, lds, c->stream>>>(data, pl, lm, TL, make_fastdiv(TL));
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// x-y planes in one pass.  A workgroup takes one z plane of one field: the real plane
// is loaded once, transformed along x (two real rows as one complex row) and along y
// in LDS, and stored as the half-complex plane; the inverse does the reverse.  The
// grid is read and written once per direction instead of twice.  Each Stockham pass
// is register-staged: every thread loads the inputs of its butterflies, the
// workgroup synchronizes, every thread writes its outputs into the same buffer, so
// one plane buffer suffices (N2/2 rows of stride N1 + 2: a row pair's unpacked
// half-spectra fit in place).

// radix-R butterfly of v (inputs already twiddled) into natural-order outputs, the
// same arithmetic as the passes above
template <int R, typename C>
__device__ __forceinline__ void bfly(C (&v)[R], int sign, const C *tw, int NR) {
  typedef decltype(C().x) Re;
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = cpx<C>(a.x + b.x, a.y + b.y);
    v[1] = cpx<C>(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    C a0 = cpx<C>(v[0].x + v[2].x, v[0].y + v[2].y);
    C a1 = cpx<C>(v[0].x - v[2].x, v[0].y - v[2].y);
    C b0 = cpx<C>(v[1].x + v[3].x, v[1].y + v[3].y);
    C b1 = cpx<C>(v[1].x - v[3].x, v[1].y - v[3].y);
    C jb1 = sign < 0 ? cpx<C>(b1.y, -b1.x) : cpx<C>(-b1.y, b1.x);
    v[0] = cpx<C>(a0.x + b0.x, a0.y + b0.y);
    v[1] = cpx<C>(a1.x + jb1.x, a1.y + jb1.y);
    v[2] = cpx<C>(a0.x - b0.x, a0.y - b0.y);
    v[3] = cpx<C>(a1.x - jb1.x, a1.y - jb1.y);
  } else if constexpr (R == 3) {
    const Re hs = (Re)sign * (Re)0.86602540378443864676;
    const C t = cpx<C>(v[1].x + v[2].x, v[1].y + v[2].y);
    const C d = cpx<C>(v[1].x - v[2].x, v[1].y - v[2].y);
    const C m = cpx<C>(v[0].x - (Re)0.5 * t.x, v[0].y - (Re)0.5 * t.y);
    const Re ux = -hs * d.y, uy = hs * d.x;
    const C y0 = cpx<C>(v[0].x + t.x, v[0].y + t.y);
    v[1] = cpx<C>(m.x + ux, m.y + uy);
    v[2] = cpx<C>(m.x - ux, m.y - uy);
    v[0] = y0;
  } else if constexpr (R == 5) {
    const Re c1 = (Re)0.30901699437494742410, c2 = (Re)-0.80901699437494742410;
    const Re sg = (Re)sign;
    const Re s1 = sg * (Re)0.95105651629515357212, s2 = sg * (Re)0.58778525229247312917;
    const C a1 = cpx<C>(v[1].x + v[4].x, v[1].y + v[4].y), b1 = cpx<C>(v[1].x - v[4].x, v[1].y - v[4].y);
    const C a2 = cpx<C>(v[2].x + v[3].x, v[2].y + v[3].y), b2 = cpx<C>(v[2].x - v[3].x, v[2].y - v[3].y);
    const C m1 = cpx<C>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const C m2 = cpx<C>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const Re u1x = -(s1 * b1.y + s2 * b2.y), u1y = s1 * b1.x + s2 * b2.x;
    const Re u2x = -(s2 * b1.y - s1 * b2.y), u2y = s2 * b1.x - s1 * b2.x;
    const C y0 = cpx<C>(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = cpx<C>(m1.x + u1x, m1.y + u1y);
    v[4] = cpx<C>(m1.x - u1x, m1.y - u1y);
    v[2] = cpx<C>(m2.x + u2x, m2.y + u2y);
    v[3] = cpx<C>(m2.x - u2x, m2.y - u2y);
    v[0] = y0;
  }
}

// one register-staged Stockham pass over L lines of length N in buf (element e of
// line l at l ls + e es).  LFAST: butterfly b is (line b % L, index b / L), consecutive
// threads on consecutive lines (lines adjacent in LDS); else (line b / NR, index
// b % NR), consecutive threads along one line (elements adjacent in LDS).
template <int R, int MAXE, bool LFAST, typename T>
__device__ __forceinline__ void rs_pass(typename CxT<T>::t *buf, int L, const FastDiv &fL, int N,
                                        int ls, int es, int Ns, const FastDiv &fNs,
                                        const FastDiv &fNR, int sign,
                                        const typename CxT<T>::t *tw) {
  typedef typename CxT<T>::t C;
  constexpr int MB = (MAXE + R - 1) / R;
  const int NR = N / R, nb = L * NR, tws = NR / Ns;
  C v[MB][R];
  int o[MB];
#pragma unroll
  for (int i = 0; i < MB; i++) {
    const int b = threadIdx.x + i * blockDim.x;
    o[i] = -1;
    if (b < nb) {
      int j, line;
      if constexpr (LFAST) {
        j = fd_div32(b, fL);
        line = b - j * L;
      } else {
        line = fd_div32(b, fNR);
        j = b - line * NR;
      }
      const int jq = fd_div32(j, fNs), k = j - jq * Ns;
      const C *src = buf + line * ls;
#pragma unroll
      for (int r = 0; r < R; r++) {
        C x = src[(j + r * NR) * es];
        if (r && k) x = cmul(x, twid(tw, r * k * tws, sign));
        v[i][r] = x;
      }
      o[i] = line * ls + (jq * Ns * R + k) * es;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MB; i++) {
    if (o[i] >= 0) {
      bfly<R>(v[i], sign, tw, NR);
#pragma unroll
      for (int q = 0; q < R; q++) buf[o[i] + q * Ns * es] = v[i][q];
    }
  }
  __syncthreads();
}

template <int MAXE, bool LFAST, typename T>
__device__ __forceinline__ void rs_fft(typename CxT<T>::t *buf, int L, const FastDiv &fL, int ls,
                                       int es, const FftPlan<T> &pl, const typename CxT<T>::t *tw) {
  int Ns = 1;
  for (int s = 0; s < pl.nrad; s++) {
    const int R = pl.rad[s];
    if (R == 2)
      rs_pass<2, MAXE, LFAST, T>(buf, L, fL, pl.N, ls, es, Ns, pl.fNs[s], pl.fNR[s], pl.sign, tw);
    else if (R == 3)
      rs_pass<3, MAXE, LFAST, T>(buf, L, fL, pl.N, ls, es, Ns, pl.fNs[s], pl.fNR[s], pl.sign, tw);
    else if (R == 4)
      rs_pass<4, MAXE, LFAST, T>(buf, L, fL, pl.N, ls, es, Ns, pl.fNs[s], pl.fNR[s], pl.sign, tw);
    else
      rs_pass<5, MAXE, LFAST, T>(buf, L, fL, pl.N, ls, es, Ns, pl.fNs[s], pl.fNR[s], pl.sign, tw);
    Ns *= R;
  }
}

#define XY_T 512    // threads per plane workgroup
#define XY_MAXE 12  // complex elements a thread holds per pass
#define XY_MAXP 10  // row-pair items a thread holds in the (un)packing steps
#ifndef XY_XLF
#define XY_XLF false  // x passes: consecutive threads along a row
#endif
struct XyArgs {
  int N1, N2, N3, H, S, nf;
  size_t ng, nc;
  FastDiv fL1, fL2, fN1, fH;  // N2 / 2 (x lines), H (y lines), N1, H
};

// forward: real fields f0.. of the spread grid -> half-complex planes (x then y)
template <typename T>
__global__ void __launch_bounds__(XY_T) k_fft_xy_fwd(XyArgs a, const T *__restrict__ in,
                                                     typename CxT<T>::t *out, FftPlan<T> px,
                                                     FftPlan<T> py) {
  typedef typename CxT<T>::t C;
  typedef decltype(C().x) Re;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);
  const int N1 = a.N1, N2 = a.N2, H = a.H, S = a.S, M = N2 / 2;
  C *twx = buf + (size_t)M * S, *twy = twx + N1;
  const int z = blockIdx.x % a.N3, f = blockIdx.x / a.N3;
  for (int m = threadIdx.x; m < N1; m += blockDim.x) twx[m] = px.tw[m];
  for (int m = threadIdx.x; m < N2; m += blockDim.x) twy[m] = py.tw[m];
  // rows 2m and 2m + 1 as the real and imaginary parts of complex row m
  const T *pin = in + a.ng * f + (size_t)N1 * N2 * z;
  for (int t = threadIdx.x; t < M * N1; t += blockDim.x) {
    const int m = fd_div32(t, a.fN1), n = t - m * N1;
    buf[m * S + n] = cpx<C>(pin[2 * m * N1 + n], pin[(2 * m + 1) * N1 + n]);
  }
  __syncthreads();
  rs_fft<XY_MAXE, XY_XLF, T>(buf, M, a.fL1, S, 1, px, twx);
  // unpack: A[k] = (Z[k] + conj Z[N1-k]) / 2 into row 2m, B[k] = (Z[k] - conj Z[N1-k]) / 2i
  // into row 2m + 1, k = 0..H-1 (rows of H complex: row y at y H)
  {
    C za[XY_MAXP], zb[XY_MAXP];
    int it[XY_MAXP];
#pragma unroll
    for (int i = 0; i < XY_MAXP; i++) {
      const int t = threadIdx.x + i * blockDim.x;
      it[i] = -1;
      if (t < M * H) {
        const int m = fd_div32(t, a.fH), k = t - m * H;
        za[i] = buf[m * S + k];
        zb[i] = buf[m * S + (k == 0 ? 0 : N1 - k)];
        it[i] = t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < XY_MAXP; i++) {
      if (it[i] >= 0) {
        const int m = fd_div32(it[i], a.fH), k = it[i] - m * H;
        const C p = za[i], q = zb[i];
        buf[m * S + k] = cpx<C>((Re)0.5 * (p.x + q.x), (Re)0.5 * (p.y - q.y));
        buf[m * S + H + k] = cpx<C>((Re)0.5 * (p.y + q.y), (Re)-0.5 * (p.x - q.x));
      }
    }
    __syncthreads();
  }
  rs_fft<XY_MAXE, true, T>(buf, H, a.fL2, 1, H, py, twy);
  C *pout = out + a.nc * f + (size_t)H * N2 * z;
  for (int t = threadIdx.x; t < H * N2; t += blockDim.x) pout[t] = buf[t];
}

// inverse: half-complex planes of fields 12..14 -> real velocity planes (y then x)
template <typename T>
__global__ void __launch_bounds__(XY_T) k_fft_xy_inv(XyArgs a, const typename CxT<T>::t *in,
                                                     T *__restrict__ out, FftPlan<T> px,
                                                     FftPlan<T> py) {
  typedef typename CxT<T>::t C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);
  const int N1 = a.N1, N2 = a.N2, H = a.H, S = a.S, M = N2 / 2;
  C *twx = buf + (size_t)M * S, *twy = twx + N1;
  const int z = blockIdx.x % a.N3, f = blockIdx.x / a.N3;
  for (int m = threadIdx.x; m < N1; m += blockDim.x) twx[m] = px.tw[m];
  for (int m = threadIdx.x; m < N2; m += blockDim.x) twy[m] = py.tw[m];
  const C *pin = in + a.nc * f + (size_t)H * N2 * z;
  for (int t = threadIdx.x; t < H * N2; t += blockDim.x) buf[t] = pin[t];
  __syncthreads();
  rs_fft<XY_MAXE, true, T>(buf, H, a.fL2, 1, H, py, twy);
  // pack: Z[k] = A[k] + i B[k] over k = 0..N1-1 with A, B extended by Hermitian
  // symmetry (the imaginary parts of the zero and Nyquist frequencies dropped: the
  // halfcomplex convention of FFTW's c2r)
  {
    C za[XY_MAXP], zb[XY_MAXP];
    int it[XY_MAXP];
#pragma unroll
    for (int i = 0; i < XY_MAXP; i++) {
      const int t = threadIdx.x + i * blockDim.x;
      it[i] = -1;
      if (t < M * N1) {
        const int m = fd_div32(t, a.fN1), k = t - m * N1;
        const int kk = k < H ? k : N1 - k;
        C A = buf[m * S + kk], B = buf[m * S + H + kk];
        if (kk == 0 || 2 * kk == N1) {
          A.y = 0;
          B.y = 0;
        }
        if (k >= H) {
          A.y = -A.y;
          B.y = -B.y;
        }
        za[i] = A;
        zb[i] = B;
        it[i] = t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < XY_MAXP; i++) {
      if (it[i] >= 0) {
        const int m = fd_div32(it[i], a.fN1), k = it[i] - m * N1;
        buf[m * S + k] = cpx<C>(za[i].x - zb[i].y, za[i].y + zb[i].x);
      }
    }
    __syncthreads();
  }
  rs_fft<XY_MAXE, XY_XLF, T>(buf, M, a.fL1, S, 1, px, twx);
  T *pout = out + a.ng * f + (size_t)N1 * N2 * z;
  for (int t = threadIdx.x; t < M * N1; t += blockDim.x) {
    const int m = fd_div32(t, a.fN1), n = t - m * N1;
    const C v = buf[m * S + n];
    pout[2 * m * N1 + n] = v.x;
    pout[(2 * m + 1) * N1 + n] = v.y;
  }
}

// whether the plane kernels apply: even N1 and N2, every pass and (un)packing step
// within the per-thread register budget, the plane within RBC_XY_LDS bytes of LDS
template <typename T>
static bool xy_plane_ok(rbc3d_ctx *c, size_t *lds) {
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], H = N1 / 2 + 1, M = N2 / 2;
  if (N1 % 2 || N2 % 2) return false;
  *lds = sizeof(typename CxT<T>::t) * ((size_t)M * (2 * H) + N1 + N2);
  if (*lds > std::min<size_t>(RBC_XY_LDS, (size_t)c->lds_max)) return false;
  auto fits = [](int L, int N, int R) { return (L * (N / R) + XY_T - 1) / XY_T <= (XY_MAXE + R - 1) / R; };
  std::vector<int> r;
  factor_radices(N1, r);
  for (int R : r)
    if (!fits(M, N1, R)) return false;
  factor_radices(N2, r);
  for (int R : r)
    if (!fits(H, N2, R)) return false;
  return (M * H + XY_T - 1) / XY_T <= XY_MAXP && (M * N1 + XY_T - 1) / XY_T <= XY_MAXP;
}

template <typename T>
static XyArgs xy_args(rbc3d_ctx *c, int nf) {
  XyArgs a;
  a.N1 = c->conf.Nb[0];
  a.N2 = c->conf.Nb[1];
  a.N3 = c->conf.Nb[2];
  a.H = a.N1 / 2 + 1;
  a.S = 2 * a.H;
  a.nf = nf;
  a.ng = (size_t)a.N1 * a.N2 * a.N3;
  a.nc = (size_t)a.H * a.N2 * a.N3;
  a.fL1 = make_fastdiv(a.N2 / 2);
  a.fL2 = make_fastdiv(a.H);
  a.fN1 = make_fastdiv(a.N1);
  a.fH = make_fastdiv(a.H);
  return a;
}

// ---------------------------------------------------------------------------
// Green's function multiply (PME_Transform :159-210), output fields 12..14
template <typename T>
struct GreenArgs {
  int N1, N2, N3, H;
  double iLb[3], alpha, vol;
  int flag_sl, flag_dl;
  const double *bb;         // bx(H), by(N2), bz(N3)
  typename CxT<T>::t *C;    // 15 complex fields
};

// Green's function at one wave vector (i, j, k): vc = G(q) applied to the spread fields
// (F(f) gives field f's value there, fields 0..2 SL, 3..11 DL), times the B-factor;
// zero at q = 0.  The k-dependent factors are formed in FP64 for either grid scalar.
template <typename T, typename F>
__device__ __forceinline__ void green_at(const GreenArgs<T> &a, int i, int j, int k, F field,
                                         double2 vc[3]) {
  vc[0] = vc[1] = vc[2] = make_double2(0., 0.);
  if (i == 0 && j == 0 && k == 0) return;
  double q[3];
  q[0] = i * a.iLb[0];
  q[1] = (j < a.N2 / 2) ? j * a.iLb[1] : (j - a.N2) * a.iLb[1];
  double q3 = (k < a.N3 / 2) ? k * a.iLb[2] : (k - a.N3) * a.iLb[2];
  q[2] = -q3;
  double spa = sqrt(RBC_PI * a.alpha), qt[3];
  for (int d = 0; d < 3; d++) qt[d] = spa * q[d];
  double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  double q2t = RBC_PI * a.alpha * q2, e = exp(-q2t);
  double phi0 = e / q2t, phi1 = (e + phi0) / q2t;
  if (a.flag_sl) {
    double2 f[3];
    for (int d = 0; d < 3; d++) f[d] = field(d);
    double2 qf = make_double2(qt[0] * f[0].x + qt[1] * f[1].x + qt[2] * f[2].x,
                              qt[0] * f[0].y + qt[1] * f[1].y + qt[2] * f[2].y);
    double s = 2 * a.alpha / a.vol * phi1;
    for (int d = 0; d < 3; d++) {
      vc[d].x += s * (q2t * f[d].x - qt[d] * qf.x);
      vc[d].y += s * (q2t * f[d].y - qt[d] * qf.y);
    }
  }
  if (a.flag_dl) {
    double2 Tm[3][3];
    for (int ii = 0; ii < 3; ii++)
      for (int jj = 0; jj < 3; jj++) Tm[ii][jj] = field(3 + ii + 3 * jj);
    double2 tr = make_double2(Tm[0][0].x + Tm[1][1].x + Tm[2][2].x, Tm[0][0].y + Tm[1][1].y + Tm[2][2].y);
    double2 qT[3], Tq[3];
    for (int jj = 0; jj < 3; jj++)
      qT[jj] = make_double2(q[0] * Tm[0][jj].x + q[1] * Tm[1][jj].x + q[2] * Tm[2][jj].x,
                            q[0] * Tm[0][jj].y + q[1] * Tm[1][jj].y + q[2] * Tm[2][jj].y);
    for (int ii = 0; ii < 3; ii++)
      Tq[ii] = make_double2(Tm[ii][0].x * q[0] + Tm[ii][1].x * q[1] + Tm[ii][2].x * q[2],
                            Tm[ii][0].y * q[0] + Tm[ii][1].y * q[1] + Tm[ii][2].y * q[2]);
    double2 qTq = make_double2(q[0] * Tq[0].x + q[1] * Tq[1].x + q[2] * Tq[2].x,
                               q[0] * Tq[0].y + q[1] * Tq[1].y + q[2] * Tq[2].y);
    double s0 = 4 * RBC_PI * a.alpha / a.vol * phi0;
    double s1 = 8 * RBC_PI * RBC_PI * a.alpha * a.alpha / a.vol * phi1;
    for (int d = 0; d < 3; d++) {
      // u = q tr + qT + Tq;  dv = i s0 u - i s1 qTq q;  vc -= dv
      double ux = q[d] * tr.x + qT[d].x + Tq[d].x, uy = q[d] * tr.y + qT[d].y + Tq[d].y;
      double wx = qTq.x * q[d], wy = qTq.y * q[d];
      // i*s0*u = (-s0*uy, s0*ux); i*s1*w = (-s1*wy, s1*wx)
      double dvx = -s0 * uy - (-s1 * wy), dvy = s0 * ux - s1 * wx;
      vc[d].x -= dvx;
      vc[d].y -= dvy;
    }
  }
  double b = ((1.0 * a.bb[i]) * a.bb[a.H + j]) * a.bb[a.H + a.N2 + k];
  for (int d = 0; d < 3; d++) {
    vc[d].x *= b;
    vc[d].y *= b;
  }
}

template <typename T>
__global__ void k_green(GreenArgs<T> a) {
  typedef typename CxT<T>::t CT;
  size_t nc = (size_t)a.H * a.N2 * a.N3;
  size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (o >= nc) return;
  int i = (int)(o % a.H), j = (int)((o / a.H) % a.N2), k = (int)(o / ((size_t)a.H * a.N2));
  double2 vc[3];
  green_at(a, i, j, k,
           [&](int f) {
             const CT v = a.C[o + nc * f];
             return make_double2(v.x, v.y);
           },
           vc);
  for (int d = 0; d < 3; d++) {
    CT v;
    v.x = (decltype(v.x))vc[d].x;
    v.y = (decltype(v.y))vc[d].y;
    a.C[o + nc * (12 + d)] = v;
  }
}

// The z-stage of PME_Transform in one pass: for TL adjacent kx at one ky, the z-lines of
// the nf spread fields (f0..f0+nf-1) are loaded, transformed forward (sign +1), taken
// through Green's function and the B-factor into the three velocity fields, transformed
// back (sign -1) and stored as fields 12..14: the grid is read and written once instead
// of twice, with Green's multiply in between.  Per line the same Stockham passes as the
// separate kernels, so the results are theirs bit for bit.
//
// Workgroups are dispatched to the 8 XCDs round-robin; the logical block order puts
// consecutive column blocks (which share 128-byte lines of the grid) on one XCD's L2:
// workgroup b takes block (b % 8) * (gridDim.x / 8) + b / 8 (the grid is a multiple of 8
// workgroups; those past the last block return at once).
template <typename T>
__global__ void __launch_bounds__(FFT_T) k_fft_zgreen(GreenArgs<T> g, FftPlan<T> pf,
                                                      FftPlan<T> pi, int f0, int nf, int TL) {
  typedef typename CxT<T>::t C;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *lds = reinterpret_cast<C *>(lds_raw);
  const int N3 = g.N3, H = g.H;
  const size_t nc = (size_t)H * g.N2 * N3;
  const int ncb = (H + TL - 1) / TL;  // column blocks per ky
  const int blk = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (blk >= ncb * g.N2) return;
  C *buf = lds, *tmp = lds + (size_t)nf * TL * N3, *stw = tmp + (size_t)nf * TL * N3;
  for (int m = threadIdx.x; m < N3; m += blockDim.x) stw[m] = pf.tw[m];
  const int j = blk / ncb, i0 = (blk % ncb) * TL;
  const int tl = min(TL, H - i0);
  // load: line (f, l) at (f * TL + l) * N3, kx fastest in memory
  for (int t = threadIdx.x; t < nf * TL * N3; t += blockDim.x) {
    const int l = t % TL, r = t / TL, z = r % N3, f = r / N3;
    C v = cpx<C>(0., 0.);
    if (l < tl) v = g.C[(i0 + l) + (size_t)H * (j + (size_t)g.N2 * z) + nc * (f0 + f)];
    buf[(f * TL + l) * N3 + z] = v;
  }
  __syncthreads();
  C *res = fft_lds<T>(buf, tmp, nf * TL, pf, stw);
  C *out = res == buf ? tmp : buf;
  for (int t = threadIdx.x; t < TL * N3; t += blockDim.x) {
    const int l = t / N3, z = t % N3;
    double2 vc[3] = {make_double2(0., 0.), make_double2(0., 0.), make_double2(0., 0.)};
    if (l < tl)
      green_at(g, i0 + l, j, z,
               [&](int f) {
                 const C v = res[((f - f0) * TL + l) * N3 + z];
                 return make_double2(v.x, v.y);
               },
               vc);
    for (int d = 0; d < 3; d++) {
      C v;
      v.x = (decltype(v.x))vc[d].x;
      v.y = (decltype(v.y))vc[d].y;
      out[(d * TL + l) * N3 + z] = v;
    }
  }
  __syncthreads();
  // the inverse z transform's twiddles (the same table, conjugated by the plan's sign)
  C *res2 = fft_lds<T>(out, res, 3 * TL, pi, stw);
  for (int t = threadIdx.x; t < 3 * TL * N3; t += blockDim.x) {
    const int l = t % TL, r = t / TL, z = r % N3, d = r / N3;
    if (l < tl) g.C[(i0 + l) + (size_t)H * (j + (size_t)g.N2 * z) + nc * (12 + d)] = res2[(d * TL + l) * N3 + z];
  }
}

// ---------------------------------------------------------------------------
// The transform at compile-time grid sizes (the tube's 96 x 96 x 180 and its reduced
// 32 x 32 x 60): x-y plane kernels and the z + Green kernel on the two-pass line FFTs
// of fft_fixed.h.  Same data layout and results (to rounding) as the general kernels.
#ifndef FX_T
#define FX_T 512  // threads of a plane workgroup
#endif
#ifndef FZ_T
#define FZ_T 256  // threads of a z + Green workgroup
#endif
#ifndef FZ_TL180
#define FZ_TL180 4  // columns per workgroup at N3 = 180 (H = 49; measured best of 2..7)
#endif

template <int N1, int N2>
struct PlaneGeom {
  static constexpr int H = N1 / 2 + 1, M = N2 / 2;
  static constexpr int S = N1 + 1;  // row-pair stride of the x stage (odd: LDS banks)
  static constexpr int P = H;       // row stride of the y stage
  static constexpr int BUF = (M * S > N2 * P ? M * S : N2 * P);
  static constexpr int LDS_ELEMS = BUF + N1 + N2;
};

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(FX_T) k_xy_fwd_fx(const T *__restrict__ in, typename CxT<T>::t *out,
                                                    int N3, size_t ng, size_t nc,
                                                    const typename CxT<T>::t *gtwx,
                                                    const typename CxT<T>::t *gtwy) {
  typedef typename CxT<T>::t C;
  typedef decltype(C().x) Re;
  typedef PlaneGeom<N1, N2> G;
  constexpr int H = G::H, M = G::M, S = G::S, P = G::P;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw), *twx = buf + G::BUF, *twy = twx + N1;
  const int z = blockIdx.x % N3, f = blockIdx.x / N3;
  for (int m = threadIdx.x; m < N1; m += FX_T) twx[m] = gtwx[m];
  for (int m = threadIdx.x; m < N2; m += FX_T) twy[m] = gtwy[m];
  const T *pin = in + ng * f + (size_t)N1 * N2 * z;
  for (int t = threadIdx.x; t < M * N1; t += FX_T) {
    const int m = t / N1, n = t % N1;
    buf[m * S + n] = cpx<C>(pin[2 * m * N1 + n], pin[(2 * m + 1) * N1 + n]);
  }
  __syncthreads();
  fft_fixed<N1, M, S, 1, true, -1, FX_T>(buf, twx);
  // unpack row pair m into rows 2m and 2m + 1 (row y at y P): A[k] = (Z[k] + conj
  // Z[N1-k]) / 2, B[k] = (Z[k] - conj Z[N1-k]) / 2i, k < H
  {
    constexpr int PER = (M * H + FX_T - 1) / FX_T;
    C za[PER], zb[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = threadIdx.x + i * FX_T;
      if (t < M * H) {
        const int m = t / H, k = t % H;
        za[i] = buf[m * S + k];
        zb[i] = buf[m * S + (k == 0 ? 0 : N1 - k)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = threadIdx.x + i * FX_T;
      if (t < M * H) {
        const int m = t / H, k = t % H;
        const C p = za[i], q = zb[i];
        buf[2 * m * P + k] = cpx<C>((Re)0.5 * (p.x + q.x), (Re)0.5 * (p.y - q.y));
        buf[(2 * m + 1) * P + k] = cpx<C>((Re)0.5 * (p.y + q.y), (Re)-0.5 * (p.x - q.x));
      }
    }
    __syncthreads();
  }
  fft_fixed<N2, H, 1, P, true, -1, FX_T>(buf, twy);
  C *pout = out + nc * f + (size_t)H * N2 * z;
  for (int t = threadIdx.x; t < H * N2; t += FX_T) pout[t] = buf[(t / H) * P + t % H];
}

template <typename T, int N1, int N2>
__global__ void __launch_bounds__(FX_T) k_xy_inv_fx(const typename CxT<T>::t *in, T *__restrict__ out,
                                                    int N3, size_t ng, size_t nc,
                                                    const typename CxT<T>::t *gtwx,
                                                    const typename CxT<T>::t *gtwy) {
  typedef typename CxT<T>::t C;
  typedef PlaneGeom<N1, N2> G;
  constexpr int H = G::H, M = G::M, S = G::S, P = G::P;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw), *twx = buf + G::BUF, *twy = twx + N1;
  const int z = blockIdx.x % N3, f = blockIdx.x / N3;
  for (int m = threadIdx.x; m < N1; m += FX_T) twx[m] = gtwx[m];
  for (int m = threadIdx.x; m < N2; m += FX_T) twy[m] = gtwy[m];
  const C *pin = in + nc * f + (size_t)H * N2 * z;
  for (int t = threadIdx.x; t < H * N2; t += FX_T) buf[(t / H) * P + t % H] = pin[t];
  __syncthreads();
  fft_fixed<N2, H, 1, P, true, +1, FX_T>(buf, twy);
  // pack: Z[k] = A[k] + i B[k], k < N1, A and B (rows 2m, 2m + 1) extended by Hermitian
  // symmetry, the imaginary parts of the zero and Nyquist frequencies dropped (FFTW's
  // c2r convention)
  {
    constexpr int PER = (M * N1 + FX_T - 1) / FX_T;
    C za[PER], zb[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = threadIdx.x + i * FX_T;
      if (t < M * N1) {
        const int m = t / N1, k = t % N1;
        const int kk = k < H ? k : N1 - k;
        C A = buf[2 * m * P + kk], B = buf[(2 * m + 1) * P + kk];
        if (kk == 0 || 2 * kk == N1) {
          A.y = 0;
          B.y = 0;
        }
        if (k >= H) {
          A.y = -A.y;
          B.y = -B.y;
        }
        za[i] = A;
        zb[i] = B;
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int t = threadIdx.x + i * FX_T;
      if (t < M * N1) {
        const int m = t / N1, k = t % N1;
        buf[m * S + k] = cpx<C>(za[i].x - zb[i].y, za[i].y + zb[i].x);
      }
    }
    __syncthreads();
  }
  fft_fixed<N1, M, S, 1, true, +1, FX_T>(buf, twx);
  T *pout = out + ng * f + (size_t)N1 * N2 * z;
  for (int t = threadIdx.x; t < M * N1; t += FX_T) {
    const int m = t / N1, n = t % N1;
    const C v = buf[m * S + n];
    pout[2 * m * N1 + n] = v.x;
    pout[(2 * m + 1) * N1 + n] = v.y;
  }
}

// z stage of the wall-only transform (fields 0..2 in, 12..14 out): TL adjacent kx at one
// ky per workgroup, forward z (sign +1), Green's function, inverse z, as k_fft_zgreen
template <typename T, int N3, int TL>
__global__ void __launch_bounds__(FZ_T) k_zg_fx(GreenArgs<T> g, const typename CxT<T>::t *gtw) {
  typedef typename CxT<T>::t C;
  constexpr int LS = N3 + 1;  // line stride (odd: LDS banks)
  extern __shared__ __align__(16) unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw), *tw = buf + 3 * TL * LS;
  const int H = g.H;
  const size_t nc = (size_t)H * g.N2 * N3;
  const int ncb = (H + TL - 1) / TL;
  const int blk = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  if (blk >= ncb * g.N2) return;
  const int j = blk / ncb, i0 = (blk % ncb) * TL, tl = min(TL, H - i0);
  for (int m = threadIdx.x; m < N3; m += FZ_T) tw[m] = gtw[m];
  for (int t = threadIdx.x; t < 3 * TL * N3; t += FZ_T) {
    const int l = t % TL, r = t / TL, z = r % N3, f = r / N3;
    C v = cpx<C>(0., 0.);
    if (l < tl) v = g.C[(i0 + l) + (size_t)H * (j + (size_t)g.N2 * z) + nc * f];
    buf[(f * TL + l) * LS + z] = v;
  }
  __syncthreads();
  fft_fixed<N3, 3 * TL, LS, 1, true, +1, FZ_T>(buf, tw);
  // Green's function point by point: each point's three outputs replace its three inputs
  for (int t = threadIdx.x; t < TL * N3; t += FZ_T) {
    const int l = t % TL, z = t / TL;
    double2 vc[3] = {make_double2(0., 0.), make_double2(0., 0.), make_double2(0., 0.)};
    if (l < tl)
      green_at(g, i0 + l, j, z,
               [&](int f) {
                 const C v = buf[(f * TL + l) * LS + z];
                 return make_double2(v.x, v.y);
               },
               vc);
    for (int d = 0; d < 3; d++) {
      C v;
      v.x = (decltype(v.x))vc[d].x;
      v.y = (decltype(v.y))vc[d].y;
      buf[(d * TL + l) * LS + z] = v;
    }
  }
  __syncthreads();
  fft_fixed<N3, 3 * TL, LS, 1, true, -1, FZ_T>(buf, tw);
  for (int t = threadIdx.x; t < 3 * TL * N3; t += FZ_T) {
    const int l = t % TL, r = t / TL, z = r % N3, d = r / N3;
    if (l < tl) g.C[(i0 + l) + (size_t)H * (j + (size_t)g.N2 * z) + nc * (12 + d)] = buf[(d * TL + l) * LS + z];
  }
}

template <typename T, int N1, int N2>
static size_t fx_plane_lds() {
  return sizeof(typename CxT<T>::t) * PlaneGeom<N1, N2>::LDS_ELEMS;
}

// the compile-time grids: x-y planes for any fields, the z stage with Green for the
// wall-only case (three single-layer fields); returns false where no instance exists
template <typename T>
static bool fx_xy(rbc3d_ctx *c, bool forward, int nf, const void *in, void *out) {
  typedef typename CxT<T>::t CT;
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2];
  const size_t ng = (size_t)N1 * N2 * N3, nc = (size_t)(N1 / 2 + 1) * N2 * N3;
  const CT *twx = make_plan<T>(c, 0, -1).tw, *twy = make_plan<T>(c, 1, -1).tw;
  hipStream_t st = c->stream;
  const unsigned nb = (unsigned)(nf * N3);
#define RBC_FX_XY(A, B)                                                                          \
  if (N1 == A && N2 == B && fx_plane_lds<T, A, B>() <= (size_t)c->lds_max) {                    \
    if (forward)                                                                                 \
      k_xy_fwd_fx<T, A, B><<<nb, FX_T, fx_plane_lds<T, A, B>(), st>>>((const T *)in, (CT *)out, \
                                                                      N3, ng, nc, twx, twy);     \
    else                                                                                         \
      k_xy_inv_fx<T, A, B><<<nb, FX_T, fx_plane_lds<T, A, B>(), st>>>((const CT *)in, (T *)out, \
                                                                      N3, ng, nc, twx, twy);     \
    return true;                                                                                 \
  }
  RBC_FX_XY(96, 96)
  RBC_FX_XY(32, 32)
#undef RBC_FX_XY
  return false;
}

template <typename T>
static bool fx_zgreen(rbc3d_ctx *c, const GreenArgs<T> *g) {  // g null: only whether it applies
  typedef typename CxT<T>::t CT;
  const int N3 = c->conf.Nb[2], H = c->conf.Nb[0] / 2 + 1, N2 = c->conf.Nb[1];
  const CT *tw = make_plan<T>(c, 2, -1).tw;
#define RBC_FX_Z(NZ, HH, TLC)                                                                    \
  if (N3 == NZ && H == HH) {                                                                     \
    const size_t lds = sizeof(CT) * (3 * (size_t)TLC * (NZ + 1) + NZ);                           \
    if (lds > (size_t)c->lds_max) return false;                                                  \
    const unsigned nblk = (unsigned)((((HH + TLC - 1) / TLC) * N2 + 7) / 8 * 8);                \
    if (g) k_zg_fx<T, NZ, TLC><<<nblk, FZ_T, lds, c->stream>>>(*g, tw);                          \
    return true;                                                                                 \
  }
  RBC_FX_Z(180, 49, FZ_TL180)
  RBC_FX_Z(60, 17, 17)
#undef RBC_FX_Z
  return false;
}

template <typename T>
static int pme_transform_t(rbc3d_ctx *c) {
  typedef typename CxT<T>::t CT;
  hipStream_t st = c->stream;
  int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], H = N1 / 2 + 1;
  size_t ng = (size_t)N1 * N2 * N3, nc = (size_t)H * N2 * N3;
  CT *C = (CT *)c->d_cgrid.p;
  int f0 = c->flag_sl ? 0 : 3, f1 = c->flag_dl ? 12 : 3;
  int nf = f1 - f0;
  rbc_stage_begin(c, ST_FFT);
  size_t xy_lds = 0;
  const bool xy = xy_plane_ok<T>(c, &xy_lds);
  bool xy_fixed_fwd = false;
  if (nf > 0 && RBC_FX_ON)
    xy_fixed_fwd = fx_xy<T>(c, true, nf, (const T *)c->d_grid.p + ng * f0, C + nc * f0);
  if (xy_fixed_fwd) {
    HIP_TRY(hipGetLastError());
  } else if (nf > 0 && xy) {
    // x and y in one pass per plane
    k_fft_xy_fwd<T><<<(unsigned)(nf * N3), XY_T, xy_lds, st>>>(
        xy_args<T>(c, nf), (const T *)c->d_grid.p + ng * f0, C + nc * f0, make_plan<T>(c, 0, -1),
        make_plan<T>(c, 1, -1));
    HIP_TRY(hipGetLastError());
  } else if (nf > 0) {
    FftPlan<T> px = make_plan<T>(c, 0, -1);
    int TL = lines_per_block<T>(N1), nrows = N2 * N3 * nf;
    size_t lds = sizeof(CT) * (2 * (size_t)TL * N1 + N1);
    k_fft_r2c<T><<<(nrows + TL - 1) / TL, FFT_T, lds, st>>>((const T *)c->d_grid.p + ng * f0,
                                                            C + nc * f0, px, nrows, TL,
                                                            make_fastdiv(N1), make_fastdiv(H));
    HIP_TRY(hipGetLastError());
    RBC_TRY(fft_c2c<T>(c, C + nc * f0, 1, -1, nf));
  }
  // the z stage: forward transform, Green's function and inverse transform in one pass
  // when two or more adjacent columns of every field fit in RBC_ZG_LDS bytes of LDS
  // (four workgroups a CU), else in three passes
  const size_t zcol = 2 * (size_t)nf * N3 * sizeof(CT);
  const size_t zlds = std::min<size_t>(RBC_ZG_LDS, (size_t)c->lds_max);
  const int TLz = nf > 0 && zlds > N3 * sizeof(CT)
                      ? (int)std::min<size_t>(8, (zlds - N3 * sizeof(CT)) / zcol)
                      : 0;
  const bool z_fixed = RBC_FX_ON && nf == 3 && f0 == 0 && fx_zgreen<T>(c, nullptr);
  const bool fused = !z_fixed && nf > 0 && TLz >= 2;
  if (nf > 0 && !fused && !z_fixed) RBC_TRY(fft_c2c<T>(c, C + nc * f0, 2, +1, nf));
  rbc_stage_end(c, ST_FFT);
  GreenArgs<T> g;
  g.N1 = N1;
  g.N2 = N2;
  g.N3 = N3;
  g.H = H;
  for (int d = 0; d < 3; d++) g.iLb[d] = c->iLb[d];
  g.alpha = c->conf.alpha;
  g.vol = c->Lb[0] * c->Lb[1] * c->Lb[2];
  g.flag_sl = c->flag_sl;
  g.flag_dl = c->flag_dl;
  g.bb = c->d_bb.p;
  g.C = C;
  if (z_fixed) {
    if ((long long)15 * nc >= (1LL << 31)) {
      rbc_set_error("PME grid too large for 32-bit FFT offsets");
      return RBC3D_EINVAL;
    }
    rbc_stage_begin(c, ST_GREEN);
    fx_zgreen<T>(c, &g);
    rbc_stage_end(c, ST_GREEN);
    HIP_TRY(hipGetLastError());
  } else if (fused) {
    if ((long long)15 * nc >= (1LL << 31)) {
      rbc_set_error("PME grid too large for 32-bit FFT offsets");
      return RBC3D_EINVAL;
    }
    FftPlan<T> pzf = make_plan<T>(c, 2, +1), pzi = make_plan<T>(c, 2, -1);
    const size_t lds = sizeof(CT) * (2 * (size_t)nf * TLz * N3 + N3);
    const unsigned nblk = (unsigned)((((H + TLz - 1) / TLz) * N2 + 7) / 8 * 8);
    rbc_stage_begin(c, ST_GREEN);
    k_fft_zgreen<T><<<nblk, FFT_T, lds, st>>>(g, pzf, pzi, f0, nf, TLz);
    rbc_stage_end(c, ST_GREEN);
    HIP_TRY(hipGetLastError());
  } else {
    rbc_stage_begin(c, ST_GREEN);
    k_green<T><<<(unsigned)((nc + 255) / 256), 256, 0, st>>>(g);
    rbc_stage_end(c, ST_GREEN);
    HIP_TRY(hipGetLastError());
  }
  rbc_stage_begin(c, ST_FFT);
  if (!fused && !z_fixed) RBC_TRY(fft_c2c<T>(c, C + nc * 12, 2, -1, 3));
  if (RBC_FX_ON && fx_xy<T>(c, false, 3, C + nc * 12, c->d_vgrid.p)) {
  } else if (xy) {
    k_fft_xy_inv<T><<<(unsigned)(3 * N3), XY_T, xy_lds, st>>>(
        xy_args<T>(c, 3), C + nc * 12, (T *)c->d_vgrid.p, make_plan<T>(c, 0, +1),
        make_plan<T>(c, 1, +1));
  } else {
    RBC_TRY(fft_c2c<T>(c, C + nc * 12, 1, +1, 3));
    FftPlan<T> pxi = make_plan<T>(c, 0, +1);
    int TL = lines_per_block<T>(N1), nrows = N2 * N3 * 3;
    size_t lds = sizeof(CT) * (2 * (size_t)TL * N1 + N1);
    k_fft_c2r<T><<<(nrows + TL - 1) / TL, FFT_T, lds, st>>>(C + nc * 12, (T *)c->d_vgrid.p, pxi,
                                                             nrows, TL, make_fastdiv(N1));
  }
  rbc_stage_end(c, ST_FFT);
  HIP_TRY(hipGetLastError());
  c->pme_grid_prec = c->prec;
  return RBC3D_OK;
}

int rbc_pme_transform(rbc3d_ctx *c) {
  return c->prec == RBC3D_PREC_SPLIT ? pme_transform_t<float>(c) : pme_transform_t<double>(c);
}

// ---------------------------------------------------------------------------
// gather: Interp_Vel (:450-489) + PME_Add_Interp_Vel (:228-248)
template <typename T>
struct GatherArgs {
  int N1, N2, N3, P, nT;
  double ih[3];
  const T *vv;
  const double *tx, *tA;
  const int *tact;
  double *v;
};


// PC: the B-spline order when it is known at compile time (0: runtime a.P)
#define GATHER_L 8  // lanes per target
template <int PC, typename T>
__global__ void __launch_bounds__(256) k_gather(GatherArgs<T> a) {
  // GATHER_L lanes per target share its P x P (y, z) stencil rows; each row is P
  // consecutive x points.  The partial sums are combined with shuffles.
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int ti = gid / GATHER_L, sub = gid % GATHER_L;
  const bool live = ti < a.nT && a.tact[ti];
  double v0 = 0, v1 = 0, v2 = 0;
  if (live) {
    double x[3] = {a.tx[ti], a.tx[ti + a.nT], a.tx[ti + 2 * (size_t)a.nT]};
    double wx[12], wy[12], wz[12];
    const int P = PC ? PC : a.P;
    int kmin = bspline_d(x[2] * a.ih[2], P, wz);
    int jmin = bspline_d(x[1] * a.ih[1], P, wy);
    int imin = bspline_d(x[0] * a.ih[0], P, wx);
    const size_t ng = (size_t)a.N1 * a.N2 * a.N3;
    for (int r = sub; r < P * P; r += GATHER_L) {
      const int j0 = r % P, k0 = r / P;
      const int k = wrap_idx(kmin + k0, a.N3), j = wrap_idx(jmin + j0, a.N2);
      const double wjk = wy[j0] * wz[k0];
      const size_t row = (size_t)a.N1 * (j + (size_t)a.N2 * k);
      for (int i0 = 0; i0 < P; i0++) {
        const int i = wrap_idx(imin + i0, a.N1);
        const double w = wx[i0] * wjk;
        const size_t o = row + i;
        v0 += w * a.vv[o];
        v1 += w * a.vv[o + ng];
        v2 += w * a.vv[o + 2 * ng];
      }
    }
  }
  for (int o = GATHER_L / 2; o > 0; o >>= 1) {
    v0 += __shfl_xor(v0, o, GATHER_L);
    v1 += __shfl_xor(v1, o, GATHER_L);
    v2 += __shfl_xor(v2, o, GATHER_L);
  }
  if (live && sub == 0) {
    const double A = a.tA[ti];
    a.v[ti] += v0 / A;
    a.v[ti + (size_t)a.nT] += v1 / A;
    a.v[ti + 2 * (size_t)a.nT] += v2 / A;
  }
}

template <typename T>
static int pme_gather_t(rbc3d_ctx *c, int nT, const double *d_tx, const int *d_act,
                        const double *d_A, double *d_v) {
  GatherArgs<T> g;
  g.N1 = c->conf.Nb[0];
  g.N2 = c->conf.Nb[1];
  g.N3 = c->conf.Nb[2];
  g.P = c->conf.P;
  g.nT = nT;
  for (int d = 0; d < 3; d++) g.ih[d] = c->conf.Nb[d] / c->Lb[d];
  g.vv = (const T *)c->d_vgrid.p;
  g.tx = d_tx;
  g.tA = d_A;
  g.tact = d_act;
  g.v = d_v;
  rbc_stage_begin(c, ST_GATHER);
  const unsigned nblk = (unsigned)(((size_t)nT * GATHER_L + 255) / 256);
  if (g.P == 6)
    k_gather<6, T><<<nblk, 256, 0, c->stream>>>(g);
  else
    k_gather<0, T><<<nblk, 256, 0, c->stream>>>(g);
  rbc_stage_end(c, ST_GATHER);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// The gather at the wall's vertices (the targets of the no-slip operator and of the
// residual wall velocity): the vertices do not move, so each one's stencil origin and
// B-spline weights are computed once per wall and grid.  GATHER_WL lanes per vertex
// share its P^3 stencil points, x fastest (a load instruction then touches a few rows
// rather than a row per lane), combined with shuffles; the division by the target's A
// closes the sum as in the general gather.
#ifndef GATHER_WL
#define GATHER_WL 32
#endif
#ifndef GATHER_PT
#define GATHER_PT 1  // lane per stencil point (measured faster than lane per row)
#endif
// OUT: the targets are the vertices last[k], k < nout, and the result goes to
// b(3 k + d) = vin(i, d) + sum / A_i; else v(i, d) += sum / A_i for every vertex i < nT
template <int PC, typename T, bool OUT = false>
__global__ void __launch_bounds__(256) k_gather_wall(int nT, int N1, int N2, int N3, int P_,
                                                      const int *__restrict__ org,
                                                      const T *__restrict__ wts,
                                                      const T *__restrict__ vv,
                                                      const double *__restrict__ tA, double *v,
                                                      const double *__restrict__ vin = nullptr,
                                                      const int *__restrict__ last = nullptr,
                                                      int nout = 0) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int tk = gid / GATHER_WL, sub = gid % GATHER_WL;
  const bool live = OUT ? tk < nout : tk < nT;
  const int ti = OUT ? (live ? last[tk] : 0) : tk;
  const int P = PC ? PC : P_;
  T v0 = 0, v1 = 0, v2 = 0;  // the split form sums in FP32, the scalar of its grid
  if (live) {
    const int imin = org[3 * ti], jmin = org[3 * ti + 1], kmin = org[3 * ti + 2];
    const T *W = wts + (size_t)3 * P * ti;
    const size_t ng = (size_t)N1 * N2 * N3;
#if GATHER_PT
    // lane per stencil point, x fastest: one load instruction covers a few rows
    for (int p = sub; p < P * P * P; p += GATHER_WL) {
      const int i0 = p % P, r = p / P, j0 = r % P, k0 = r / P;
      const int k = wrap_idx(kmin + k0, N3), j = wrap_idx(jmin + j0, N2), i = wrap_idx(imin + i0, N1);
      const T w = W[i0] * (W[P + j0] * W[2 * P + k0]);
      const size_t o = i + (size_t)N1 * (j + (size_t)N2 * k);
      v0 += w * vv[o];
      v1 += w * vv[o + ng];
      v2 += w * vv[o + 2 * ng];
    }
#else
    for (int r = sub; r < P * P; r += GATHER_WL) {
      const int j0 = r % P, k0 = r / P;
      const int k = wrap_idx(kmin + k0, N3), j = wrap_idx(jmin + j0, N2);
      const T wjk = W[P + j0] * W[2 * P + k0];
      const size_t row = (size_t)N1 * (j + (size_t)N2 * k);
#pragma unroll
      for (int i0 = 0; i0 < (PC ? PC : 16); i0++) {
        if (!PC && i0 >= P) break;
        const int i = wrap_idx(imin + i0, N1);
        const T w = W[i0] * wjk;
        const size_t o = row + i;
        v0 += w * vv[o];
        v1 += w * vv[o + ng];
        v2 += w * vv[o + 2 * ng];
      }
    }
#endif
  }
  for (int o = GATHER_WL / 2; o > 0; o >>= 1) {
    v0 += __shfl_xor(v0, o, GATHER_WL);
    v1 += __shfl_xor(v1, o, GATHER_WL);
    v2 += __shfl_xor(v2, o, GATHER_WL);
  }
  if (live && sub == 0) {
    const double A = tA[ti];
    const double u0 = v0, u1 = v1, u2 = v2;
    if constexpr (OUT) {
      v[3 * (size_t)tk] = vin[ti] + u0 / A;
      v[3 * (size_t)tk + 1] = vin[ti + (size_t)nT] + u1 / A;
      v[3 * (size_t)tk + 2] = vin[ti + 2 * (size_t)nT] + u2 / A;
    } else {
      v[ti] += u0 / A;
      v[ti + (size_t)nT] += u1 / A;
      v[ti + 2 * (size_t)nT] += u2 / A;
    }
  }
}

static int wall_gather_setup(rbc3d_ctx *c) {
  WallData &w = c->wall;
  if (w.wgt_ready) return RBC3D_OK;
  const int nv = w.nvert, P = c->conf.P;
  if (P > 16) {
    rbc_set_error("B-spline order above 16");
    return RBC3D_EINVAL;
  }
  std::vector<double> x((size_t)3 * nv), wts((size_t)3 * P * nv);
  std::vector<int> org((size_t)3 * nv);
  HIP_TRY(hipMemcpy(x.data(), w.d_x.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToHost));
  for (int i = 0; i < nv; i++)
    for (int d = 0; d < 3; d++)
      org[3 * i + d] = bspline_d(x[i + (size_t)nv * d] * (c->conf.Nb[d] / c->Lb[d]), P,
                                 &wts[(size_t)3 * P * i + d * P]);
  std::vector<float> wtsf(wts.begin(), wts.end());
  RBC_TRY(w.d_wg_org.ensure(org.size()));
  RBC_TRY(w.d_wg_w.ensure(wts.size()));
  RBC_TRY(w.d_wg_wf.ensure(wts.size()));
  HIP_TRY(hipMemcpy(w.d_wg_org.p, org.data(), org.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_wg_w.p, wts.data(), wts.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(w.d_wg_wf.p, wtsf.data(), wtsf.size() * sizeof(float), hipMemcpyHostToDevice));
  w.wgt_ready = true;
  return RBC3D_OK;
}

int rbc_pme_gather_wall(rbc3d_ctx *c, const double *d_A, double *d_v) {
  WallData &w = c->wall;
  const int nT = w.nvert;
  if (nT <= 0) return RBC3D_OK;
  RBC_TRY(wall_gather_setup(c));
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], P = c->conf.P;
  const unsigned nblk = (unsigned)(((size_t)nT * GATHER_WL + 255) / 256);
  rbc_stage_begin(c, ST_GATHER);
  if (c->pme_grid_prec == RBC3D_PREC_SPLIT) {
    const float *vv = (const float *)c->d_vgrid.p;
    if (P == 6)
      k_gather_wall<6, float><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_wf.p, vv, d_A, d_v);
    else
      k_gather_wall<0, float><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_wf.p, vv, d_A, d_v);
  } else {
    const double *vv = (const double *)c->d_vgrid.p;
    if (P == 6)
      k_gather_wall<6, double><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_w.p, vv, d_A, d_v);
    else
      k_gather_wall<0, double><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_w.p, vv, d_A, d_v);
  }
  rbc_stage_end(c, ST_GATHER);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

int rbc_pme_gather_wall_out(rbc3d_ctx *c, const double *d_A, const double *d_v, const int *d_last,
                            int nu, double *d_b) {
  WallData &w = c->wall;
  const int nT = w.nvert;
  if (nT <= 0 || nu <= 0) return RBC3D_OK;
  RBC_TRY(wall_gather_setup(c));
  const int N1 = c->conf.Nb[0], N2 = c->conf.Nb[1], N3 = c->conf.Nb[2], P = c->conf.P;
  const unsigned nblk = (unsigned)(((size_t)nu * GATHER_WL + 255) / 256);
  rbc_stage_begin(c, ST_GATHER);
  if (c->pme_grid_prec == RBC3D_PREC_SPLIT) {
    const float *vv = (const float *)c->d_vgrid.p;
    if (P == 6)
      k_gather_wall<6, float, true><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_wf.p, vv, d_A, d_b, d_v, d_last, nu);
    else
      k_gather_wall<0, float, true><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_wf.p, vv, d_A, d_b, d_v, d_last, nu);
  } else {
    const double *vv = (const double *)c->d_vgrid.p;
    if (P == 6)
      k_gather_wall<6, double, true><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_w.p, vv, d_A, d_b, d_v, d_last, nu);
    else
      k_gather_wall<0, double, true><<<nblk, 256, 0, c->stream>>>(nT, N1, N2, N3, P, w.d_wg_org.p, w.d_wg_w.p, vv, d_A, d_b, d_v, d_last, nu);
  }
  rbc_stage_end(c, ST_GATHER);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

int rbc_pme_gather(rbc3d_ctx *c, int nT, const double *d_tx, const int *d_act, const double *d_A,
                   double *d_v) {
  if (nT <= 0) return RBC3D_OK;
  // the velocity grid holds the scalar of the transform that wrote it
  return c->pme_grid_prec == RBC3D_PREC_SPLIT ? pme_gather_t<float>(c, nT, d_tx, d_act, d_A, d_v)
                                              : pme_gather_t<double>(c, nT, d_tx, d_act, d_A, d_v);
}
