// Per-cell spherical-harmonic building blocks shared by the transform, spline and
// membrane-mechanics kernels: longitude DFT, Legendre analysis and synthesis sums and
// the longitude synthesis, each a workgroup-wide loop over LDS-resident data.
#pragma once
#include <hip/hip_runtime.h>

// The per-cell transforms below sum in the reference's order (longitude DFT, then
// (2/nlon) sum_i (Pbar w_i) C_i, then the synthesis sums over n and m) with FMA
// contraction off, so the surface splines built from them equal the CPU restatement's
// bit for bit: the double-layer self patch subtracts spline positions of magnitude ~L
// that lie ~0.01 apart, and one ulp there moves a velocity by ~1e-12.

// C[m][i] = sum_j f[i + ld*j] cos(m phi_j),  S[m][i] = sum_j f sin(m phi_j)
__device__ inline void lon_dft(const double *f, int ld, int nl, int nn, int mmax, const double *cs,
                        int cs_ld_m, double *C, double *S) {
#pragma clang fp contract(off)
  const double *cosT = cs, *sinT = cs + (size_t)cs_ld_m * nn;
  for (int t = threadIdx.x; t < mmax * nl; t += blockDim.x) {
    int m = t / nl, i = t % nl;
    double a = 0, b = 0;
    for (int j = 0; j < nn; j++) {
      double v = f[i + ld * j];
      a += v * cosT[m * nn + j];
      b += v * sinT[m * nn + j];
    }
    C[t] = a;
    S[t] = b;
  }
}

// a[m + lda*n] = scale sum_i pw[m][n][i] C[m][i];  b = -scale sum_i pw S  (n in [m, nmax)),
// pw = Pbar(m,n,th_i) w_i and scale = 2/nlon
__device__ inline void leg_anal(const double *C, const double *S, const double *pw, int nl, int mmax,
                         int nmax, double *a, double *b, int lda, double scale) {
#pragma clang fp contract(off)
  for (int t = threadIdx.x; t < mmax * nmax; t += blockDim.x) {
    int m = t / nmax, n = t % nmax;
    double sa = 0, sb = 0;
    if (n >= m) {
      const double *p = pw + ((size_t)m * nl + n) * nl;
      for (int i = 0; i < nl; i++) {
        sa += p[i] * C[m * nl + i];
        sb += p[i] * S[m * nl + i];
      }
    }
    a[m + lda * n] = scale * sa;
    b[m + lda * n] = -scale * sb;
  }
}

// G[m][i] = sum_{n=m}^{nmax-1} pt[m][n][i] a[m + lda n]  (pt includes the 1/2 at m = 0)
__device__ inline void leg_synth(const double *a, const double *b, int lda, const double *pt, int nlt,
                          int nrow, int mmax, int nmax, double *Ga, double *Gb) {
#pragma clang fp contract(off)
  for (int t = threadIdx.x; t < mmax * nrow; t += blockDim.x) {
    int m = t / nrow, i = t % nrow;
    double sa = 0, sb = 0;
    for (int n = m; n < nmax; n++) {
      double p = pt[((size_t)m * nlt + n) * nrow + i];
      sa += p * a[m + lda * n];
      sb += p * b[m + lda * n];
    }
    Ga[t] = sa;
    Gb[t] = sb;
  }
}

// g[i + ld*j] = sum_m Ga[m][i] cos(m phi_j) - Gb[m][i] sin(m phi_j)
__device__ inline void lon_synth(const double *Ga, const double *Gb, int nrow, int nn, int mmax,
                          const double *cs, int cs_ld_m, double *g, int ld) {
#pragma clang fp contract(off)
  const double *cosT = cs, *sinT = cs + (size_t)cs_ld_m * nn;
  for (int t = threadIdx.x; t < nrow * nn; t += blockDim.x) {
    int i = t % nrow, j = t / nrow;
    double v = 0;
    for (int m = 0; m < mmax; m++) v += Ga[m * nrow + i] * cosT[m * nn + j] - Gb[m * nrow + i] * sinT[m * nn + j];
    g[i + ld * j] = v;
  }
}

