// Spherical-harmonic transforms on the cell grids (SPHEREPACK conventions, as
// used through common/ModSphpk.F90), the per-cell surface splines built from
// them (Rbc_BuildSurfaceSource, common/ModRbc.F90:716-807; Spline_Build_on_Sphere,
// common/ModSpline.F90:121-142; FFT_Diff, common/ModFFT.F90:25-93) and the global
// packing of Glob_Sph_Trans (common/ModVelSolver.F90:641-719).
//
// One workgroup per cell.  A cell's data is a few KB, so each stage (longitude
// DFT, Legendre sums, equi-spaced synthesis, derivative operators) runs out of LDS
// as small dense contractions with precomputed tables.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

#define SH_THREADS 256

#include "sh_core.h"

// ---------------------------------------------------------------------------
// Surface splines of a cell: which = 0 -> x, a3, detJ; 1 -> f detJ; 2 -> g detJ
// The grid-type tables of one launch over cells of up to SH_MAXSEG grid types
#define SH_MAXSEG 8
struct SplTabs {
  int ncell;  // cells of the launch (cell_list entries from cell0)
  int gt[SH_MAXSEG];
  const double *pw[SH_MAXSEG], *pe[SH_MAXSEG], *cs[SH_MAXSEG], *Dm[SH_MAXSEG], *Dn[SH_MAXSEG];
};

// One persistent launch over the cells of every grid type: each workgroup takes cells
// in turn (cells of different sizes side by side) with its type's tables.
__global__ void __launch_bounds__(SH_THREADS) k_build_splines(
    const CellRec *__restrict__ cells, int cell0, SplTabs T, const double *__restrict__ x,
    const double *__restrict__ a3, const double *__restrict__ detj, const double *__restrict__ fd,
    int np, int which, double *__restrict__ spl, const int *__restrict__ cell_list) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  for (int w = blockIdx.x; w < T.ncell; w += gridDim.x) {
  __syncthreads();  // the previous cell is done with the LDS buffers
  int ci = cell_list[cell0 + w];
  const CellRec cr = cells[ci];
  int k = 0;
  while (k < SH_MAXSEG - 1 && T.gt[k] != cr.gtype) k++;
  const double *pw = T.pw[k], *pe = T.pe[k], *cs = T.cs[k], *Dm = T.Dm[k], *Dn = T.Dn[k];
  int nl = cr.nlat, nn = cr.nlon, n0 = cr.nlat0, ne = nl + 1, m2 = 2 * nl;
  int mmax = min(nl, nn / 2 + 1);
  int R1 = ne * nn, R2 = 2 * ne * nl;
  double *r1 = sm, *r2 = sm + R1, *u = r2 + R2, *u1 = u + m2 * nn;
  int nvar = which == 0 ? 7 : 3;
  for (int v = 0; v < nvar; v++) {
    // field values on the Gauss grid
    const double *src;
    int comp, var;
    bool mul = false;
    if (which == 0) {
      if (v < 3) {
        src = x;
        comp = v;
        var = SPL_X + v;
      } else if (v < 6) {
        src = a3;
        comp = v - 3;
        var = SPL_A3 + v - 3;
      } else {
        src = detj;
        comp = 0;
        var = SPL_DETJ;
      }
    } else {
      src = fd;
      comp = v;
      var = (which == 1 ? SPL_F : SPL_G) + v;
      mul = true;
    }
    for (int t = threadIdx.x; t < nl * nn; t += blockDim.x) {
      double val = src[cr.off + t + (size_t)np * comp];
      if (mul) val *= detj[cr.off + t];
      r1[t] = val;  // (ilat, ilon) at ilat + nl*ilon
    }
    __syncthreads();
    // analysis: C,S in r2
    lon_dft(r1, nl, nl, nn, mmax, cs, nl + 1, r2, r2 + mmax * nl);
    __syncthreads();
    // a, b (n < nlat0: ShFilter) in r1, lda = n0
    leg_anal(r2, r2 + mmax * nl, pw, nl, min(mmax, n0), n0, r1, r1 + n0 * n0, n0, 2.0 / nn);
    __syncthreads();
    // equi-spaced synthesis on nl+1 rows: Ga, Gb in r2
    int mme = min(n0, min(ne, nn / 2 + 1));
    leg_synth(r1, r1 + n0 * n0, n0, pe, ne, ne, mme, n0, r2, r2 + mme * ne);
    __syncthreads();
    lon_synth(r2, r2 + mme * ne, ne, nn, mme, cs, nl + 1, r1, ne);  // g_e[ie + ne*j] in r1
    __syncthreads();
    // doubled sphere: u(i, j), i < 2 nl
    for (int t = threadIdx.x; t < m2 * nn; t += blockDim.x) {
      int i = t % m2, j = t / m2;
      u[t] = i < nl ? r1[i + ne * j] : r1[(m2 - i) + ne * ((j + nn / 2) % nn)];
    }
    __syncthreads();
    // u1 = D_m u (first index)
    for (int t = threadIdx.x; t < m2 * nn; t += blockDim.x) {
      int i = t % m2, j = t / m2;
      double s = 0;
      for (int q = 0; q < m2; q++) s += Dm[i * m2 + q] * u[q + m2 * j];
      u1[t] = s;
    }
    __syncthreads();
    // u2 = derivative of u along j, u12 = derivative of u1 along j; write block
    double *out = spl + cr.spl + (size_t)var * m2 * nn * 4;
    for (int t = threadIdx.x; t < m2 * nn; t += blockDim.x) {
      int i = t % m2, j = t / m2;
      double s2 = 0, s12 = 0;
      for (int q = 0; q < nn; q++) {
        double d = Dn[j * nn + q];
        s2 += d * u[i + m2 * q];
        s12 += d * u1[i + m2 * q];
      }
      double4 o;
      o.x = u[t];
      o.y = u1[t];
      o.z = s2;
      o.w = s12;
      *(double4 *)(out + 4 * (size_t)t) = o;
    }
    __syncthreads();
  }
  }
}

static size_t spline_lds(int nl) {
  int ne = nl + 1, nn = 2 * nl;
  return sizeof(double) * ((size_t)ne * nn + 2 * (size_t)ne * nl + 2 * (size_t)(2 * nl) * nn);
}

// grouped launches: one per grid type (cells of one type share the tables)
static int cells_of_type(rbc3d_ctx *c, int gt, std::vector<int> &ids) {
  ids.clear();
  for (int i = 0; i < c->nrbc; i++)
    if (c->cells[i].gtype == gt) ids.push_back(i);
  return (int)ids.size();
}

static int upload_ids(rbc3d_ctx *c) {
  // d_tmp_i holds, per grid type, the list of cells of that type (concatenated)
  std::vector<int> all;
  std::vector<int> ids;
  for (size_t gt = 0; gt < c->gtypes.size(); gt++) {
    cells_of_type(c, (int)gt, ids);
    all.insert(all.end(), ids.begin(), ids.end());
  }
  RBC_TRY(c->d_tmp_i.ensure(all.size() + 1));
  if (!all.empty())
    HIP_TRY(hipMemcpyAsync(c->d_tmp_i.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

// workgroups of a persistent launch over ncell cells
static unsigned persistent_grid(const rbc3d_ctx *c, int nwork, int per_cu) {
  return (unsigned)std::max(1, std::min(nwork, per_cu * c->ncu));
}

int rbc_sh_build_splines(rbc3d_ctx *c, int which) {
  if (c->nrbc == 0) return RBC3D_OK;
  if (which == 0) RBC_TRY(upload_ids(c));
  std::vector<int> ids;
  const double *fd = which == 1 ? c->d_f.p : which == 2 ? c->d_g.p : nullptr;
  // the grid types in groups of SH_MAXSEG, one launch per group (one launch in all for
  // up to SH_MAXSEG distinct grids); cell_list holds the cells type by type
  int base = 0;
  for (size_t g0 = 0; g0 < c->gtypes.size(); g0 += SH_MAXSEG) {
    SplTabs T{};
    int nseg = 0, ncell = 0;
    size_t lds = 0;
    for (size_t gt = g0; gt < std::min(c->gtypes.size(), g0 + SH_MAXSEG); gt++) {
      const int n = cells_of_type(c, (int)gt, ids);
      if (!n) continue;
      GridType *g = c->gtypes[gt];
      T.gt[nseg] = (int)gt;
      T.pw[nseg] = g->d_pws.p;
      T.pe[nseg] = g->d_pe.p;
      T.cs[nseg] = g->d_cs.p;
      T.Dm[nseg] = g->d_Dm.p;
      T.Dn[nseg] = g->d_Dn.p;
      lds = std::max(lds, spline_lds(g->nlat));
      nseg++;
      ncell += n;
    }
    if (!ncell) continue;
    for (int k = nseg; k < SH_MAXSEG; k++) T.gt[k] = -1;
    T.ncell = ncell;
    rbc_stage_begin(c, ST_SPLINE);
    hipLaunchKernelGGL(k_build_splines, dim3(persistent_grid(c, ncell, 4)), dim3(SH_THREADS), lds,
                       c->stream, c->d_cells.p, base, T, c->d_x.p, c->d_a3.p, c->d_detj.p, fd, c->np,
                       which, c->d_spl.p, c->d_tmp_i.p);
    rbc_stage_end(c, ST_SPLINE);
    HIP_TRY(hipGetLastError());
    base += ncell;
  }
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// Glob_Sph_Trans: analysis (v -> packed coefficients) and synthesis.
__device__ __forceinline__ int pack_a(int m, int n, int n0) {
  // a(m,n), m = 0..n0-1, n = m..n0-1, in (m outer, n inner) order
  return m * n0 - m * (m - 1) / 2 + (n - m);
}
__device__ __forceinline__ int pack_b(int m, int n, int n0) {
  int nb0 = n0 * (n0 + 1) / 2;
  int mm = m - 1;  // m = 1..n0-1
  return nb0 + mm * (n0 - 1) - mm * (mm - 1) / 2 + (n - m);
}

// ---------------------------------------------------------------------------
// Glob_Sph_Trans with the Legendre stage as batched GEMMs on the matrix cores.
// For every m the Legendre sums of all cells of one grid type form one product:
//   analysis   Y_m(n, col) = sum_i pw[m][n][i] X_m(i, col)
//   synthesis  G_m(i, col) = sum_n pg[m][n][i] Z_m(n, col)
// with col = 6 * cell + 2 * component + (0: cos, 1: sin).  The longitude DFTs and the
// coefficient (un)packing stay per cell.  Matrices are (m, row, col) with col fastest,
// ld = ncolp (6 * ncell rounded up to 16).
typedef double rbc_d4 __attribute__((ext_vector_type(4)));

// One launch of each stage covers the cells of every grid type: a plan holds, per grid
// type (segment), its sizes, tables and the offset of its operands in the staging
// buffer; cell-wise kernels take cells (workgroups) and the GEMMs take 16 x 16 output
// tiles (wavefronts) from one range over all segments, persistently.
struct ShSeg {
  int nl, nn, mmax, ncell, cell0, ncolp, RT, CT, tile0;
  long long xoff, mat;  // X at xoff, Y at xoff + mat (doubles)
  const double *cs, *P;          // longitude table; Legendre table of the direction
  const unsigned short *P3;      // its three BF16 terms (split form)
};
struct ShPlan {
  int nseg, ncell, ntile;
  ShSeg s[SH_MAXSEG];
};
__device__ __forceinline__ int seg_of_cell(const ShPlan &P, int w) {
  int k = 0;
  while (k + 1 < P.nseg && w >= P.s[k + 1].cell0) k++;
  return k;
}
__device__ __forceinline__ int seg_of_tile(const ShPlan &P, int w) {
  int k = 0;
  while (k + 1 < P.nseg && w >= P.s[k + 1].tile0) k++;
  return k;
}

// X_m(i, col) = sum_j v(i, j) cos(m phi_j) / sin(m phi_j) for the 3 components
__global__ void __launch_bounds__(SH_THREADS) k_sh_lon_anal(
    const CellRec *__restrict__ cells, const int *__restrict__ cell_list, int cell0, ShPlan P,
    const double *__restrict__ v, int np, double *Xb) {
  for (int w = blockIdx.x; w < P.ncell; w += gridDim.x) {
    const ShSeg &S = P.s[seg_of_cell(P, w)];
    const int c = w - S.cell0, mmax = S.mmax, ncolp = S.ncolp;
    const CellRec cr = cells[cell_list[cell0 + w]];
    const int nl = cr.nlat, nn = cr.nlon;
    const double *cosT = S.cs, *sinT = S.cs + (size_t)(nl + 1) * nn;
    double *X = Xb + S.xoff;
    for (int t = threadIdx.x; t < 3 * mmax * nl; t += blockDim.x) {
      const int d = t / (mmax * nl), r = t - d * mmax * nl, m = r / nl, i = r - m * nl;
      const double *f = v + cr.off + (size_t)np * d;
      double a = 0, b = 0;
      for (int j = 0; j < nn; j++) {
        const double x = f[i + nl * j];
        a += x * cosT[m * nn + j];
        b += x * sinT[m * nn + j];
      }
      double *Xm = X + ((size_t)m * nl + i) * ncolp + 6 * c + 2 * d;
      Xm[0] = a;
      Xm[1] = b;
    }
  }
}

// One wavefront per 16 x 16 output tile of one m: D += A(16 x 4) B(4 x 16) with the
// FP64 MFMA (v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 i + l / 16][l % 16], i < 4).
// TRANS_A: A(row, k) = P[m][k][row] (synthesis) instead of P[m][row][k] (analysis).
template <bool TRANS_A>
__global__ void __launch_bounds__(256) k_sh_leg_mfma(ShPlan Pl, double *XY) {
  const int lane = threadIdx.x & 63;
  const int nwave = (int)(gridDim.x * 4);
  for (int wid = (int)((blockIdx.x * 256 + threadIdx.x) >> 6); wid < Pl.ntile; wid += nwave) {
    const ShSeg &S = Pl.s[seg_of_tile(Pl, wid)];
    const int nl = S.nl, ncolp = S.ncolp, RT = S.RT, CT = S.CT, tw = wid - S.tile0;
    const int ct = tw % CT, rt = (tw / CT) % RT, m = tw / (CT * RT);
    const int row = rt * 16 + (lane & 15), kq = lane >> 4, col = ct * 16 + (lane & 15);
    const double *Pm = S.P + (size_t)m * nl * nl;
    const double *Xm = XY + S.xoff + (size_t)m * nl * ncolp;
    rbc_d4 acc = {0., 0., 0., 0.};
    for (int k0 = 0; k0 < nl; k0 += 4) {
      const int k = k0 + kq;
      double a = 0., b = 0.;
      if (k < nl) {
        if (row < nl) a = TRANS_A ? Pm[(size_t)k * nl + row] : Pm[(size_t)row * nl + k];
        b = Xm[(size_t)k * ncolp + col];
      }
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    double *Ym = XY + S.xoff + S.mat + (size_t)m * nl * ncolp;
    for (int i = 0; i < 4; i++) {
      const int r = rt * 16 + 4 * i + (lane >> 4);
      if (r < nl) Ym[(size_t)r * ncolp + col] = acc[i];
    }
  }
}

// Split form of the Legendre GEMM: three-way split BF16 products with FP32
// accumulation on the BF16 matrix cores.  Each operand is carried as three BF16
// terms (x = x0 + x1 + x2, 24 significant bits); the product keeps the six terms
// x_i y_j with i + j <= 2, the largest in one FP32 accumulator and the five smaller
// ones in another, so the result is FP32-accurate.  The Legendre tables are split on
// the host; the staged X operand (FP64) is split on the fly.
// v_mfma_f32_16x16x16bf16_1k: lane l supplies A[l % 16][4 (l / 16) + j] and
// B[4 (l / 16) + j][l % 16], j < 4, and holds D[4 (l / 16) + i][l % 16], i < 4.
typedef short rbc_s4 __attribute__((ext_vector_type(4)));
typedef float rbc_f4 __attribute__((ext_vector_type(4)));

// round to nearest even FP32 -> BF16 (finite inputs)
__host__ __device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_to_f(unsigned short h) {
  unsigned u = (unsigned)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ __forceinline__ void bf16_split3(double x, unsigned short &h0, unsigned short &h1,
                                                     unsigned short &h2) {
  h0 = bf16_rne((float)x);
  const double r1 = x - (double)bf16_to_f(h0);
  h1 = bf16_rne((float)r1);
  const double r2 = r1 - (double)bf16_to_f(h1);
  h2 = bf16_rne((float)r2);
}

template <bool TRANS_A>
__global__ void __launch_bounds__(256) k_sh_leg_bf16x3(ShPlan Pl, double *XY) {
  const int lane = threadIdx.x & 63;
  const int nwave = (int)(gridDim.x * 4);
  for (int wid = (int)((blockIdx.x * 256 + threadIdx.x) >> 6); wid < Pl.ntile; wid += nwave) {
    const ShSeg &S = Pl.s[seg_of_tile(Pl, wid)];
    const int nl = S.nl, mmax = S.mmax, ncolp = S.ncolp, RT = S.RT, CT = S.CT, tw = wid - S.tile0;
    const int ct = tw % CT, rt = (tw / CT) % RT, m = tw / (CT * RT);
    const int row = rt * 16 + (lane & 15), kq = lane >> 4, col = ct * 16 + (lane & 15);
    const size_t psz = (size_t)mmax * nl * nl;
    const unsigned short *Pm = S.P3 + (size_t)m * nl * nl;
    const double *Xm = XY + S.xoff + (size_t)m * nl * ncolp;
    rbc_f4 hi = {0.f, 0.f, 0.f, 0.f}, lo = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < nl; k0 += 16) {
      rbc_s4 a0, a1, a2, b0, b1, b2;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int k = k0 + 4 * kq + j;
        unsigned short p0 = 0, p1 = 0, p2 = 0, x0 = 0, x1 = 0, x2 = 0;
        if (k < nl) {
          if (row < nl) {
            const size_t o = TRANS_A ? (size_t)k * nl + row : (size_t)row * nl + k;
            p0 = Pm[o];
            p1 = Pm[psz + o];
            p2 = Pm[2 * psz + o];
          }
          bf16_split3(Xm[(size_t)k * ncolp + col], x0, x1, x2);
        }
        a0[j] = (short)p0;
        a1[j] = (short)p1;
        a2[j] = (short)p2;
        b0[j] = (short)x0;
        b1[j] = (short)x1;
        b2[j] = (short)x2;
      }
      hi = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, hi, 0, 0, 0);
      lo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b2, lo, 0, 0, 0);
      lo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, lo, 0, 0, 0);
      lo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b0, lo, 0, 0, 0);
      lo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, lo, 0, 0, 0);
      lo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, lo, 0, 0, 0);
    }
    double *Ym = XY + S.xoff + S.mat + (size_t)m * nl * ncolp;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = rt * 16 + 4 * (lane >> 4) + i;
      if (r < nl) Ym[(size_t)r * ncolp + col] = (double)hi[i] + (double)lo[i];
    }
  }
}

// the split tables of one grid type: [3][mmax][nlat][nlat] BF16 terms of pw and pg
static int split_tables(GridType *g) {
  if (g->d_pw3.p) return RBC3D_OK;
  const size_t n = (size_t)g->mmax * g->nlat * g->nlat;
  std::vector<double> h(n);
  std::vector<unsigned short> t(3 * n);
  for (int which = 0; which < 2; which++) {
    DBuf<double> &src = which ? g->d_pg : g->d_pw;
    DBuf<unsigned short> &dst = which ? g->d_pg3 : g->d_pw3;
    HIP_TRY(hipMemcpy(h.data(), src.p, n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) bf16_split3(h[i], t[i], t[n + i], t[2 * n + i]);
    RBC_TRY(dst.ensure(3 * n));
    HIP_TRY(hipMemcpy(dst.p, t.data(), 3 * n * 2, hipMemcpyHostToDevice));
  }
  return RBC3D_OK;
}

// coefficients of Y into the Glob_Sph_Trans vector: a = Y(cos), b = -Y(sin)
__global__ void k_sh_pack(const CellRec *__restrict__ cells, const int *__restrict__ cell_list,
                          int cell0, ShPlan P, const double *__restrict__ XY, double *cf) {
  for (int w = blockIdx.x; w < P.ncell; w += gridDim.x) {
    const ShSeg &S = P.s[seg_of_cell(P, w)];
    const int c = w - S.cell0, nl = S.nl, ncolp = S.ncolp;
    const double *Y = XY + S.xoff + S.mat;
    const CellRec cr = cells[cell_list[cell0 + w]];
    const int n0 = cr.nlat0, mm = min(S.mmax, n0);
    for (int t = threadIdx.x; t < 3 * n0 * n0; t += blockDim.x) {
      const int d = t / (n0 * n0), r = t - d * n0 * n0, m = r % n0, n = r / n0;
      if (m >= mm || n < m) continue;
      const double *y = Y + ((size_t)m * nl + n) * ncolp + 6 * c + 2 * d;
      cf[cr.coff + 3 * pack_a(m, n, n0) + d] = y[0];
      if (m >= 1) cf[cr.coff + 3 * pack_b(m, n, n0) + d] = -y[1];
    }
  }
}

// Z_m(n, col) from the Glob_Sph_Trans vector (zero outside m <= n < nlat0, m < mm)
__global__ void k_sh_unpack(const CellRec *__restrict__ cells, const int *__restrict__ cell_list,
                            int cell0, ShPlan P, const double *__restrict__ cf, double *XY) {
  for (int w = blockIdx.x; w < P.ncell; w += gridDim.x) {
    const ShSeg &S = P.s[seg_of_cell(P, w)];
    const int c = w - S.cell0, nl = S.nl, mmax = S.mmax, ncolp = S.ncolp;
    double *Z = XY + S.xoff;
    const CellRec cr = cells[cell_list[cell0 + w]];
    const int n0 = cr.nlat0, mm = min(n0, mmax);
    for (int t = threadIdx.x; t < 3 * mmax * nl; t += blockDim.x) {
      const int d = t / (mmax * nl), r = t - d * mmax * nl, m = r / nl, n = r - m * nl;
      double a = 0., b = 0.;
      if (m < mm && n >= m && n < n0) {
        a = cf[cr.coff + 3 * pack_a(m, n, n0) + d];
        if (m >= 1) b = cf[cr.coff + 3 * pack_b(m, n, n0) + d];
      }
      double *z = Z + ((size_t)m * nl + n) * ncolp + 6 * c + 2 * d;
      z[0] = a;
      z[1] = b;
    }
  }
}

// v(i, j) = sum_m G_m(i, cos) cos(m phi_j) - G_m(i, sin) sin(m phi_j)
__global__ void __launch_bounds__(SH_THREADS) k_sh_lon_synth(
    const CellRec *__restrict__ cells, const int *__restrict__ cell_list, int cell0, ShPlan P,
    const double *__restrict__ XY, int np, double *v) {
  for (int w = blockIdx.x; w < P.ncell; w += gridDim.x) {
    const ShSeg &S = P.s[seg_of_cell(P, w)];
    const int c = w - S.cell0, ncolp = S.ncolp;
    const double *G = XY + S.xoff + S.mat;
    const CellRec cr = cells[cell_list[cell0 + w]];
    const int nl = cr.nlat, nn = cr.nlon, mm = min(cr.nlat0, S.mmax);
    const double *cosT = S.cs, *sinT = S.cs + (size_t)(nl + 1) * nn;
    for (int t = threadIdx.x; t < 3 * nl * nn; t += blockDim.x) {
      const int d = t / (nl * nn), q = t - d * nl * nn, i = q % nl, j = q / nl;
      double x = 0;
      for (int m = 0; m < mm; m++) {
        const double *g = G + ((size_t)m * nl + i) * ncolp + 6 * c + 2 * d;
        x += g[0] * cosT[m * nn + j] - g[1] * sinT[m * nn + j];
      }
      v[cr.off + q + (size_t)np * d] = x;
    }
  }
}

// Glob_Sph_Trans (common/ModVelSolver.F90:641-719): direction 1 analysis (node values
// -> packed coefficients), -1 synthesis.  Each stage is one persistent launch over
// the cells of every grid type (up to SH_MAXSEG distinct grids per launch).
int rbc_sh_glob(rbc3d_ctx *c, const double *d_in, double *d_out, int direction) {
  std::vector<int> ids;
  const bool split = c->prec == RBC3D_PREC_SPLIT;
  int base = 0;
  for (size_t g0 = 0; g0 < c->gtypes.size(); g0 += SH_MAXSEG) {
    ShPlan P{};
    long long off = 0;
    for (size_t gt = g0; gt < std::min(c->gtypes.size(), g0 + SH_MAXSEG); gt++) {
      const int n = cells_of_type(c, (int)gt, ids);
      if (!n) continue;
      GridType *g = c->gtypes[gt];
      if (split && !g->d_pw3.p) {  // once per grid type
        HIP_TRY(hipStreamSynchronize(c->stream));
        RBC_TRY(split_tables(g));
      }
      ShSeg &S = P.s[P.nseg];
      S.nl = g->nlat;
      S.nn = g->nlon;
      S.mmax = g->mmax;
      S.ncell = n;
      S.cell0 = P.ncell;
      S.ncolp = (6 * n + 15) / 16 * 16;
      S.RT = (S.nl + 15) / 16;
      S.CT = S.ncolp / 16;
      S.tile0 = P.ntile;
      S.mat = (long long)S.mmax * S.nl * S.ncolp;
      S.xoff = off;
      S.cs = g->d_cs.p;
      S.P = direction == 1 ? g->d_pw.p : g->d_pg.p;
      S.P3 = direction == 1 ? g->d_pw3.p : g->d_pg3.p;
      off += 2 * S.mat;
      P.ncell += n;
      P.ntile += S.mmax * S.RT * S.CT;
      P.nseg++;
    }
    if (!P.ncell) continue;
    RBC_TRY(c->d_shx.ensure((size_t)off));
    double *XY = c->d_shx.p;
    const unsigned gc = persistent_grid(c, P.ncell, 4);
    const unsigned gl = persistent_grid(c, (P.ntile + 3) / 4, 8);
    rbc_stage_begin(c, ST_SH_GLOB);
    if (direction == 1) {
      k_sh_lon_anal<<<gc, SH_THREADS, 0, c->stream>>>(c->d_cells.p, c->d_tmp_i.p, base, P, d_in,
                                                     c->np, XY);
      if (split)
        k_sh_leg_bf16x3<false><<<gl, 256, 0, c->stream>>>(P, XY);
      else
        k_sh_leg_mfma<false><<<gl, 256, 0, c->stream>>>(P, XY);
      k_sh_pack<<<gc, 256, 0, c->stream>>>(c->d_cells.p, c->d_tmp_i.p, base, P, XY, d_out);
    } else {
      k_sh_unpack<<<gc, 256, 0, c->stream>>>(c->d_cells.p, c->d_tmp_i.p, base, P, d_in, XY);
      if (split)
        k_sh_leg_bf16x3<true><<<gl, 256, 0, c->stream>>>(P, XY);
      else
        k_sh_leg_mfma<true><<<gl, 256, 0, c->stream>>>(P, XY);
      k_sh_lon_synth<<<gc, SH_THREADS, 0, c->stream>>>(c->d_cells.p, c->d_tmp_i.p, base, P, XY,
                                                      c->np, d_out);
    }
    rbc_stage_end(c, ST_SH_GLOB);
    HIP_TRY(hipGetLastError());
    base += P.ncell;
  }
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
// Standalone ShAnalGau / ShSynthGau / ShSynthEqu on nt fields with the reference's
// leading dimensions (general nlat, nlon).  Tables are built per call.
void host_gauss(int n, std::vector<double> &th, std::vector<double> &w) {
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

void host_pbar(int nl, int mmax, const std::vector<double> &th, std::vector<double> &out) {
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

__global__ void __launch_bounds__(SH_THREADS) k_anal_batch(int nl, int nn, int mmax,
                                                         const double *__restrict__ pw,
                                                         const double *__restrict__ cs,
                                                         const double *__restrict__ g, int idg,
                                                         int jdg, double *a, double *b, int ida,
                                                         int jda) {
  extern __shared__ double sm[];
  int k = blockIdx.x;
  double *f = sm, *C = f + nl * nn, *ab = C + 2 * mmax * nl;
  for (int t = threadIdx.x; t < nl * nn; t += blockDim.x) {
    int i = t % nl, j = t / nl;
    f[t] = g[(size_t)k * idg * jdg + i + (size_t)idg * j];
  }
  __syncthreads();
  lon_dft(f, nl, nl, nn, mmax, cs, mmax, C, C + mmax * nl);
  __syncthreads();
  leg_anal(C, C + mmax * nl, pw, nl, mmax, nl, ab, ab + mmax * nl, mmax, 2.0 / nn);
  __syncthreads();
  for (int t = threadIdx.x; t < mmax * nl; t += blockDim.x) {
    int m = t % mmax, n = t / mmax;
    if (n < m) continue;
    a[(size_t)k * ida * jda + m + (size_t)ida * n] = ab[t];
    b[(size_t)k * ida * jda + m + (size_t)ida * n] = ab[mmax * nl + t];
  }
}

__global__ void __launch_bounds__(SH_THREADS) k_synth_batch(int nl, int nn, int mmax, int nmax,
                                                          const double *__restrict__ pt,
                                                          const double *__restrict__ cs, double *g,
                                                          int idg, int jdg, const double *a,
                                                          const double *b, int ida, int jda) {
  extern __shared__ double sm[];
  int k = blockIdx.x;
  double *ab = sm, *G = ab + 2 * mmax * nmax, *gg = G + 2 * mmax * nl;
  for (int t = threadIdx.x; t < mmax * nmax; t += blockDim.x) {
    int m = t % mmax, n = t / mmax;
    ab[t] = n >= m ? a[(size_t)k * ida * jda + m + (size_t)ida * n] : 0.;
    ab[mmax * nmax + t] = n >= m ? b[(size_t)k * ida * jda + m + (size_t)ida * n] : 0.;
  }
  __syncthreads();
  leg_synth(ab, ab + mmax * nmax, mmax, pt, nl, nl, mmax, nmax, G, G + mmax * nl);
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, cs, mmax, gg, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < nl * nn; t += blockDim.x) {
    int i = t % nl, j = t / nl;
    g[(size_t)k * idg * jdg + i + (size_t)idg * j] = gg[t];
  }
}

// ShGradGau: v = dg/dth = sum dPbar/dth (a cos m phi - b sin m phi), w = dg/dphi =
// sum m Pbar (-a sin m phi - b cos m phi), from a, b (n < nmax) on the Gauss grid.
// pd holds dPbar/dth and pt Pbar, both with the 1/2 of m = 0.
__global__ void __launch_bounds__(SH_THREADS) k_grad_batch(int nl, int nn, int mmax, int nmax,
                                                         const double *__restrict__ pd,
                                                         const double *__restrict__ pt,
                                                         const double *__restrict__ cs, double *v,
                                                         double *w, int idv, int jdv,
                                                         const double *a, const double *b,
                                                         int ida, int jda) {
  extern __shared__ double sm[];
  int k = blockIdx.x;
  double *ab = sm, *G = ab + 2 * mmax * nmax, *gg = G + 2 * mmax * nl;
  for (int t = threadIdx.x; t < mmax * nmax; t += blockDim.x) {
    int m = t % mmax, n = t / mmax;
    ab[t] = n >= m ? a[(size_t)k * ida * jda + m + (size_t)ida * n] : 0.;
    ab[mmax * nmax + t] = n >= m ? b[(size_t)k * ida * jda + m + (size_t)ida * n] : 0.;
  }
  __syncthreads();
  leg_synth(ab, ab + mmax * nmax, mmax, pd, nl, nl, mmax, nmax, G, G + mmax * nl);
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, cs, mmax, gg, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < nl * nn; t += blockDim.x) {
    int i = t % nl, j = t / nl;
    v[(size_t)k * idv * jdv + i + (size_t)idv * j] = gg[t];
  }
  __syncthreads();
  leg_synth(ab, ab + mmax * nmax, mmax, pt, nl, nl, mmax, nmax, G, G + mmax * nl);
  __syncthreads();
  // d/dphi of (Ga cos m phi - Gb sin m phi) is (-m Gb) cos m phi - (m Ga) sin m phi
  for (int t = threadIdx.x; t < mmax * nl; t += blockDim.x) {
    const double mm = (double)(t / nl), ga = G[t], gb = G[mmax * nl + t];
    G[t] = -mm * gb;
    G[mmax * nl + t] = mm * ga;
  }
  __syncthreads();
  lon_synth(G, G + mmax * nl, nl, nn, mmax, cs, mmax, gg, nl);
  __syncthreads();
  for (int t = threadIdx.x; t < nl * nn; t += blockDim.x) {
    int i = t % nl, j = t / nl;
    w[(size_t)k * idv * jdv + i + (size_t)idv * j] = gg[t];
  }
}

static int batch_tables(int nl, int nn, int mmax, bool equi, bool anal, double **d_p, double **d_cs) {
  std::vector<double> th, w, P;
  if (equi) {
    th.resize(nl);
    for (int i = 0; i < nl; i++) th[i] = RBC_PI * i / (nl - 1);
  } else {
    host_gauss(nl, th, w);
  }
  host_pbar(nl, mmax, th, P);
  for (int m = 0; m < mmax; m++)
    for (int n = 0; n < nl; n++)
      for (int i = 0; i < nl; i++) {
        size_t o = ((size_t)m * nl + n) * nl + i;
        P[o] = anal ? P[o] * w[i] : (m == 0 ? 0.5 : 1.0) * P[o];
      }
  std::vector<double> cs((size_t)2 * mmax * nn);
  for (int m = 0; m < mmax; m++)
    for (int j = 0; j < nn; j++) {
      double ph = RBC_TWO_PI * j / nn;
      cs[(size_t)m * nn + j] = cos(m * ph);
      cs[(size_t)mmax * nn + m * nn + j] = sin(m * ph);
    }
  HIP_TRY(hipMalloc(d_p, P.size() * 8));
  HIP_TRY(hipMalloc(d_cs, cs.size() * 8));
  HIP_TRY(hipMemcpy(*d_p, P.data(), P.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(*d_cs, cs.data(), cs.size() * 8, hipMemcpyHostToDevice));
  return RBC3D_OK;
}

int rbc_sh_anal_batch(int nlat, int nlon, int nt, const double *d_g, int idg, int jdg, double *d_a,
                      double *d_b, int ida, int jda, hipStream_t st) {
  int mmax = std::min(nlat, nlon / 2 + 1);
  double *p, *cs;
  RBC_TRY(batch_tables(nlat, nlon, mmax, false, true, &p, &cs));
  size_t lds = sizeof(double) * ((size_t)nlat * nlon + 4 * (size_t)mmax * nlat);
  if (lds > 160 * 1024) return RBC3D_EINVAL;
  hipLaunchKernelGGL(k_anal_batch, dim3(nt), dim3(SH_THREADS), lds, st, nlat, nlon, mmax, p, cs, d_g,
                     idg, jdg, d_a, d_b, ida, jda);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  (void)hipFree(p);
  (void)hipFree(cs);
  return RBC3D_OK;
}

// dPbar/dth on the Gauss colatitudes, [m][n][i] with the 1/2 of m = 0, from
// (1 - x^2) dP_n^m/dx = (n + m) P_{n-1}^m - n x P_n^m in the normalised functions:
// dPbar_n^m/dth = (n cos th Pbar_n^m - sqrt((2n+1)(n^2-m^2)/(2n-1)) Pbar_{n-1}^m) / sin th
int rbc_sh_grad_batch(int nlat, int nlon, int nt, double *d_v, double *d_w, int idv, int jdv,
                      const double *d_a, const double *d_b, int ida, int jda, hipStream_t st) {
  int mmax = std::min(std::min(nlat, nlon / 2 + 1), ida);
  int nmax = std::min(nlat, jda);
  std::vector<double> th, wt, P;
  host_gauss(nlat, th, wt);
  host_pbar(nlat, mmax, th, P);
  std::vector<double> pd(P.size()), pt(P.size());
  for (int m = 0; m < mmax; m++)
    for (int n = 0; n < nlat; n++)
      for (int i = 0; i < nlat; i++) {
        size_t o = ((size_t)m * nlat + n) * nlat + i;
        double d = (double)n * cos(th[i]) * P[o];
        if (n > m) {
          const double nn = n, mm = m;
          d -= sqrt((2 * nn + 1) * (nn * nn - mm * mm) / (2 * nn - 1)) * P[o - nlat];
        }
        const double h = m == 0 ? 0.5 : 1.0;
        pd[o] = h * (d / sin(th[i]));
        pt[o] = h * P[o];
      }
  std::vector<double> cs((size_t)2 * mmax * nlon);
  for (int m = 0; m < mmax; m++)
    for (int j = 0; j < nlon; j++) {
      double ph = RBC_TWO_PI * j / nlon;
      cs[(size_t)m * nlon + j] = cos(m * ph);
      cs[(size_t)mmax * nlon + m * nlon + j] = sin(m * ph);
    }
  size_t lds = sizeof(double) * (2 * (size_t)mmax * nmax + 2 * (size_t)mmax * nlat + (size_t)nlat * nlon);
  if (lds > 160 * 1024) return RBC3D_EINVAL;
  DBuf<double> dpd, dpt, dcs;
  RBC_TRY(dpd.ensure(pd.size()));
  RBC_TRY(dpt.ensure(pt.size()));
  RBC_TRY(dcs.ensure(cs.size()));
  HIP_TRY(hipMemcpy(dpd.p, pd.data(), pd.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dpt.p, pt.data(), pt.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dcs.p, cs.data(), cs.size() * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_grad_batch, dim3(nt), dim3(SH_THREADS), lds, st, nlat, nlon, mmax, nmax,
                     dpd.p, dpt.p, dcs.p, d_v, d_w, idv, jdv, d_a, d_b, ida, jda);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

int rbc_sh_synth_batch(int nlat, int nlon, int nt, double *d_g, int idg, int jdg, const double *d_a,
                       const double *d_b, int ida, int jda, int equi, hipStream_t st) {
  int mmax = std::min(std::min(nlat, nlon / 2 + 1), ida);
  int nmax = std::min(nlat, jda);
  double *p, *cs;
  RBC_TRY(batch_tables(nlat, nlon, mmax, equi, false, &p, &cs));
  size_t lds = sizeof(double) * (2 * (size_t)mmax * nmax + 2 * (size_t)mmax * nlat + (size_t)nlat * nlon);
  if (lds > 160 * 1024) return RBC3D_EINVAL;
  hipLaunchKernelGGL(k_synth_batch, dim3(nt), dim3(SH_THREADS), lds, st, nlat, nlon, mmax, nmax, p,
                     cs, d_g, idg, jdg, d_a, d_b, ida, jda);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  (void)hipFree(p);
  (void)hipFree(cs);
  return RBC3D_OK;
}
