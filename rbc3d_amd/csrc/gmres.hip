// Restarted GMRES on the device with PETSc KSPGMRES semantics (the solver both
// reference solves use: common/ModVelSolver.F90:74-90, common/ModNoSlip.F90:70-87):
// restart 30 unless set, classical Gram-Schmidt without refinement, no preconditioner, the
// residual estimate from the Givens-rotated Hessenberg system, convergence when the
// estimate drops to max(rtol * ||r0||, atol) -- with ||r0|| = ||b|| when the initial
// guess is nonzero, as KSPConvergedDefault does -- explicit residual b - A x at
// restarts.
//
// Vectors stay in HBM; the (restart+1) x restart Hessenberg matrix and the
// rotations live on the host (a few hundred doubles).
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "internal.h"

#define G_T 256
#define G_MAXV (RBC3D_GMRES_MAX_RESTART + 1)

// partial dot products of w with V[0..k) (column stride ld): blockIdx.y takes eight
// vectors, blockIdx.x a slice of the rows; each thread accumulates the eight dot
// products over its elements, the wavefronts reduce them with shuffles and the block
// adds its wavefronts in order into part[j][blockIdx.x]
__global__ void __launch_bounds__(G_T) k_mdot(int n, int k, const double *__restrict__ V, size_t ld,
                                              const double *__restrict__ w, double *part) {
  __shared__ double red[8][G_T / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j0 = blockIdx.y * 8;
  const int nj = min(8, k - j0);
  double s[8];
#pragma unroll
  for (int q = 0; q < 8; q++) s[q] = 0.;
  for (int i = blockIdx.x * G_T + threadIdx.x; i < n; i += gridDim.x * G_T) {
    const double wi = w[i];
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q < nj) s[q] += V[(size_t)(j0 + q) * ld + i] * wi;
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    s[q] = v;
  }
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 8; q++) red[q][wv] = s[q];
  __syncthreads();
  if (threadIdx.x < nj) {
    double t = 0.;
    for (int u = 0; u < G_T / 64; u++) t += red[threadIdx.x][u];
    part[(size_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

// The partial sums of vector j over nb blocks, by one wavefront: lanes stride the
// blocks in order, then a shuffle tree (the order of k_mdot_final)
__device__ __forceinline__ double sum_parts(const double *part, int nb, int lane) {
  double s = 0;
  for (int b = lane; b < nb; b += 64) s += part[b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// w -= sum_j h_j V_j (j < k), with the block's partial sum of the new |w|^2 in
// part[blockIdx.x] (wavefront shuffle trees, then the block's wavefronts in order)
__global__ void __launch_bounds__(G_T) k_maxpy_part(int n, int k, const double *__restrict__ V,
                                                    size_t ld, const double *__restrict__ h,
                                                    double *w, double *part) {
  __shared__ double red[G_T / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double q = 0.;
  if (i < n) {
    double s = w[i];
    for (int j = 0; j < k; j++) s -= h[j] * V[j * ld + i];
    w[i] = s;
    q = s * s;
  }
  for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o, 64);
  if (lane == 0) red[wv] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.;
    for (int u = 0; u < G_T / 64; u++) t += red[u];
    part[blockIdx.x] = t;
  }
}

// t2 = sum of the nb partials (every block forms the same sum, in the same order) and
// w /= sqrt(t2) unless it is zero; block 0 stores t2.  PETSc skips the scaling at a
// happy breakdown (|w| <= 1e-30 |g|), which the host decides from the same value
// afterwards; the vector is then never used, so scaling it regardless changes nothing.
// Block 0 also copies the Hessenberg column (h_0..h_{k-1} from hd, then t2) to the
// host's coherent pinned slot col, which the host reads once the launch is done.
__global__ void __launch_bounds__(G_T) k_scale_norm_sum(int n, int nb, const double *__restrict__ part,
                                                        double *t2, double *w, int k,
                                                        const double *__restrict__ hd, double *col) {
  __shared__ double tot;
  if (threadIdx.x < 64) {
    const double t = sum_parts(part, nb, threadIdx.x);
    if (threadIdx.x == 0) {
      tot = t;
      if (blockIdx.x == 0) *t2 = t;
    }
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    for (int q = threadIdx.x; q < k; q += blockDim.x) col[q] = hd[q];
    if (threadIdx.x == 0) col[k] = tot;
    __threadfence_system();
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double tt = sqrt(tot);
  if (i < n && tt > 0.) w[i] = (1.0 / tt) * w[i];
}

// out[j] = sum over the nb partial sums of vector j: one wavefront per vector
__global__ void k_mdot_final(int k, int nb, const double *__restrict__ part, double *out) {
  const int j = blockIdx.x, lane = threadIdx.x;
  if (j >= k) return;
  double s = 0;
  for (int b = lane; b < nb; b += 64) s += part[(size_t)j * nb + b];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[j] = s;
}

// w -= sum_j h_j V_j
__global__ void k_maxpy(int n, int k, const double *__restrict__ V, size_t ld,
                        const double *__restrict__ h, double *w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int j = 0; j < k; j++) s -= h[j] * V[j * ld + i];
  w[i] = s;
}

// x += sum_j y_j V_j
__global__ void k_update(int n, int k, const double *__restrict__ V, size_t ld,
                         const double *__restrict__ y, double *x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = x[i];
  for (int j = 0; j < k; j++) s += y[j] * V[j * ld + i];
  x[i] = s;
}

__global__ void k_scale(int n, double a, const double *__restrict__ x, double *y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ void k_sub(int n, const double *__restrict__ b, double *r) {  // r = b - r
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}


// out[j] = <V_j, w>, j < k, left on the device
static int dots_dev(hipStream_t st, GmresWork &ws, int n, int k, const double *V, size_t ld,
                    const double *w, double *out) {
  // about 2048 workgroups over the row slices and the groups of eight vectors
  const int groups = (k + 7) / 8;
  const int nb = std::max(1, std::min((n + G_T - 1) / G_T, std::max(32, 2048 / groups)));
  RBC_TRY(ws.part.ensure((size_t)nb * G_MAXV + 1));
  k_mdot<<<dim3(nb, groups), G_T, 0, st>>>(n, k, V, ld, w, ws.part.p);
  k_mdot_final<<<k, 64, 0, st>>>(k, nb, ws.part.p, out);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

static int dots(hipStream_t st, GmresWork &ws, int n, int k, const double *V, size_t ld,
                const double *w, double *h_host) {
  RBC_TRY(ws.hd.ensure(2 * G_MAXV + 2));
  RBC_TRY(dots_dev(st, ws, n, k, V, ld, w, ws.hd.p));
  HIP_TRY(hipMemcpyAsync(h_host, ws.hd.p, sizeof(double) * k, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return RBC3D_OK;
}

// w /= sqrt(t2[0]) unless it is zero.  PETSc skips the scaling at a happy breakdown
// (|w| <= 1e-30 |g|), which the host decides from the same value afterwards; the
// vector is then never used (the cycle ends and the update takes the vectors before
// it), so scaling it regardless changes no result.
__global__ void k_scale_norm(int n, const double *__restrict__ t2, double *w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tt = sqrt(t2[0]);
  if (tt > 0.) w[i] = (1.0 / tt) * w[i];
}

int rbc_gmres(rbc3d_ctx *c, int n, const MatvecFn &matvec, const double *d_b, double *d_x,
              bool x0_nonzero, double rtol, double atol, double dtol, int maxit, int restart,
              int *niter, double *rnorm, int *reason, std::vector<double> *history,
              const MatvecFn *matvec_lo, double lo_below) {
  hipStream_t st = c->stream;
  if (restart < 1 || restart > G_MAXV - 1) return RBC3D_EINVAL;
  GmresWork &ws = c->gmres;
  size_t ld = (size_t)n + 64;
  RBC_TRY(ws.V.ensure(ld * (restart + 1)));
  RBC_TRY(ws.hd.ensure(2 * G_MAXV + 2));
  RBC_TRY(ws.part2.ensure((size_t)(n + G_T - 1) / G_T + 1));
  if (!ws.hpin) HIP_TRY(hipHostMalloc((void **)&ws.hpin, sizeof(double) * 3 * (G_MAXV + 2)));
  if (!ws.hcoh) {
    HIP_TRY(hipHostMalloc((void **)&ws.hcoh, sizeof(double) * 2 * (G_MAXV + 2), hipHostMallocCoherent));
    HIP_TRY(hipHostGetDevicePointer((void **)&ws.hcoh_dev, ws.hcoh, 0));
  }
  for (auto &e : ws.ev)
    if (!e) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  double *V = ws.V.p;
  const int nb = (n + 255) / 256;
  int its = 0;
  double ttol = 0, rnorm0 = 0, res = 0;
  *reason = 0;
  if (!x0_nonzero) HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * n, st));
  std::vector<double> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1);
  bool first = true;
  while (true) {
    // r = b - A x  (KSPInitialResidual); with x = 0 this is b
    if (x0_nonzero || !first) {
      RBC_TRY(matvec(d_x, V));
      k_sub<<<nb, 256, 0, st>>>(n, d_b, V);
    } else {
      HIP_TRY(hipMemcpyAsync(V, d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    double nrm2;
    RBC_TRY(dots(st, ws, n, 1, V, ld, V, &nrm2));
    res = sqrt(nrm2);
    if (first) {
      rnorm0 = res;
      double bnorm = res;
      if (x0_nonzero) {
        // KSPConvergedDefault with a nonzero initial guess: the tolerance is relative
        // to the norm of the right-hand side (PCNONE: B b = b)
        double b2;
        RBC_TRY(dots(st, ws, n, 1, d_b, (size_t)n, d_b, &b2));
        bnorm = sqrt(b2);
      }
      ttol = std::max(rtol * bnorm, atol);
      if (history) history->push_back(res);
    }
    first = false;
    if (!std::isfinite(res)) {  // KSP_DIVERGED_NANORINF
      *reason = -4;
      break;
    }
    if (res == 0.0 || res <= ttol) {
      *reason = 1;
      break;
    }
    k_scale<<<nb, 256, 0, st>>>(n, 1.0 / res, V, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = res;
    int it = 0;
    bool happy = false;
    // One Arnoldi step on the device: w = A V_j, classical Gram-Schmidt h = V^T w,
    // w -= V h, |w|^2 and the normalisation, then (h, |w|^2) copied to a pinned slot
    // and marked by that slot's event.  The step needs nothing from the host, so step
    // j + 1 is queued before the host reads step j's column: the device never waits
    // for the Givens update.  A step queued past the end of the cycle (convergence,
    // breakdown) only writes V_{j+2} and the device scratch, which nothing reads.
    // With a second, cheaper operator (an inexact one), an Arnoldi step uses it once the
    // latest residual estimate is below lo_below times the initial residual: the error a
    // step's product may carry grows as the residual falls (relaxed inexact GMRES), so the
    // first steps take the exact operator and the rest the cheap one.  Residuals at
    // restarts always use the exact operator.
    auto enqueue = [&](int j) -> int {
      double *vj = V + ld * j, *vn = V + ld * (j + 1);
      if (matvec_lo && res <= lo_below * rnorm0)
        RBC_TRY((*matvec_lo)(vj, vn));
      else
        RBC_TRY(matvec(vj, vn));
      // h = V^T w; w -= V h with the partial sums of |w|^2; |w|^2 and w / |w|
      RBC_TRY(dots_dev(st, ws, n, j + 1, V, ld, vn, ws.hd.p));
      k_maxpy_part<<<nb, G_T, 0, st>>>(n, j + 1, V, ld, ws.hd.p, vn, ws.part2.p);
      k_scale_norm_sum<<<nb, G_T, 0, st>>>(n, nb, ws.part2.p, ws.hd.p + j + 1, vn, j + 1, ws.hd.p,
                                           ws.hcoh_dev + (size_t)(j & 1) * (G_MAXV + 2));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipEventRecord(ws.ev[j & 1], st));
      return RBC3D_OK;
    };
    if (its < maxit) RBC_TRY(enqueue(0));
    while (it < restart && its < maxit) {
      if (it + 1 < restart && its + 1 < maxit) RBC_TRY(enqueue(it + 1));
      HIP_TRY(hipEventSynchronize(ws.ev[it & 1]));
      const volatile double *col = ws.hcoh + (size_t)(it & 1) * (G_MAXV + 2);
      // happy breakdown bound as in PETSc: haptol * |g(it)|
      const double hapbnd = fabs(g[it]) * 1.0e-30;
      const double tt = sqrt(col[it + 1]);
      // a non-finite Hessenberg column ends the solve before it reaches x
      // (KSP_DIVERGED_NANORINF): x keeps the last finite update
      bool finite = std::isfinite(tt);
      for (int i = 0; i <= it; i++) finite = finite && std::isfinite(col[i]);
      if (!finite) {
        *reason = -4;
        break;
      }
      for (int i = 0; i <= it; i++) H[(size_t)i * restart + it] = col[i];
      H[(size_t)(it + 1) * restart + it] = tt;
      if (!(tt > hapbnd)) happy = true;
      // apply the previous rotations and form a new one (KSPGMRESUpdateHessenberg)
      for (int i = 0; i < it; i++) {
        double a = H[(size_t)i * restart + it], b = H[(size_t)(i + 1) * restart + it];
        H[(size_t)i * restart + it] = cs[i] * a + sn[i] * b;
        H[(size_t)(i + 1) * restart + it] = -sn[i] * a + cs[i] * b;
      }
      double a = H[(size_t)it * restart + it], b = H[(size_t)(it + 1) * restart + it];
      double tden = sqrt(a * a + b * b);
      if (tden != 0.0) {
        cs[it] = a / tden;
        sn[it] = b / tden;
        H[(size_t)it * restart + it] = tden;
        H[(size_t)(it + 1) * restart + it] = 0.;
        g[it + 1] = -sn[it] * g[it];
        g[it] = cs[it] * g[it];
      } else {
        cs[it] = 1.;
        sn[it] = 0.;
        g[it + 1] = 0.;
      }
      res = fabs(g[it + 1]);
      if (!std::isfinite(res)) {
        *reason = -4;
        break;
      }
      it++;
      its++;
      if (history) history->push_back(res);
      if (res <= ttol) {
        *reason = 1;
        break;
      }
      if (dtol > 0 && res >= dtol * rnorm0) {
        *reason = -1;
        break;
      }
      if (happy) {
        *reason = 1;
        break;
      }
    }
    if (*reason == -4) break;
    // x += V y with H y = g (upper triangular); y goes through the pinned buffer
    double *y = ws.hpin + 2 * (size_t)(G_MAXV + 2);
    for (int i = it - 1; i >= 0; i--) {
      double s = g[i];
      for (int k = i + 1; k < it; k++) s -= H[(size_t)i * restart + k] * y[k];
      y[i] = s / H[(size_t)i * restart + i];
    }
    if (it > 0) {
      HIP_TRY(hipMemcpyAsync(ws.hd.p, y, sizeof(double) * it, hipMemcpyHostToDevice, st));
      k_update<<<nb, 256, 0, st>>>(n, it, V, ld, ws.hd.p, d_x);
      HIP_TRY(hipGetLastError());
      // the copy has read y before the host writes it again: the next cycle starts
      // with a residual norm that synchronises the stream
    }
    if (*reason != 0 || its >= maxit) break;
  }
  HIP_TRY(hipStreamSynchronize(st));
  *niter = its;
  *rnorm = res;
  if (*reason == 0) *reason = -3;  // KSP_DIVERGED_ITS
  // the Krylov basis and the scratch stay allocated for the next solve
  return RBC3D_OK;
}

// ---------------------------------------------------------------------------
int rbc_wall_matmult_dev(rbc3d_ctx *c, const double *d_u, double *d_b);
int rbc_wall_residual_dev(rbc3d_ctx *c);

__global__ void k_neg_assemble(int nuniq, int nvert, const int *__restrict__ last,
                               const double *__restrict__ v, double *u1d) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nuniq) return;
  int i = last[g];
  for (int d = 0; d < 3; d++) u1d[3 * g + d] = -v[i + (size_t)nvert * d];
}
__global__ void k_add_assembled(int nvert, const int *__restrict__ glb, const double *__restrict__ f0,
                                const double *__restrict__ u1d, double *f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvert) return;
  int g = glb[i] - 1;
  for (int d = 0; d < 3; d++) f[i + (size_t)nvert * d] = f0[i + (size_t)nvert * d] + u1d[3 * g + d];
}
__global__ void k_absmax(int n, const double *__restrict__ v, double *out) {
  __shared__ double red[256];
  double m = 0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(v[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Status of a solve from the GMRES stop reason: the outputs are written either way
// (the reference prints the residual and goes on, common/ModNoSlip.F90:120-126); a
// caller that keeps the reference's behaviour ignores RBC3D_ENOCONV.
static int gmres_status(int reason, const char *who) {
  if (reason > 0) return RBC3D_OK;
  if (reason == -4) {
    rbc_set_error(std::string(who) + ": non-finite residual in GMRES (the input or the operator "
                  "holds NaN or Inf)");
    return RBC3D_ENONFINITE;
  }
  rbc_set_error(std::string(who) + (reason == -1 ? ": GMRES diverged"
                                                 : ": GMRES reached maxit before the tolerance"));
  return RBC3D_ENOCONV;
}

__global__ void k_axpy1(int n, const double *__restrict__ d, double *x) {  // x += d
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

// Runs the enclosed operator applications in the given precision form, then restores.
struct PrecScope {
  rbc3d_ctx *c;
  int saved;
  PrecScope(rbc3d_ctx *ctx, int mode) : c(ctx), saved(ctx->prec) { c->prec = mode; }
  ~PrecScope() { c->prec = saved; }
};

#define RBC_IR_MAX_OUTER 12
#ifndef RBC_IR_INNER_MIN
#define RBC_IR_INNER_MIN 5e-9
#endif
#ifndef RBC_IR_SWITCH
#define RBC_IR_SWITCH 1e-2
#endif

// The wall solve of both NoSlipWall entries (x += the solution of L x = b, from x when
// x0nz).  FP64 form: restarted GMRES on the FP64 operator, as the reference.  Split form:
// the FP32 operator is exact only to ~1e-7 relative, so GMRES on it alone drives its own
// residual estimate below 1e-8 while the residual of the FP64 system stalls near 1e-7.
// Here one GMRES builds its Krylov space with the FP64 operator while the residual
// estimate is above 1e-2 |r| and with the split operator below (relaxed inexact GMRES,
// Simoncini & Szyld, SIAM J. Sci. Comput. 25 (2003) 454; Bouras & Fraysse, SIAM J.
// Matrix Anal. Appl. 26 (2005) 660: the operator error a step may carry grows as the
// residual falls), so the split operator's ~1e-7 error enters only in steps whose
// contribution to the solution is already small.  The FP64 residual r = b - L_64 x is
// then computed; while it is above rtol |b|, one more such solve of L d = r (to
// 0.9 rtol |b| / |r|) refines x += d (iterative refinement, Carson & Higham, SIAM J. Sci.
// Comput. 40 (2018) A817).  The reported residual is the FP64 one; the iteration count
// sums the solves.
static int wall_gmres(rbc3d_ctx *c, int nd, const double *d_b, double *d_x, bool x0nz,
                      double rtol, int maxit, int *niter, double *rnorm, int *reason) {
  MatvecFn mv = [c](const double *u, double *b) { return rbc_wall_matmult_dev(c, u, b); };
  GmresWork &gw = c->gmres;
  if (c->prec != RBC3D_PREC_SPLIT)
    return rbc_gmres(c, nd, mv, d_b, d_x, x0nz, rtol, 1e-50, 1e5, maxit, c->gmres_restart, niter,
                     rnorm, reason, &gw.history);
  MatvecFn mv64 = [c](const double *u, double *b) {
    PrecScope fp64(c, RBC3D_PREC_FP64);
    return rbc_wall_matmult_dev(c, u, b);
  };
  hipStream_t st = c->stream;
  RBC_TRY(gw.ir_r.ensure(nd + 1));
  RBC_TRY(gw.ir_d.ensure(nd + 1));
  const int nb = (nd + 255) / 256;
  if (!x0nz) HIP_TRY(hipMemsetAsync(d_x, 0, sizeof(double) * nd, st));
  double b2;
  RBC_TRY(dots(st, gw, nd, 1, d_b, (size_t)nd, d_b, &b2));
  const double ttol = rtol * sqrt(b2);
  int its = 0;
  double res = 0, prev = 0;
  *reason = 0;
  for (int k = 0;; k++) {
    if (k == 0 && !x0nz) {
      HIP_TRY(hipMemcpyAsync(gw.ir_r.p, d_b, sizeof(double) * nd, hipMemcpyDeviceToDevice, st));
    } else {
      {
        PrecScope fp64(c, RBC3D_PREC_FP64);
        RBC_TRY(rbc_wall_matmult_dev(c, d_x, gw.ir_r.p));
      }
      k_sub<<<nb, 256, 0, st>>>(nd, d_b, gw.ir_r.p);
      HIP_TRY(hipGetLastError());
    }
    double r2;
    RBC_TRY(dots(st, gw, nd, 1, gw.ir_r.p, (size_t)nd, gw.ir_r.p, &r2));
    res = sqrt(r2);
    if (!std::isfinite(res)) {
      *reason = -4;
      break;
    }
    if (res <= ttol) {
      *reason = 1;
      break;
    }
    // out of iterations, or refinement no longer contracts (the split operator is too
    // inexact for this system): not converged
    if (its >= maxit || k >= RBC_IR_MAX_OUTER || (k > 0 && !(res < 0.9 * prev))) {
      *reason = -3;
      break;
    }
    prev = res;
    const double tau = std::min(0.5, std::max(0.9 * ttol / res, RBC_IR_INNER_MIN));
    int in_its = 0, in_reason = 0;
    double in_res = 0;
    RBC_TRY(rbc_gmres(c, nd, mv64, gw.ir_r.p, gw.ir_d.p, false, tau, 1e-50, 1e5, maxit - its,
                      c->gmres_restart, &in_its, &in_res, &in_reason, &gw.history, &mv,
                      RBC_IR_SWITCH));
    its += in_its;
    if (in_reason == -4) {
      *reason = -4;
      break;
    }
    k_axpy1<<<nb, 256, 0, st>>>(nd, gw.ir_d.p, d_x);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(st));
  *niter = its;
  *rnorm = res;
  return RBC3D_OK;
}

// The right-hand side of NoSlipWall, -(residual velocity) on the unique vertices, from
// FP64 sums in either form (the split form refines against it).
static int wall_rhs_fp64(rbc3d_ctx *c, double *rhs) {
  WallData &w = c->wall;
  const int nv = w.nvert, nu = w.ndof / 3;
  {
    PrecScope fp64(c, RBC3D_PREC_FP64);
    RBC_TRY(rbc_wall_residual_dev(c));
  }
  k_neg_assemble<<<(nu + 255) / 256, 256, 0, c->stream>>>(nu, nv, w.d_glb.p + nv, c->d_v.p, rhs);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

// NoSlipWall (common/ModNoSlip.F90:45-150)
extern "C" int rbc3d_noslip_wall(rbc3d_ctx *c, double rtol, int maxit, const double *finc0,
                                 int *niter, double *rnorm, int *converged, double *vres_max) {
  if (!c || !niter || !rnorm || !converged) return RBC3D_EINVAL;
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  RBC_TRY(rbc_refresh_moved(c));
  hipStream_t st = c->stream;
  int nv = w.nvert, nd = w.ndof, nu = nd / 3;
  DBuf<double> &f0 = c->gmres.f0, &rhs = c->gmres.rhs, &x = c->gmres.x, &tmpv = c->gmres.tmp;
  RBC_TRY(f0.ensure((size_t)3 * nv));
  RBC_TRY(rhs.ensure(nd + 1));
  RBC_TRY(x.ensure(nd + 1));
  HIP_TRY(hipMemcpyAsync(f0.p, w.d_f.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToDevice, st));
  // rhs = -(residual velocity), assembled on the unique vertices
  RBC_TRY(wall_rhs_fp64(c, rhs.p));
  (void)nu;
  bool x0nz = finc0 != nullptr;
  if (x0nz) HIP_TRY(hipMemcpyAsync(x.p, finc0, sizeof(double) * nd, hipMemcpyHostToDevice, st));
  int reason = 0;
  c->gmres.history.clear();
  RBC_TRY(wall_gmres(c, nd, rhs.p, x.p, x0nz, rtol, maxit, niter, rnorm, &reason));
  *converged = reason > 0;
  // f = f0 + finc (AssembleArray(finc, finc1D, -1))
  k_add_assembled<<<(nv + 255) / 256, 256, 0, st>>>(nv, w.d_glb.p, f0.p, x.p, w.d_f.p);
  HIP_TRY(hipGetLastError());
  if (vres_max) {
    RBC_TRY(rbc_wall_residual_dev(c));
    RBC_TRY(tmpv.ensure(1));
    k_absmax<<<1, 256, 0, st>>>(3 * nv, c->d_v.p, tmpv.p);
    HIP_TRY(hipMemcpyAsync(vres_max, tmpv.p, sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return gmres_status(reason, "rbc3d_noslip_wall");
}

// ---------------------------------------------------------------------------
// NoSlipWall with the learned warm start (no counterpart in the reference): the
// per-wall-node MLP predicts the converged traction from the cells' velocity at the
// vertex, the previous traction and the vertex normal.  GMRES's initial guess is the
// combination of the predicted increment p and the three uniform tractions e_x, e_y,
// e_z with the least residual, x0 = c0 p + sum_d c_d e_d minimising |b - L x0| (the
// images L e_d are computed once per wall and precision): a per-vertex network is
// accurate locally, while a small global bias of its traction would show up in the
// residual at full strength.  Plain GMRES from zero is the fallback when that
// residual is not below |b| or the warm-started solve does not converge.
int rbc_mlp_forward(rbc3d_mlp *m, int n, const float *d_x, float *d_y, hipStream_t st);
int rbc_mlp_wait(rbc3d_mlp *m, hipStream_t st);
int rbc_wall_cell_vel_dev(rbc3d_ctx *c, double *d_out);


__global__ void k_mlp_features(int nv, const double *__restrict__ uc, const double *__restrict__ f0,
                               const double *__restrict__ vn, float *X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  for (int d = 0; d < 3; d++) {
    X[(size_t)9 * i + d] = (float)uc[i + (size_t)nv * d];
    X[(size_t)9 * i + 3 + d] = (float)f0[i + (size_t)nv * d];
    X[(size_t)9 * i + 6 + d] = (float)vn[i + (size_t)nv * d];
  }
}
// x0 on the unique unknowns: the last vertex of each global index supplies it
__global__ void k_mlp_x0(int nu, const int *__restrict__ last, const float *__restrict__ Y,
                         double *x0) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nu) return;
  const int i = last[g];
  for (int d = 0; d < 3; d++) x0[3 * g + d] = (double)Y[(size_t)3 * i + d];
}
__global__ void k_fill_uniform(int nu, double *e) {  // e_d: unit traction along d
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nu) return;
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 3; k++) e[(size_t)d * 3 * nu + 3 * g + k] = k == d ? 1. : 0.;
}
// x = c0 x + c_d on component d (the uniform modes), r = b - (c0 Lx + sum c_d L e_d)
__global__ void k_combine_x0(int nu, double c0, double c1, double c2, double c3, double *x,
                             const double *__restrict__ b, const double *__restrict__ V,
                             size_t ld, double *r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nu) return;
  const int d = i % 3;
  x[i] = c0 * x[i] + (d == 0 ? c1 : (d == 1 ? c2 : c3));
  r[i] = b[i] - (c0 * V[i] + c1 * V[ld + i] + c2 * V[2 * ld + i] + c3 * V[3 * ld + i]);
}
// solve the k x k symmetric system G c = h by Gaussian elimination with partial
// pivoting; directions whose pivot is below tol times the largest diagonal get c = 0
static void small_solve(int k, std::vector<double> G, std::vector<double> h, std::vector<double> &c) {
  double dmax = 0;
  for (int i = 0; i < k; i++) dmax = std::max(dmax, fabs(G[i * k + i]));
  std::vector<int> alive(k, 1);
  for (int i = 0; i < k; i++)
    if (!(fabs(G[i * k + i]) > 1e-12 * dmax)) alive[i] = 0;
  std::vector<int> idx;
  for (int i = 0; i < k; i++)
    if (alive[i]) idx.push_back(i);
  const int m = (int)idx.size();
  std::vector<double> A(m * m), y(m);
  for (int a = 0; a < m; a++) {
    y[a] = h[idx[a]];
    for (int b = 0; b < m; b++) A[a * m + b] = G[idx[a] * k + idx[b]];
  }
  for (int col = 0; col < m; col++) {
    int piv = col;
    for (int r = col + 1; r < m; r++)
      if (fabs(A[r * m + col]) > fabs(A[piv * m + col])) piv = r;
    if (!(fabs(A[piv * m + col]) > 1e-14 * dmax)) continue;
    if (piv != col) {
      for (int b = 0; b < m; b++) std::swap(A[col * m + b], A[piv * m + b]);
      std::swap(y[col], y[piv]);
    }
    for (int r = col + 1; r < m; r++) {
      const double f = A[r * m + col] / A[col * m + col];
      for (int b = col; b < m; b++) A[r * m + b] -= f * A[col * m + b];
      y[r] -= f * y[col];
    }
  }
  std::vector<double> z(m, 0.);
  for (int r = m - 1; r >= 0; r--) {
    if (!(fabs(A[r * m + r]) > 1e-14 * dmax)) continue;
    double s = y[r];
    for (int b = r + 1; b < m; b++) s -= A[r * m + b] * z[b];
    z[r] = s / A[r * m + r];
  }
  c.assign(k, 0.);
  for (int a = 0; a < m; a++) c[idx[a]] = z[a];
}

__global__ void k_add3(int n, double *v, const double *__restrict__ a, double s0, double s1,
                       double s2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] += a[i] + s0;
  v[i + n] += a[i + n] + s1;
  v[i + 2 * (size_t)n] += a[i + 2 * (size_t)n] + s2;
}

extern "C" int rbc3d_noslip_wall_learned(rbc3d_ctx *c, rbc3d_mlp *m, double rtol, int maxit,
                                         int *niter, double *rnorm, int *converged,
                                         double *vres_max, int *fallback) {
  if (!c || !m || !niter || !rnorm || !converged || !fallback) return RBC3D_EINVAL;
  WallData &w = c->wall;
  if (!w.nvert) return RBC3D_ESTATE;
  RBC_TRY(rbc_refresh_moved(c));
  if (!w.prepared) RBC_TRY(rbc_wall_prepare(c));
  hipStream_t st = c->stream;
  const int nv = w.nvert, nd = w.ndof, nu = nd / 3;
  GmresWork &gw = c->gmres;
  DBuf<double> &f0 = gw.f0, &uc = gw.uc, &rhs = gw.rhs, &x = gw.x, &r0 = gw.r0, &LV = gw.LV;
  DBuf<float> &X = gw.X, &Y = gw.Y;
  RBC_TRY(f0.ensure((size_t)3 * nv));
  RBC_TRY(uc.ensure((size_t)3 * nv));
  RBC_TRY(rhs.ensure(nd + 1));
  RBC_TRY(x.ensure(nd + 1));
  RBC_TRY(r0.ensure(nd + 1));
  RBC_TRY(X.ensure((size_t)9 * nv));
  RBC_TRY(Y.ensure((size_t)3 * nv));
  HIP_TRY(hipMemcpyAsync(f0.p, w.d_f.p, sizeof(double) * 3 * nv, hipMemcpyDeviceToDevice, st));
  // residual velocity = cells' velocity + walls' velocity from f0 + background flow,
  // the cells' part kept as the network's input
  RBC_TRY(rbc_wall_cell_vel_dev(c, uc.p));
  HIP_TRY(hipMemsetAsync(c->d_v.p, 0, sizeof(double) * 3 * nv, st));
  const double c1 = 1. / (4. * RBC_PI);
  RBC_TRY(rbc_add_int_on_walls_dev(c, c1, nv, c->d_tx.p, c->d_tindx.p, c->d_tact.p, c->d_tA.p,
                                   c->d_v.p, true));
  RBC_TRY(rbc_pme_spread(c, c1, 0., 0, 1));
  RBC_TRY(rbc_pme_transform(c));
  RBC_TRY(rbc_pme_gather_wall(c, c->d_tA.p, c->d_v.p));
  k_add3<<<(nv + 255) / 256, 256, 0, st>>>(nv, c->d_v.p, uc.p, c->conf.vBkg[0], c->conf.vBkg[1],
                                           c->conf.vBkg[2]);
  k_neg_assemble<<<(nu + 255) / 256, 256, 0, st>>>(nu, nv, w.d_glb.p + nv, c->d_v.p, rhs.p);
  // the split form refines against the FP64 right-hand side; the features above keep the
  // cells' velocity of the current form
  if (c->prec == RBC3D_PREC_SPLIT) RBC_TRY(wall_rhs_fp64(c, rhs.p));
  // the network's prediction of the increment
  k_mlp_features<<<(nv + 255) / 256, 256, 0, st>>>(nv, uc.p, f0.p, w.d_vnorm.p, X.p);
  HIP_TRY(hipGetLastError());
  // the network's parameters may still be written by work queued on its own stream
  // (an Adam step): order the forward after it
  RBC_TRY(rbc_mlp_wait(m, st));
  RBC_TRY(rbc_mlp_forward(m, nv, X.p, Y.p, st));
  k_mlp_x0<<<(nu + 255) / 256, 256, 0, st>>>(nu, w.d_glb.p + nv, Y.p, x.p);
  HIP_TRY(hipGetLastError());
  // least-residual combination of the prediction and the uniform tractions:
  // V = [L p, L e_x, L e_y, L e_z], G = V^T V, h = V^T b, x0 = [p e] c
  const size_t ld = (size_t)nd;
  if (w.d_lmodes.n < 3 * ld || w.lmodes_prec != c->prec || !w.lmodes_ready) {
    DBuf<double> e;
    RBC_TRY(e.ensure(3 * ld));
    RBC_TRY(w.d_lmodes.ensure(3 * ld));
    k_fill_uniform<<<(nu + 255) / 256, 256, 0, st>>>(nu, e.p);
    HIP_TRY(hipGetLastError());
    for (int d = 0; d < 3; d++) RBC_TRY(rbc_wall_matmult_dev(c, e.p + d * ld, w.d_lmodes.p + d * ld));
    HIP_TRY(hipStreamSynchronize(st));
    w.lmodes_prec = c->prec;
    w.lmodes_ready = true;
  }
  RBC_TRY(LV.ensure(4 * ld));
  RBC_TRY(rbc_wall_matmult_dev(c, x.p, LV.p));
  HIP_TRY(hipMemcpyAsync(LV.p + ld, w.d_lmodes.p, sizeof(double) * 3 * ld, hipMemcpyDeviceToDevice,
                         st));
  std::vector<double> G(16), h(4), cc;
  for (int j = 0; j < 4; j++) {
    double col[4];
    RBC_TRY(dots(st, c->gmres, nd, 4, LV.p, ld, LV.p + j * ld, col));
    for (int i = 0; i < 4; i++) G[i * 4 + j] = col[i];
  }
  {
    double hb[4];
    RBC_TRY(dots(st, c->gmres, nd, 4, LV.p, ld, rhs.p, hb));
    for (int i = 0; i < 4; i++) h[i] = hb[i];
  }
  small_solve(4, G, h, cc);
  k_combine_x0<<<(nd + 255) / 256, 256, 0, st>>>(nu, cc[0], cc[1], cc[2], cc[3], x.p, rhs.p, LV.p, ld,
                                                 r0.p);
  HIP_TRY(hipGetLastError());
  double b2, r2;
  RBC_TRY(dots(st, c->gmres, nd, 1, rhs.p, (size_t)nd, rhs.p, &b2));
  RBC_TRY(dots(st, c->gmres, nd, 1, r0.p, (size_t)nd, r0.p, &r2));
  int reason = 0, its = 0;
  *fallback = 0;
  c->gmres.history.clear();
  if (r2 < b2) RBC_TRY(wall_gmres(c, nd, rhs.p, x.p, true, rtol, maxit, &its, rnorm, &reason));
  *niter = its;
  if (reason <= 0) {  // plain GMRES from zero
    *fallback = 1;
    int its2 = 0;
    RBC_TRY(wall_gmres(c, nd, rhs.p, x.p, false, rtol, maxit, &its2, rnorm, &reason));
    *niter += its2;
  }
  *converged = reason > 0;
  k_add_assembled<<<(nv + 255) / 256, 256, 0, st>>>(nv, w.d_glb.p, f0.p, x.p, w.d_f.p);
  HIP_TRY(hipGetLastError());
  if (vres_max) {
    DBuf<double> &tmpv = gw.tmp;
    RBC_TRY(rbc_wall_residual_dev(c));
    RBC_TRY(tmpv.ensure(1));
    k_absmax<<<1, 256, 0, st>>>(3 * nv, c->d_v.p, tmpv.p);
    HIP_TRY(hipMemcpyAsync(vres_max, tmpv.p, sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  return gmres_status(reason, "rbc3d_noslip_wall_learned");
}

extern "C" int rbc3d_gmres_history(rbc3d_ctx *c, double *h, int maxn, int *n) {
  if (!c || !n || maxn < 0 || (maxn > 0 && !h)) return RBC3D_EINVAL;
  const std::vector<double> &hv = c->gmres.history;
  *n = (int)hv.size();
  for (int i = 0; i < std::min(maxn, *n); i++) h[i] = hv[i];
  return RBC3D_OK;
}

extern "C" int rbc3d_get_wall_density(rbc3d_ctx *c, double *f) {
  if (!c || !f) return RBC3D_EINVAL;
  if (!c->wall.nvert) return RBC3D_ESTATE;
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(f, c->wall.d_f.p, sizeof(double) * 3 * c->wall.nvert, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

// Solve_RBC_Vel (common/ModVelSolver.F90:44-132): v(nPoint,3) on return; the cell
// double-layer density is then set to v (common/ModTimeInt.F90:773-790)
extern "C" int rbc3d_cell_rhs_dev(rbc3d_ctx *c, double *d_rhs);
extern "C" int rbc3d_cell_matmult_dev(rbc3d_ctx *c, const double *d_u, double *d_b);
int rbc_set_g_nodes(rbc3d_ctx *c, const double *d_gnodes);

// ---------------------------------------------------------------------------
// Deflation of the cell BIE (common/ModVelSolver.F90:144-449, the reference's Deflate
// switch): per cell the six solid-body modes (DeflateGetSBModes), the right-hand side
// and matvec corrections (DeflateRHS, DeflateMatMult) and the recovery of the velocity
// (DeflateRecoverSol).  One workgroup per cell; the surface integrals sum(. * ds) are
// workgroup reductions over the cell's nodes (ds = detJ w, the d_ds array).  Node
// vectors are [3][np] over the concatenated cells, each cell at its own offset.
#define DF_T 256

__device__ __forceinline__ double df_block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0;
  for (int k = 0; k < DF_T / 64; k++) t += red[k];
  return t;
}

// <a, b> over cell cr: sum_q sum_d a[q + lda d] b[q + ldb d] ds[q]
__device__ double df_dot(const CellRec &cr, const double *a, size_t lda, const double *b,
                         size_t ldb, const double *ds, double *red) {
  const int npt = cr.nlat * cr.nlon;
  double s = 0;
  for (int q = threadIdx.x; q < npt; q += DF_T) {
    double t = 0;
    for (int d = 0; d < 3; d++) t += a[q + lda * d] * b[q + ldb * d];
    s += t * ds[q];
  }
  return df_block_sum(s, red);
}

// DeflateGetSBModes (:354-449): translations and rotations about the centroid,
// Gram-Schmidt orthonormalised (Proj, :603-621), into sbq[(m * 3 + d) * np + node]
__global__ void __launch_bounds__(DF_T) k_defl_modes(int nrbc, const CellRec *__restrict__ cells,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ ds, int np,
                                                     double *sbq) {
  __shared__ double red[DF_T / 64];
  const size_t NP = np;
  for (int c = blockIdx.x; c < nrbc; c += gridDim.x) {
    const CellRec cr = cells[c];
    const int npt = cr.nlat * cr.nlon;
    const double *X = x + cr.off, *S = ds + cr.off;
    double ar = 0, c3[3] = {0, 0, 0};
    for (int q = threadIdx.x; q < npt; q += DF_T) {
      ar += S[q];
      for (int d = 0; d < 3; d++) c3[d] += S[q] * X[q + NP * d];
    }
    ar = df_block_sum(ar, red);
    for (int d = 0; d < 3; d++) c3[d] = df_block_sum(c3[d], red) / ar;
    double *Q = sbq + cr.off;  // mode m, component d at Q + (m * 3 + d) * NP
    for (int q = threadIdx.x; q < npt; q += DF_T) {
      const double X0 = X[q] - c3[0], X1 = X[q + NP] - c3[1], X2 = X[q + 2 * NP] - c3[2];
      const double v[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, -X2, X1}, {X2, 0, -X0}, {-X1, X0, 0}};
      for (int m = 0; m < 6; m++)
        for (int d = 0; d < 3; d++) Q[q + NP * (3 * m + d)] = v[m][d];
    }
    __syncthreads();
    // qq_m = pp_m - sum_{n < m} <pp_m, qq_n> / <qq_n, qq_n> qq_n (pp_m is Q_m before the
    // subtractions, kept in step by subtracting the projections all at the end)
    for (int m = 1; m < 6; m++) {
      double coef[6];
      for (int n = 0; n < m; n++) {
        const double fn = df_dot(cr, Q + NP * 3 * m, NP, Q + NP * 3 * n, NP, S, red);
        const double fd = df_dot(cr, Q + NP * 3 * n, NP, Q + NP * 3 * n, NP, S, red);
        coef[n] = fn / fd;
      }
      __syncthreads();
      for (int q = threadIdx.x; q < npt; q += DF_T)
        for (int d = 0; d < 3; d++) {
          double t = Q[q + NP * (3 * m + d)];
          for (int n = 0; n < m; n++) t = t - coef[n] * Q[q + NP * (3 * n + d)];
          Q[q + NP * (3 * m + d)] = t;
        }
      __syncthreads();
    }
    double nrm[6];
    for (int m = 0; m < 6; m++) nrm[m] = sqrt(df_dot(cr, Q + NP * 3 * m, NP, Q + NP * 3 * m, NP, S, red));
    __syncthreads();
    for (int q = threadIdx.x; q < npt; q += DF_T)
      for (int m = 0; m < 6; m++)
        for (int d = 0; d < 3; d++) Q[q + NP * (3 * m + d)] = Q[q + NP * (3 * m + d)] / nrm[m];
    __syncthreads();
  }
}

// DeflateRHS (:219-274): v += alpha/(1 - alpha) a3 <v, a3> / area, alpha = Bcoef/Acoef
__global__ void __launch_bounds__(DF_T) k_defl_rhs(int nrbc, const CellRec *__restrict__ cells,
                                                   const double *__restrict__ a3,
                                                   const double *__restrict__ ds, int np, double *v) {
  __shared__ double red[DF_T / 64];
  const size_t NP = np;
  for (int c = blockIdx.x; c < nrbc; c += gridDim.x) {
    const CellRec cr = cells[c];
    const int npt = cr.nlat * cr.nlon;
    const double alpha = cr.Bcoef / cr.Acoef;
    double *V = v + cr.off;
    const double *A = a3 + cr.off;
    const double fac = df_dot(cr, V, NP, A, NP, ds + cr.off, red);
    __syncthreads();
    for (int q = threadIdx.x; q < npt; q += DF_T)
      for (int d = 0; d < 3; d++)
        V[q + NP * d] = V[q + NP * d] + alpha / (1. - alpha) * A[q + NP * d] * fac / cr.area;
    __syncthreads();
  }
}

// DeflateMatMult (:281-349): v += alpha a3 <g, a3> / area - alpha sum_m qq_m <g, qq_m>
__global__ void __launch_bounds__(DF_T) k_defl_matmult(int nrbc, const CellRec *__restrict__ cells,
                                                       const double *__restrict__ a3,
                                                       const double *__restrict__ ds,
                                                       const double *__restrict__ sbq,
                                                       const double *__restrict__ g, int np,
                                                       double *v) {
  __shared__ double red[DF_T / 64];
  const size_t NP = np;
  for (int c = blockIdx.x; c < nrbc; c += gridDim.x) {
    const CellRec cr = cells[c];
    const int npt = cr.nlat * cr.nlon;
    const double alpha = cr.Bcoef / cr.Acoef;
    const double *S = ds + cr.off, *A = a3 + cr.off, *G = g + cr.off, *Q = sbq + cr.off;
    const double f0 = df_dot(cr, G, NP, A, NP, S, red);
    double fm[6];
    for (int m = 0; m < 6; m++) fm[m] = df_dot(cr, G, NP, Q + NP * 3 * m, NP, S, red);
    __syncthreads();
    double *V = v + cr.off;
    for (int q = threadIdx.x; q < npt; q += DF_T)
      for (int d = 0; d < 3; d++) {
        double t = V[q + NP * d] + alpha * A[q + NP * d] * f0 / cr.area;
        for (int m = 0; m < 6; m++) t = t - alpha * Q[q + NP * (3 * m + d)] * fm[m];
        V[q + NP * d] = t;
      }
    __syncthreads();
  }
}

// DeflateRecoverSol (:144-214) in place on the node velocities v; a rigid cell
// (viscRat <= 0) keeps its solution before the recovery as the double-layer potential g
__global__ void __launch_bounds__(DF_T) k_defl_recover(int nrbc, const CellRec *__restrict__ cells,
                                                       const double *__restrict__ visc,
                                                       const double *__restrict__ ds,
                                                       const double *__restrict__ sbq, int np,
                                                       double *v, double *g) {
  __shared__ double red[DF_T / 64];
  const size_t NP = np;
  for (int c = blockIdx.x; c < nrbc; c += gridDim.x) {
    const CellRec cr = cells[c];
    const int npt = cr.nlat * cr.nlon;
    const double alpha = cr.Bcoef / cr.Acoef;
    const bool finite = visc[cr.celltype - 1] > 0;
    const double afac = finite ? alpha / (alpha + 1.) : -1. / 2.;
    double *V = v + cr.off;
    const double *S = ds + cr.off, *Q = sbq + cr.off;
    double fm[6];
    for (int m = 0; m < 6; m++) fm[m] = df_dot(cr, V, NP, Q + NP * 3 * m, NP, S, red);
    __syncthreads();
    for (int q = threadIdx.x; q < npt; q += DF_T)
      for (int d = 0; d < 3; d++) {
        const double vv = V[q + NP * d];
        if (!finite) g[cr.off + q + NP * d] = vv;
        double t = finite ? vv : 0.;
        for (int m = 0; m < 6; m++) t = t - afac * Q[q + NP * (3 * m + d)] * fm[m];
        V[q + NP * d] = t;
      }
    __syncthreads();
  }
}

// g := v on the nodes of cells of finite viscosity (common/ModTimeInt.F90:780-784)
__global__ void k_g_from_v_finite(int np, const int *__restrict__ node_cell,
                                  const CellRec *__restrict__ cells, const double *__restrict__ visc,
                                  const double *__restrict__ v, double *g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (!(visc[cells[node_cell[p]].celltype - 1] > 0)) return;
  for (int d = 0; d < 3; d++) g[p + (size_t)np * d] = v[p + (size_t)np * d];
}

int rbc_defl_rhs(rbc3d_ctx *c, double *d_v) {
  if (!c->deflate || !c->nrbc) return RBC3D_OK;
  k_defl_rhs<<<std::min(c->nrbc, 4096), DF_T, 0, c->stream>>>(c->nrbc, c->d_cells.p, c->d_a3.p,
                                                               c->d_ds.p, c->np, d_v);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}
int rbc_defl_matmult(rbc3d_ctx *c, const double *d_g, double *d_v) {
  if (!c->deflate || !c->nrbc) return RBC3D_OK;
  if (!c->sbq_ready) {
    rbc_set_error("deflated matvec before the solid-body modes (rbc3d_solve_rbc_vel builds them)");
    return RBC3D_ESTATE;
  }
  k_defl_matmult<<<std::min(c->nrbc, 4096), DF_T, 0, c->stream>>>(
      c->nrbc, c->d_cells.p, c->d_a3.p, c->d_ds.p, c->d_sbq.p, d_g, c->np, d_v);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_set_deflation(rbc3d_ctx *c, int on) {
  if (!c) return RBC3D_EINVAL;
  c->deflate = on != 0;
  return RBC3D_OK;
}

extern "C" int rbc3d_sb_modes(rbc3d_ctx *c, double *modes) {
  if (!c || !modes) return RBC3D_EINVAL;
  if (!c->sbq_ready) return RBC3D_ESTATE;
  HIP_TRY(hipMemcpyAsync(modes, c->d_sbq.p, sizeof(double) * 18 * (size_t)c->np,
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_solve_rbc_vel(rbc3d_ctx *c, double *v, int *niter, double *rnorm) {
  if (!c || !niter || !rnorm) return RBC3D_EINVAL;
  if (c->nrbc == 0) return RBC3D_ESTATE;
  RBC_TRY(rbc_refresh_moved(c));
  hipStream_t st = c->stream;
  int dof;
  RBC_TRY(rbc3d_cell_dof(c, &dof));
  DBuf<double> &rhs = c->gmres.rhs;
  RBC_TRY(rhs.ensure(dof + 1));
  if (c->d_cell_sol.n < (size_t)dof + 1) c->cell_sol_init = false;  // a new buffer
  RBC_TRY(c->d_cell_sol.ensure(dof + 1));
  if (!c->cell_sol_init) {
    HIP_TRY(hipMemsetAsync(c->d_cell_sol.p, 0, sizeof(double) * dof, st));
    c->cell_sol_init = true;
  }
  RBC_TRY(rbc3d_cell_rhs_dev(c, rhs.p));
  if (c->deflate) {  // DeflateGetSBModes after Compute_Rhs (common/ModVelSolver.F90:98-101)
    RBC_TRY(c->d_sbq.ensure(18 * (size_t)c->np + 1));
    k_defl_modes<<<std::min(c->nrbc, 4096), DF_T, 0, st>>>(c->nrbc, c->d_cells.p, c->d_x.p,
                                                           c->d_ds.p, c->np, c->d_sbq.p);
    HIP_TRY(hipGetLastError());
    c->sbq_ready = true;
  }
  bool all_one = true;
  for (int l = 0; l < c->conf.ntypes; l++) all_one = all_one && c->conf.viscRat[l] == 1.0;
  *niter = 0;
  *rnorm = 0;
  int reason = 1;
  if (all_one) {
    HIP_TRY(hipMemcpyAsync(c->d_cell_sol.p, rhs.p, sizeof(double) * dof, hipMemcpyDeviceToDevice, st));
  } else {
    auto mv = [c](const double *u, double *b) { return rbc3d_cell_matmult_dev(c, u, b); };
    RBC_TRY(rbc_gmres(c, dof, mv, rhs.p, c->d_cell_sol.p, true, 1e-11, 1e-50, 1e5, 200, c->gmres_restart, niter,
                      rnorm, &reason, nullptr));
  }
  RBC_TRY(c->d_shbuf.ensure((size_t)3 * c->np + 1));
  RBC_TRY(rbc_sh_glob(c, c->d_cell_sol.p, c->d_shbuf.p, -1));
  bool all_finite = true;
  for (int l = 0; l < c->conf.ntypes; l++) all_finite = all_finite && c->conf.viscRat[l] > 0;
  if (c->deflate || !all_finite) {
    RBC_TRY(c->d_visc.ensure(8));
    HIP_TRY(hipMemcpyAsync(c->d_visc.p, c->conf.viscRat, sizeof(double) * c->conf.ntypes,
                           hipMemcpyHostToDevice, st));
  }
  if (c->deflate) {  // DeflateRecoverSol (common/ModVelSolver.F90:122-125)
    k_defl_recover<<<std::min(c->nrbc, 4096), DF_T, 0, st>>>(c->nrbc, c->d_cells.p, c->d_visc.p,
                                                             c->d_ds.p, c->d_sbq.p, c->np,
                                                             c->d_shbuf.p, c->d_g.p);
    HIP_TRY(hipGetLastError());
  }
  if (all_finite) {
    RBC_TRY(rbc_set_g_nodes(c, c->d_shbuf.p));
  } else {  // a rigid cell keeps its potential (the recovery's, or the last one set)
    k_g_from_v_finite<<<(c->np + 255) / 256, 256, 0, st>>>(c->np, c->d_node_cell.p, c->d_cells.p,
                                                           c->d_visc.p, c->d_shbuf.p, c->d_g.p);
    HIP_TRY(hipGetLastError());
    RBC_TRY(rbc_update_density(c, 2));
  }
  if (v) HIP_TRY(hipMemcpyAsync(v, c->d_shbuf.p, sizeof(double) * 3 * c->np, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return gmres_status(reason, "rbc3d_solve_rbc_vel");
}
