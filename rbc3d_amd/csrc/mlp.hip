// The learned warm start of the no-slip wall solve: a per-wall-node MLP
// (9 -> 64 -> 64 -> 64 -> 3, GELU) with BF16 weights and activations and FP32
// accumulation on the BF16 matrix cores, its training kernels (forward, backward,
// gradient reduction, Adam) and its inference kernel.
//
// Inputs per wall vertex: the cells' velocity there (3), the previous traction (3)
// and the vertex normal (3), standardised with a mean and 1/std per feature.  The
// network's output, scaled per component, is the traction increment over the
// previous traction; the predicted traction is f0 + scale * y.
//
// MFMA formulation.  A layer computes Z^T = W H^T + b with v_mfma_f32_16x16x16bf16_1k:
// A = W (rows: output features, k: input features), B = H^T (k: input features,
// columns: 16 nodes).  Lane l supplies A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j]
// [l % 16] and holds D[4 (l / 16) + i][l % 16]: for node l % 16 it holds the features
// 16 mt + 4 (l / 16) + i of output tile mt, which is exactly the B operand of K-step mt
// of the next layer, so activations never leave the registers.  The backward pass
// uses the same layout with W^T; the weight gradients dW = dZ^T H (k: nodes) take
// both operands through an LDS transpose and accumulate in LDS, one workgroup of one
// wavefront per partial sum, reduced over workgroups in a fixed order.
#include <cmath>
#include <vector>

#include "internal.h"

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

#define MLP_NIN 9
#define MLP_H 64
#define MLP_NOUT 3
// FP32 parameter vector: W1[64][9] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
enum {
  MP_W1 = 0,
  MP_B1 = MP_W1 + MLP_H * MLP_NIN,
  MP_W2 = MP_B1 + MLP_H,
  MP_B2 = MP_W2 + MLP_H * MLP_H,
  MP_W3 = MP_B2 + MLP_H,
  MP_B3 = MP_W3 + MLP_H * MLP_H,
  MP_W4 = MP_B3 + MLP_H,
  MP_B4 = MP_W4 + MLP_NOUT * MLP_H,
  MP_N = MP_B4 + MLP_NOUT
};
// BF16 weights padded to MFMA tiles: W1p[64][16], W2[64][64], W3[64][64], W4p[16][64]
enum { WB_1 = 0, WB_2 = WB_1 + 64 * 16, WB_3 = WB_2 + 64 * 64, WB_4 = WB_3 + 64 * 64, WB_N = WB_4 + 16 * 64 };
// gradient accumulators in 16 x 16 tiles: layer 1 (4 x 1 tiles), 2 (4 x 4), 3 (4 x 4),
// 4 (1 x 4), then the biases b1 b2 b3 (64 each) and b4 (16, 3 used)
enum { GA_1 = 0, GA_2 = GA_1 + 4 * 256, GA_3 = GA_2 + 16 * 256, GA_4 = GA_3 + 16 * 256,
       GA_B = GA_4 + 4 * 256, GA_N = GA_B + 3 * 64 + 16 };
#define MLP_TS 72  // LDS row stride (bf16) of the transpose tiles: 64 + 8

struct rbc3d_mlp {
  hipStream_t stream = nullptr;  // own stream, or `ext` once rbc3d_mlp_set_stream was
  hipStream_t ext = nullptr;     // called (which may name the null stream)
  bool use_ext = false;
  hipStream_t st() const { return use_ext ? ext : stream; }
  DBuf<float> p, m1, m2, grad, part, norm;
  DBuf<unsigned short> wb;
  DBuf<double> lpart;
  hipEvent_t ev = nullptr;  // marks the last work queued on st() (rbc_mlp_wait)
  int t = 0;
};

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// GELU, tanh form (Hendrycks and Gimpel, 2016), and its derivative
#define GELU_C 0.7978845608028654f  // sqrt(2/pi)
#define GELU_A 0.044715f
__device__ __forceinline__ float mlp_tanh(float u) {
  const float e = __expf(2.f * u);
  return 1.f - 2.f / (e + 1.f);
}
__device__ __forceinline__ float gelu(float x) {
  const float t = mlp_tanh(GELU_C * (x + GELU_A * x * x * x));
  return 0.5f * x * (1.f + t);
}
__device__ __forceinline__ float gelu_d(float x) {
  const float t = mlp_tanh(GELU_C * (x + GELU_A * x * x * x));
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * GELU_C * (1.f + 3.f * GELU_A * x * x);
}

__device__ __forceinline__ mlp_f4 mfma(mlp_s4 a, mlp_s4 b, mlp_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A operand from a row-major BF16 matrix W[rows][ld]: lane supplies W[r0 + l%16][k0 + 4q + j]
__device__ __forceinline__ mlp_s4 ldA(const unsigned short *W, int ld, int r0, int k0, int lane) {
  const unsigned short *p = W + (size_t)(r0 + (lane & 15)) * ld + k0 + 4 * (lane >> 4);
  mlp_s4 a;
  a[0] = (short)p[0];
  a[1] = (short)p[1];
  a[2] = (short)p[2];
  a[3] = (short)p[3];
  return a;
}
// A operand of W^T: lane supplies W^T[r0 + l%16][k0 + 4q + j] = W[k0 + 4q + j][r0 + l%16]
__device__ __forceinline__ mlp_s4 ldAT(const unsigned short *W, int ld, int r0, int k0, int lane) {
  const unsigned short *p = W + (size_t)(k0 + 4 * (lane >> 4)) * ld + r0 + (lane & 15);
  mlp_s4 a;
  a[0] = (short)p[0];
  a[1] = (short)p[ld];
  a[2] = (short)p[2 * ld];
  a[3] = (short)p[3 * ld];
  return a;
}
__device__ __forceinline__ mlp_s4 tobf(mlp_f4 v) {
  mlp_s4 r;
  r[0] = (short)f2bf(v[0]);
  r[1] = (short)f2bf(v[1]);
  r[2] = (short)f2bf(v[2]);
  r[3] = (short)f2bf(v[3]);
  return r;
}

// Forward pass of one tile of 16 nodes (node = node0 + lane % 16); keeps the
// pre-activations z and the BF16 activations h (B-operand layout) for the backward.
struct MlpTile {
  mlp_s4 x;         // inputs (features 4q + j)
  mlp_f4 z[3][4];   // pre-activations of the hidden layers, tile mt
  mlp_s4 h[3][4];   // BF16 activations of the hidden layers, K-step mt
  mlp_f4 y;         // outputs 4q + i (valid: q == 0, i < 3)
};

__device__ __forceinline__ void mlp_forward(const unsigned short *sW, const float *p,
                                            const float *norm, const float *X, int n, int node0,
                                            int lane, MlpTile &T) {
  const int q = lane >> 4, node = node0 + (lane & 15);
  #pragma unroll
  for (int j = 0; j < 4; j++) {
    const int f = 4 * q + j;
    float v = 0.f;
    if (f < MLP_NIN && node < n) v = (X[(size_t)node * MLP_NIN + f] - norm[f]) * norm[MLP_NIN + f];
    T.x[j] = (short)f2bf(v);
  }
  // layer 1: W1p[64][16]
  #pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ldA(sW + WB_1, 16, 16 * mt, 0, lane), T.x, acc);
    #pragma unroll
    for (int i = 0; i < 4; i++) acc[i] += p[MP_B1 + 16 * mt + 4 * q + i];
    T.z[0][mt] = acc;
    mlp_f4 g;
    #pragma unroll
    for (int i = 0; i < 4; i++) g[i] = gelu(acc[i]);
    T.h[0][mt] = tobf(g);
  }
  // layers 2 and 3: 64 x 64
  #pragma unroll
  for (int L = 1; L < 3; L++) {
    const unsigned short *W = sW + (L == 1 ? WB_2 : WB_3);
    const float *b = p + (L == 1 ? MP_B2 : MP_B3);
    #pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
      #pragma unroll
      for (int kt = 0; kt < 4; kt++) acc = mfma(ldA(W, 64, 16 * mt, 16 * kt, lane), T.h[L - 1][kt], acc);
      #pragma unroll
      for (int i = 0; i < 4; i++) acc[i] += b[16 * mt + 4 * q + i];
      T.z[L][mt] = acc;
      mlp_f4 g;
      #pragma unroll
      for (int i = 0; i < 4; i++) g[i] = gelu(acc[i]);
      T.h[L][mt] = tobf(g);
    }
  }
  // layer 4: W4p[16][64]
  mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
  #pragma unroll
  for (int kt = 0; kt < 4; kt++) acc = mfma(ldA(sW + WB_4, 64, 0, 16 * kt, lane), T.h[2][kt], acc);
  #pragma unroll
  for (int i = 0; i < 4; i++) acc[i] += (q == 0 && i < MLP_NOUT) ? p[MP_B4 + i] : 0.f;
  T.y = acc;
}

__device__ __forceinline__ void load_weights(unsigned short *sW, const unsigned short *wb) {
  for (int t = threadIdx.x; t < WB_N / 4; t += blockDim.x)
    reinterpret_cast<uint2 *>(sW)[t] = reinterpret_cast<const uint2 *>(wb)[t];
}

// inference: Y[node][c] = scale_c * y_c (the predicted traction increment)
__global__ void __launch_bounds__(256) k_mlp_forward(int n, const float *__restrict__ X,
                                                     const unsigned short *__restrict__ wb,
                                                     const float *__restrict__ p,
                                                     const float *__restrict__ norm, float *Y) {
  __shared__ __align__(16) unsigned short sW[WB_N];
  load_weights(sW, wb);
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ntile = (n + 15) / 16, nw = gridDim.x * (blockDim.x >> 6);
  for (int tile = blockIdx.x * (blockDim.x >> 6) + wv; tile < ntile; tile += nw) {
    MlpTile T;
    mlp_forward(sW, p, norm, X, n, 16 * tile, lane, T);
    const int node = 16 * tile + (lane & 15);
    if (lane < 16 && node < n)
      #pragma unroll
      for (int i = 0; i < MLP_NOUT; i++) Y[(size_t)node * MLP_NOUT + i] = T.y[i] * norm[2 * MLP_NIN + i];
  }
}

// transpose tile: S[node][feature] in LDS, row stride MLP_TS
__device__ __forceinline__ void st_tile(unsigned short *S, int f0, mlp_s4 v, int lane) {
  // lane holds node l%16, features f0 + 4q + j
  unsigned short *r = S + (lane & 15) * MLP_TS + f0 + 4 * (lane >> 4);
  #pragma unroll
  for (int j = 0; j < 4; j++) r[j] = (unsigned short)v[j];
}
// operand from a transposed tile: lane supplies M[node 4q + j][c0 + l%16]
__device__ __forceinline__ mlp_s4 ld_tile(const unsigned short *S, int c0, int lane) {
  const unsigned short *r = S + 4 * (lane >> 4) * MLP_TS + c0 + (lane & 15);
  mlp_s4 v;
  #pragma unroll
  for (int j = 0; j < 4; j++) v[j] = (short)r[j * MLP_TS];
  return v;
}
// dW tile (mt, nt) += dZ^T H over the 16 nodes of the tile, accumulated in LDS
__device__ __forceinline__ void acc_dw(float *acc, const unsigned short *SZ, int z0,
                                       const unsigned short *SH, int h0, int lane) {
  float *t = acc + 16 * (4 * (lane >> 4)) + (lane & 15);
  mlp_f4 c;
  #pragma unroll
  for (int i = 0; i < 4; i++) c[i] = t[16 * i];
  c = mfma(ld_tile(SZ, z0, lane), ld_tile(SH, h0, lane), c);
  #pragma unroll
  for (int i = 0; i < 4; i++) t[16 * i] = c[i];
}
// bias gradient: sum over the 16 nodes held by the lanes of one quarter
__device__ __forceinline__ void acc_db(float *db, int f0, mlp_f4 v, int lane) {
  #pragma unroll
  for (int i = 0; i < 4; i++) {
    float s = v[i];
    for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 16);
    v[i] = s;
  }
  if ((lane & 15) == 0)
    #pragma unroll
    for (int i = 0; i < 4; i++) db[f0 + 4 * (lane >> 4) + i] += v[i];
}

// training: loss and gradient partial sums of one workgroup (one wavefront) over the
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...; targets are raw increments T[n][3]
__global__ void __launch_bounds__(64) k_mlp_train(int n, const float *__restrict__ X,
                                                  const float *__restrict__ Tg,
                                                  const unsigned short *__restrict__ wb,
                                                  const float *__restrict__ p,
                                                  const float *__restrict__ norm, float *part,
                                                  double *lpart) {
  __shared__ __align__(16) unsigned short sW[WB_N];
  __shared__ __align__(16) unsigned short SZ[16 * MLP_TS], SH[16 * MLP_TS];
  __shared__ float acc[GA_N];
  load_weights(sW, wb);
  for (int t = threadIdx.x; t < GA_N; t += 64) acc[t] = 0.f;
  __syncthreads();
  const int lane = threadIdx.x, q = lane >> 4;
  const int ntile = (n + 15) / 16;
  const float gs = 2.f / (3.f * (float)n);
  double lsum = 0.;
  for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    MlpTile T;
    mlp_forward(sW, p, norm, X, n, 16 * tile, lane, T);
    const int node = 16 * tile + (lane & 15);
    // dL/dy of the normalised outputs
    mlp_f4 dy = {0.f, 0.f, 0.f, 0.f};
    if (q == 0 && node < n)
      #pragma unroll
      for (int i = 0; i < MLP_NOUT; i++) {
        const float e = T.y[i] - Tg[(size_t)node * MLP_NOUT + i] / norm[2 * MLP_NIN + i];
        lsum += (double)e * e;
        dy[i] = gs * e;
      }
    const mlp_s4 dyb = tobf(dy);
    // layer 4: db4, dW4 = dY^T H3, dH3 = W4^T dY
    acc_db(acc + GA_B + 192, 0, dy, lane);
    st_tile(SZ, 0, dyb, lane);
    #pragma unroll
    for (int kt = 0; kt < 4; kt++) st_tile(SH, 16 * kt, T.h[2][kt], lane);
    __syncthreads();
    #pragma unroll
    for (int nt = 0; nt < 4; nt++) acc_dw(acc + GA_4 + 256 * nt, SZ, 0, SH, 16 * nt, lane);
    mlp_f4 dz[4];
    mlp_s4 dzb[4];
    #pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      mlp_f4 d = {0.f, 0.f, 0.f, 0.f};
      d = mfma(ldAT(sW + WB_4, 64, 16 * mt, 0, lane), dyb, d);
      #pragma unroll
      for (int i = 0; i < 4; i++) d[i] *= gelu_d(T.z[2][mt][i]);
      dz[mt] = d;
      dzb[mt] = tobf(d);
    }
    // layers 3 and 2: db, dW = dZ^T H_prev, dH_prev = W^T dZ
    #pragma unroll
    for (int L = 2; L >= 1; L--) {
      const unsigned short *W = sW + (L == 2 ? WB_3 : WB_2);
      float *dW = acc + (L == 2 ? GA_3 : GA_2);
      #pragma unroll
      for (int mt = 0; mt < 4; mt++) acc_db(acc + GA_B + 64 * L, 16 * mt, dz[mt], lane);
      __syncthreads();
      #pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        st_tile(SZ, 16 * mt, dzb[mt], lane);
        st_tile(SH, 16 * mt, T.h[L - 1][mt], lane);
      }
      __syncthreads();
      for (int mt = 0; mt < 4; mt++)
        #pragma unroll
        for (int nt = 0; nt < 4; nt++) acc_dw(dW + 256 * (4 * mt + nt), SZ, 16 * mt, SH, 16 * nt, lane);
      mlp_f4 nz[4];
      mlp_s4 nzb[4];
      #pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        mlp_f4 d = {0.f, 0.f, 0.f, 0.f};
        #pragma unroll
        for (int kt = 0; kt < 4; kt++) d = mfma(ldAT(W, 64, 16 * mt, 16 * kt, lane), dzb[kt], d);
        #pragma unroll
        for (int i = 0; i < 4; i++) d[i] *= gelu_d(T.z[L - 1][mt][i]);
        nz[mt] = d;
        nzb[mt] = tobf(d);
      }
      #pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        dz[mt] = nz[mt];
        dzb[mt] = nzb[mt];
      }
    }
    // layer 1: db1, dW1 = dZ1^T X
    #pragma unroll
    for (int mt = 0; mt < 4; mt++) acc_db(acc + GA_B, 16 * mt, dz[mt], lane);
    __syncthreads();
    #pragma unroll
    for (int mt = 0; mt < 4; mt++) st_tile(SZ, 16 * mt, dzb[mt], lane);
    st_tile(SH, 0, T.x, lane);
    __syncthreads();
    #pragma unroll
    for (int mt = 0; mt < 4; mt++) acc_dw(acc + GA_1 + 256 * mt, SZ, 16 * mt, SH, 0, lane);
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  __syncthreads();
  for (int t = threadIdx.x; t < GA_N; t += 64) part[(size_t)blockIdx.x * GA_N + t] = acc[t];
  if (lane == 0) lpart[blockIdx.x] = lsum;
}

// accumulator slot of parameter p (the inverse of the tile layout above)
__device__ __forceinline__ int ga_slot(int p) {
  if (p < MP_B1) {
    const int o = p / MLP_NIN, i = p % MLP_NIN;
    return GA_1 + 256 * (o / 16) + 16 * (o % 16) + i;
  }
  if (p < MP_W2) return GA_B + (p - MP_B1);
  if (p < MP_B2 || (p >= MP_W3 && p < MP_B3)) {
    const int base = p < MP_B2 ? MP_W2 : MP_W3, g = p < MP_B2 ? GA_2 : GA_3;
    const int o = (p - base) / 64, i = (p - base) % 64;
    return g + 256 * (4 * (o / 16) + i / 16) + 16 * (o % 16) + i % 16;
  }
  if (p < MP_W3) return GA_B + 64 + (p - MP_B2);
  if (p < MP_W4) return GA_B + 128 + (p - MP_B3);
  if (p < MP_B4) {
    const int o = (p - MP_W4) / 64, i = (p - MP_W4) % 64;
    return GA_4 + 256 * (i / 16) + 16 * o + i % 16;
  }
  return GA_B + 192 + (p - MP_B4);
}

// grad[p] = sum over workgroups of the partials, in workgroup order
__global__ void k_mlp_reduce(int nwg, const float *__restrict__ part, float *grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MP_N) return;
  const int s = ga_slot(p);
  float g = 0.f;
  for (int w = 0; w < nwg; w++) g += part[(size_t)w * GA_N + s];
  grad[p] = g;
}

// BF16 copy of the weights in the padded tile layout
__device__ __forceinline__ int wb_slot(int p) {
  if (p < MP_B1) return WB_1 + 16 * (p / MLP_NIN) + p % MLP_NIN;
  if (p >= MP_W2 && p < MP_B2) return WB_2 + (p - MP_W2);
  if (p >= MP_W3 && p < MP_B3) return WB_3 + (p - MP_W3);
  if (p >= MP_W4 && p < MP_B4) return WB_4 + (p - MP_W4);
  return -1;
}
__global__ void k_mlp_to_bf16(const float *__restrict__ p, unsigned short *wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MP_N) return;
  const int s = wb_slot(i);
  if (s >= 0) wb[s] = f2bf(p[i]);
}

// Adam (Kingma and Ba, 2015) on the FP32 master parameters, then their BF16 copy
__global__ void k_mlp_adam(float *p, float *m1, float *m2, const float *__restrict__ g, float lr,
                           float b1, float b2, float eps, float c1, float c2, unsigned short *wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MP_N) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi, b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float v = p[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
  p[i] = v;
  const int s = wb_slot(i);
  if (s >= 0) wb[s] = f2bf(v);
}

// ---------------------------------------------------------------------------
extern "C" int rbc3d_mlp_create(rbc3d_mlp **out) {
  if (!out) return RBC3D_EINVAL;
  rbc3d_mlp *m = new rbc3d_mlp();
  int r = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess ? 0 : RBC3D_EHIP;
  if (!r) r = m->p.ensure(MP_N);
  if (!r) r = m->m1.ensure(MP_N);
  if (!r) r = m->m2.ensure(MP_N);
  if (!r) r = m->grad.ensure(MP_N);
  if (!r) r = m->norm.ensure(2 * MLP_NIN + MLP_NOUT);
  if (!r) r = m->wb.ensure(WB_N);
  if (!r && hipMemset(m->wb.p, 0, WB_N * 2) != hipSuccess) r = RBC3D_EHIP;
  if (!r && hipMemset(m->p.p, 0, MP_N * 4) != hipSuccess) r = RBC3D_EHIP;
  if (!r) {
    std::vector<float> nm(2 * MLP_NIN + MLP_NOUT, 1.f);
    #pragma unroll
    for (int i = 0; i < MLP_NIN; i++) nm[i] = 0.f;
    if (hipMemcpy(m->norm.p, nm.data(), nm.size() * 4, hipMemcpyHostToDevice) != hipSuccess) r = RBC3D_EHIP;
  }
  if (r) {
    rbc_set_error("rbc3d_mlp_create: device allocation");
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return r;
  }
  *out = m;
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_destroy(rbc3d_mlp *m) {
  if (!m) return RBC3D_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->ev) (void)hipEventDestroy(m->ev);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_nparams(void) { return MP_N; }

extern "C" int rbc3d_mlp_set_stream(rbc3d_mlp *m, void *stream) {
  if (!m) return RBC3D_EINVAL;
  HIP_TRY(hipStreamSynchronize(m->st()));
  m->ext = (hipStream_t)stream;
  m->use_ext = true;
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_set_params(rbc3d_mlp *m, const float *params) {
  if (!m || !params) return RBC3D_EINVAL;
  for (int i = 0; i < MP_N; i++)
    if (!std::isfinite(params[i])) return RBC3D_ENONFINITE;
  HIP_TRY(hipStreamSynchronize(m->st()));
  HIP_TRY(hipMemcpy(m->p.p, params, MP_N * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(m->m1.p, 0, MP_N * 4));
  HIP_TRY(hipMemset(m->m2.p, 0, MP_N * 4));
  m->t = 0;
  k_mlp_to_bf16<<<(MP_N + 255) / 256, 256, 0, m->stream>>>(m->p.p, m->wb.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(m->stream));
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_get_params(rbc3d_mlp *m, float *params) {
  if (!m || !params) return RBC3D_EINVAL;
  HIP_TRY(hipStreamSynchronize(m->st()));
  HIP_TRY(hipMemcpy(params, m->p.p, MP_N * 4, hipMemcpyDeviceToHost));
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_set_norm(rbc3d_mlp *m, const float *mean_in, const float *std_in,
                                  const float *scale_out) {
  if (!m || !mean_in || !std_in || !scale_out) return RBC3D_EINVAL;
  std::vector<float> nm(2 * MLP_NIN + MLP_NOUT);
  #pragma unroll
  for (int i = 0; i < MLP_NIN; i++) {
    if (!(std_in[i] > 0.f)) return RBC3D_EINVAL;
    nm[i] = mean_in[i];
    nm[MLP_NIN + i] = 1.f / std_in[i];
  }
  #pragma unroll
  for (int i = 0; i < MLP_NOUT; i++) {
    if (!(scale_out[i] > 0.f)) return RBC3D_EINVAL;
    nm[2 * MLP_NIN + i] = scale_out[i];
  }
  HIP_TRY(hipStreamSynchronize(m->st()));
  HIP_TRY(hipMemcpy(m->norm.p, nm.data(), nm.size() * 4, hipMemcpyHostToDevice));
  return RBC3D_OK;
}

int rbc_mlp_forward(rbc3d_mlp *m, int n, const float *d_x, float *d_y, hipStream_t st) {
  if (n <= 0) return RBC3D_OK;
  const int ntile = (n + 15) / 16;
  const unsigned nblk = (unsigned)std::min(1024, (ntile + 3) / 4);
  k_mlp_forward<<<nblk, 256, 0, st>>>(n, d_x, m->wb.p, m->p.p, m->norm.p, d_y);
  HIP_TRY(hipGetLastError());
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_forward_dev(rbc3d_mlp *m, int n, const float *d_x, float *d_y) {
  if (!m || n < 0 || (n > 0 && (!d_x || !d_y))) return RBC3D_EINVAL;
  RBC_TRY(rbc_mlp_forward(m, n, d_x, d_y, m->st()));
  HIP_TRY(hipStreamSynchronize(m->st()));
  return RBC3D_OK;
}

static int mlp_grad(rbc3d_mlp *m, int n, const float *d_x, const float *d_t, double *loss) {
  hipStream_t st = m->st();
  const int ntile = (n + 15) / 16;
  const int nwg = std::min(512, ntile);
  RBC_TRY(m->part.ensure((size_t)nwg * GA_N));
  RBC_TRY(m->lpart.ensure(nwg));
  k_mlp_train<<<nwg, 64, 0, m->st()>>>(n, d_x, d_t, m->wb.p, m->p.p, m->norm.p, m->part.p,
                                         m->lpart.p);
  k_mlp_reduce<<<(MP_N + 255) / 256, 256, 0, m->st()>>>(nwg, m->part.p, m->grad.p);
  HIP_TRY(hipGetLastError());
  if (loss) {
    std::vector<double> lp(nwg);
    HIP_TRY(hipMemcpyAsync(lp.data(), m->lpart.p, nwg * 8, hipMemcpyDeviceToHost, m->st()));
    HIP_TRY(hipStreamSynchronize(m->st()));
    double s = 0;
    #pragma unroll
    for (int i = 0; i < nwg; i++) s += lp[i];
    *loss = s / (3.0 * n);
  }
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_grad_dev(rbc3d_mlp *m, int n, const float *d_x, const float *d_t,
                                  float *d_grad, double *loss) {
  if (!m || n <= 0 || !d_x || !d_t) return RBC3D_EINVAL;
  RBC_TRY(mlp_grad(m, n, d_x, d_t, loss));
  if (d_grad)
    HIP_TRY(hipMemcpyAsync(d_grad, m->grad.p, MP_N * 4, hipMemcpyDeviceToDevice, m->st()));
  HIP_TRY(hipStreamSynchronize(m->st()));
  return RBC3D_OK;
}

extern "C" int rbc3d_mlp_adam_step_dev(rbc3d_mlp *m, int n, const float *d_x, const float *d_t,
                                       float lr, double *loss) {
  if (!m || n <= 0 || !d_x || !d_t || !(lr > 0.f)) return RBC3D_EINVAL;
  RBC_TRY(mlp_grad(m, n, d_x, d_t, loss));
  m->t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, m->t)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, m->t)));
  k_mlp_adam<<<(MP_N + 255) / 256, 256, 0, m->st()>>>(m->p.p, m->m1.p, m->m2.p, m->grad.p, lr, b1,
                                                        b2, 1e-8f, c1, c2, m->wb.p);
  HIP_TRY(hipGetLastError());
  // the update is complete when the call returns, so a caller may read or use the
  // parameters from any stream
  HIP_TRY(hipStreamSynchronize(m->st()));
  return RBC3D_OK;
}

// Make stream st wait for the work queued on the network's stream (an event, no host
// wait): a forward on another stream then reads the parameters after every update.
int rbc_mlp_wait(rbc3d_mlp *m, hipStream_t st) {
  if (m->st() == st) return RBC3D_OK;
  if (!m->ev) HIP_TRY(hipEventCreateWithFlags(&m->ev, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(m->ev, m->st()));
  HIP_TRY(hipStreamWaitEvent(st, m->ev, 0));
  return RBC3D_OK;
}
