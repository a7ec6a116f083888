// fft_fixed.h -- line FFTs of compile-time length in two register passes.
//
// A length N = R0 R1 splits into two Stockham passes of radix R0 and R1 (each at most
// 16), so a line is read from and written to LDS twice instead of once per radix-2..5
// factor, and every index is a compile-time constant.  The radix-R transforms run in
// registers: R = 2, 3, 4, 5 directly, composite R = A B (6, 8, 9, 10, 12, 15, 16) as
// B transforms of length A, twiddles exp(s 2 pi i n2 k1 / R), then A transforms of
// length B (Cooley-Tukey within the thread; the roots of unity are exact constants).
// Passes are register-staged: every thread loads the inputs of its butterflies, the
// workgroup synchronizes, every thread writes its outputs into the same buffer.
#ifndef RBC3D_FFT_FIXED_H
#define RBC3D_FFT_FIXED_H

// cos and sin of 2 pi m / R, m = 0..R-1
__device__ constexpr double RT6_C[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
__device__ constexpr double RT6_S[6] = {0.0, 0.866025403784438646764, 0.866025403784438646764, 0.0, -0.866025403784438646764, -0.866025403784438646764};
__device__ constexpr double RT8_C[8] = {1.0, 0.707106781186547524401, 0.0, -0.707106781186547524401, -1.0, -0.707106781186547524401, 0.0, 0.707106781186547524401};
__device__ constexpr double RT8_S[8] = {0.0, 0.707106781186547524401, 1.0, 0.707106781186547524401, 0.0, -0.707106781186547524401, -1.0, -0.707106781186547524401};
__device__ constexpr double RT9_C[9] = {1.0, 0.766044443118978035202, 0.173648177666930348852, -0.5, -0.939692620785908384054, -0.939692620785908384054, -0.5, 0.173648177666930348852, 0.766044443118978035202};
__device__ constexpr double RT9_S[9] = {0.0, 0.642787609686539326323, 0.984807753012208059367, 0.866025403784438646764, 0.342020143325668733044, -0.342020143325668733044, -0.866025403784438646764, -0.984807753012208059367, -0.642787609686539326323};
__device__ constexpr double RT10_C[10] = {1.0, 0.809016994374947424102, 0.309016994374947424102, -0.309016994374947424102, -0.809016994374947424102, -1.0, -0.809016994374947424102, -0.309016994374947424102, 0.309016994374947424102, 0.809016994374947424102};
__device__ constexpr double RT10_S[10] = {0.0, 0.587785252292473129169, 0.951056516295153572116, 0.951056516295153572116, 0.587785252292473129169, 0.0, -0.587785252292473129169, -0.951056516295153572116, -0.951056516295153572116, -0.587785252292473129169};
__device__ constexpr double RT12_C[12] = {1.0, 0.866025403784438646764, 0.5, 0.0, -0.5, -0.866025403784438646764, -1.0, -0.866025403784438646764, -0.5, 0.0, 0.5, 0.866025403784438646764};
__device__ constexpr double RT12_S[12] = {0.0, 0.5, 0.866025403784438646764, 1.0, 0.866025403784438646764, 0.5, 0.0, -0.5, -0.866025403784438646764, -1.0, -0.866025403784438646764, -0.5};
__device__ constexpr double RT15_C[15] = {1.0, 0.913545457642600895502, 0.669130606358858213826, 0.309016994374947424102, -0.1045284632676534714, -0.5, -0.809016994374947424102, -0.978147600733805637929, -0.978147600733805637929, -0.809016994374947424102, -0.5, -0.1045284632676534714, 0.309016994374947424102, 0.669130606358858213826, 0.913545457642600895502};
__device__ constexpr double RT15_S[15] = {0.0, 0.406736643075800207754, 0.743144825477394235015, 0.951056516295153572116, 0.994521895368273336923, 0.866025403784438646764, 0.587785252292473129169, 0.207911690817759337102, -0.207911690817759337102, -0.587785252292473129169, -0.866025403784438646764, -0.994521895368273336923, -0.951056516295153572116, -0.743144825477394235015, -0.406736643075800207754};
__device__ constexpr double RT16_C[16] = {1.0, 0.923879532511286756128, 0.707106781186547524401, 0.382683432365089771728, 0.0, -0.382683432365089771728, -0.707106781186547524401, -0.923879532511286756128, -1.0, -0.923879532511286756128, -0.707106781186547524401, -0.382683432365089771728, 0.0, 0.382683432365089771728, 0.707106781186547524401, 0.923879532511286756128};
__device__ constexpr double RT16_S[16] = {0.0, 0.382683432365089771728, 0.707106781186547524401, 0.923879532511286756128, 1.0, 0.923879532511286756128, 0.707106781186547524401, 0.382683432365089771728, 0.0, -0.382683432365089771728, -0.707106781186547524401, -0.923879532511286756128, -1.0, -0.923879532511286756128, -0.707106781186547524401, -0.382683432365089771728};
template <int R>
__device__ __forceinline__ double root_c(int m) {
  if constexpr (R == 6) return RT6_C[m];
  else if constexpr (R == 8) return RT8_C[m];
  else if constexpr (R == 9) return RT9_C[m];
  else if constexpr (R == 10) return RT10_C[m];
  else if constexpr (R == 12) return RT12_C[m];
  else if constexpr (R == 15) return RT15_C[m];
  else return RT16_C[m];
}
template <int R>
__device__ __forceinline__ double root_s(int m) {
  if constexpr (R == 6) return RT6_S[m];
  else if constexpr (R == 8) return RT8_S[m];
  else if constexpr (R == 9) return RT9_S[m];
  else if constexpr (R == 10) return RT10_S[m];
  else if constexpr (R == 12) return RT12_S[m];
  else if constexpr (R == 15) return RT15_S[m];
  else return RT16_S[m];
}

// the radix split of N used here: R0 R1 = N with both at most 16 and the larger as
// small as possible (0 when there is none)
template <int N>
struct Split2 {
  static constexpr int pick() {
    int best = 0;
    for (int a = 2; a <= 16; a++)
      if (N % a == 0 && N / a >= 2 && N / a <= 16 && (best == 0 || (a > N / a ? a : N / a) < (best > N / best ? best : N / best)))
        best = a;
    return best;
  }
  static constexpr int R0 = pick();
  static constexpr int R1 = R0 ? N / R0 : 0;
  static_assert(R0 == 0 || (R0 != 7 && R0 != 11 && R0 != 13 && R0 != 14 && R1 != 7 && R1 != 11 &&
                            R1 != 13 && R1 != 14),
                "radix with a factor above 5");
};

template <typename C>
__device__ __forceinline__ C cpx2(decltype(C().x) x, decltype(C().x) y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}

// x times exp(SG 2 pi i m / R); m is a constant once the callers' loops are unrolled
template <int R, int SG, typename C>
__device__ __forceinline__ C rot(C x, int m) {
  typedef decltype(C().x) Re;
  m %= R;
  if (m == 0) return x;
  if (2 * m == R) return cpx2<C>(-x.x, -x.y);
  if (4 * m == R) return SG > 0 ? cpx2<C>(-x.y, x.x) : cpx2<C>(x.y, -x.x);
  if (4 * m == 3 * R) return SG > 0 ? cpx2<C>(x.y, -x.x) : cpx2<C>(-x.y, x.x);
  const Re c = (Re)root_c<R>(m), s = (Re)(SG * root_s<R>(m));
  return cpx2<C>(x.x * c - x.y * s, x.x * s + x.y * c);
}

// in-place DFT of length R in registers, natural order in and out:
// X[k] = sum_n v[n] exp(SG 2 pi i n k / R)
template <int R, int SG, typename C>
__device__ __forceinline__ void dftr(C (&v)[R]) {
  typedef decltype(C().x) Re;
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = cpx2<C>(a.x + b.x, a.y + b.y);
    v[1] = cpx2<C>(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 3) {
    const Re hs = (Re)SG * (Re)0.86602540378443864676;
    const C t = cpx2<C>(v[1].x + v[2].x, v[1].y + v[2].y);
    const C d = cpx2<C>(v[1].x - v[2].x, v[1].y - v[2].y);
    const C m = cpx2<C>(v[0].x - (Re)0.5 * t.x, v[0].y - (Re)0.5 * t.y);
    const Re ux = -hs * d.y, uy = hs * d.x;
    v[0] = cpx2<C>(v[0].x + t.x, v[0].y + t.y);
    v[1] = cpx2<C>(m.x + ux, m.y + uy);
    v[2] = cpx2<C>(m.x - ux, m.y - uy);
  } else if constexpr (R == 4) {
    const C a0 = cpx2<C>(v[0].x + v[2].x, v[0].y + v[2].y);
    const C a1 = cpx2<C>(v[0].x - v[2].x, v[0].y - v[2].y);
    const C b0 = cpx2<C>(v[1].x + v[3].x, v[1].y + v[3].y);
    const C b1 = cpx2<C>(v[1].x - v[3].x, v[1].y - v[3].y);
    const C jb1 = SG < 0 ? cpx2<C>(b1.y, -b1.x) : cpx2<C>(-b1.y, b1.x);
    v[0] = cpx2<C>(a0.x + b0.x, a0.y + b0.y);
    v[1] = cpx2<C>(a1.x + jb1.x, a1.y + jb1.y);
    v[2] = cpx2<C>(a0.x - b0.x, a0.y - b0.y);
    v[3] = cpx2<C>(a1.x - jb1.x, a1.y - jb1.y);
  } else if constexpr (R == 5) {
    const Re c1 = (Re)0.30901699437494742410, c2 = (Re)-0.80901699437494742410;
    const Re s1 = (Re)SG * (Re)0.95105651629515357212, s2 = (Re)SG * (Re)0.58778525229247312917;
    const C a1 = cpx2<C>(v[1].x + v[4].x, v[1].y + v[4].y), b1 = cpx2<C>(v[1].x - v[4].x, v[1].y - v[4].y);
    const C a2 = cpx2<C>(v[2].x + v[3].x, v[2].y + v[3].y), b2 = cpx2<C>(v[2].x - v[3].x, v[2].y - v[3].y);
    const C m1 = cpx2<C>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const C m2 = cpx2<C>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    const Re u1x = -(s1 * b1.y + s2 * b2.y), u1y = s1 * b1.x + s2 * b2.x;
    const Re u2x = -(s2 * b1.y - s1 * b2.y), u2y = s2 * b1.x - s1 * b2.x;
    v[0] = cpx2<C>(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = cpx2<C>(m1.x + u1x, m1.y + u1y);
    v[4] = cpx2<C>(m1.x - u1x, m1.y - u1y);
    v[2] = cpx2<C>(m2.x + u2x, m2.y + u2y);
    v[3] = cpx2<C>(m2.x - u2x, m2.y - u2y);
  } else {
    constexpr int A = (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : (R % 3 == 0) ? 3 : 5, B = R / A;
    C u[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; n2++) {
#pragma unroll
      for (int n1 = 0; n1 < A; n1++) u[n2][n1] = v[B * n1 + n2];
      dftr<A, SG>(u[n2]);
    }
    // twiddles exp(SG 2 pi i n2 k1 / R)
#pragma unroll
    for (int n2 = 1; n2 < B; n2++)
#pragma unroll
      for (int k1 = 1; k1 < A; k1++) u[n2][k1] = rot<R, SG>(u[n2][k1], n2 * k1);
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
      C w[B];
#pragma unroll
      for (int n2 = 0; n2 < B; n2++) w[n2] = u[n2][k1];
      dftr<B, SG>(w);
#pragma unroll
      for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = w[k2];
    }
  }
}

// One register-staged pass over L lines (element e of line l at l LS + e ES): NB
// butterflies of radix R per line; butterfly (line, j) reads elements j + r SIN,
// multiplies them by the table twiddles tw[r j] (TW; the table holds exp(-2 pi i m/N)),
// and writes output q to j OM + q OS.  LF: consecutive threads take consecutive lines.
template <int R, int NB, int L, int LS, int ES, int SIN, int OM, int OS, bool TW, bool LF, int SG,
          int NT, typename C>
__device__ __forceinline__ void fpass(C *buf, const C *tw) {
  constexpr int TOT = L * NB, PER = (TOT + NT - 1) / NT;
  C v[PER][R];
  int o[PER];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int b = threadIdx.x + i * NT;
    o[i] = -1;
    if (TOT % NT == 0 || b < TOT) {
      const int line = LF ? b % L : b / NB, j = LF ? b / L : b % NB;
      const C *src = buf + line * LS + j * ES;
#pragma unroll
      for (int r = 0; r < R; r++) {
        C x = src[r * SIN * ES];
        if (TW && r) {
          if (j) {
            C t = tw[r * j];
            if (SG > 0) t.y = -t.y;
            x = cpx2<C>(x.x * t.x - x.y * t.y, x.x * t.y + x.y * t.x);
          }
        }
        v[i][r] = x;
      }
      o[i] = line * LS + j * OM * ES;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; i++) {
    if (TOT % NT == 0 || o[i] >= 0) {
      dftr<R, SG>(v[i]);
#pragma unroll
      for (int q = 0; q < R; q++) buf[o[i] + q * OS * ES] = v[i][q];
    }
  }
  __syncthreads();
}

// the transform of L lines of length N (sign SG) in place
template <int N, int L, int LS, int ES, bool LF, int SG, int NT, typename C>
__device__ __forceinline__ void fft_fixed(C *buf, const C *tw) {
  constexpr int R0 = Split2<N>::R0, R1 = Split2<N>::R1;
  static_assert(R0 > 0 && R0 <= 16 && R1 <= 16, "no two-pass split");
  // pass 0: radix R0, j < R1, inputs j + r R1, outputs j R0 + q
  fpass<R0, R1, L, LS, ES, R1, R0, 1, false, LF, SG, NT>(buf, tw);
  // pass 1: radix R1, j < R0, inputs j + r R0 times w^(r j), outputs j + q R0
  fpass<R1, R0, L, LS, ES, R0, 1, R0, true, LF, SG, NT>(buf, tw);
}

#endif
