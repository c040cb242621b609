// fft2s.h -- two-stage register FFT for CDNA4 (gfx950), FP32.
//
// A line of length L = R1 * R2 is split as n = R2*n1 + n2, k = k1 + R1*k2:
//   stage 1: for each n2, a radix-R1 DFT over n1 (in registers),
//            twiddle W_L^{n2 k1},
//   exchange: one padded LDS transpose,
//   stage 2: for each k1, a radix-R2 DFT over n2 (in registers).
// NT threads cooperate on a line; thread t owns the stage-1 columns n2 = t + b*NT and the stage-2
// rows k1 = t + b*NT.  Inputs arrive in registers straight from global memory and outputs leave
// from registers, so a transform costs one LDS write, one LDS read and one barrier.  The same
// register layout that one plan produces, (k1 own, k2 = 0..R2), is the input layout of the plan
// with R1 and R2 swapped, so an inverse transform can be followed by a pointwise product and a
// forward transform without another exchange.
//
// Twiddles W_L^m come from a table tw[m] = exp(-2 pi i m / L) computed in FP64 on the host.
#pragma once
#include <hip/hip_runtime.h>

namespace glia {
namespace dev {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }    // i * a
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }  // -i * a
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// cos(2 pi j / 48); the radices 2, 3, 4, 6, 8, 12, 16 and 24 divide 48.  Non-recursive so
// that it folds to a constant wherever j is a compile-time constant.
__device__ __forceinline__ constexpr float c48q(int q) {  // q in [0, 12]
  return q == 0 ? 1.0f : q == 1 ? 0.99144486137381041f : q == 2 ? 0.96592582628906829f
       : q == 3 ? 0.92387953251128674f : q == 4 ? 0.86602540378443865f : q == 5 ? 0.79335334029123517f
       : q == 6 ? 0.70710678118654752f : q == 7 ? 0.60876142900872066f : q == 8 ? 0.5f
       : q == 9 ? 0.38268343236508977f : q == 10 ? 0.25881904510252076f : q == 11 ? 0.13052619222005160f : 0.0f;
}
__device__ __forceinline__ constexpr float c48(int j) {
  const int m = ((j % 48) + 48) % 48;
  const int q = (m <= 24) ? m : 48 - m;          // cos is even
  return (q <= 12) ? c48q(q) : -c48q(24 - q);    // cos(pi - x) = -cos(x)
}
__device__ __forceinline__ constexpr float s48(int j) { return c48(j - 12); }

// cos(2 pi j / 20) for the radices 5, 10 and 20
__device__ __forceinline__ constexpr float c20q(int q) {  // q in [0, 5]
  return q == 0 ? 1.0f : q == 1 ? 0.95105651629515357f : q == 2 ? 0.80901699437494742f
       : q == 3 ? 0.58778525229247313f : q == 4 ? 0.30901699437494742f : 0.0f;
}
__device__ __forceinline__ constexpr float c20(int j) {
  const int m = ((j % 20) + 20) % 20;
  const int q = (m <= 10) ? m : 20 - m;
  return (q <= 5) ? c20q(q) : -c20q(10 - q);
}
__device__ __forceinline__ constexpr float s20(int j) { return c20(j - 5); }

// W_N^m with sign S: exp(S * 2 pi i m / N), for N dividing 48 or 20
template <int N, int S>
__device__ __forceinline__ constexpr float2 wn(int m) {
  static_assert(48 % N == 0 || 20 % N == 0, "no twiddle constants for this radix");
  if constexpr (48 % N == 0)
    return make_float2(c48(m * (48 / N)), (S < 0 ? -1.0f : 1.0f) * s48(m * (48 / N)));
  else
    return make_float2(c20(m * (20 / N)), (S < 0 ? -1.0f : 1.0f) * s20(m * (20 / N)));
}

// ---------------------------------------------------------------------------------------------
// register DFTs in natural order; S = -1 forward (exp(-2 pi i nk/R)), S = +1 inverse
// ---------------------------------------------------------------------------------------------
template <int R, int S>
struct Dft;

template <int S>
struct Dft<1, S> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int S>
struct Dft<2, S> {
  __device__ __forceinline__ static void run(float2* v) {
    const float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};

template <int S>
struct Dft<3, S> {
  __device__ __forceinline__ static void run(float2* v) {
    const float s3 = 0.86602540378443865f;
    const float2 t1 = cadd(v[1], v[2]);
    const float2 t2 = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    const float2 n = (S < 0) ? mul_mi(cscale(t2, s3)) : mul_i(cscale(t2, s3));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, n);
    v[2] = csub(m, n);
  }
};

template <int S>
struct Dft<4, S> {
  __device__ __forceinline__ static void run(float2* v) {
    const float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const float2 c = cadd(v[1], v[3]);
    float2 d = csub(v[1], v[3]);
    d = (S < 0) ? mul_mi(d) : mul_i(d);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

// radix 5: with t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3 and c_k, s_k the cosine and
// sine of 2 pi k / 5, X1,4 = x0 + c1 t1 + c2 t2 -/+ S' i (s1 t3 + s2 t4) and
// X2,3 = x0 + c2 t1 + c1 t2 -/+ S' i (s2 t3 - s1 t4), S' = 1 forward and -1 inverse
template <int S>
struct Dft<5, S> {
  __device__ __forceinline__ static void run(float2* v) {
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    b1 = (S < 0) ? mul_mi(b1) : mul_i(b1);
    b2 = (S < 0) ? mul_mi(b2) : mul_i(b2);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// N = R1 * R2 in registers: n = R2*n1 + n2, k = k1 + R1*k2
template <int R1, int R2, int S>
__device__ __forceinline__ void dft_split(float2* v) {
  constexpr int N = R1 * R2;
  float2 a[R2][R1];
#pragma unroll
  for (int n2 = 0; n2 < R2; n2++) {
    float2 t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++) t[n1] = v[R2 * n1 + n2];
    Dft<R1, S>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) a[n2][k1] = ((n2 * k1) % N == 0) ? t[k1] : cmul(t[k1], wn<N, S>(n2 * k1));
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) {
    float2 t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++) t[n2] = a[n2][k1];
    Dft<R2, S>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = t[k2];
  }
}

template <int S>
struct Dft<6, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<3, 2, S>(v); }
};
template <int S>
struct Dft<8, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<4, 2, S>(v); }
};
template <int S>
struct Dft<12, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<3, 4, S>(v); }
};
template <int S>
struct Dft<16, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<4, 4, S>(v); }
};
template <int S>
struct Dft<24, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<3, 8, S>(v); }
};
template <int S>
struct Dft<10, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<5, 2, S>(v); }
};
template <int S>
struct Dft<20, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<5, 4, S>(v); }
};

// ---------------------------------------------------------------------------------------------
// plans: L = R1 * R2 with NT threads per line.  NT divides R2; when it does not divide R1 the
// stage-2 rows are split unevenly: 160 = 20 x 8 gives 8 threads 3, 3, 3, 3, 2, 2, 2, 2 rows, and
// 192 = 12 x 16 gives 12 of 16 threads one row.
// ---------------------------------------------------------------------------------------------
template <int L>
struct Plan2;
#define GLIA_PLAN2(LL, A, B, T)                        \
  template <>                                          \
  struct Plan2<LL> {                                   \
    static constexpr int R1 = A, R2 = B, NT = T;       \
  };
GLIA_PLAN2(16, 4, 4, 4)
GLIA_PLAN2(24, 6, 4, 2)
GLIA_PLAN2(32, 4, 8, 4)
GLIA_PLAN2(48, 12, 4, 4)
GLIA_PLAN2(64, 8, 8, 8)
GLIA_PLAN2(80, 20, 4, 4)
GLIA_PLAN2(96, 12, 8, 4)
GLIA_PLAN2(128, 8, 16, 8)
GLIA_PLAN2(160, 20, 8, 8)
GLIA_PLAN2(192, 12, 16, 16)
GLIA_PLAN2(256, 16, 16, 16)
#undef GLIA_PLAN2

// padded exchange index (one slot of padding per 16 entries)
__device__ __forceinline__ int xpad(int i) { return i + (i >> 4); }
// exchange pitch of one line: L + L/16 + 2 keeps the transposes free of LDS bank conflicts
template <int L>
struct XPitch {
  static constexpr int value = L + L / 16 + 2;
};

// One transform of a line of length R1*R2.  Input v[b*R1 + n1] = x[R2*n1 + n2] for the thread's
// n2 = t + b*NT (b < R2/NT, NT divides R2); output v[b*R2 + k2] = X[k1 + R1*k2] for k1 = t + b*NT
// (b < ceil(R1/NT), k1 < R1: when NT does not divide R1 the last row is absent on some threads and
// its slots are left as they are).
// 'ex' is the line's exchange area (xpad layout); one barrier inside, and the caller must place a
// barrier between two uses of the same exchange area.
template <int R1, int R2, int NT, int S>
__device__ __forceinline__ void fft2s(float2* v, float2* ex, int t, const float2* __restrict__ tw) {
  constexpr int L = R1 * R2;
  constexpr int B1 = R2 / NT, B2 = (R1 + NT - 1) / NT;
  static_assert(B1 * NT == R2, "NT must divide R2");
  constexpr bool EVEN = B2 * NT == R1;  // otherwise the last stage-2 row exists on t < R1 % NT only
#pragma unroll
  for (int b = 0; b < B1; b++) {
    Dft<R1, S>::run(v + b * R1);
    const int n2 = t + b * NT;
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      float2 a = v[b * R1 + k1];
      if (k1 > 0) {
        const float2 w = tw[n2 * k1];  // n2*k1 < L
        a = (S < 0) ? cmul(a, w) : cmulc(a, w);
      }
      ex[xpad(k1 * R2 + n2)] = a;
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < B2; b++) {
    const int k1 = t + b * NT;
    if (EVEN || k1 < R1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) v[b * R2 + n2] = ex[xpad(k1 * R2 + n2)];
      Dft<R2, S>::run(v + b * R2);
    }
  }
  (void)L;
}

}  // namespace dev
}  // namespace glia
