// fft_engine.h -- LDS-staged mixed-radix Stockham FFT engine for CDNA4 (gfx950), FP32.
//
// A line of length L lives in LDS (element i at buf[i * bs]).  NT = L/16 threads cooperate on a
// line; every pass reads R values per butterfly into registers, synchronises the workgroup,
// applies the twiddles and a register radix-R DFT and writes the autosorted result back
// (Stockham: no bit reversal).  Radices are 2, 3, 4, 8 and 16; the plan for each supported length
// is fixed at compile time so all butterflies unroll into straight-line code.
//
// Twiddles come from a per-length table tw[m] = exp(-2 pi i m / L) computed in FP64 on the host,
// so the transform's accuracy does not depend on device transcendental functions.
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

// cos and sin of 2 pi m / 16
__device__ __forceinline__ constexpr float c16(int m) {
  switch (m & 15) {
    case 0: return 1.0f;
    case 1: return 0.92387953251128674f;
    case 2: return 0.70710678118654752f;
    case 3: return 0.38268343236508977f;
    case 4: return 0.0f;
    case 5: return -0.38268343236508977f;
    case 6: return -0.70710678118654752f;
    case 7: return -0.92387953251128674f;
    case 8: return -1.0f;
    case 9: return -0.92387953251128674f;
    case 10: return -0.70710678118654752f;
    case 11: return -0.38268343236508977f;
    case 12: return 0.0f;
    case 13: return 0.38268343236508977f;
    case 14: return 0.70710678118654752f;
    default: return 0.92387953251128674f;
  }
}
__device__ __forceinline__ constexpr float s16(int m) { return c16(m - 4); }

// ---------------------------------------------------------------------------------------------
// Register DFTs in natural order.  S = -1: forward (exp(-2 pi i nk/R)), S = +1: inverse.
// ---------------------------------------------------------------------------------------------
template <int R, int S>
struct Dft;

template <int S>
struct Dft<2, S> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};

template <int S>
struct Dft<3, S> {
  __device__ __forceinline__ static void run(float2* v) {
    const float s3 = 0.86602540378443865f;
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
    // n = i * S * s3 * t2
    float2 n = (S < 0) ? mul_mi(cscale(t2, s3)) : mul_i(cscale(t2, s3));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, n);
    v[2] = csub(m, n);
  }
};

template <int S>
struct Dft<4, S> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    float2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    d = (S < 0) ? mul_mi(d) : mul_i(d);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

// N = R1 * R2 (N divides 16): n = R2*n1 + n2, k = k1 + R1*k2
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
    for (int k1 = 0; k1 < R1; k1++) {
      const int m = (n2 * k1 * (16 / N)) & 15;
      a[n2][k1] = (m == 0) ? t[k1] : cmul(t[k1], make_float2(c16(m), (S < 0 ? -1.0f : 1.0f) * s16(m)));
    }
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
struct Dft<8, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<4, 2, S>(v); }
};

template <int S>
struct Dft<16, S> {
  __device__ __forceinline__ static void run(float2* v) { dft_split<4, 4, S>(v); }
};

// ---------------------------------------------------------------------------------------------
// Compile-time plans.  The last radix is 16 wherever L allows, so NT = L/16 threads per line.
// ---------------------------------------------------------------------------------------------
template <int L>
struct Plan;
#define GLIA_PLAN(LL, NPASS, R0, R1, R2)                                            \
  template <>                                                                       \
  struct Plan<LL> {                                                                 \
    static constexpr int np = NPASS;                                                \
    __device__ __forceinline__ static constexpr int r(int p) { return p == 0 ? R0 : (p == 1 ? R1 : R2); } \
  };
GLIA_PLAN(16, 1, 16, 1, 1)
GLIA_PLAN(24, 2, 3, 8, 1)
GLIA_PLAN(32, 2, 2, 16, 1)
GLIA_PLAN(48, 2, 3, 16, 1)
GLIA_PLAN(64, 2, 4, 16, 1)
GLIA_PLAN(96, 3, 3, 2, 16)
GLIA_PLAN(128, 2, 8, 16, 1)
GLIA_PLAN(192, 3, 3, 4, 16)
GLIA_PLAN(256, 2, 16, 16, 1)
#undef GLIA_PLAN

template <int L>
struct LineCfg {
  static constexpr int NT = (L >= 16) ? L / 16 : 1;  // threads per line
};
template <>
struct LineCfg<24> {
  static constexpr int NT = 3;  // 24 = 3 * 8: last pass has 3 radix-8 butterflies
};

// One Stockham pass of radix R with NS = product of previous radices.
// Every thread of the workgroup must call this (it contains barriers).
template <int L, int R, int NS, int NT, int S>
__device__ __forceinline__ void stockham_pass(float2* buf, int bs, int t, const float2* __restrict__ tw, bool active) {
  constexpr int LR = L / R;
  constexpr int NB = (LR + NT - 1) / NT;
  float2 v[NB][R];
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = t + b * NT;
      if (j < LR) {
#pragma unroll
        for (int r = 0; r < R; r++) v[b][r] = buf[(j + r * LR) * bs];
        if (NS > 1) {
          const int k = j % NS;
#pragma unroll
          for (int r = 1; r < R; r++) {
            const int m = k * r * (L / (NS * R));
            const float2 w = tw[m];
            v[b][r] = (S < 0) ? cmul(v[b][r], w) : cmulc(v[b][r], w);
          }
        }
        Dft<R, S>::run(v[b]);
      }
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = t + b * NT;
      if (j < LR) {
        const int k = j % NS;
        const int idxD = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) buf[(idxD + r * NS) * bs] = v[b][r];
      }
    }
  }
  __syncthreads();
}

template <int L, int NT, int S, int P, int NS>
__device__ __forceinline__ void fft_passes(float2* buf, int bs, int t, const float2* __restrict__ tw, bool active) {
  if constexpr (P < Plan<L>::np) {
    constexpr int R = Plan<L>::r(P);
    stockham_pass<L, R, NS, NT, S>(buf, bs, t, tw, active);
    fft_passes<L, NT, S, P + 1, NS * R>(buf, bs, t, tw, active);
  }
}

// In-place unnormalized FFT of one line in LDS.  Contains __syncthreads(): all threads of the
// workgroup must call it the same number of times; 'active' masks threads without a line.
template <int L, int S>
__device__ __forceinline__ void fft_line(float2* buf, int bs, int t, const float2* __restrict__ tw, bool active) {
  fft_passes<L, LineCfg<L>::NT, S, 0, 1>(buf, bs, t, tw, active);
}

}  // namespace dev
}  // namespace glia
