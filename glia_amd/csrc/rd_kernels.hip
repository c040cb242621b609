// rd_kernels.hip -- CDNA4 kernels of the spectral reaction-diffusion solver.
//
// Spectral layout ("packed half spectrum"): per member, n0*n1*(n2/2) complex float2, index
// (kx*n1 + ky)*m2 + kz' with m2 = n2/2.  For kz' >= 1 the entry is the R2C coefficient
// X[kx,ky,kz']; the kz' = 0 entry packs the two real-along-z planes kz = 0 and kz = n2/2 as
// X0 + i*X_{n2/2} (both planes have Hermitian symmetry in (kx, ky)).  Every multiplier used here
// (i*w_d, the preconditioner) has w_z = 0 on both planes (the Nyquist mode is zeroed), so it
// acts identically on both halves of the packed entry, and Parseval sums over the packed column
// give the sum over both planes.  The array has exactly n0*n1*n2 floats, no padding.
//
// Stages of a 3D transform (inverse order shown; forward is the reverse):
//   X pass  (column kernel, lines along axis 0, stride n1*m2)   k_x... kernels
//   Y pass  (column kernel, lines along axis 1, stride m2)      k_y... kernels
//   Z pass  (line kernel, contiguous n2 reals <-> m2 packed complex) k_z... kernels
// Column kernels stage a tile of 16 adjacent columns (128-byte rows) in LDS; line kernels stage
// NL whole lines.  Pointwise work (wave-number multipliers, k(x) products, reaction, PCG vector
// updates, FP64 reductions) is fused into these passes so no kernel exists only for it.
#include "rd_launch.h"

#include <hip/hip_runtime.h>

#include "fft2s.h"

// minimum waves per SIMD requested from the compiler for the transform kernels (caps VGPRs); the
// per-kernel values apply to plans of at most 16 values per thread, longer plans use 2
#ifndef GLIA_COL_MINW
#define GLIA_COL_MINW 2
#endif
#ifndef GLIA_XI_MINW
#define GLIA_XI_MINW 3
#endif
#ifndef GLIA_YI_MINW
#define GLIA_YI_MINW 3
#endif
#ifndef GLIA_YF_MINW
#define GLIA_YF_MINW 2
#endif
#ifndef GLIA_XF_MINW
#define GLIA_XF_MINW 2
#endif
// xf_div loads p before its transforms (1) or in its output stage (0)
// xf_div loads p at kernel start up to this many values per thread (the 192-point plan has 24)
#ifndef GLIA_XF_PF_VPT
#define GLIA_XF_PF_VPT 24
#endif
#ifndef GLIA_XF_PF
#define GLIA_XF_PF 1
#endif
#ifndef GLIA_Z_MINW
#define GLIA_Z_MINW 2
#endif
// threads per block of the line (Z pass) kernels
#ifndef GLIA_Z_TPB
#define GLIA_Z_TPB 256
#endif

namespace glia {
using namespace dev;

// Non-temporal bits of the passes (bit 0 loads, bit 1 stores; see GRefT), chosen by timing: on the
// 8 x 256^3 evaluation the CG passes gain about 5% over none (both X passes stream their loads only:
// the next pass reads their outputs at once); on the 160^3 elasticity solve the plain column
// transforms and the pointwise Z passes gain about 4% from streamed stores alone.
#ifndef GLIA_NT_XI
#define GLIA_NT_XI 1
#endif
#ifndef GLIA_NT_YI
#define GLIA_NT_YI 3
#endif
#ifndef GLIA_NT_ZD
#define GLIA_NT_ZD 2
#endif
#ifndef GLIA_NT_YF
#define GLIA_NT_YF 3
#endif
#ifndef GLIA_NT_XF
#define GLIA_NT_XF 1
#endif
#ifndef GLIA_NT_PU
#define GLIA_NT_PU 3
#endif
// the plain column transforms and the Z passes with a pointwise stage
#ifndef GLIA_NT_COL
#define GLIA_NT_COL 2
#endif
#ifndef GLIA_NT_ZP
#define GLIA_NT_ZP 2
#endif

// ---------------------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------------------
// pointers into device global memory (address space 1): loads and stores through them compile to
// global_* instructions with a uniform base and 32-bit lane offsets instead of flat accesses
#define GLIA_GLOBAL __attribute__((address_space(1)))
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
// Cache policy of a kernel's global accesses, a template argument of its pointers: bit 0 makes
// the loads, bit 1 the stores non-temporal (the nt bit: the line streams through the caches instead
// of being kept for reuse).  Each pass sets its own (GLIA_NT_* below).
template <int H>
struct GRefT {
  GLIA_GLOBAL v2f* p;
  __device__ __forceinline__ operator float2() const {
    v2f a;
    if constexpr ((H & 1) != 0) {
      a = __builtin_nontemporal_load(p);
    } else {
      a = *p;
    }
    return make_float2(a.x, a.y);
  }
  __device__ __forceinline__ GRefT& operator=(float2 v) {
    v2f a;
    a.x = v.x;
    a.y = v.y;
    if constexpr ((H & 2) != 0) {
      __builtin_nontemporal_store(a, p);
    } else {
      *p = a;
    }
    return *this;
  }
};
template <int H>
struct GPtrT {
  GLIA_GLOBAL v2f* p;
  __device__ __forceinline__ GRefT<H> operator[](unsigned i) const { return GRefT<H>{p + i}; }
  __device__ __forceinline__ GPtrT operator+(int64_t o) const { return GPtrT{p + o}; }
  __device__ __forceinline__ explicit operator bool() const { return p != nullptr; }
};
using GPtr = GPtrT<0>;
// 16-byte accesses of the elementwise passes under the same policy bits
template <int H>
__device__ __forceinline__ v4f ld4(GLIA_GLOBAL v4f* p) {
  if constexpr ((H & 1) != 0) {
    return __builtin_nontemporal_load(p);
  } else {
    return *p;
  }
}
template <int H>
__device__ __forceinline__ void st4(GLIA_GLOBAL v4f* p, v4f v) {
  if constexpr ((H & 2) != 0) {
    __builtin_nontemporal_store(v, p);
  } else {
    *p = v;
  }
}
template <int H = 0>
__device__ __forceinline__ GPtrT<H> gptr(const void* q) { return GPtrT<H>{(GLIA_GLOBAL v2f*)q}; }

// complex slot s of a member: work slots below CS_NUM, state-pool entries (history/checkpoints) above
template <int H = 0>
__device__ __forceinline__ GPtrT<H> cptr(const MemberDev& M, int s) {
  return gptr<H>((s < CS_NUM) ? M.c[s] : M.hist[s - CS_NUM]);
}
// the same slot as a plain pointer (elementwise kernels)
__device__ __forceinline__ float2* craw(const MemberDev& M, int s) { return (s < CS_NUM) ? M.c[s] : M.hist[s - CS_NUM]; }
// a real field viewed as float2 pairs (x[2n], x[2n+1])
template <int H = 0>
__device__ __forceinline__ GPtrT<H> rptr2(const MemberDev& M, int r) { return gptr<H>(M.r[r]); }

// wave number with the Nyquist mode zeroed (reference src/cuda/SpectralOperators.cu:63-72)
__device__ __forceinline__ float wnum(int k, int n) {
  const int h = n >> 1;
  return (k < h) ? (float)k : ((k == h) ? 0.f : (float)(k - n));
}

// spectral preconditioner factor times N^3 (reference src/cuda/DiffCoef.cu:11-63):
// FFT(B r) = r_hat / (1 + 0.25 dt kbar |w|^2), or r_hat when the denominator is 0
__device__ __forceinline__ float precf(float pc, float wx, float wy, float wz) {
  const float v = 1.0f + pc * (wx * wx + wy * wy + wz * wz);
  return (v == 0.0f) ? 1.0f : 1.0f / v;
}

// block-wide FP64 reduction of NQ values; thread 0 writes out[0..NQ)
template <int NQ>
__device__ __forceinline__ void block_reduce_store(double (&v)[NQ], double* out) {
  __shared__ double red[NQ][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[q][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += red[q][w];
      out[q] = s;
    }
  }
}

// block-wide sum of one value, result broadcast to every thread
__device__ __forceinline__ double block_sum_all(double x) {
  __shared__ double red[16];
  __shared__ double res;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  if (lane == 0) red[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < nw; w++) s += red[w];
    res = s;
  }
  __syncthreads();
  const double r = res;
  __syncthreads();
  return r;
}

// sum of column q of part[nb][nq], identical on every thread (fixed order: deterministic)
__device__ __forceinline__ double sum_partials(const double* part, int nb, int nq, int q) {
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(int64_t)b * nq + q];
  return block_sum_all(s);
}

// ---------------------------------------------------------------------------------------------
// column kernels (X and Y passes): a tile of 16 adjacent columns (128-byte rows), NT threads per
// column; thread (col = tid & 15, t = tid >> 4) loads its stage-1 rows straight from HBM, the
// two-stage transform exchanges once through LDS, results are stored from registers.
// ---------------------------------------------------------------------------------------------
template <int L>
struct CK {
  static constexpr int R1 = Plan2<L>::R1, R2 = Plan2<L>::R2, NT = Plan2<L>::NT;
  static constexpr int TPB = 16 * NT;
  // stage 1 gives each thread R2 / NT columns (NT divides R2); stage 2 gives it ceil(R1 / NT) rows,
  // the last of which is missing on some threads when NT does not divide R1 (uneven plans)
  static constexpr int B1 = R2 / NT, B2 = (R1 + NT - 1) / NT;
  static constexpr bool EVEN = B2 * NT == R1;
  static constexpr int VIN = R1 * B1, VOUT = R2 * B2;  // input and output values per thread
  static constexpr int VPT = VIN > VOUT ? VIN : VOUT;
  // output slot j = b * R2 + k2 holds stage-2 row k1 = t + b * NT, which exists when k1 < R1
  __device__ __forceinline__ static bool out_ok(int t, int j) { return EVEN || t + (j / R2) * NT < R1; }
  static constexpr int CP = XPitch<L>::value;
  // a kernel's later operands are loaded at kernel start when the extra registers fit
  static constexpr bool PF = VPT <= 16;
};

// row of stage-1 input slot j of thread t: j = b*R1 + n1 -> R2*n1 + t + b*NT
template <int R1, int R2, int NT>
__device__ __forceinline__ int in_row(int t, int j) {
  return R2 * (j % R1) + t + (j / R1) * NT;
}
// row of output slot j of thread t: j = b*R2 + k2 -> t + b*NT + R1*k2
template <int R1, int R2, int NT>
__device__ __forceinline__ int out_row(int t, int j) {
  return t + (j / R2) * NT + R1 * (j % R2);
}

struct Tile {
  int64_t base;   // element offset of the tile's first entry (uniform)
  unsigned rs;    // row stride in elements
  int o, c0;
};

__device__ __forceinline__ Tile col_tile_at(int dir, const Dims& d, int idx) {
  const int ntc = d.m2 >> 4;
  Tile t;
  t.o = idx / ntc;
  t.c0 = (idx % ntc) << 4;
  if (dir == 0) {
    t.base = (int64_t)t.o * d.m2 + t.c0;
    t.rs = (unsigned)(d.n1 * d.m2);
  } else {
    t.base = (int64_t)t.o * d.n1 * d.m2 + t.c0;
    t.rs = (unsigned)d.m2;
  }
  return t;
}
__device__ __forceinline__ Tile col_tile(int dir, const Dims& d) { return col_tile_at(dir, d, (int)blockIdx.x); }

// Twiddle table of a block, staged in LDS after the block's data loads have been issued.  Read
// from LDS after a barrier, the twiddles of each transform are loaded where they are used instead
// of being held in registers across the whole kernel.
template <int L, int TPB>
__device__ __forceinline__ void stage_twiddles(float2* twl, const float2* __restrict__ tw) {
#pragma unroll
  for (int i = 0; i < (L + TPB - 1) / TPB; i++) {
    const int e = (int)threadIdx.x + i * TPB;
    if (e < L) twl[e] = tw[e];
  }
  __syncthreads();
}

#define GLIA_COL_PRELUDE(L)                                  \
  using K = CK<L>;                                           \
  constexpr int R1 = K::R1, R2 = K::R2, NT = K::NT;          \
  __shared__ float2 ex[16 * K::CP];                          \
  const int col = threadIdx.x & 15, t = threadIdx.x >> 4;    \
  float2* exc = ex + col * K::CP;

// terms of the cross-lane spectral combinations (members used as lanes, at most 3): output lane l
// gets i sum_t w_{dir[l][t]} X_t with X_t the slot slot[l][t] of lane lane[l][t]
struct LaneTerms {
  int n[3];
  int lane[3][4], slot[3][4], dir[3][4];
};

// X pass, inverse, of a cross-lane combination formed as the columns are loaded: lane l = blockIdx.y
// transforms i sum_t w_{dir[l][t]} X_t (NTM terms, summed in k_lane_lin's order) into s_out.  It
// replaces a k_lane_lin launch followed by k_col_plain<L, 0, +1> without the intermediate field.
template <int L, int NTM>
__global__ void __launch_bounds__(CK<L>::TPB, GLIA_COL_MINW)
k_col_lin(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, LaneTerms T, int s_out, int honor) {
  constexpr int H = GLIA_NT_COL;
  const int l = blockIdx.y;
  const MemberDev& M = mem[l];
  if (honor && M.st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(0, d);
  GPtrT<H> src[NTM];
  int dir[NTM];
#pragma unroll
  for (int q = 0; q < NTM; q++) {
    src[q] = cptr<H>(mem[T.lane[l][q]], T.slot[l][q]) + tl.base;
    dir[q] = T.dir[l][q];
  }
  GPtrT<H> O = cptr<H>(M, s_out) + tl.base;
  // this tile's columns: ky = tile index o, kz = c0 + col; the rows run over kx
  const int ky = tl.o, kz = tl.c0 + col;
  const float wy = (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1), wz = (float)kz;
  float2 a[K::VPT];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const int row = in_row<R1, R2, NT>(t, j);
    const unsigned g = (unsigned)row * tl.rs + col;
    const float wx = (2 * row == d.n0) ? 0.0f : (float)(2 * row < d.n0 ? row : row - d.n0);
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int q = 0; q < NTM; q++) {
      const float2 x = src[q][g];
      const float wt = dir[q] == 0 ? wx : (dir[q] == 1 ? wy : wz);
      re += wt * x.x;
      im += wt * x.y;
    }
    a[j] = make_float2(-im, re);
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, 1>(a, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) O[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
}

// X pass, inverse: U = IFFT_x(p), V = IFFT_x(i w_x p).  With PUPD the search direction is
// first updated in place: p = B r (i == 0) or p = B r + (beta/betaold) p.
template <int L, bool PUPD>
__global__ void __launch_bounds__(CK<L>::TPB, CK<L>::VPT <= 16 ? GLIA_XI_MINW : 2)
k_xi_grad(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_in, int s_r, int s_u, int s_v) {
  constexpr int H = GLIA_NT_XI;
  const MemberDev& M = mem[blockIdx.y];
  const PcgState* st = M.st;
  if (st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(0, d);
  const float wy = wnum(tl.o, d.n1), wz = (float)(tl.c0 + col);
  GPtrT<H> pin = cptr<H>(M, s_in) + tl.base;
  GPtrT<H> rr = cptr<H>(M, s_r) + tl.base;
  GPtrT<H> U = cptr<H>(M, s_u) + tl.base;
  GPtrT<H> V = cptr<H>(M, s_v) + tl.base;
  bool first = false;
  float bc = 0.f, pc = 0.f;
  if (PUPD) {
    first = (st->i == 0);
    bc = (float)st->bcoef;
    pc = (float)st->pc;
  }
  float2 a[K::VPT], b[K::VPT];
  // 'first' is uniform: the two load loops are kept separate so that every load of a loop is
  // issued before the first use
  if (PUPD && first) {
#pragma unroll
    for (int j = 0; j < K::VIN; j++) a[j] = rr[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  } else if (PUPD) {
#pragma unroll
    for (int j = 0; j < K::VIN; j++) {
      const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
      a[j] = rr[g];
      b[j] = pin[g];
    }
  } else {
#pragma unroll
    for (int j = 0; j < K::VIN; j++) a[j] = pin[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  }
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const int row = in_row<R1, R2, NT>(t, j);
    const float wx = wnum(row, d.n0);
    float2 p = a[j];
    if (PUPD) {
      p = cscale(p, precf(pc, wx, wy, wz));
      if (!first) p = make_float2(p.x + bc * b[j].x, p.y + bc * b[j].y);
      pin[(unsigned)row * tl.rs + col] = p;
    }
    a[j] = p;
    b[j] = mul_i(cscale(p, wx));
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, +1>(a, exc, t, twl);
  __syncthreads();
  fft2s<R1, R2, NT, +1>(b, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) {
    if (!K::out_ok(t, j)) continue;
    const unsigned g = (unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col;
    U[g] = a[j];
    V[g] = b[j];
  }
}

// Y pass, inverse, gradient: Gx = IFFT_y(V), Gy = IFFT_y(i w_y U), Gz = i w_z IFFT_y(U),
// and optionally G0 = IFFT_y(U).  Gx must not alias U.
template <int L, bool WITH0>
__global__ void __launch_bounds__(CK<L>::TPB, CK<L>::VPT <= 16 ? GLIA_YI_MINW : 2)
k_yi_grad(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_u, int s_v, int s_gx, int s_gy,
          int s_gz, int s_g0) {
  constexpr int H = GLIA_NT_YI;
  const MemberDev& M = mem[blockIdx.y];
  if (M.st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(1, d);
  const float wz = (float)(tl.c0 + col);
  GPtrT<H> U = cptr<H>(M, s_u) + tl.base;
  GPtrT<H> V = cptr<H>(M, s_v) + tl.base;
  GPtrT<H> Gx = cptr<H>(M, s_gx) + tl.base;
  GPtrT<H> Gy = cptr<H>(M, s_gy) + tl.base;
  GPtrT<H> Gz = cptr<H>(M, s_gz) + tl.base;
  GPtrT<H> G0 = WITH0 ? cptr<H>(M, s_g0) + tl.base : gptr<H>(nullptr);
  float2 a[K::VPT], b[K::VPT], u[K::PF ? K::VPT : 1];
  // both inputs are loaded up front (when registers allow) so that their latencies overlap
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
    a[j] = V[g];
    if (K::PF) u[j] = U[g];
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, +1>(a, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) Gx[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const int row = in_row<R1, R2, NT>(t, j);
    a[j] = K::PF ? u[j] : U[(unsigned)row * tl.rs + col];
    b[j] = mul_i(cscale(a[j], wnum(row, d.n1)));
  }
  __syncthreads();
  fft2s<R1, R2, NT, +1>(a, exc, t, twl);
  __syncthreads();
  fft2s<R1, R2, NT, +1>(b, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) {
    if (!K::out_ok(t, j)) continue;
    const unsigned g = (unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col;
    Gz[g] = mul_i(cscale(a[j], wz));
    if (WITH0) G0[g] = a[j];
    Gy[g] = b[j];
  }
}


// Y pass, forward, divergence: Yx = FFT_y(Hx), Rr = FFT_y(Q) + i w_y FFT_y(Hy).
// Yx must not alias Q or Hy.
template <int L>
__global__ void __launch_bounds__(CK<L>::TPB, CK<L>::VPT <= 16 ? GLIA_YF_MINW : 2)
k_yf_div(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_hx, int s_hy, int s_q, int s_yx,
         int s_rr) {
  constexpr int H = GLIA_NT_YF;
  const MemberDev& M = mem[blockIdx.y];
  if (M.st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(1, d);
  GPtrT<H> Hx = cptr<H>(M, s_hx) + tl.base;
  GPtrT<H> Hy = cptr<H>(M, s_hy) + tl.base;
  GPtrT<H> Q = cptr<H>(M, s_q) + tl.base;
  GPtrT<H> Yx = cptr<H>(M, s_yx) + tl.base;
  GPtrT<H> Rr = cptr<H>(M, s_rr) + tl.base;
  float2 a[K::VPT], b[K::VPT], q[K::PF ? K::VPT : 1];
  // all three inputs are loaded up front (when registers allow) so that their latencies overlap
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
    a[j] = Hx[g];
    if (K::PF) {
      q[j] = Q[g];
      b[j] = Hy[g];
    }
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) Yx[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
    a[j] = K::PF ? q[j] : Q[g];
    if (!K::PF) b[j] = Hy[g];
  }
  __syncthreads();
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  __syncthreads();
  fft2s<R1, R2, NT, -1>(b, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) {
    if (!K::out_ok(t, j)) continue;
    const int row = out_row<R1, R2, NT>(t, j);
    Rr[(unsigned)row * tl.rs + col] = cadd(a[j], mul_i(cscale(b[j], wnum(row, d.n1))));
  }
}


// X pass, forward, closing the divergence: D = i w_x FFT_x(Yx) + FFT_x(Rr) = FFT(div(k grad p)).
//   XF_MATVEC: w = p - (tau/2) D; partial <p, w>
//   XF_SETUP : r = tau D (= b - A x for x0 = c); partials ||B b||^2, ||B r||^2, <B r, r>
//              with b = x + (tau/2) D
//   XF_D     : out = D
enum { XF_MATVEC = 0, XF_SETUP = 1, XF_D = 2 };

template <int L, int MODE>
__global__ void __launch_bounds__(CK<L>::TPB, CK<L>::VPT <= 16 ? GLIA_XF_MINW : 2)
k_xf_div(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_yx, int s_rr, int s_p, int s_out) {
  constexpr int H = GLIA_NT_XF;
  const MemberDev& M = mem[blockIdx.y];
  const PcgState* st = M.st;
  if (st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(0, d);
  GPtrT<H> Yx = cptr<H>(M, s_yx) + tl.base;
  GPtrT<H> Rr = cptr<H>(M, s_rr) + tl.base;
  GPtrT<H> P = cptr<H>(M, s_p) + tl.base;
  GPtrT<H> O = cptr<H>(M, s_out) + tl.base;
  const float tau = (float)st->tau;
  const float pc = (float)st->pc;
  const float wy = wnum(tl.o, d.n1), wz = (float)(tl.c0 + col);
  const double wgt = (tl.c0 + col == 0) ? 1.0 : 2.0;
  constexpr bool PF = K::VPT <= GLIA_XF_PF_VPT && GLIA_XF_PF;
  float2 a[K::VPT], b[K::VPT], pv[PF ? K::VPT : 1];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
    a[j] = Yx[g];
    b[j] = Rr[g];
  }
  // the operand of the output stage is loaded before the transforms (a barrier would otherwise
  // keep its loads behind them)
  if (MODE != XF_D && PF) {
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) pv[j] = P[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col];
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  __syncthreads();
  fft2s<R1, R2, NT, -1>(b, exc, t, twl);
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) {
    if (!K::out_ok(t, j)) continue;
    const int row = out_row<R1, R2, NT>(t, j);
    const unsigned g = (unsigned)row * tl.rs + col;
    const float wx = wnum(row, d.n0);
    const float2 D = cadd(mul_i(cscale(a[j], wx)), b[j]);
    if (MODE == XF_MATVEC) {
      const float2 p = PF ? pv[j] : P[g];
      const float2 w = make_float2(p.x - 0.5f * tau * D.x, p.y - 0.5f * tau * D.y);
      O[g] = w;
      acc[0] += wgt * ((double)p.x * w.x + (double)p.y * w.y);
    } else if (MODE == XF_SETUP) {
      const float2 x = PF ? pv[j] : P[g];
      const float2 bb = make_float2(x.x + 0.5f * tau * D.x, x.y + 0.5f * tau * D.y);
      const float2 r = cscale(D, tau);
      O[g] = r;
      const float f = precf(pc, wx, wy, wz);
      const double r2 = (double)r.x * r.x + (double)r.y * r.y;
      acc[0] += wgt * (double)f * f * ((double)bb.x * bb.x + (double)bb.y * bb.y);
      acc[1] += wgt * (double)f * f * r2;
      acc[2] += wgt * (double)f * r2;
    } else {
      O[g] = D;
    }
  }
  if (MODE == XF_SETUP) block_reduce_store<3>(acc, M.part + (int64_t)blockIdx.x * 3);
  if (MODE == XF_MATVEC) {
    double a1[1] = {acc[0]};
    block_reduce_store<1>(a1, M.part + blockIdx.x);
  }
}

// ---------------------------------------------------------------------------------------------
// Split-group variants of the two divergence passes: each field's transform runs on its own group
// of 16 * NT threads with its own exchange area, so a thread holds one field's values instead of
// two or three; the groups meet once more in LDS to combine.  The arithmetic of every output value
// is that of k_yf_div / k_xf_div (same operations in the same order), so the results are the same
// bit for bit.  Groups are whole waves (16 * NT is a multiple of 64 for the plans used), and every
// group passes the same barriers.
// ---------------------------------------------------------------------------------------------
template <int L>
struct CKS {
  using K = CK<L>;
  static constexpr bool OK = (K::TPB % 64) == 0;  // groups must be whole waves
};

// Y pass, forward, divergence in three groups: Hx -> Yx = FFT_y(Hx); Hy -> i w_y FFT_y(Hy) into
// LDS; Q -> Rr = FFT_y(Q) + (that term).
template <int L>
__global__ void __launch_bounds__(3 * CK<L>::TPB)
k_yf_div_s(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_hx, int s_hy, int s_q, int s_yx,
           int s_rr) {
  constexpr int H = GLIA_NT_YF;
  static_assert(CKS<L>::OK, "thread groups must be whole waves");
  const MemberDev& M = mem[blockIdx.y];
  if (M.st->done) return;
  using K = CK<L>;
  constexpr int R1 = K::R1, R2 = K::R2, NT = K::NT;
  __shared__ float2 ex[3][16 * K::CP];
  const int grp = (int)threadIdx.x / K::TPB;
  const int lt = (int)threadIdx.x - grp * K::TPB;
  const int col = lt & 15, t = lt >> 4;
  float2* exc = ex[grp] + col * K::CP;
  const Tile tl = col_tile(1, d);
  GPtrT<H> In = cptr<H>(M, grp == 0 ? s_hx : (grp == 1 ? s_hy : s_q)) + tl.base;
  float2 a[K::VPT];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) a[j] = In[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  __shared__ float2 twl[L];
  stage_twiddles<L, 3 * K::TPB>(twl, tw);
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  __syncthreads();  // every exchange area is free again
  if (grp == 0) {
    GPtrT<H> Yx = cptr<H>(M, s_yx) + tl.base;
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) Yx[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
  } else if (grp == 1) {
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      exc[xpad(row)] = mul_i(cscale(a[j], wnum(row, d.n1)));
    }
  }
  __syncthreads();
  if (grp == 2) {
    GPtrT<H> Rr = cptr<H>(M, s_rr) + tl.base;
    const float2* exh = ex[1] + col * K::CP;
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      Rr[(unsigned)row * tl.rs + col] = cadd(a[j], exh[xpad(row)]);
    }
  }
}

// Y pass, forward, divergence split over blocks: block 2 i transforms Hx -> Yx = FFT_y(Hx) for tile
// i, block 2 i + 1 computes Rr = FFT_y(Q) + i w_y FFT_y(Hy) for the same tile.  A thread holds at
// most two fields' values (k_yf_div holds three), so more waves fit per SIMD, and no exchange
// between the halves is needed.  Each output value takes the operations of k_yf_div.  MINW is the
// occupancy asked of the compiler (at 256 points, 3 waves per SIMD costs a few spilled registers;
// the autotuner times both bounds).
template <int L, int MINW>
__global__ void __launch_bounds__(CK<L>::TPB, CK<L>::VPT <= 16 ? MINW : 2)
k_yf_div_b(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_hx, int s_hy, int s_q, int s_yx,
           int s_rr) {
  constexpr int H = GLIA_NT_YF;
  const MemberDev& M = mem[blockIdx.y];
  if (M.st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile_at(1, d, (int)(blockIdx.x >> 1));
  __shared__ float2 twl[L];
  if ((blockIdx.x & 1) == 0) {  // uniform per block
    GPtrT<H> Hx = cptr<H>(M, s_hx) + tl.base;
    GPtrT<H> Yx = cptr<H>(M, s_yx) + tl.base;
    float2 a[K::VPT];
#pragma unroll
    for (int j = 0; j < K::VIN; j++) a[j] = Hx[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
    stage_twiddles<L, K::TPB>(twl, tw);
    fft2s<R1, R2, NT, -1>(a, exc, t, twl);
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) Yx[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
    return;
  }
  GPtrT<H> Hy = cptr<H>(M, s_hy) + tl.base;
  GPtrT<H> Q = cptr<H>(M, s_q) + tl.base;
  GPtrT<H> Rr = cptr<H>(M, s_rr) + tl.base;
  float2 a[K::VPT], b[K::VPT];
  // Hy is loaded up front for short plans, else after the transform of Q (fewer live values)
  constexpr bool early = K::VPT <= 8;
#pragma unroll
  for (int j = 0; j < K::VIN; j++) {
    const unsigned g = (unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col;
    a[j] = Q[g];
    if (early) b[j] = Hy[g];
  }
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  if (!early) {
#pragma unroll
    for (int j = 0; j < K::VIN; j++) b[j] = Hy[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  }
  __syncthreads();
  fft2s<R1, R2, NT, -1>(b, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) {
    if (!K::out_ok(t, j)) continue;
    const int row = out_row<R1, R2, NT>(t, j);
    Rr[(unsigned)row * tl.rs + col] = cadd(a[j], mul_i(cscale(b[j], wnum(row, d.n1))));
  }
}

// X pass, forward, closing the divergence in two groups: Yx -> i w_x FFT_x(Yx) into LDS; Rr ->
// D = (that term) + FFT_x(Rr), then the MODE's output and partial sums as in k_xf_div.
template <int L, int MODE>
__global__ void __launch_bounds__(2 * CK<L>::TPB)
k_xf_div_s(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_yx, int s_rr, int s_p,
           int s_out) {
  constexpr int H = GLIA_NT_XF;
  static_assert(CKS<L>::OK, "thread groups must be whole waves");
  const MemberDev& M = mem[blockIdx.y];
  const PcgState* st = M.st;
  if (st->done) return;
  using K = CK<L>;
  constexpr int R1 = K::R1, R2 = K::R2, NT = K::NT;
  __shared__ float2 ex[2][16 * K::CP];
  const int grp = (int)threadIdx.x / K::TPB;
  const int lt = (int)threadIdx.x - grp * K::TPB;
  const int col = lt & 15, t = lt >> 4;
  float2* exc = ex[grp] + col * K::CP;
  const Tile tl = col_tile(0, d);
  GPtrT<H> In = cptr<H>(M, grp == 0 ? s_yx : s_rr) + tl.base;
  GPtrT<H> P = cptr<H>(M, s_p) + tl.base;
  GPtrT<H> O = cptr<H>(M, s_out) + tl.base;
  const float tau = (float)st->tau;
  const float pc = (float)st->pc;
  const float wy = wnum(tl.o, d.n1), wz = (float)(tl.c0 + col);
  const double wgt = (tl.c0 + col == 0) ? 1.0 : 2.0;
  float2 a[K::VPT], pv[K::VPT];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) a[j] = In[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  if (MODE != XF_D && grp == 1) {
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) pv[j] = P[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col];
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, 2 * K::TPB>(twl, tw);
  fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  __syncthreads();  // both exchange areas are free again
  if (grp == 0) {
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      exc[xpad(row)] = mul_i(cscale(a[j], wnum(row, d.n0)));
    }
  }
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  if (grp == 1) {
    const float2* exa = ex[0] + col * K::CP;
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      const unsigned g = (unsigned)row * tl.rs + col;
      const float2 D = cadd(exa[xpad(row)], a[j]);
      if (MODE == XF_MATVEC) {
        const float2 p = pv[j];
        const float2 w = make_float2(p.x - 0.5f * tau * D.x, p.y - 0.5f * tau * D.y);
        O[g] = w;
        acc[0] += wgt * ((double)p.x * w.x + (double)p.y * w.y);
      } else if (MODE == XF_SETUP) {
        const float2 x = pv[j];
        const float2 bb = make_float2(x.x + 0.5f * tau * D.x, x.y + 0.5f * tau * D.y);
        const float2 r = cscale(D, tau);
        O[g] = r;
        const float f = precf(pc, wnum(row, d.n0), wy, wz);
        const double r2 = (double)r.x * r.x + (double)r.y * r.y;
        acc[0] += wgt * (double)f * f * ((double)bb.x * bb.x + (double)bb.y * bb.y);
        acc[1] += wgt * (double)f * f * r2;
        acc[2] += wgt * (double)f * r2;
      } else {
        O[g] = D;
      }
    }
  }
  if (MODE == XF_SETUP) block_reduce_store<3>(acc, M.part + (int64_t)blockIdx.x * 3);
  if (MODE == XF_MATVEC) {
    double a1[1] = {acc[0]};
    block_reduce_store<1>(a1, M.part + blockIdx.x);
  }
}

// X pass, forward, closing the divergence in three groups: Yx -> i w_x FFT_x(Yx) into LDS; Rr ->
// FFT_x(Rr) into LDS; the third group loads the operand of the output stage at kernel start and
// forms D = (first) + (second) and the MODE's output and partial sums as in k_xf_div_s.  A thread
// holds one field's values, so more waves fit per SIMD than in the two-group kernel; every output
// value takes the same operations in the same order (the same results bit for bit).
template <int L, int MODE>
__global__ void __launch_bounds__(3 * CK<L>::TPB)
k_xf_div_s3(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_yx, int s_rr, int s_p,
            int s_out) {
  static_assert(CKS<L>::OK, "thread groups must be whole waves");
  constexpr int H = GLIA_NT_XF;
  const MemberDev& M = mem[blockIdx.y];
  const PcgState* st = M.st;
  if (st->done) return;
  using K = CK<L>;
  constexpr int R1 = K::R1, R2 = K::R2, NT = K::NT;
  __shared__ float2 ex[2][16 * K::CP];
  const int grp = (int)threadIdx.x / K::TPB;
  const int lt = (int)threadIdx.x - grp * K::TPB;
  const int col = lt & 15, t = lt >> 4;
  float2* exc = ex[grp == 1 ? 1 : 0] + col * K::CP;
  const Tile tl = col_tile(0, d);
  const float tau = (float)st->tau;
  const float pc = (float)st->pc;
  const float wy = wnum(tl.o, d.n1), wz = (float)(tl.c0 + col);
  const double wgt = (tl.c0 + col == 0) ? 1.0 : 2.0;
  float2 a[K::VPT];
  if (grp < 2) {
    GPtrT<H> In = cptr<H>(M, grp == 0 ? s_yx : s_rr) + tl.base;
#pragma unroll
    for (int j = 0; j < K::VIN; j++) a[j] = In[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  } else if (MODE != XF_D) {
    GPtrT<H> P = cptr<H>(M, s_p) + tl.base;
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) a[j] = P[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col];
  }
  __shared__ float2 twl[L];
  stage_twiddles<L, 3 * K::TPB>(twl, tw);
  if (grp < 2)
    fft2s<R1, R2, NT, -1>(a, exc, t, twl);
  else
    __syncthreads();  // the one barrier of fft2s
  __syncthreads();    // both exchange areas are free again
  if (grp < 2) {
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      exc[xpad(row)] = grp == 0 ? mul_i(cscale(a[j], wnum(row, d.n0))) : a[j];
    }
  }
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  if (grp == 2) {
    GPtrT<H> O = cptr<H>(M, s_out) + tl.base;
    const float2* exa = ex[0] + col * K::CP;
    const float2* exb = ex[1] + col * K::CP;
#pragma unroll
    for (int j = 0; j < K::VOUT; j++) {
      if (!K::out_ok(t, j)) continue;
      const int row = out_row<R1, R2, NT>(t, j);
      const unsigned g = (unsigned)row * tl.rs + col;
      const float2 D = cadd(exa[xpad(row)], exb[xpad(row)]);
      if (MODE == XF_MATVEC) {
        const float2 p = a[j];
        const float2 w = make_float2(p.x - 0.5f * tau * D.x, p.y - 0.5f * tau * D.y);
        O[g] = w;
        acc[0] += wgt * ((double)p.x * w.x + (double)p.y * w.y);
      } else if (MODE == XF_SETUP) {
        const float2 x = a[j];
        const float2 bb = make_float2(x.x + 0.5f * tau * D.x, x.y + 0.5f * tau * D.y);
        const float2 r = cscale(D, tau);
        O[g] = r;
        const float f = precf(pc, wnum(row, d.n0), wy, wz);
        const double r2 = (double)r.x * r.x + (double)r.y * r.y;
        acc[0] += wgt * (double)f * f * ((double)bb.x * bb.x + (double)bb.y * bb.y);
        acc[1] += wgt * (double)f * f * r2;
        acc[2] += wgt * (double)f * r2;
      } else {
        O[g] = D;
      }
    }
  }
  if (MODE == XF_SETUP) block_reduce_store<3>(acc, M.part + (int64_t)blockIdx.x * 3);
  if (MODE == XF_MATVEC) {
    double a1[1] = {acc[0]};
    block_reduce_store<1>(a1, M.part + blockIdx.x);
  }
}

// plain column transform: out = FFT_dir(in) along axis DIR (S = -1 forward, +1 inverse)
template <int L, int DIR, int S>
__global__ void __launch_bounds__(CK<L>::TPB, GLIA_COL_MINW)
k_col_plain(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ tw, int s_in, int s_out, int honor_done) {
  constexpr int H = GLIA_NT_COL;
  const MemberDev& M = mem[blockIdx.y];
  if (honor_done && M.st->done) return;
  GLIA_COL_PRELUDE(L)
  const Tile tl = col_tile(DIR, d);
  GPtrT<H> I = cptr<H>(M, s_in) + tl.base;
  GPtrT<H> O = cptr<H>(M, s_out) + tl.base;
  float2 a[K::VPT];
#pragma unroll
  for (int j = 0; j < K::VIN; j++) a[j] = I[(unsigned)in_row<R1, R2, NT>(t, j) * tl.rs + col];
  __shared__ float2 twl[L];
  stage_twiddles<L, K::TPB>(twl, tw);
  fft2s<R1, R2, NT, S>(a, exc, t, twl);
#pragma unroll
  for (int j = 0; j < K::VOUT; j++) if (K::out_ok(t, j)) O[(unsigned)out_row<R1, R2, NT>(t, j) * tl.rs + col] = a[j];
}

// ---------------------------------------------------------------------------------------------
// line kernels (Z pass): real n2 <-> packed m2 = n2/2 complex through one m2-point complex FFT.
// NL lines per block, NT threads per line.  The inverse uses plan (R1, R2); its register output
// is the input layout of the forward plan (R2, R1), so pointwise work between the two needs no
// exchange.  Thread t of a line holds z[n] for n = out_row(t, j).
// ---------------------------------------------------------------------------------------------
// threads per Z block; 96-point lines (24 values per thread, 149 VGPRs) take 16 lines per block:
// with 64 lines only two blocks fit in LDS, with 16 the register file sets occupancy (3 waves/SIMD)
#ifndef GLIA_Z_TPB96
#define GLIA_Z_TPB96 64
#endif
template <int M>
struct ZTpb {
  static constexpr int value = GLIA_Z_TPB;
};
template <>
struct ZTpb<96> {
  static constexpr int value = GLIA_Z_TPB96;
};
template <int M>
struct ZK {
  static constexpr int R1 = Plan2<M>::R1, R2 = Plan2<M>::R2, NT = Plan2<M>::NT;
  static constexpr int NL = ZTpb<M>::value / NT;
  static constexpr int TPB = NL * NT;
  static constexpr int LP = XPitch<M>::value;
  static constexpr int VPT = M / NT;
};

// coalesced load of the block's NL packed lines into the padded line buffers, in two halves:
// zl_fetch issues every global load of the block into registers (a compile-time trip count keeps
// the loop unrolled), zl_put writes them to LDS.  Between the two the caller may do other work.
template <int M>
struct ZIt {
  static constexpr int value = ZK<M>::NL * M / ZK<M>::TPB;
  static_assert(value * ZK<M>::TPB == ZK<M>::NL * M, "block must tile its lines");
};
template <int M, int H>
__device__ __forceinline__ void zl_fetch(float2 (&tmp)[ZIt<M>::value], GPtrT<H> src, int64_t base) {
  GPtrT<H> sb = src + base;
#pragma unroll
  for (int i = 0; i < ZIt<M>::value; i++) tmp[i] = sb[(int)threadIdx.x + i * ZK<M>::TPB];
}
template <int M>
__device__ __forceinline__ void zl_put(float2* lb, const float2 (&tmp)[ZIt<M>::value]) {
#pragma unroll
  for (int i = 0; i < ZIt<M>::value; i++) {
    const int e = (int)threadIdx.x + i * ZK<M>::TPB;
    const int li = e / M, k = e - li * M;
    lb[li * ZK<M>::LP + xpad(k)] = tmp[i];
  }
}
template <int M, int H>
__device__ __forceinline__ void zl_load(float2* lb, GPtrT<H> src, int64_t base) {
  float2 tmp[ZIt<M>::value];
  zl_fetch<M>(tmp, src, base);
  zl_put<M>(lb, tmp);
}

// both twiddle tables of a line kernel, staged in LDS (see stage_twiddles)
template <int M>
__device__ __forceinline__ void stage_line_twiddles(float2* twml, float2* tw2l, const float2* __restrict__ twm,
                                                    const float2* __restrict__ tw2) {
  constexpr int TPB = ZK<M>::TPB;
#pragma unroll
  for (int i = 0; i < (M + TPB - 1) / TPB; i++) {
    const int e = (int)threadIdx.x + i * TPB;
    if (e < M) {
      twml[e] = twm[e];
      tw2l[e] = tw2[e];
    }
  }
}

// packed spectrum entry k -> Z'[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/n2} (X[k] - conj X[M-k]),
// with X[0] = Re P[0], X[M] = Im P[0]; an inverse M-point FFT of Z' gives z[n] = x[2n] + i x[2n+1]
// (unnormalized C2R of length n2).
template <int M>
__device__ __forceinline__ float2 zprep(const float2* line, int k, const float2* __restrict__ tw2) {
  // branch-free: the k = 0 entry reads itself as its mirror and is replaced by a select
  const float2 a = line[xpad(k)];
  const float2 bc = cconj(line[xpad(k == 0 ? 0 : M - k)]);
  const float2 g = cadd(cadd(a, bc), mul_i(cmulc(csub(a, bc), tw2[k])));
  return (k == 0) ? make_float2(a.x + a.y, a.x - a.y) : g;
}

// after a forward M-point FFT of z[n] = x[2n] + i x[2n+1]: packed spectrum entry k
template <int M>
__device__ __forceinline__ float2 zpost(const float2* line, int k, const float2* __restrict__ tw2) {
  const float2 a = line[xpad(k)];
  const float2 bc = cconj(line[xpad(k == 0 ? 0 : M - k)]);
  const float2 e = cscale(cadd(a, bc), 0.5f);
  const float2 o = cscale(mul_mi(csub(a, bc)), 0.5f);
  const float2 g = cadd(e, cmul(o, tw2[k]));
  return (k == 0) ? make_float2(a.x + a.y, a.x - a.y) : g;
}

// inverse: the block's packed lines (already in lb) -> v[j] = z[out_row(t, j)] of line li
template <int M>
__device__ __forceinline__ void zl_inverse(float2* v, float2* line, int t, const float2* __restrict__ twm,
                                           const float2* __restrict__ tw2) {
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT;
#pragma unroll
  for (int j = 0; j < ZK<M>::VPT; j++) v[j] = zprep<M>(line, in_row<R1, R2, NT>(t, j), tw2);
  __syncthreads();
  fft2s<R1, R2, NT, +1>(v, line, t, twm);
}

// forward: v[j] = z[out_row(t, j)] -> packed spectrum of every line, stored block-cooperatively
// to dst[base ...] (multiplied by i w_z = i k when mulz)
// real output of a Z line kernel (C2R times scale), stored coalesced: the values go back to the
// line buffer at their line positions and the block writes its lines in order.  With SCREEN the
// stored value is c - a u, a and u read in the same order (the screened elasticity operator).
template <int M, bool SCREEN, int NDOT, int H>
__device__ __forceinline__ void zl_real_store(const float2* v, float2* lb, float2* line, int t, float scale, GPtrT<H> out,
                                              GPtrT<H> ap, GPtrT<H> up, double* acc) {
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT;
  __syncthreads();  // every thread's inverse has read the line buffer
#pragma unroll
  for (int j = 0; j < ZK<M>::VPT; j++) line[xpad(out_row<R1, R2, NT>(t, j))] = cscale(v[j], scale);
  __syncthreads();
  constexpr int TPB = ZK<M>::TPB, IT = ZK<M>::NL * M / TPB;
#pragma unroll 8
  for (int i = 0; i < IT; i++) {
    const int e = (int)threadIdx.x + i * TPB;
    const int li = e / M, k = e - li * M;
    const float2 c = lb[li * ZK<M>::LP + xpad(k)];
    if (SCREEN) {
      const float2 a = ap[e], u = up[e];
      const float2 w = make_float2(c.x - a.x * u.x, c.y - a.y * u.y);
      out[e] = w;
      if (NDOT == 1) acc[0] += (double)u.x * (double)w.x + (double)u.y * (double)w.y;  // <u, w>
    } else {
      out[e] = c;
      if (NDOT == 2) {  // <c, c> and <c, r> with r from up
        const float2 r = up[e];
        acc[0] += (double)c.x * (double)c.x + (double)c.y * (double)c.y;
        acc[1] += (double)c.x * (double)r.x + (double)c.y * (double)r.y;
      }
    }
  }
}

template <int M, int H>
__device__ __forceinline__ void zl_forward_store(float2* v, float2* lb, float2* line, int t, const float2* __restrict__ twm,
                                                 const float2* __restrict__ tw2, GPtrT<H> dst, int64_t base,
                                                 bool mulz) {
  GPtrT<H> db = dst + base;
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT;
  __syncthreads();
  fft2s<R2, R1, NT, -1>(v, line, t, twm);
  __syncthreads();
  // output slot j = b*R1 + k2 of the (R2, R1) plan holds Z[t + b*NT + R2*k2]
#pragma unroll
  for (int j = 0; j < ZK<M>::VPT; j++) line[xpad(out_row<R2, R1, NT>(t, j))] = v[j];
  __syncthreads();
  constexpr int TPB = ZK<M>::TPB, IT = ZK<M>::NL * M / TPB;
  // unrolled by 8 at most: a longer unroll gathers all its LDS loads up front and spills
#pragma unroll 8
  for (int i = 0; i < IT; i++) {
    const int e = (int)threadIdx.x + i * TPB;
    const int li = e / M, k = e - li * M;
    float2 x = zpost<M>(lb + li * ZK<M>::LP, k, tw2);
    if (mulz) x = mul_i(cscale(x, (float)k));  // w_z = k' (0 at the packed column)
    db[e] = x;
  }
}

// Z pass of the matvec: for each component d, h_d = k * C2R(G_d)/N^3, out_d = R2C(h_d), the z
// component multiplied by i w_z; in place on the three slots.  One component per block (3 blocks
// per line group): the three blocks of a line group get ids b, b+8, b+16 so that they run on the
// same XCD (blocks are dispatched to the 8 XCDs round-robin) at about the same time and share the
// k(x) tile through that XCD's L2.
#ifndef GLIA_ZD_MINW
#define GLIA_ZD_MINW 3
#endif
// minimum blocks per CU of z_diff for plans with more than 16 values per thread
#ifndef GLIA_ZD_MINW24
#define GLIA_ZD_MINW24 2
#endif
template <int M>
__global__ void __launch_bounds__(ZK<M>::TPB, ZK<M>::VPT <= 16 ? GLIA_ZD_MINW : GLIA_ZD_MINW24)
k_z_diff(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ twm, const float2* __restrict__ tw2,
          int s_gx, int s_gy, int s_gz, int k_slot) {
  constexpr int H = GLIA_NT_ZD;
  const MemberDev& Mb = mem[blockIdx.y];
  if (Mb.st->done) return;
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT, VPT = ZK<M>::VPT;
  __shared__ float2 lb[ZK<M>::NL * ZK<M>::LP];
  __shared__ float2 twml[M], tw2l[M];
  const int ngroups = (int)gridDim.x / 3;
  int comp, grp;
  if ((ngroups & 7) == 0) {
    const int x = (int)blockIdx.x & 7, q = (int)blockIdx.x >> 3;
    comp = q % 3;
    grp = (q / 3) * 8 + x;
  } else {
    comp = (int)blockIdx.x % 3;
    grp = (int)blockIdx.x / 3;
  }
  const int li = threadIdx.x / NT, t = threadIdx.x % NT;
  const int64_t base = (int64_t)grp * ZK<M>::NL * M;
  float2* line = lb + li * ZK<M>::LP;
  GPtrT<H> kf = rptr2<H>(Mb, k_slot) + base;
  const int lo = li * M;
  const float s = (float)d.inv_ng;
  const int slot = comp == 0 ? s_gx : (comp == 1 ? s_gy : s_gz);
  float2 nxt[ZIt<M>::value];
  zl_fetch<M>(nxt, cptr<H>(Mb, slot), base);
  float2 kv[VPT], v[VPT];
#pragma unroll
  for (int j = 0; j < VPT; j++) {
    const float2 k2 = kf[lo + out_row<R1, R2, NT>(t, j)];
    kv[j] = make_float2(k2.x * s, k2.y * s);
  }
  stage_line_twiddles<M>(twml, tw2l, twm, tw2);
  zl_put<M>(lb, nxt);
  __syncthreads();
  zl_inverse<M>(v, line, t, twml, tw2l);
#pragma unroll
  for (int j = 0; j < VPT; j++) v[j] = make_float2(kv[j].x * v[j].x, kv[j].y * v[j].y);
  zl_forward_store<M>(v, lb, line, t, twml, tw2l, cptr<H>(Mb, slot), base, comp == 2);
}

// Z pass for a real-space pointwise stage between an inverse and a forward transform.
//   ZR_REACT    : c = C2R/N^3; [store c -> chalf]; c = c e^{rho dt} / (c e^{rho dt} + 1 - c)
//                 (the exact logistic step of src/pde/PdeOperators.cpp:140-190)
//   ZR_REACT_ADJ: a = C2R/N^3; a = a e^{rho dt} / (ch e^{rho dt} + 1 - ch)^2 with ch from chalf
//                 (logisticReactionLinearized, src/cuda/PdeOperators.cu:4-13)
//   ZR_C2R      : real out = C2R/N^3 (no forward transform)
//   ZR_R2C      : packed out = R2C(real in)
//   ZR_MISFIT   : c = C2R/N^3 -> CT; t = O c - d1; partial sum t^2; a = -O t; out = R2C(a)
//   ZR_OBS2     : c = C2R/N^3; a = -O (O c); out = R2C(a)  (final condition of the incremental adjoint)
//   ZR_COEF     : c = C2R/N^3; out = R2C(m c) with m the real slot r_slot (a Lame coefficient times a
//                 strain term of the elasticity operator)
//   ZR_SCREEN   : c = C2R/N^3; real slot r_slot = c - s u with s in RS_RHO and u the real slot s_out
//                 (the screened elasticity operator's output; no forward transform); block partials
//                 of <u, out> (the CG's <p, A p>)
//   ZR_C2R_DOT  : real slot r_slot = c = C2R/N^3, with block partials of <c, c> and <c, r>, r the real
//                 slot s_out (the preconditioned residual's norm and <z, r>)
enum { ZR_REACT = 0, ZR_REACT_ADJ = 1, ZR_C2R = 2, ZR_R2C = 3, ZR_MISFIT = 4, ZR_OBS2 = 5, ZR_COEF = 6,
       ZR_SCREEN = 7, ZR_C2R_DOT = 8 };

template <int M, int MODE>
__global__ void __launch_bounds__(ZK<M>::TPB, GLIA_Z_MINW)
k_z_point(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ twm, const float2* __restrict__ tw2,
          int s_in, int s_out, int r_slot, int ch_idx, float dt, int honor_done) {
  constexpr int H = GLIA_NT_ZP;
  const MemberDev& Mb = mem[blockIdx.y];
  if (honor_done && Mb.st->done) return;
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT, VPT = ZK<M>::VPT;
  // pointwise operands are loaded before the transform when the registers fit
  constexpr bool PF = VPT <= 16;
  constexpr bool TWO = (MODE == ZR_REACT_ADJ || MODE == ZR_MISFIT);
  constexpr bool REAL_OUT = (MODE == ZR_C2R || MODE == ZR_SCREEN || MODE == ZR_C2R_DOT);  // no forward transform
  constexpr int NDOT = MODE == ZR_SCREEN ? 1 : (MODE == ZR_C2R_DOT ? 2 : 0);
  __shared__ float2 lb[ZK<M>::NL * ZK<M>::LP];
  __shared__ float2 twml[M], tw2l[M];
  const int li = threadIdx.x / NT, t = threadIdx.x % NT;
  const int64_t base = (int64_t)blockIdx.x * ZK<M>::NL * M;
  const int lo = li * M;
  float2* line = lb + li * ZK<M>::LP;
  // real field: a half-step history entry, a real slot, or none (block-base pointers, uniform)
  GPtrT<H> rf = gptr<H>(ch_idx >= 0 ? (const void*)Mb.chalf[ch_idx] : (r_slot >= 0 ? (const void*)Mb.r[r_slot] : nullptr));
  if (rf) rf = rf + base;
  const float s = (float)d.inv_ng;
  double acc[2] = {0.0, 0.0};
  float2 v[VPT];
  GPtrT<H> op1 = MODE == ZR_COEF ? rf : rptr2<H>(Mb, (MODE == ZR_MISFIT || MODE == ZR_OBS2) ? RS_OBS : RS_RHO) + base;
  GPtrT<H> op2 = MODE == ZR_MISFIT ? rptr2<H>(Mb, RS_D1) + base
                               : ((MODE == ZR_SCREEN || MODE == ZR_C2R_DOT) ? rptr2<H>(Mb, s_out) + base : rf);
  GPtrT<H> ctp = rptr2<H>(Mb, RS_CT) + base;
  float2 o1[PF ? VPT : 1], o2[PF && TWO ? VPT : 1];
  if (MODE == ZR_R2C) {
    zl_load<M>(lb, rf, 0);  // the real lines, coalesced, into the line buffer
  } else {
    zl_load<M>(lb, cptr<H>(Mb, s_in), base);
    if (PF && !REAL_OUT) {
#pragma unroll
      for (int j = 0; j < VPT; j++) {
        const int gi = lo + out_row<R1, R2, NT>(t, j);
        o1[j] = op1[gi];
        if (TWO) o2[j] = op2[gi];
      }
    }
  }
  stage_line_twiddles<M>(twml, tw2l, twm, tw2);
  __syncthreads();
  if (MODE == ZR_R2C) {
#pragma unroll
    for (int j = 0; j < VPT; j++) v[j] = line[xpad(out_row<R1, R2, NT>(t, j))];
  }
  if (REAL_OUT) {
    zl_inverse<M>(v, line, t, twml, tw2l);
    zl_real_store<M, MODE == ZR_SCREEN, NDOT>(v, lb, line, t, s, rf, op1, op2, acc);
    if (NDOT == 1) {
      double a1[1] = {acc[0]};
      block_reduce_store<1>(a1, Mb.part + blockIdx.x);
    } else if (NDOT == 2) {
      block_reduce_store<2>(acc, Mb.part + (int64_t)blockIdx.x * 2);
    }
  } else if (MODE != ZR_R2C) {
    zl_inverse<M>(v, line, t, twml, tw2l);
#pragma unroll
    for (int j = 0; j < VPT; j++) {
      const int gi = lo + out_row<R1, R2, NT>(t, j);
      float2 c = cscale(v[j], s);
      {
        const float2 a1 = PF ? o1[j] : op1[gi];
        if (MODE == ZR_REACT) {
          if (rf) rf[gi] = c;
          const float fx = expf(a1.x * dt), fy = expf(a1.y * dt);
          c.x = c.x * fx / (c.x * fx + 1.0f - c.x);
          c.y = c.y * fy / (c.y * fy + 1.0f - c.y);
          v[j] = c;
        } else if (MODE == ZR_REACT_ADJ) {
          const float2 ch = PF ? o2[j] : op2[gi];
          const float fx = expf(a1.x * dt), fy = expf(a1.y * dt);
          const float ax = ch.x * fx + 1.0f - ch.x, ay = ch.y * fy + 1.0f - ch.y;
          c.x = c.x * fx / (ax * ax);
          c.y = c.y * fy / (ay * ay);
          v[j] = c;
        } else if (MODE == ZR_OBS2) {
          v[j] = make_float2(-(a1.x * (a1.x * c.x)), -(a1.y * (a1.y * c.y)));
        } else if (MODE == ZR_COEF) {
          v[j] = make_float2(a1.x * c.x, a1.y * c.y);
        } else if (MODE == ZR_MISFIT) {
          ctp[gi] = c;
          const float2 dd = PF ? o2[j] : op2[gi];
          const float tx = a1.x * c.x - dd.x, ty = a1.y * c.y - dd.y;
          acc[0] += (double)tx * tx + (double)ty * ty;
          v[j] = make_float2(-(a1.x * tx), -(a1.y * ty));
        }
      }
    }
  }
  if (!REAL_OUT) zl_forward_store<M>(v, lb, line, t, twml, tw2l, cptr<H>(Mb, s_out), base, false);
  if (MODE == ZR_MISFIT) {
    double a1[1] = {acc[0]};
    block_reduce_store<1>(a1, Mb.part + blockIdx.x);
  }
}

// Z pass of a forward transform whose real input is first updated in place: the elasticity CG's
// vector work folded into the transform that reads the vector next.  The CG scalars come from lane
// 0's PcgState (the elasticity CG's state, kept by k_el_cg_begin / k_el_cg_fin) and sc:
//   ZU_AXPBY: p = z + b p, b = beta / betaold (p = z in the first iteration), out = R2C(p)
//             [ra = p, rb = z]
//   ZU_CGUPD: x += t p, r -= t w with t = beta / q, q = sc[0] (<p, A p>); out = R2C(r)
//             [ra = x, rb = r, rc = p, rd = w].  No update when q is 0, not finite, or of the sign
//             opposite to the previous <p, A p> after the first iteration (the CG stops there).
// With honor set, a lane whose CG has finished returns at once.
enum { ZU_AXPBY = 0, ZU_CGUPD = 1 };

template <int M, int MODE>
__global__ void __launch_bounds__(ZK<M>::TPB, GLIA_Z_MINW)
k_z_upd_r2c(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ twm, const float2* __restrict__ tw2,
            int ra, int rb, int rc, int rd, const double* __restrict__ sc, int s_out, int honor) {
  constexpr int H = GLIA_NT_ZP;
  const MemberDev& Mb = mem[blockIdx.y];
  if (honor && Mb.st->done) return;
  const PcgState* S = mem[0].st;
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT, VPT = ZK<M>::VPT, TPB = ZK<M>::TPB;
  constexpr int IT = ZIt<M>::value;
  __shared__ float2 lb[ZK<M>::NL * ZK<M>::LP];
  __shared__ float2 twml[M], tw2l[M];
  const int li = threadIdx.x / NT, t = threadIdx.x % NT;
  const int64_t base = (int64_t)blockIdx.x * ZK<M>::NL * M;
  float2* line = lb + li * ZK<M>::LP;
  float cf = 0.0f;
  bool upd = true;
  double beta = 0.0;
  if (MODE == ZU_AXPBY) {
    beta = S->i ? S->beta / S->betaold : 0.0;
    cf = (float)beta;
  } else {
    const double q = sc[0], dpiold = S->dpi;
    upd = !(isnan(q) || isinf(q) || q == 0.0 || (S->i != 0 && q * dpiold <= 0.0));
    if (upd) cf = (float)(S->beta / q);
  }
  GPtrT<H> A = rptr2<H>(Mb, ra) + base, B = rptr2<H>(Mb, rb) + base;
  float2 tmp[IT];
  if (MODE == ZU_AXPBY) {
    float2 za[IT], pa[IT];
#pragma unroll
    for (int i = 0; i < IT; i++) {
      za[i] = B[(int)threadIdx.x + i * TPB];
      if (beta != 0.0) pa[i] = A[(int)threadIdx.x + i * TPB];
    }
#pragma unroll
    for (int i = 0; i < IT; i++) {
      float2 v = za[i];
      if (beta != 0.0) v = make_float2(za[i].x + cf * pa[i].x, za[i].y + cf * pa[i].y);
      A[(int)threadIdx.x + i * TPB] = v;
      tmp[i] = v;
    }
  } else {
    GPtrT<H> C = rptr2<H>(Mb, rc) + base, D = rptr2<H>(Mb, rd) + base;
    float2 xa[IT], pa[IT], wa[IT];
#pragma unroll
    for (int i = 0; i < IT; i++) {
      const int e = (int)threadIdx.x + i * TPB;
      tmp[i] = B[e];
      if (upd) {
        xa[i] = A[e];
        pa[i] = C[e];
        wa[i] = D[e];
      }
    }
    if (upd) {
#pragma unroll
      for (int i = 0; i < IT; i++) {
        const int e = (int)threadIdx.x + i * TPB;
        A[e] = make_float2(xa[i].x + cf * pa[i].x, xa[i].y + cf * pa[i].y);
        tmp[i] = make_float2(tmp[i].x - cf * wa[i].x, tmp[i].y - cf * wa[i].y);
        B[e] = tmp[i];
      }
    }
  }
  stage_line_twiddles<M>(twml, tw2l, twm, tw2);
  zl_put<M>(lb, tmp);
  __syncthreads();
  float2 v[VPT];
#pragma unroll
  for (int j = 0; j < VPT; j++) v[j] = line[xpad(out_row<R1, R2, NT>(t, j))];
  zl_forward_store<M>(v, lb, line, t, twml, tw2l, cptr<H>(Mb, s_out), base, false);
}

// Z pass of the gradient quadrature at one time point: from the y-transformed fields of c
// (G0, Gx, Gy, Gz in slots s_c) and of alpha (s_a):
//   q_k   = sum (wm + r_k gm) (grad c . grad alpha)
//   q_rho = sum (wm + r_rho gm) (c^2 - c) alpha
// (the integrands of gradDiffusion / gradReaction, src/grad/DerivativeOperators.cpp:189-321)
template <int M>
__global__ void __launch_bounds__(ZK<M>::TPB, GLIA_Z_MINW)
k_z_gradq(const MemberDev* __restrict__ mem, Dims d, const float2* __restrict__ twm, const float2* __restrict__ tw2,
          int4 s_c, int4 s_a) {
  const MemberDev& Mb = mem[blockIdx.y];
  constexpr int R1 = ZK<M>::R1, R2 = ZK<M>::R2, NT = ZK<M>::NT, VPT = ZK<M>::VPT;
  __shared__ float2 lb[ZK<M>::NL * ZK<M>::LP];
  const int li = threadIdx.x / NT, t = threadIdx.x % NT;
  const int64_t base = (int64_t)blockIdx.x * ZK<M>::NL * M;
  const int lo = li * M;
  float2* line = lb + li * ZK<M>::LP;
  const float s = (float)d.inv_ng;
  GPtr wmp = rptr2(Mb, RS_WM) + base;
  GPtr gmp = rptr2(Mb, RS_GM) + base;
  float2 dot[VPT], cv[VPT], v[VPT];
#pragma unroll
  for (int j = 0; j < VPT; j++) dot[j] = make_float2(0.f, 0.f);
  const int sc[4] = {s_c.x, s_c.y, s_c.z, s_c.w};
  const int sa[4] = {s_a.x, s_a.y, s_a.z, s_a.w};
  double acc[2] = {0.0, 0.0};
  __shared__ float2 twml[M], tw2l[M];
  // the eight line loads (c and alpha for each of the 4 fields) are pipelined when registers allow
  constexpr bool PIPE = VPT <= 8;
  float2 nxt[ZIt<M>::value];
  zl_fetch<M>(nxt, cptr(Mb, sc[0]), base);
  stage_line_twiddles<M>(twml, tw2l, twm, tw2);
  for (int comp = 0; comp < 4; comp++) {
    for (int which = 0; which < 2; which++) {
      const int step = comp * 2 + which;
      if (!PIPE && step > 0) zl_fetch<M>(nxt, cptr(Mb, which == 0 ? sc[comp] : sa[comp]), base);
      zl_put<M>(lb, nxt);
      __syncthreads();
      if (PIPE && step < 7) {
        const int c2 = (step + 1) >> 1, w2 = (step + 1) & 1;
        zl_fetch<M>(nxt, cptr(Mb, w2 == 0 ? sc[c2] : sa[c2]), base);
      }
      zl_inverse<M>(v, line, t, twml, tw2l);
#pragma unroll
      for (int j = 0; j < VPT; j++) {
        const float2 x = cscale(v[j], s);
        if (which == 0) {
          cv[j] = x;
        } else if (comp > 0) {
          dot[j].x += cv[j].x * x.x;
          dot[j].y += cv[j].y * x.y;
        } else {
          const int gi = lo + out_row<R1, R2, NT>(t, j);
          const float2 wm = wmp[gi];
          const float2 gm = gmp[gi];
          const float rx = cv[j].x * cv[j].x - cv[j].x, ry = cv[j].y * cv[j].y - cv[j].y;
          acc[1] += (double)(wm.x + Mb.rr_gm * gm.x) * (double)(rx * x.x) +
                    (double)(wm.y + Mb.rr_gm * gm.y) * (double)(ry * x.y);
        }
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int j = 0; j < VPT; j++) {
    const int gi = lo + out_row<R1, R2, NT>(t, j);
    const float2 wm = wmp[gi];
    const float2 gm = gmp[gi];
    acc[0] += (double)(wm.x + Mb.kr_gm * gm.x) * (double)dot[j].x + (double)(wm.y + Mb.kr_gm * gm.y) * (double)dot[j].y;
  }
  block_reduce_store<2>(acc, Mb.part + (int64_t)blockIdx.x * 2);
}

// ---------------------------------------------------------------------------------------------
// PCG vector update and scalar state machine (PETSc 3.11 KSPSolve_CG + KSPConvergedDefault)
// ---------------------------------------------------------------------------------------------
constexpr int UPD_BLOCKS = 1024;

// x += alpha p; r -= alpha w; partials ||B r||^2 and <B r, r> (Parseval weights)
__global__ void __launch_bounds__(256)
k_pcg_update(const MemberDev* __restrict__ mem, Dims d, int s_x) {
  const MemberDev& M = mem[blockIdx.y];
  const PcgState* st = M.st;
  if (st->done) return;
  // a lane handles a pair of adjacent entries along z (m2 is even, so both share kx and ky) with
  // 16-byte accesses, and two pairs per loop trip so that eight loads are in flight per lane
  constexpr int H = GLIA_NT_PU;
  GLIA_GLOBAL v4f* X = (GLIA_GLOBAL v4f*)craw(M, s_x);
  GLIA_GLOBAL v4f* R = (GLIA_GLOBAL v4f*)M.c[CS_R];
  GLIA_GLOBAL v4f* P = (GLIA_GLOBAL v4f*)M.c[CS_P];
  GLIA_GLOBAL v4f* W = (GLIA_GLOBAL v4f*)M.c[CS_W];
  const float a = (float)st->alpha;
  const float pc = (float)st->pc;
  const unsigned npairs = (unsigned)(d.nspec >> 1), hm = (unsigned)(d.m2 >> 1), n1 = (unsigned)d.n1;
  double acc[2] = {0.0, 0.0};
  auto finish = [&](unsigned i, v4f x, v4f r, v4f p, v4f w) {
    x += a * p;
    r -= a * w;
    st4<H>(X + i, x);
    st4<H>(R + i, r);
    const unsigned line = i / hm, kz = 2u * (i - line * hm);
    const unsigned kx = line / n1, ky = line - kx * n1;
    const float wx = wnum((int)kx, d.n0), wy = wnum((int)ky, d.n1);
    const float f0 = precf(pc, wx, wy, (float)kz), f1 = precf(pc, wx, wy, (float)(kz + 1));
    const double w0 = (kz == 0) ? 1.0 : 2.0;
    const double r0 = (double)r.x * r.x + (double)r.y * r.y, r1 = (double)r.z * r.z + (double)r.w * r.w;
    acc[0] += w0 * (double)f0 * f0 * r0 + 2.0 * (double)f1 * f1 * r1;
    acc[1] += w0 * (double)f0 * r0 + 2.0 * (double)f1 * r1;
  };
  const unsigned stride = gridDim.x * blockDim.x;
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + stride < npairs; i += 2 * stride) {
    const unsigned j = i + stride;
    const v4f p0 = ld4<H>(P + i), w0 = ld4<H>(W + i), x0 = ld4<H>(X + i), r0 = ld4<H>(R + i);
    const v4f p1 = ld4<H>(P + j), w1 = ld4<H>(W + j), x1 = ld4<H>(X + j), r1 = ld4<H>(R + j);
    finish(i, x0, r0, p0, w0);
    finish(j, x1, r1, p1, w1);
  }
  if (i < npairs) finish(i, ld4<H>(X + i), ld4<H>(R + i), ld4<H>(P + i), ld4<H>(W + i));
  block_reduce_store<2>(acc, M.part + (int64_t)blockIdx.x * 2);
}

__device__ __forceinline__ int ksp_conv_test(const PcgState* s, double rnorm) {
  if (isnan(rnorm) || isinf(rnorm)) return KSPR_DIVERGED_NANORINF;
  if (rnorm <= s->ttol) return (rnorm < s->atol) ? KSPR_CONVERGED_ATOL : KSPR_CONVERGED_RTOL;
  if (rnorm >= s->dtol * s->rnorm0) return KSPR_DIVERGED_DTOL;
  return 0;
}

__device__ __forceinline__ void ksp_finish(PcgState* s, int reason, int its) {
  s->reason = reason;
  s->its = its;
  s->done = 1;
  s->its_total += its;
  s->solves += 1;
  if (reason < s->worst_reason) s->worst_reason = reason;
}

// start of every CN solve: parameters and a clean iteration state
__global__ void k_pcg_begin(const MemberDev* __restrict__ mem, double tau) {
  if (threadIdx.x != 0) return;
  PcgState* s = mem[blockIdx.x].st;
  s->tau = tau;
  s->beta = s->betaold = s->dpi = s->dpiold = s->alpha = s->bcoef = 0;
  s->dp = s->rnorm0 = s->ttol = 0;
  s->i = 0;
  s->its = 0;
  s->reason = 0;
  s->done = 0;
  if (!s->active) {  // DiffusionSolver::solve returns at once when k_scale == 0
    s->done = 1;
    s->solves += 1;
  }
}

// after the setup matvec: the n = 0 convergence test and the first beta
__global__ void k_pcg_fin_setup(const MemberDev* __restrict__ mem, Dims d, int nb, int* hflag, int hslot, int nm) {
  const MemberDev& M = mem[blockIdx.x];
  PcgState* s = M.st;
  const bool was_done = s->done;
  const double bb = sum_partials(M.part, nb, 3, 0) * d.inv_ng;
  const double rr = sum_partials(M.part, nb, 3, 1) * d.inv_ng;
  const double zr = sum_partials(M.part, nb, 3, 2) * d.inv_ng;
  if (threadIdx.x == 0) {
    if (!was_done) {
      s->dp = sqrt(rr);
      double snorm = sqrt(bb);
      if (snorm == 0) snorm = s->dp;
      s->rnorm0 = snorm;
      s->ttol = fmax(s->rtol * snorm, s->atol);
      const int reason = ksp_conv_test(s, s->dp);
      if (reason) {
        ksp_finish(s, reason, 0);
      } else {
        s->beta = zr;
        s->i = 0;
        if (isnan(zr) || isinf(zr)) {
          ksp_finish(s, KSPR_DIVERGED_NANORINF, 1);
        } else if (zr == 0.0) {
          ksp_finish(s, KSPR_CONVERGED_ATOL, 1);
        }
      }
    }
    if (hflag) {  // 0 while the solve runs, 1 + its CG iterations once it is done
      hflag[hslot * nm + blockIdx.x] = s->done ? s->its + 1 : 0;
      __threadfence_system();
    }
  }
}

// after w = A p: dpi = <p, w>, alpha = beta / dpi
__global__ void k_pcg_fin_dpi(const MemberDev* __restrict__ mem, Dims d, int nb) {
  const MemberDev& M = mem[blockIdx.x];
  PcgState* s = M.st;
  if (s->done) return;
  const double dpi = sum_partials(M.part, nb, 1, 0) * d.inv_ng;
  if (threadIdx.x == 0) {
    s->dpiold = s->dpi;
    s->dpi = dpi;
    s->betaold = s->beta;
    if (isnan(dpi) || isinf(dpi)) {
      ksp_finish(s, KSPR_DIVERGED_NANORINF, s->i + 1);
    } else if (dpi == 0.0 || (s->i > 0 && dpi * s->dpiold <= 0.0)) {
      ksp_finish(s, KSPR_DIVERGED_INDEFINITE_MAT, s->i + 1);
    } else {
      s->alpha = s->beta / dpi;
    }
  }
}

// after the x/r update: convergence test, next beta and the loop-head checks of KSPSolve_CG
__global__ void k_pcg_fin_update(const MemberDev* __restrict__ mem, Dims d, int nb, int* hflag, int hslot, int nm) {
  const MemberDev& M = mem[blockIdx.x];
  PcgState* s = M.st;
  const bool was_done = s->done;
  double rr = 0, zr = 0;
  if (!was_done) {
    rr = sum_partials(M.part, nb, 2, 0) * d.inv_ng;
    zr = sum_partials(M.part, nb, 2, 1) * d.inv_ng;
  }
  if (threadIdx.x == 0) {
    if (!was_done) {
      s->dp = sqrt(rr);
      const int reason = ksp_conv_test(s, s->dp);
      if (reason) {
        ksp_finish(s, reason, s->i + 1);
      } else if (isnan(zr) || isinf(zr)) {
        ksp_finish(s, KSPR_DIVERGED_NANORINF, s->i + 1);
      } else {
        s->beta = zr;
        s->i += 1;
        if (s->i >= s->maxit) {
          ksp_finish(s, KSPR_DIVERGED_ITS, s->i);
        } else if (s->beta == 0.0) {
          ksp_finish(s, KSPR_CONVERGED_ATOL, s->i + 1);
        } else if (s->beta * s->betaold < 0.0) {
          ksp_finish(s, KSPR_DIVERGED_INDEFINITE_PC, s->i + 1);
        } else {
          s->bcoef = s->beta / s->betaold;
        }
      }
    }
    if (hflag) {  // 0 while the solve runs, 1 + its CG iterations once it is done
      hflag[hslot * nm + blockIdx.x] = s->done ? s->its + 1 : 0;
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------------------------
// elementwise / setup kernels
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_copy(const MemberDev* __restrict__ mem, Dims d, int s_in, int s_out) {
  const MemberDev& M = mem[blockIdx.y];
  const float4* I = reinterpret_cast<const float4*>(craw(M, s_in));
  float4* O = reinterpret_cast<float4*>(craw(M, s_out));
  const int64_t n4 = d.nspec / 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) O[i] = I[i];
}

// MatProp::setValuesCustom (src/mat/MatProp.cpp:134-203): clip, mask rule; partial sum(filter)
__global__ void __launch_bounds__(256) k_tissue(const MemberDev* __restrict__ mem, Dims d) {
  const MemberDev& M = mem[blockIdx.y];
  float* wm = M.r[RS_WM];
  float* gm = M.r[RS_GM];
  float* csf = M.r[RS_CSF];
  float* vt = M.r[RS_VT];
  float* fl = M.r[RS_FILTER];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) {
    float w = wm[i], g = gm[i], c = csf[i], v = vt[i];
    w = (w <= 0.f) ? 0.f : w;
    g = (g <= 0.f) ? 0.f : g;
    c = (c <= 0.f) ? 0.f : c;
    v = (v <= 0.f) ? 0.f : v;
    wm[i] = w;
    gm[i] = g;
    csf[i] = c;
    vt[i] = v;
    const float f = ((w > 0.1f || g > 0.1f) && v < 0.8f) ? 1.f : 0.f;
    fl[i] = f;
    acc[0] += f;
  }
  block_reduce_store<1>(acc, M.part + blockIdx.x);
}

// DiffCoef::setValues / ReacCoef::setValues: field = s_gm gm + s_wm wm + s_csf csf;
// partial sum of the field (for the mask average kbar)
__global__ void __launch_bounds__(256) k_coef(const MemberDev* __restrict__ mem, Dims d, int r_out, const double* coefs) {
  const MemberDev& M = mem[blockIdx.y];
  const float* wm = M.r[RS_WM];
  const float* gm = M.r[RS_GM];
  const float* csf = M.r[RS_CSF];
  float* o = M.r[r_out];
  const float a = (float)coefs[3 * blockIdx.y + 0], b = (float)coefs[3 * blockIdx.y + 1], c = (float)coefs[3 * blockIdx.y + 2];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) {
    float v = 0.f;
    v += a * gm[i];
    v += b * wm[i];
    v += c * csf[i];
    o[i] = v;
    acc[0] += v;
  }
  block_reduce_store<1>(acc, M.part + blockIdx.x);
}

// Obs::setDefaultFilter (src/mat/Obs.cpp:26-55): O = 1[d1 > thr]
__global__ void __launch_bounds__(256) k_obs_default(const MemberDev* __restrict__ mem, Dims d, float thr) {
  const MemberDev& M = mem[blockIdx.y];
  const float* d1 = M.r[RS_D1];
  float* ob = M.r[RS_OBS];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) ob[i] = (d1[i] > thr) ? 1.f : 0.f;
}

// finalize kinds that fold a reduction into the state
enum { FIN_FILTER = 0, FIN_KSUM = 1, FIN_MISFIT = 2, FIN_C0SQ = 3, FIN_QUAD = 4 };
__global__ void k_fin_state(const MemberDev* __restrict__ mem, int nb, int kind, double w) {
  const MemberDev& M = mem[blockIdx.x];
  PcgState* s = M.st;
  const int nq = (kind == FIN_QUAD) ? 2 : 1;
  const double v0 = sum_partials(M.part, nb, nq, 0);
  const double v1 = (nq > 1) ? sum_partials(M.part, nb, nq, 1) : 0.0;
  if (threadIdx.x == 0) {
    if (kind == FIN_FILTER) s->filter_sum = v0;
    else if (kind == FIN_KSUM) { s->ksum = v0; s->kavg = v0 * (1.0 / s->filter_sum); }
    else if (kind == FIN_MISFIT) s->m1 = v0;
    else if (kind == FIN_C0SQ) s->c0sq = v0;
    else { s->gk += w * v0; s->gr += w * v1; }
  }
}

// DiffusionSolver::precFactor: the preconditioner is fixed by (prec_dt, kbar) at call time
__global__ void k_prec_factor(const MemberDev* __restrict__ mem, double prec_dt) {
  if (threadIdx.x != 0) return;
  PcgState* s = mem[blockIdx.x].st;
  s->pc = 0.25 * prec_dt * s->kavg;
}

__global__ void k_state_reset(const MemberDev* __restrict__ mem, int what) {
  if (threadIdx.x != 0) return;
  PcgState* s = mem[blockIdx.x].st;
  if (what & 1) { s->its_total = 0; s->solves = 0; s->worst_reason = 0; }
  if (what & 2) { s->gk = 0; s->gr = 0; s->m1 = 0; }
  if (what & 4) {  // snapshot at the start of a captured step
    s->gk_snap = s->gk;
    s->gr_snap = s->gr;
    s->its_snap = s->its_total;
    s->solves_snap = s->solves;
    s->worst_snap = s->worst_reason;
  }
  if (what & 8) {  // back to the snapshot before the step runs again
    s->gk = s->gk_snap;
    s->gr = s->gr_snap;
    s->its_total = s->its_snap;
    s->solves = s->solves_snap;
    s->worst_reason = s->worst_snap;
  }
}

// ---------------------------------------------------------------------------------------------
// Gaussian basis Phi as compact bricks
// ---------------------------------------------------------------------------------------------
// source of the smoothing convolution on the extended brick: mask * exp(-r^2 / (2 sigma^2)) with
// the non-periodic distance of Phi::initialize (src/mat/Phi.cpp:264-311); indices wrap (mod n)
// for the periodic convolution.


__global__ void k_brick_src(const BrickJob* __restrict__ jobs, Dims d, const MemberDev* __restrict__ mem, double h0,
                            double h1, double h2, double sigma, double* __restrict__ ext) {
  // no FMA contraction: the Gaussian argument is computed with the same IEEE operations as the
  // FP64 restatement of Phi::initialize
#pragma clang fp contract(off)
  const BrickJob J = jobs[blockIdx.y];
  const int e0 = J.b[0] + 2 * J.kr[0], e1 = J.b[1] + 2 * J.kr[1], e2 = J.b[2] + 2 * J.kr[2];
  const int64_t ne = (int64_t)e0 * e1 * e2;
  const float* fl = mem[J.member].r[RS_FILTER];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % e2);
    const int y = (int)((i / e2) % e1);
    const int x = (int)(i / ((int64_t)e1 * e2));
    int X = J.o[0] - J.kr[0] + x, Y = J.o[1] - J.kr[1] + y, Z = J.o[2] - J.kr[2] + z;
    X = ((X % d.n0) + d.n0) % d.n0;
    Y = ((Y % d.n1) + d.n1) % d.n1;
    Z = ((Z % d.n2) + d.n2) % d.n2;
    const double dx = h0 * X - J.ctr[0], dy = h1 * Y - J.ctr[1], dz = h2 * Z - J.ctr[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double R = sqrt(2.0) * sigma;
    const double m = fl[((int64_t)X * d.n1 + Y) * d.n2 + Z];
    ext[J.ext_off + i] = exp(-r2 / (R * R)) * m;
  }
}

// One axis of the separable periodic smoothing of a real field (the reference's
// weierstrassSmoother, src/grad/SpectralOperators.cpp:264-381, whose 8-image Gaussian kernel of
// src/cuda/SpectralOperators.cu:12-51 is a product of per-axis kernels): out(x) = sum_s taps[s]
// in(x - s e_axis), indices periodic.  Taps beyond 12 sigma (< e^-72) are dropped.
__global__ void __launch_bounds__(256) k_smooth_axis(const float* __restrict__ in, float* __restrict__ out, Dims d,
                                                     int axis, const double* __restrict__ taps, int kr) {
  // block: 64 points along z times 4 rows along y of one x-plane (blockIdx.z); 32-bit offsets
  const int z = (int)(blockIdx.x * 64 + threadIdx.x), y = (int)(blockIdx.y * 4 + threadIdx.y), x = (int)blockIdx.z;
  if (z >= d.n2 || y >= d.n1) return;
  const int n = axis == 0 ? d.n0 : (axis == 1 ? d.n1 : d.n2);
  const unsigned st = axis == 0 ? (unsigned)(d.n1 * d.n2) : (axis == 1 ? (unsigned)d.n2 : 1u);
  const int ca = axis == 0 ? x : (axis == 1 ? y : z);
  const unsigned i = ((unsigned)x * (unsigned)d.n1 + (unsigned)y) * (unsigned)d.n2 + (unsigned)z;
  const unsigned base = i - (unsigned)ca * st;
  double acc = 0.0;
  for (int s = -kr; s <= kr; s++) {
    int q = ca - s;
    q += (q < 0) ? n : 0;
    q -= (q >= n) ? n : 0;
    acc += taps[s + kr] * (double)in[base + (unsigned)q * st];
  }
  out[i] = (float)acc;
}

// data-driven basis support (Phi::setGaussians(Vec data), src/mat/Phi.cpp:457-522): one thread per
// lattice candidate (step apart on every axis, l0 x l1 x l2 of them in index order) counts the
// voxels with f > thr at integer offsets |d|^2 <= d2max, |d_a| <= R, clipped to the grid (no wrap)
__global__ void __launch_bounds__(256) k_support_count(const float* __restrict__ f, Dims d, int step, int l0, int l1,
                                                       int l2, int R, int d2max, float thr, int* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)l0 * l1 * l2) return;
  const int c2 = (int)(t % l2);
  const int64_t t1 = t / l2;
  const int c1 = (int)(t1 % l1), c0 = (int)(t1 / l1);
  const int x = c0 * step, y = c1 * step, z = c2 * step;
  int num = 0;
  for (int i = -R; i <= R; i++) {
    const int xi = x + i;
    if (xi < 0 || xi >= d.n0) continue;
    for (int j = -R; j <= R; j++) {
      const int yj = y + j, rem = d2max - i * i - j * j;
      if (yj < 0 || yj >= d.n1 || rem < 0) continue;
      const float* row = f + ((int64_t)xi * d.n1 + yj) * d.n2;
      for (int k = -R; k <= R; k++) {
        const int zk = z + k;
        if (k * k <= rem && zk >= 0 && zk < d.n2) num += row[zk] > thr ? 1 : 0;
      }
    }
  }
  cnt[t] = num;
}

// the same with the tap radius known at compile time: every load of a point is issued before the
// sum, which runs in the same order as above (the same results bit for bit)
template <int KR>
__global__ void __launch_bounds__(256) k_smooth_axis_r(const float* __restrict__ in, float* __restrict__ out, Dims d,
                                                       int axis, const double* __restrict__ taps) {
  const int z = (int)(blockIdx.x * 64 + threadIdx.x), y = (int)(blockIdx.y * 4 + threadIdx.y), x = (int)blockIdx.z;
  if (z >= d.n2 || y >= d.n1) return;
  const int n = axis == 0 ? d.n0 : (axis == 1 ? d.n1 : d.n2);
  const unsigned st = axis == 0 ? (unsigned)(d.n1 * d.n2) : (axis == 1 ? (unsigned)d.n2 : 1u);
  const int ca = axis == 0 ? x : (axis == 1 ? y : z);
  const unsigned i = ((unsigned)x * (unsigned)d.n1 + (unsigned)y) * (unsigned)d.n2 + (unsigned)z;
  const unsigned base = i - (unsigned)ca * st;
  float v[2 * KR + 1];
#pragma unroll
  for (int s = -KR; s <= KR; s++) {
    int q = ca - s;
    q += (q < 0) ? n : 0;
    q -= (q >= n) ? n : 0;
    v[s + KR] = in[base + (unsigned)q * st];
  }
  double acc = 0.0;
#pragma unroll
  for (int s = -KR; s <= KR; s++) acc += taps[s + KR] * (double)v[s + KR];
  out[i] = (float)acc;
}

// one axis of the separable smoothing: the kept extent shrinks by 2*kr[axis] along 'axis'; the
// axis's taps sit at tp[s + kr[axis]], |s| <= kr[axis]
__global__ void k_brick_conv(const BrickJob* __restrict__ jobs, int axis, const double* __restrict__ taps, int tap_stride,
                             const double* __restrict__ src, double* __restrict__ dst) {
  const BrickJob J = jobs[blockIdx.y];
  const int kr = J.kr[axis];
  int ein[3], eout[3];
  for (int a = 0; a < 3; a++) {
    // axes < axis already reduced to b, axes >= axis still extended
    ein[a] = (a < axis) ? J.b[a] : J.b[a] + 2 * J.kr[a];
    eout[a] = (a <= axis) ? J.b[a] : J.b[a] + 2 * J.kr[a];
  }
  const int64_t no = (int64_t)eout[0] * eout[1] * eout[2];
  const double* tp = taps + (int64_t)blockIdx.y * tap_stride * 3 + axis * tap_stride;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < no; i += (int64_t)gridDim.x * blockDim.x) {
    int c[3];
    c[2] = (int)(i % eout[2]);
    c[1] = (int)((i / eout[2]) % eout[1]);
    c[0] = (int)(i / ((int64_t)eout[1] * eout[2]));
    double acc = 0;
    for (int s = -kr; s <= kr; s++) {
      int ci[3] = {c[0], c[1], c[2]};
      ci[axis] = c[axis] + kr - s;  // input index of (out - s) in the extended frame
      acc += tp[s + kr] * src[J.ext_off + ((int64_t)ci[0] * ein[1] + ci[1]) * ein[2] + ci[2]];
    }
    dst[J.ext_off + i] = acc;
  }
}

// truncation at 5 sigma (Phi::truncate, src/mat/Phi.cpp:238-260) and FP32 store
__global__ void k_brick_store(const BrickJob* __restrict__ jobs, Dims d, double h0, double h1, double h2, double sigma,
                              const double* __restrict__ src, float* __restrict__ store) {
  // Centers on grid points put many voxels at exactly r = 5 sigma; computing r without FMA
  // contraction makes the r/sigma <= 5 decision identical to the FP64 restatement.
#pragma clang fp contract(off)
  const BrickJob J = jobs[blockIdx.y];
  const int64_t nb = (int64_t)J.b[0] * J.b[1] * J.b[2];
  float* out = store + J.out_off;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    const int z = (int)(i % J.b[2]);
    const int y = (int)((i / J.b[2]) % J.b[1]);
    const int x = (int)(i / ((int64_t)J.b[1] * J.b[2]));
    const int X = J.o[0] + x, Y = J.o[1] + y, Z = J.o[2] + z;
    const double dx = h0 * X - J.ctr[0], dy = h1 * Y - J.ctr[1], dz = h2 * Z - J.ctr[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    out[i] = (r / sigma <= 5) ? (float)src[J.ext_off + i] : 0.0f;
  }
}

// per-brick maximum (one block per basis) into out[job]
__global__ void k_brick_max(const BrickJob* __restrict__ jobs, const float* __restrict__ store, double* __restrict__ out) {
  const BrickJob J = jobs[blockIdx.x];
  const int64_t nb = (int64_t)J.b[0] * J.b[1] * J.b[2];
  const float* v = store + J.out_off;
  float m = -INFINITY;
  for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) m = fmaxf(m, v[i]);
  __shared__ float red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// c0 = sum_i scale_i * brick_i (gathered per voxel: deterministic, no atomics); partial ||c0||^2
__global__ void __launch_bounds__(256) k_phi_apply(const MemberDev* __restrict__ mem, Dims d) {
  const MemberDev& M = mem[blockIdx.y];
  float* c0 = M.r[RS_C0];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) {
    const int z = (int)(i % d.n2);
    const int y = (int)((i / d.n2) % d.n1);
    const int x = (int)(i / ((int64_t)d.n1 * d.n2));
    double v = 0.0;
    for (int b = 0; b < M.np; b++) {
      const double sc = M.phi_scale[b];
      if (sc == 0.0) continue;
      const Brick bk = M.bricks[b];
      const int lx = x - bk.o[0], ly = y - bk.o[1], lz = z - bk.o[2];
      if (lx >= 0 && lx < bk.b[0] && ly >= 0 && ly < bk.b[1] && lz >= 0 && lz < bk.b[2])
        v += sc * (double)M.brick_data[bk.off + ((int64_t)lx * bk.b[1] + ly) * bk.b[2] + lz];
    }
    const float vf = (float)v;
    c0[i] = vf;
    acc[0] += (double)vf * vf;
  }
  block_reduce_store<1>(acc, M.part + blockIdx.x);
}

// Phi^T v: out_i = scale * <brick_i, v> (one block per basis, FP64)
__global__ void __launch_bounds__(256) k_phi_applyT(const MemberDev* __restrict__ mem, Dims d, int r_in, double scale) {
  const MemberDev& M = mem[blockIdx.y];
  const int b = blockIdx.x;
  if (b >= M.np) return;
  const Brick bk = M.bricks[b];
  const float* v = M.r[r_in];
  const int64_t nb = (int64_t)bk.b[0] * bk.b[1] * bk.b[2];
  double acc = 0;
  for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) {
    const int lz = (int)(i % bk.b[2]);
    const int ly = (int)((i / bk.b[2]) % bk.b[1]);
    const int lx = (int)(i / ((int64_t)bk.b[1] * bk.b[2]));
    const int64_t g = ((int64_t)(bk.o[0] + lx) * d.n1 + (bk.o[1] + ly)) * d.n2 + (bk.o[2] + lz);
    acc += (double)M.brick_data[bk.off + i] * (double)v[g];
  }
  const double s = block_sum_all(acc);
  if (threadIdx.x == 0) M.phit_out[b] = s * scale;
}

// partial <x, y> of two real fields
__global__ void __launch_bounds__(256) k_dot_real(const MemberDev* __restrict__ mem, Dims d, int r_x, int r_y) {
  const MemberDev& M = mem[blockIdx.y];
  const float* x = M.r[r_x];
  const float* y = M.r[r_y];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) acc[0] += (double)x[i] * y[i];
  block_reduce_store<1>(acc, M.part + blockIdx.x);
}

__global__ void k_set_done(const MemberDev* __restrict__ mem, int v) {
  if (threadIdx.x == 0) mem[blockIdx.x].st->done = v;
}

// real-space helper: out = a*x + b*y (used for alpha0 - beta c0)
__global__ void __launch_bounds__(256) k_axpby_real(const MemberDev* __restrict__ mem, Dims d, int r_x, double a, int r_y,
                                                     double b, int r_out) {
  const MemberDev& M = mem[blockIdx.y];
  const float* x = M.r[r_x];
  const float* y = M.r[r_y];
  float* o = M.r[r_out];
  const float fa = (float)a, fb = (float)b;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.ng; i += stride) o[i] = fa * x[i] + fb * y[i];
}

// out = a x + b y on packed spectra (all members; slots >= CS_NUM address the state pool)
__global__ void __launch_bounds__(256) k_axpby_c(const MemberDev* __restrict__ mem, Dims d, int s_x, double a, int s_y,
                                                 double b, int s_out) {
  const MemberDev& M = mem[blockIdx.y];
  const float2* x = craw(M, s_x);
  const float2* y = craw(M, s_y);
  float2* o = craw(M, s_out);
  const float fa = (float)a, fb = (float)b;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nspec; i += stride) {
    const float2 u = x[i], v = y[i];
    o[i] = make_float2(fa * u.x + fb * v.x, fa * u.y + fb * v.y);
  }
}

// out = i (w_x X + w_y Y + w_z Z) on packed spectra: the spectrum of div (x, y, z) up to the 1/N^3
// of the inverse (computeDivergence, src/grad/SpectralOperators.cpp:179-261).  w = k below n/2, k - n
// above, 0 at n/2; the packed column kz = 0 holds the kz = 0 and Nyquist planes, both with w_z = 0.
__global__ void __launch_bounds__(256) k_spec_div(const MemberDev* __restrict__ mem, Dims d, int s_x, int s_y, int s_z,
                                                  int s_out) {
  const MemberDev& M = mem[blockIdx.y];
  const float2* x = craw(M, s_x);
  const float2* y = craw(M, s_y);
  const float2* z = craw(M, s_z);
  float2* o = craw(M, s_out);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nspec; i += stride) {
    const int kz = (int)(i % d.m2);
    const int64_t r = i / d.m2;
    const int ky = (int)(r % d.n1), kx = (int)(r / d.n1);
    const float wx = (2 * kx == d.n0) ? 0.0f : (float)(2 * kx < d.n0 ? kx : kx - d.n0);
    const float wy = (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1);
    const float wz = (float)kz;
    const float2 a = x[i], b = y[i], c = z[i];
    const float re = wx * a.x + wy * b.x + wz * c.x, im = wx * a.y + wy * b.y + wz * c.y;
    o[i] = make_float2(-im, re);
  }
}

// out = i sum_k w_{d_k} X_k over up to 4 packed spectra (gradient, divergence and strain terms of the
// elasticity operator, src/pde/ElasticitySolver.cpp:268-374); wave numbers as in k_spec_div
struct SpecTerms {
  int slot[4], dir[4], n;
};

__global__ void __launch_bounds__(256) k_spec_lin(const MemberDev* __restrict__ mem, Dims d, SpecTerms T, int s_out) {
  const MemberDev& M = mem[blockIdx.y];
  float2* o = craw(M, s_out);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nspec; i += stride) {
    const int kz = (int)(i % d.m2);
    const int64_t r = i / d.m2;
    const int ky = (int)(r % d.n1), kx = (int)(r / d.n1);
    const float w[3] = {(2 * kx == d.n0) ? 0.0f : (float)(2 * kx < d.n0 ? kx : kx - d.n0),
                        (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1), (float)kz};
    float re = 0.0f, im = 0.0f;
    for (int t = 0; t < T.n; t++) {
      const float2 a = craw(M, T.slot[t])[i];
      const float wt = w[T.dir[t]];
      re += wt * a.x;
      im += wt * a.y;
    }
    o[i] = make_float2(-im, re);
  }
}

// the same combination across members used as lanes (one displacement component per lane): output
// lane l = blockIdx.y gets i sum_k w_{dir[l][k]} X_k with X_k the slot slot[l][k] of lane lane[l][k];
// the terms are summed in k_spec_lin's order.  NT terms (every output lane of a launch has as many);
// the source pointers are taken once per block and the spectral index is split in 32 bits.
template <int NT>
__global__ void __launch_bounds__(256) k_lane_lin(const MemberDev* __restrict__ mem, Dims d, LaneTerms T, int s_out) {
  const int l = blockIdx.y;
  float2* __restrict__ o = craw(mem[l], s_out);
  const float2* __restrict__ src[NT];
  int dir[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    src[t] = craw(mem[T.lane[l][t]], T.slot[l][t]);
    dir[t] = T.dir[l][t];
  }
  const unsigned m2 = (unsigned)d.m2, n1 = (unsigned)d.n1, nspec = (unsigned)d.nspec;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < nspec; i += stride) {
    const unsigned r = i / m2;
    const int kz = (int)(i - r * m2);
    const unsigned kxu = r / n1;
    const int ky = (int)(r - kxu * n1), kx = (int)kxu;
    const float w[3] = {(2 * kx == d.n0) ? 0.0f : (float)(2 * kx < d.n0 ? kx : kx - d.n0),
                        (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1), (float)kz};
    float re = 0.0f, im = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const float2 a = src[t][i];
      const float wt = w[dir[t]];
      re += wt * a.x;
      im += wt * a.y;
    }
    o[i] = make_float2(-im, re);
  }
}

// precFactorElasticity (src/cuda/ElasticitySolver.cu:36-106) on packed spectra, without the 1/N^3
// that the inverse transform applies: s1 = -screen - mu |w|^2, s2 = lam + mu,
// s3 = 1 / (1 - |w|^2 s2 / s1), U_d = F_d / s1 + (s2 s3 / s1^2) w_d (w.F).  Out of place.
__global__ void __launch_bounds__(256) k_el_prec(const MemberDev* __restrict__ mem, Dims d, int4 sf, int4 su, float mu,
                                                 float lam, float scr) {
  const MemberDev& M = mem[blockIdx.y];
  const float2* F[3] = {craw(M, sf.x), craw(M, sf.y), craw(M, sf.z)};
  float2* U[3] = {craw(M, su.x), craw(M, su.y), craw(M, su.z)};
  const float s2 = lam + mu;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nspec; i += stride) {
    const int kz = (int)(i % d.m2);
    const int64_t r = i / d.m2;
    const int ky = (int)(r % d.n1), kx = (int)(r / d.n1);
    const float w[3] = {(2 * kx == d.n0) ? 0.0f : (float)(2 * kx < d.n0 ? kx : kx - d.n0),
                        (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1), (float)kz};
    const float ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const float s1 = -scr - mu * ww;
    const float s3 = 1.0f / (1.0f - ww * s2 / s1);
    const float c = s2 * s3 / (s1 * s1);
    const float2 f0 = F[0][i], f1 = F[1][i], f2 = F[2][i];
    const float wfx = w[0] * f0.x + w[1] * f1.x + w[2] * f2.x, wfy = w[0] * f0.y + w[1] * f1.y + w[2] * f2.y;
    const float2 f[3] = {f0, f1, f2};
#pragma unroll
    for (int k = 0; k < 3; k++) U[k][i] = make_float2(f[k].x / s1 + c * w[k] * wfx, f[k].y / s1 + c * w[k] * wfy);
  }
}

// fixed-order sums of the Z-pass block partials of the first nl members: out[q] = sum over lanes l
// and blocks b < nb of part_l[nq b + q] (one block)
__global__ void __launch_bounds__(256) k_lane_part_sum(const MemberDev* __restrict__ mem, int nl, int nb, int nq,
                                                       double* __restrict__ out) {
  __shared__ double red[256];
  for (int q = 0; q < nq; q++) {
    double v = 0.0;
    for (int l = 0; l < nl; l++)
      for (int b = threadIdx.x; b < nb; b += 256) v += mem[l].part[(int64_t)nq * b + q];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

// k_el_prec across lanes: F_d in slot sf of lane d, U_d into slot su of lane d (one pass, grid.y = 1)
__global__ void __launch_bounds__(256) k_el_prec_lanes(const MemberDev* __restrict__ mem, Dims d, int sf, int su,
                                                       float mu, float lam, float scr, int honor) {
  if (honor && mem[0].st->done) return;
  const float2* F[3] = {craw(mem[0], sf), craw(mem[1], sf), craw(mem[2], sf)};
  float2* U[3] = {craw(mem[0], su), craw(mem[1], su), craw(mem[2], su)};
  const float s2 = lam + mu;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < d.nspec; i += stride) {
    const int kz = (int)(i % d.m2);
    const int64_t r = i / d.m2;
    const int ky = (int)(r % d.n1), kx = (int)(r / d.n1);
    const float w[3] = {(2 * kx == d.n0) ? 0.0f : (float)(2 * kx < d.n0 ? kx : kx - d.n0),
                        (2 * ky == d.n1) ? 0.0f : (float)(2 * ky < d.n1 ? ky : ky - d.n1), (float)kz};
    const float ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const float s1 = -scr - mu * ww;
    const float s3 = 1.0f / (1.0f - ww * s2 / s1);
    const float c = s2 * s3 / (s1 * s1);
    const float2 f0 = F[0][i], f1 = F[1][i], f2 = F[2][i];
    const float wfx = w[0] * f0.x + w[1] * f1.x + w[2] * f2.x, wfy = w[0] * f0.y + w[1] * f1.y + w[2] * f2.y;
    const float2 f[3] = {f0, f1, f2};
#pragma unroll
    for (int k = 0; k < 3; k++) U[k][i] = make_float2(f[k].x / s1 + c * w[k] * wfx, f[k].y / s1 + c * w[k] * wfy);
  }
}

// ---------------------------------------------------------------------------------------------
// launchers (size dispatch)
// ---------------------------------------------------------------------------------------------
#define GLIA_CHECK_LAUNCH() \
  do {                      \
    hipError_t e_ = hipGetLastError(); \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

template <template <int> class F, typename... Args>
static int dispatch_L(int L, Args&&... args) {
  switch (L) {
    case 32: return F<32>::run(args...);
    case 48: return F<48>::run(args...);
    case 64: return F<64>::run(args...);
    case 80: return F<80>::run(args...);
    case 96: return F<96>::run(args...);
    case 128: return F<128>::run(args...);
    case 160: return F<160>::run(args...);
    case 192: return F<192>::run(args...);
    case 256: return F<256>::run(args...);
    default: return -1;
  }
}

template <template <int> class F, typename... Args>
static int dispatch_M(int M, Args&&... args) {
  switch (M) {
    case 16: return F<16>::run(args...);
    case 24: return F<24>::run(args...);
    case 32: return F<32>::run(args...);
    case 48: return F<48>::run(args...);
    case 64: return F<64>::run(args...);
    case 80: return F<80>::run(args...);
    case 96: return F<96>::run(args...);
    case 128: return F<128>::run(args...);
    default: return -1;
  }
}

bool supported_line(int L) {
  return L == 32 || L == 48 || L == 64 || L == 80 || L == 96 || L == 128 || L == 160 || L == 192 || L == 256;
}

template <int L>
struct RunXiGrad {
  static int run(const LaunchCtx& c, int s_in, int s_r, int s_u, int s_v, bool pupd) {
    dim3 g(c.d.n1 * (c.d.m2 / 16), c.nm);
    if (pupd)
      k_xi_grad<L, true><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, s_in, s_r, s_u, s_v);
    else
      k_xi_grad<L, false><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, s_in, s_r, s_u, s_v);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_xi_grad(const LaunchCtx& c, int s_in, int s_r, int s_u, int s_v, bool pupd) {
  return dispatch_L<RunXiGrad>(c.d.n0, c, s_in, s_r, s_u, s_v, pupd);
}

template <int L>
struct RunYiGrad {
  static int run(const LaunchCtx& c, int s_u, int s_v, int s_gx, int s_gy, int s_gz, int s_g0) {
    dim3 g(c.d.n0 * (c.d.m2 / 16), c.nm);
    if (s_g0 >= 0)
      k_yi_grad<L, true><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_u, s_v, s_gx, s_gy, s_gz, s_g0);
    else
      k_yi_grad<L, false><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_u, s_v, s_gx, s_gy, s_gz, 0);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_yi_grad(const LaunchCtx& c, int s_u, int s_v, int s_gx, int s_gy, int s_gz, int s_g0) {
  return dispatch_L<RunYiGrad>(c.d.n1, c, s_u, s_v, s_gx, s_gy, s_gz, s_g0);
}

template <int L>
struct RunYfDiv {
  static int run(const LaunchCtx& c, int s_hx, int s_hy, int s_q, int s_yx, int s_rr) {
    dim3 g(c.d.n0 * (c.d.m2 / 16), c.nm);
    if (c.yf_split == 1) {
      k_yf_div_s<L><<<g, 3 * CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_hx, s_hy, s_q, s_yx, s_rr);
    } else if (c.yf_split == 2) {
      dim3 g2(2 * g.x, g.y);
      k_yf_div_b<L, 3><<<g2, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_hx, s_hy, s_q, s_yx, s_rr);
    } else if (c.yf_split == 3) {
      dim3 g2(2 * g.x, g.y);
      k_yf_div_b<L, 2><<<g2, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_hx, s_hy, s_q, s_yx, s_rr);
    } else
      k_yf_div<L><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw1, s_hx, s_hy, s_q, s_yx, s_rr);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_yf_div(const LaunchCtx& c, int s_hx, int s_hy, int s_q, int s_yx, int s_rr) {
  return dispatch_L<RunYfDiv>(c.d.n1, c, s_hx, s_hy, s_q, s_yx, s_rr);
}

template <int L>
struct RunXfDiv {
  static int run(const LaunchCtx& c, int mode, int s_yx, int s_rr, int s_p, int s_out) {
    dim3 g(c.d.n1 * (c.d.m2 / 16), c.nm);
    const int tb = 2 * CK<L>::TPB, tb3 = 3 * CK<L>::TPB;
    if (c.xf_split == 2 && mode == XF_MATVEC)
      k_xf_div_s3<L, XF_MATVEC><<<g, tb3, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (c.xf_split == 2 && mode == XF_SETUP)
      k_xf_div_s3<L, XF_SETUP><<<g, tb3, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (c.xf_split == 2)
      k_xf_div_s3<L, XF_D><<<g, tb3, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (c.xf_split && mode == XF_MATVEC)
      k_xf_div_s<L, XF_MATVEC><<<g, tb, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (c.xf_split && mode == XF_SETUP)
      k_xf_div_s<L, XF_SETUP><<<g, tb, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (c.xf_split)
      k_xf_div_s<L, XF_D><<<g, tb, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (mode == XF_MATVEC)
      k_xf_div<L, XF_MATVEC><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else if (mode == XF_SETUP)
      k_xf_div<L, XF_SETUP><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    else
      k_xf_div<L, XF_D><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, s_yx, s_rr, s_p, s_out);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_xf_div(const LaunchCtx& c, int mode, int s_yx, int s_rr, int s_p, int s_out) {
  return dispatch_L<RunXfDiv>(c.d.n0, c, mode, s_yx, s_rr, s_p, s_out);
}
int xf_div_blocks(const Dims& d) { return d.n1 * (d.m2 / 16); }

template <int L>
struct RunColPlain {
  static int run(const LaunchCtx& c, int dir, int sign, int s_in, int s_out, int honor) {
    const int other = (dir == 0) ? c.d.n1 : c.d.n0;
    const float2* tw = (dir == 0) ? c.tw0 : c.tw1;
    dim3 g(other * (c.d.m2 / 16), c.nm);
    if (dir == 0 && sign < 0) k_col_plain<L, 0, -1><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, tw, s_in, s_out, honor);
    else if (dir == 0) k_col_plain<L, 0, 1><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, tw, s_in, s_out, honor);
    else if (sign < 0) k_col_plain<L, 1, -1><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, tw, s_in, s_out, honor);
    else k_col_plain<L, 1, 1><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, tw, s_in, s_out, honor);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
template <int L>
struct RunColLin {
  static int run(const LaunchCtx& c, const LaneTerms& T, int s_out, int honor) {
    dim3 g(c.d.n1 * (c.d.m2 / 16), c.nm);
    switch (T.n[0]) {
      case 1: k_col_lin<L, 1><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, T, s_out, honor); break;
      case 2: k_col_lin<L, 2><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, T, s_out, honor); break;
      case 3: k_col_lin<L, 3><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, T, s_out, honor); break;
      default: k_col_lin<L, 4><<<g, CK<L>::TPB, 0, c.stream>>>(c.mem, c.d, c.tw0, T, s_out, honor); break;
    }
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
static int lane_terms(const LaunchCtx& c, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                      const int (*dir)[4], LaneTerms* T) {
  if (c.nm < 1 || c.nm > 3) return (int)hipErrorInvalidValue;
  *T = LaneTerms{};
  for (int l = 0; l < c.nm; l++) {
    if (nterms[l] < 1 || nterms[l] > 4 || nterms[l] != nterms[0]) return (int)hipErrorInvalidValue;
    T->n[l] = nterms[l];
    for (int t = 0; t < nterms[l]; t++) {
      T->lane[l][t] = lane[l][t];
      T->slot[l][t] = slot[l][t];
      T->dir[l][t] = dir[l][t];
    }
  }
  return 0;
}
int launch_col_lin(const LaunchCtx& c, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                   const int (*dir)[4], int s_out, bool honor_done) {
  LaneTerms T;
  const int rc = lane_terms(c, nterms, lane, slot, dir, &T);
  if (rc) return rc;
  return dispatch_L<RunColLin>(c.d.n0, c, T, s_out, honor_done ? 1 : 0);
}

int launch_col_plain(const LaunchCtx& c, int dir, int sign, int s_in, int s_out, bool honor_done) {
  return dispatch_L<RunColPlain>(dir == 0 ? c.d.n0 : c.d.n1, c, dir, sign, s_in, s_out, honor_done ? 1 : 0);
}

template <int M>
struct RunZDiff {
  static int run(const LaunchCtx& c, int s_gx, int s_gy, int s_gz, int k_slot) {
    dim3 g((unsigned)((int64_t)c.d.n0 * c.d.n1 / ZK<M>::NL * 3), c.nm);
    k_z_diff<M><<<g, ZK<M>::TPB, 0, c.stream>>>(c.mem, c.d, c.twm, c.tw2, s_gx, s_gy, s_gz, k_slot);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_z_diff(const LaunchCtx& c, int s_gx, int s_gy, int s_gz, int k_slot) {
  return dispatch_M<RunZDiff>(c.d.m2, c, s_gx, s_gy, s_gz, k_slot);
}

template <int M>
struct RunZPoint {
  static int run(const LaunchCtx& c, int mode, int s_in, int s_out, int r_slot, int ch_idx, float dt, int honor) {
    dim3 g((unsigned)((int64_t)c.d.n0 * c.d.n1 / ZK<M>::NL), c.nm);
#define GLIA_ZP(MD) k_z_point<M, MD><<<g, ZK<M>::TPB, 0, c.stream>>>(c.mem, c.d, c.twm, c.tw2, s_in, s_out, r_slot, ch_idx, dt, honor)
    switch (mode) {
      case ZR_REACT: GLIA_ZP(ZR_REACT); break;
      case ZR_REACT_ADJ: GLIA_ZP(ZR_REACT_ADJ); break;
      case ZR_C2R: GLIA_ZP(ZR_C2R); break;
      case ZR_R2C: GLIA_ZP(ZR_R2C); break;
      case ZR_MISFIT: GLIA_ZP(ZR_MISFIT); break;
      case ZR_OBS2: GLIA_ZP(ZR_OBS2); break;
      case ZR_COEF: GLIA_ZP(ZR_COEF); break;
      case ZR_SCREEN: GLIA_ZP(ZR_SCREEN); break;
      case ZR_C2R_DOT: GLIA_ZP(ZR_C2R_DOT); break;
      default: return -1;
    }
#undef GLIA_ZP
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_z_point(const LaunchCtx& c, int mode, int s_in, int s_out, int r_slot, int ch_idx, float dt, bool honor_done) {
  return dispatch_M<RunZPoint>(c.d.m2, c, mode, s_in, s_out, r_slot, ch_idx, dt, honor_done ? 1 : 0);
}
template <int M>
struct RunZUpd {
  static int run(const LaunchCtx& c, int mode, int ra, int rb, int rc, int rd, const double* sc, int s_out, int honor) {
    dim3 g((unsigned)((int64_t)c.d.n0 * c.d.n1 / ZK<M>::NL), c.nm);
    if (mode == ZU_AXPBY)
      k_z_upd_r2c<M, ZU_AXPBY><<<g, ZK<M>::TPB, 0, c.stream>>>(c.mem, c.d, c.twm, c.tw2, ra, rb, rc, rd, sc, s_out,
                                                                 honor);
    else if (mode == ZU_CGUPD)
      k_z_upd_r2c<M, ZU_CGUPD><<<g, ZK<M>::TPB, 0, c.stream>>>(c.mem, c.d, c.twm, c.tw2, ra, rb, rc, rd, sc, s_out,
                                                                 honor);
    else
      return -1;
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_z_upd_r2c(const LaunchCtx& c, int mode, int ra, int rb, int rc, int rd, const double* sc, int s_out,
                     bool honor_done) {
  if (mode == ZU_CGUPD && !sc) return -1;
  return dispatch_M<RunZUpd>(c.d.m2, c, mode, ra, rb, rc, rd, sc, s_out, honor_done ? 1 : 0);
}
int z_blocks(const Dims& d) {
  switch (d.m2) {
    case 16: return (int)((int64_t)d.n0 * d.n1 / ZK<16>::NL);
    case 24: return (int)((int64_t)d.n0 * d.n1 / ZK<24>::NL);
    case 32: return (int)((int64_t)d.n0 * d.n1 / ZK<32>::NL);
    case 48: return (int)((int64_t)d.n0 * d.n1 / ZK<48>::NL);
    case 64: return (int)((int64_t)d.n0 * d.n1 / ZK<64>::NL);
    case 80: return (int)((int64_t)d.n0 * d.n1 / ZK<80>::NL);
    case 96: return (int)((int64_t)d.n0 * d.n1 / ZK<96>::NL);
    case 128: return (int)((int64_t)d.n0 * d.n1 / ZK<128>::NL);
    default: return -1;
  }
}
int z_lines_per_block(int m2) {
  switch (m2) {
    case 16: return ZK<16>::NL;
    case 24: return ZK<24>::NL;
    case 32: return ZK<32>::NL;
    case 48: return ZK<48>::NL;
    case 64: return ZK<64>::NL;
    case 80: return ZK<80>::NL;
    case 96: return ZK<96>::NL;
    case 128: return ZK<128>::NL;
    default: return -1;
  }
}

template <int M>
struct RunZGradq {
  static int run(const LaunchCtx& c, int4 sc, int4 sa) {
    dim3 g((unsigned)((int64_t)c.d.n0 * c.d.n1 / ZK<M>::NL), c.nm);
    k_z_gradq<M><<<g, ZK<M>::TPB, 0, c.stream>>>(c.mem, c.d, c.twm, c.tw2, sc, sa);
    GLIA_CHECK_LAUNCH();
    return 0;
  }
};
int launch_z_gradq(const LaunchCtx& c, int4 sc, int4 sa) { return dispatch_M<RunZGradq>(c.d.m2, c, sc, sa); }

int launch_support_count(const LaunchCtx& c, const float* f, int step, int l0, int l1, int l2, int R, int d2max,
                         float thr, int* cnt) {
  const int64_t nl = (int64_t)l0 * l1 * l2;
  if (nl <= 0 || nl > c.d.ng || step < 1 || R < 0) return -1;
  k_support_count<<<(unsigned)((nl + 255) / 256), 256, 0, c.stream>>>(f, c.d, step, l0, l1, l2, R, d2max, thr, cnt);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_smooth_axis(const LaunchCtx& c, const float* in, float* out, int axis, const double* taps, int kr) {
  const dim3 g((unsigned)((c.d.n2 + 63) / 64), (unsigned)((c.d.n1 + 3) / 4), (unsigned)c.d.n0), b(64, 4);
#define GLIA_SMK(K) \
  case K: k_smooth_axis_r<K><<<g, b, 0, c.stream>>>(in, out, c.d, axis, taps); break;
  switch (kr) {
    GLIA_SMK(1) GLIA_SMK(2) GLIA_SMK(3) GLIA_SMK(4) GLIA_SMK(5) GLIA_SMK(6) GLIA_SMK(7) GLIA_SMK(8)
    GLIA_SMK(9) GLIA_SMK(10) GLIA_SMK(11) GLIA_SMK(12) GLIA_SMK(13) GLIA_SMK(14) GLIA_SMK(15) GLIA_SMK(16)
    default: k_smooth_axis<<<g, b, 0, c.stream>>>(in, out, c.d, axis, taps, kr); break;
  }
#undef GLIA_SMK
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_pcg_update(const LaunchCtx& c, int s_x) {
  dim3 g(UPD_BLOCKS, c.nm);
  k_pcg_update<<<g, 256, 0, c.stream>>>(c.mem, c.d, s_x);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int pcg_update_blocks() { return UPD_BLOCKS; }

int launch_pcg_begin(const LaunchCtx& c, double tau) {
  k_pcg_begin<<<c.nm, 64, 0, c.stream>>>(c.mem, tau);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_pcg_fin_setup(const LaunchCtx& c, int nb, int* hflag, int hslot) {
  k_pcg_fin_setup<<<c.nm, 256, 0, c.stream>>>(c.mem, c.d, nb, hflag, hslot, c.nm);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_pcg_fin_dpi(const LaunchCtx& c, int nb) {
  k_pcg_fin_dpi<<<c.nm, 256, 0, c.stream>>>(c.mem, c.d, nb);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_pcg_fin_update(const LaunchCtx& c, int nb, int* hflag, int hslot) {
  k_pcg_fin_update<<<c.nm, 256, 0, c.stream>>>(c.mem, c.d, nb, hflag, hslot, c.nm);
  GLIA_CHECK_LAUNCH();
  return 0;
}

constexpr int EW_BLOCKS = 1024;
int ew_blocks() { return EW_BLOCKS; }

int launch_copy(const LaunchCtx& c, int s_in, int s_out) {
  k_copy<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, s_in, s_out);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_tissue(const LaunchCtx& c) {
  k_tissue<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_coef(const LaunchCtx& c, int r_out, const double* coefs_dev) {
  k_coef<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, r_out, coefs_dev);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_obs_default(const LaunchCtx& c, float thr) {
  k_obs_default<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, thr);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_fin_state(const LaunchCtx& c, int nb, int kind, double w) {
  k_fin_state<<<c.nm, 256, 0, c.stream>>>(c.mem, nb, kind, w);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_prec_factor(const LaunchCtx& c, double prec_dt) {
  k_prec_factor<<<c.nm, 64, 0, c.stream>>>(c.mem, prec_dt);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_state_reset(const LaunchCtx& c, int what) {
  k_state_reset<<<c.nm, 64, 0, c.stream>>>(c.mem, what);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_phi_apply(const LaunchCtx& c) {
  k_phi_apply<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_phi_applyT(const LaunchCtx& c, int np_max, int r_in, double scale) {
  if (np_max <= 0) return 0;
  k_phi_applyT<<<dim3(np_max, c.nm), 256, 0, c.stream>>>(c.mem, c.d, r_in, scale);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_dot_real(const LaunchCtx& c, int r_x, int r_y) {
  k_dot_real<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, r_x, r_y);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_set_done(const LaunchCtx& c, int v) {
  k_set_done<<<c.nm, 64, 0, c.stream>>>(c.mem, v);
  GLIA_CHECK_LAUNCH();
  return 0;
}
int launch_axpby_real(const LaunchCtx& c, int r_x, double a, int r_y, double b, int r_out) {
  k_axpby_real<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, r_x, a, r_y, b, r_out);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_spec_lin(const LaunchCtx& c, int nterms, const int* slots, const int* dirs, int s_out) {
  if (nterms < 1 || nterms > 4) return (int)hipErrorInvalidValue;
  SpecTerms T{};
  T.n = nterms;
  for (int t = 0; t < nterms; t++) {
    T.slot[t] = slots[t];
    T.dir[t] = dirs[t];
  }
  k_spec_lin<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, T, s_out);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_lane_lin(const LaunchCtx& c, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                    const int (*dir)[4], int s_out) {
  if (c.nm < 1 || c.nm > 3) return (int)hipErrorInvalidValue;
  LaneTerms T{};
  for (int l = 0; l < c.nm; l++) {
    if (nterms[l] < 1 || nterms[l] > 4) return (int)hipErrorInvalidValue;
    T.n[l] = nterms[l];
    for (int t = 0; t < nterms[l]; t++) {
      T.lane[l][t] = lane[l][t];
      T.slot[l][t] = slot[l][t];
      T.dir[l][t] = dir[l][t];
    }
  }
  for (int l = 1; l < c.nm; l++)
    if (T.n[l] != T.n[0]) return (int)hipErrorInvalidValue;  // one term count per launch
  if ((int64_t)c.d.nspec >= (int64_t)1 << 31) return (int)hipErrorInvalidValue;
  const dim3 g(EW_BLOCKS, c.nm);
  switch (T.n[0]) {
    case 1: k_lane_lin<1><<<g, 256, 0, c.stream>>>(c.mem, c.d, T, s_out); break;
    case 2: k_lane_lin<2><<<g, 256, 0, c.stream>>>(c.mem, c.d, T, s_out); break;
    case 3: k_lane_lin<3><<<g, 256, 0, c.stream>>>(c.mem, c.d, T, s_out); break;
    default: k_lane_lin<4><<<g, 256, 0, c.stream>>>(c.mem, c.d, T, s_out); break;
  }
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_lane_part_sum(const LaunchCtx& c, int nq, double* out) {
  k_lane_part_sum<<<1, 256, 0, c.stream>>>(c.mem, c.nm, z_blocks(c.d), nq, out);
  GLIA_CHECK_LAUNCH();
  return 0;
}

// The elasticity CG's state machine (KSPSolve_CG with KSP_NORM_PRECONDITIONED and
// KSPConvergedDefault, as the host loop of el_capi.cpp ran it) in lane 0's PcgState, one thread.
// k_el_cg_begin starts it from the host's initial scalars; k_el_cg_fin closes an iteration from the
// summed partials sc = (<p, A p>, <z, z>, <z, r>) and makes the checks of the next iteration's
// start.  When it stops, every lane is marked done, and the mapped host flag is set.
__global__ void k_el_cg_begin(const MemberDev* __restrict__ mem, int nl, double beta0, double dp0, double rnorm0,
                              double ttol, double atol, double dtol, int maxit) {
  if (threadIdx.x != 0) return;
  PcgState* S = mem[0].st;
  S->i = 0;
  S->its = 1;
  S->beta = beta0;
  S->betaold = 0.0;
  S->dpi = 0.0;
  S->dpiold = 0.0;
  S->dp = dp0;
  S->rnorm0 = rnorm0;
  S->ttol = ttol;
  S->atol = atol;
  S->dtol = dtol;
  S->maxit = maxit;
  S->reason = (beta0 == 0.0) ? KSPR_CONVERGED_ATOL : 0;
  const int done = S->reason != 0;
  for (int l = 0; l < nl; l++) mem[l].st->done = done;
}

__global__ void k_el_cg_fin(const MemberDev* __restrict__ mem, int nl, const double* __restrict__ sc, int* hflag) {
  if (threadIdx.x != 0) return;
  PcgState* S = mem[0].st;
  if (S->done) return;
  const double q = sc[0], zz = sc[1], zr = sc[2];
  const int i = S->i;
  const double dpiold = S->dpi;
  S->betaold = S->beta;
  S->dpiold = dpiold;
  S->dpi = q;
  int reason = 0;
  if (isnan(q) || isinf(q)) {
    reason = KSPR_DIVERGED_NANORINF;
  } else if (q == 0.0 || (i > 0 && q * dpiold <= 0.0)) {
    reason = KSPR_DIVERGED_INDEFINITE_MAT;
  } else {
    S->dp = sqrt(zz);
    reason = ksp_conv_test(S, S->dp);
    if (!reason) {
      S->beta = zr;
      if (isnan(zr) || isinf(zr)) {
        reason = KSPR_DIVERGED_NANORINF;
      } else {
        S->i = i + 1;
        if (i + 1 >= S->maxit) {
          reason = KSPR_DIVERGED_ITS;
        } else {
          S->its = i + 2;
          if (zr == 0.0)
            reason = KSPR_CONVERGED_ATOL;
          else if (zr * S->betaold < 0.0)
            reason = KSPR_DIVERGED_INDEFINITE_PC;
        }
      }
    }
  }
  if (reason) {
    S->reason = reason;
    for (int l = 0; l < nl; l++) mem[l].st->done = 1;
    if (hflag) *hflag = 1;
  }
}

int launch_el_cg_begin(const LaunchCtx& c, double beta0, double dp0, double rnorm0, double ttol, double atol,
                       double dtol, int maxit) {
  k_el_cg_begin<<<1, 64, 0, c.stream>>>(c.mem, c.nm, beta0, dp0, rnorm0, ttol, atol, dtol, maxit);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_el_cg_fin(const LaunchCtx& c, const double* sc, int* hflag) {
  k_el_cg_fin<<<1, 64, 0, c.stream>>>(c.mem, c.nm, sc, hflag);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_el_prec_lanes(const LaunchCtx& c, int sf, int su, double mu, double lam, double scr, bool honor_done) {
  k_el_prec_lanes<<<dim3(EW_BLOCKS, 1), 256, 0, c.stream>>>(c.mem, c.d, sf, su, (float)mu, (float)lam, (float)scr,
                                                             honor_done ? 1 : 0);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_el_prec(const LaunchCtx& c, int4 sf, int4 su, double mu, double lam, double scr) {
  k_el_prec<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, sf, su, (float)mu, (float)lam, (float)scr);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_spec_div(const LaunchCtx& c, int s_x, int s_y, int s_z, int s_out) {
  k_spec_div<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, s_x, s_y, s_z, s_out);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_axpby_c(const LaunchCtx& c, int s_x, double a, int s_y, double b, int s_out) {
  k_axpby_c<<<dim3(EW_BLOCKS, c.nm), 256, 0, c.stream>>>(c.mem, c.d, s_x, a, s_y, b, s_out);
  GLIA_CHECK_LAUNCH();
  return 0;
}

int launch_bricks(const LaunchCtx& c, const BrickJob* jobs_host, int njobs, const double* taps_host, int tap_stride,
                  double h0, double h1, double h2, double sigma, float* store_dev, double* max_out_host) {
  if (njobs <= 0) return 0;
  int64_t ext_total = 0;
  for (int i = 0; i < njobs; i++) {
    const BrickJob& J = jobs_host[i];
    const int64_t ne = (int64_t)(J.b[0] + 2 * J.kr[0]) * (J.b[1] + 2 * J.kr[1]) * (J.b[2] + 2 * J.kr[2]);
    if (jobs_host[i].ext_off + ne > ext_total) ext_total = jobs_host[i].ext_off + ne;
  }
  BrickJob* jd = nullptr;
  double *e0 = nullptr, *e1 = nullptr, *taps = nullptr, *mx = nullptr;
  hipError_t err = hipSuccess;
#define GLIA_TRY(x) do { err = (x); if (err != hipSuccess) goto fail; } while (0)
  GLIA_TRY(hipMalloc(&jd, sizeof(BrickJob) * njobs));
  GLIA_TRY(hipMalloc(&e0, sizeof(double) * ext_total));
  GLIA_TRY(hipMalloc(&e1, sizeof(double) * ext_total));
  GLIA_TRY(hipMalloc(&taps, sizeof(double) * (size_t)njobs * 3 * tap_stride));
  GLIA_TRY(hipMalloc(&mx, sizeof(double) * njobs));
  GLIA_TRY(hipMemcpyAsync(jd, jobs_host, sizeof(BrickJob) * njobs, hipMemcpyHostToDevice, c.stream));
  GLIA_TRY(hipMemcpyAsync(taps, taps_host, sizeof(double) * (size_t)njobs * 3 * tap_stride, hipMemcpyHostToDevice, c.stream));
  k_brick_src<<<dim3(256, njobs), 256, 0, c.stream>>>(jd, c.d, c.mem, h0, h1, h2, sigma, e0);
  k_brick_conv<<<dim3(256, njobs), 256, 0, c.stream>>>(jd, 0, taps, tap_stride, e0, e1);
  k_brick_conv<<<dim3(256, njobs), 256, 0, c.stream>>>(jd, 1, taps, tap_stride, e1, e0);
  k_brick_conv<<<dim3(256, njobs), 256, 0, c.stream>>>(jd, 2, taps, tap_stride, e0, e1);
  k_brick_store<<<dim3(256, njobs), 256, 0, c.stream>>>(jd, c.d, h0, h1, h2, sigma, e1, store_dev);
  k_brick_max<<<njobs, 256, 0, c.stream>>>(jd, store_dev, mx);
  GLIA_TRY(hipGetLastError());
  GLIA_TRY(hipMemcpyAsync(max_out_host, mx, sizeof(double) * njobs, hipMemcpyDeviceToHost, c.stream));
  GLIA_TRY(hipStreamSynchronize(c.stream));
fail:
#undef GLIA_TRY
  (void)hipFree(jd);
  (void)hipFree(e0);
  (void)hipFree(e1);
  (void)hipFree(taps);
  (void)hipFree(mx);
  return (int)err;
}

}  // namespace glia
