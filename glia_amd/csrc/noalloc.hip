// noalloc.hip -- glia_rd_smooth_noalloc (include/glia_rd.h): the periodic Gaussian smoother of
// glia_rd_smooth_field on caller-owned device fields, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops).  Its scratch memory -- the taps and the
// intermediate field of the three axis passes -- is one file-scope device array per GPU, so a call
// allocates nothing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/glia_rd.h"
#include "rd_launch.h"

namespace {

// scratch layout: the taps of the three axes (3 x (2 kr + 1) doubles, kr <= kMaxRadius), then the
// intermediate field (at most kMaxPoints floats)
constexpr int kMaxRadius = 127;
constexpr int kTapStride = 2 * kMaxRadius + 1;
constexpr size_t kTapBytes = ((sizeof(double) * 3 * kTapStride + 255) / 256) * 256;
constexpr int64_t kMaxPoints = int64_t(1) << 23;  // 8,388,608 grid points (e.g. 192^3, 128 x 256^2)
constexpr size_t kScratchBytes = kTapBytes + sizeof(float) * kMaxPoints;  // 32 MiB + 6 KiB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// the device address of this GPU's scratch array (the host-side symbol is only a stand-in)
hipError_t scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

int fail(char* err, int len, int rc, const char* msg) {
  if (err && len > 0) {
    std::strncpy(err, msg, (size_t)len - 1);
    err[len - 1] = '\0';
  }
  return rc;
}

// The taps of glia_rd_smooth_field's smoother on axis blockIdx.x (length n[a], spacing 2 pi / n[a]):
// f(i) = exp(-X^2 / 2 sigma^2) + exp(-(X - 2 pi)^2 / 2 sigma^2) at X = i h, normalized by its sum
// over the axis; entry s + kr_a of the axis's row (stride doubles apart) is f(s mod n) / sum for
// |s| <= kr_a.  One block per axis; the sum is reduced in a fixed order, so every call gives the
// same taps.
__global__ void k_smooth_taps(int n0, int n1, int n2, double sigma, int kr0, int kr1, int kr2, int stride,
                              double* __restrict__ taps) {
  const int a = blockIdx.x;
  const int n = (a == 0) ? n0 : (a == 1) ? n1 : n2;
  const int kr = (a == 0) ? kr0 : (a == 1) ? kr1 : kr2;
  const double h = 2.0 * M_PI / n;
  const double den = 2.0 * sigma * sigma;
  __shared__ double part[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double X = i * h, Xp = X - 2.0 * M_PI;
    double v = exp(-X * X / den) + exp(-Xp * Xp / den);
    if (v != v) v = 0.0;
    s += v;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  const double sum = part[0];
  for (int t = threadIdx.x; t < 2 * kr + 1; t += blockDim.x) {
    const int i = (((t - kr) % n) + n) % n;
    const double X = i * h, Xp = X - 2.0 * M_PI;
    double v = exp(-X * X / den) + exp(-Xp * Xp / den);
    if (v != v) v = 0.0;
    taps[a * stride + t] = v / sum;
  }
}

}  // namespace

extern "C" int glia_rd_smooth_noalloc(const int n[3], const float* in, float* out, double sigma, void* stream,
                                      char* err, int err_len) {
  if (!n || !in || !out) return fail(err, err_len, GLIA_RD_ERR_ARG, "null grid or field");
  if (in == out) return fail(err, err_len, GLIA_RD_ERR_ARG, "in and out must be distinct fields");
  for (int a = 0; a < 3; a++)
    if (n[a] < 4) return fail(err, err_len, GLIA_RD_ERR_ARG, "every grid length must be at least 4");
  if (!(sigma > 0) || !std::isfinite(sigma)) return fail(err, err_len, GLIA_RD_ERR_ARG, "sigma must be positive");
  const int64_t ng = (int64_t)n[0] * n[1] * n[2];
  if (ng > kMaxPoints) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "grid of %lld points exceeds the scratch limit of %lld points",
                  (long long)ng, (long long)kMaxPoints);
    return fail(err, err_len, GLIA_RD_ERR_UNSUPPORTED, msg);
  }
  // the smoother's radius on each axis: 12 sigma in grid units, at most half the axis less one
  int kra[3], kr = 0;
  for (int a = 0; a < 3; a++) {
    const double h = 2.0 * M_PI / n[a];
    kra[a] = std::min(n[a] / 2 - 1, (int)std::ceil(12.0 * sigma / h));
    kr = std::max(kr, kra[a]);
  }
  if (kr > kMaxRadius) return fail(err, err_len, GLIA_RD_ERR_UNSUPPORTED, "sigma too wide: the smoother radius exceeds 127 points");
  void* base = nullptr;
  hipError_t e = scratch_address(&base);
  if (e != hipSuccess) return fail(err, err_len, (int)e, "hipGetSymbolAddress of the scratch array failed");
  double* taps = reinterpret_cast<double*>(base);
  float* mid = reinterpret_cast<float*>(static_cast<unsigned char*>(base) + kTapBytes);
  glia::LaunchCtx c{};
  c.d.n0 = n[0];
  c.d.n1 = n[1];
  c.d.n2 = n[2];
  c.d.m2 = n[2] / 2;
  c.d.ng = ng;
  c.d.nspec = (int64_t)n[0] * n[1] * (n[2] / 2);
  c.d.inv_ng = 1.0 / (double)ng;
  c.stream = (hipStream_t)stream;
  const int stride = 2 * kr + 1;
  hipLaunchKernelGGL(k_smooth_taps, dim3(3), dim3(256), 0, c.stream, n[0], n[1], n[2], sigma, kra[0], kra[1], kra[2],
                     stride, taps);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(err, err_len, (int)e, "launch of the taps kernel failed");
  int rc = glia::launch_smooth_axis(c, in, out, 0, taps, kra[0]);
  if (!rc) rc = glia::launch_smooth_axis(c, out, mid, 1, taps + stride, kra[1]);
  if (!rc) rc = glia::launch_smooth_axis(c, mid, out, 2, taps + 2 * stride, kra[2]);
  if (rc) return fail(err, err_len, rc, "launch of the smoothing passes failed");
  return 0;
}
