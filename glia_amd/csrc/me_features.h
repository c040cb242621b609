// me_features.h -- launchers of the biophysical-feature kernels of the model-4 loop (me_features.hip).
// Fields are FP32 arrays of n values on the device; every launcher enqueues on st and returns 0 or a
// hipError_t.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace glia {
namespace mef {

// per-field layout of the sums in acc (doubles), kPerField slots per field from its slot:
//   0 sum x, 1 sum x tc, 2 sum x healthy, 3 sum I x, 4 tc voxels, 5 healthy voxels,
//   6..8 sums of squared deviations (grid, tc, healthy),
//   9 + 3 r + k: quantile k (0.875, 0.95, 0.99) of region r (grid, tc, healthy)
constexpr int kPerField = 18;

struct StressIn {
  const float* d[9];  // d[3 a + b] = d u_a / d x_b
};

// bytes of device scratch that launch_cstar / launch_stats need for n values
size_t scratch_bytes(int64_t n);

int launch_mag3(hipStream_t st, int64_t n, const float* x, const float* y, const float* z, float* out);
// o_d = c v_d
int launch_flux(hipStream_t st, int64_t n, const float* c, const float* const v[3], float* const o[3]);
// out = (dc + rho (c - c^2)) - dv
int launch_dcdt(hipStream_t st, int64_t n, const float* dc, const float* c, const float* rho, const float* dv,
                float* out);
int launch_stress(hipStream_t st, int64_t n, const StressIn& g, const float* mu, const float* lam, float* jac,
                  float* tr, float* tau);
// *cstar = c at the first largest |grad c| (1 - tc)
int launch_cstar(hipStream_t st, int64_t n, const float* gradc, const float* seg, const float* c, void* scratch,
                 float* cstar);
// the kPerField sums and order statistics of x into acc[slot..]
int launch_stats(hipStream_t st, int64_t n, const float* x, const float* seg, const float* c, const float* cstar,
                 void* scratch, float* keys_in, float* keys_out, double* acc, int slot);

}  // namespace mef
}  // namespace glia
