// sl_launch.h -- launchers of the semi-Lagrangian kernels (sl_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace glia {
namespace sl {

struct SLDims {
  int n0, n1, n2;
  int64_t ng;
};

// fields advected by one launch on one set of departure offsets; in[f] and out[f] must not alias
constexpr int SL_MAX_FIELDS = 8;
struct SLFields {
  const float* in[SL_MAX_FIELDS];
  float* out[SL_MAX_FIELDS];
  int mode[SL_MAX_FIELDS];  // 1 mass conservation, 2 pure advection
  int n;
  int any_mass;
};

int launch_sl_trajectories(const SLDims& d, hipStream_t st, const float* vx, const float* vy, const float* vz,
                           float sx, float sy, float sz, float* qx, float* qy, float* qz);
int launch_sl_apply(const SLDims& d, hipStream_t st, const SLFields& F, const float* qx, const float* qy,
                    const float* qz, const float* dv, const float* dvi, float dt);

}  // namespace sl
}  // namespace glia
