// el_launch.h -- launchers of the elasticity kernels (el_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace glia {
namespace el {

// Lame parameters of the tissue classes and the screening coefficients
struct ElMaterial {
  float mu_h, mu_t, mu_bg, mu_c, lam_h, lam_t, lam_bg, lam_c, screen_low, screen_high;
};
struct Vec3 {
  float* p[3];
};
struct Vec3C {
  const float* p[3];
};

int partial_count();
int launch_material(hipStream_t st, int64_t ng, const float* c, const float* wm, const float* gm, const float* csf,
                    const float* vt, const float* bg, const ElMaterial& m, const Vec3& mu, float* lam, const Vec3& scr,
                    double* part);
int launch_force_scale(hipStream_t st, int64_t ng, const float* g, const float* c, float ff, int use_tanh, float* out);
int launch_axpby3(hipStream_t st, int64_t ng, const Vec3& p, const Vec3C& z, double b, int sub);

}  // namespace el
}  // namespace glia
