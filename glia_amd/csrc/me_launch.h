// me_launch.h -- launchers of the pointwise kernels of the mass-effect forward (me_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace glia {
namespace me {

int partial_count();
// MatProp::filterTumor: wm *= 1 - c, gm *= 1 - c
int launch_filter(hipStream_t st, int64_t ng, const float* c, float* wm, float* gm);
// per-block maxima of |v| (partial_count() floats)
int launch_vmax(hipStream_t st, int64_t ng, const float* vx, const float* vy, const float* vz, float* part);
// PdeOperatorsRD::reaction(0): the logistic step of c with the reaction field rho over dt
int launch_reaction(hipStream_t st, int64_t ng, float* c, const float* rho, float dt);
// v = (u - u_old) / dt for the three components, then u_old = u
int launch_velocity(hipStream_t st, int64_t ng, float* const v[3], const float* const u[3], float* const uold[3],
                    float inv_dt);

// the multispecies start: p = max(c0, 0), i = 0, ox = 1 (src/pde/PdeOperatorsMultiSpecies.cpp:253-266)
int launch_species_init(hipStream_t st, int64_t ng, const float* c0, float* p, float* i, float* ox);
// out = a + b (+ c when c is not null)
int launch_sum(hipStream_t st, int64_t ng, const float* a, const float* b, const float* c, float* out);

}  // namespace me
}  // namespace glia
