/*
 * glia_adv.h -- C ABI of the semi-Lagrangian advection solver: the reference's
 * SemiLagrangianSolver (include/pde/AdvectionSolver.h, src/pde/AdvectionSolver.cpp:279-704), which
 * the mass-effect models use to carry tissue and tumor fields along the velocity of the brain.
 *
 * Conventions are those of glia_rd.h: fields are FP32 n0*n1*n2 arrays with z fastest, mem_kind is
 * GLIA_RD_HOST or GLIA_RD_DEVICE, calls return 0, a negative GLIA_RD_ERR_* or a positive hipError_t,
 * and glia_sl_last_error() gives the message.  All device work of a handle is ordered on its
 * stream; use one handle per host thread.
 *
 * Interpolation: the value at a departure point is the tensor-product cubic Lagrange interpolant
 * of the 4 x 4 x 4 surrounding grid values on the periodic grid, the interpolant of the reference's
 * GPU path (cubicTex3DFastLagrange, src/cuda/InterpolationTexture.cu:270-366, wrap addressing).
 * The reference evaluates its middle taps with hardware-filtered texture fetches; here every tap is
 * exact FP32.
 */
#ifndef GLIA_ADV_H_
#define GLIA_ADV_H_

#include "glia_rd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glia_sl_s* glia_sl_t;

/* SemiLagrangianSolver::advection_mode_ */
enum { GLIA_SL_MASS_CONSERVATION = 1, GLIA_SL_PURE_ADVECTION = 2 };

/* SemiLagrangianSolver ctor (src/pde/AdvectionSolver.cpp:279-345): a solver for an n0 x n1 x n2
   grid (the grids glia_rd_create accepts).  It holds the velocity, the departure offsets, div v at
   the grid and departure points, two work fields and a one-member spectral engine for div v
   (SpectralOperators::computeDivergence).  stream: a hipStream_t, or NULL for a private
   non-blocking stream at default priority. */
int glia_sl_create(const int n[3], void* stream, glia_sl_t* out);
int glia_sl_destroy(glia_sl_t h);
const char* glia_sl_last_error(glia_sl_t h);
int glia_sl_sync(glia_sl_t h);

/* the velocity (Tumor::velocity_) of the next solves; like every change of velocity in the
   reference's callers (src/pde/PdeOperatorsMassEffect.cpp:607), it sets trajectoryIsComputed_ =
   false */
int glia_sl_set_velocity(glia_sl_t h, const float* vx, const float* vy, const float* vz, int mem_kind);
/* trajectoryIsComputed_ = false without a new velocity */
int glia_sl_invalidate(glia_sl_t h);
/* computeTrajectories (src/pde/AdvectionSolver.cpp:470-590): RK2 departure points of the grid for
   the time step dt, X = x - dt (v(x) + v(X*)) / 2 with the Euler points X* = x - dt v(x) */
int glia_sl_compute_trajectories(glia_sl_t h, double dt);
/* X - x of the computed trajectories, in grid units (x_d / h_d, h_d = 2 pi / n_d) */
int glia_sl_get_offsets(glia_sl_t h, float* qx, float* qy, float* qz, int mem_kind);
/* interpolate(Vec output, Vec input) (src/pde/AdvectionSolver.cpp:364-417): out(x) = in(X(x)) at
   the computed departure points; in and out must not alias */
int glia_sl_interpolate(glia_sl_t h, const float* in, float* out, int mem_kind);
/* solve(Vec scalar, VecField velocity, dt) (src/pde/AdvectionSolver.cpp:592-660): one step of the
   field in place, with mode GLIA_SL_PURE_ADVECTION (c <- c(X)) or GLIA_SL_MASS_CONSERVATION
   (d_t c + v.grad c = -c div v: ci = c(X), s = -ci (div v)(X), c* = ci + dt s,
   c <- ci + dt/2 (s - c* div v)).  Stale trajectories are computed first with this dt; fresh ones
   are reused whatever dt is, as in the reference. */
int glia_sl_solve(glia_sl_t h, float* scalar, double dt, int mode, int mem_kind);
/* nfields solves on one set of trajectories in one pass over the grid (device arrays): out[f] =
   one step of in[f] with modes[f].  No out[f] may alias any in[g].  The mass-effect step advects
   gm, wm, vt, csf, c and the MRI this way (src/pde/PdeOperatorsMassEffect.cpp:605-620); this entry
   point has no single counterpart in the reference. */
int glia_sl_advect(glia_sl_t h, int nfields, const float* const* in, float* const* out, const int* modes, double dt);

#ifdef __cplusplus
}
#endif

#endif /* GLIA_ADV_H_ */
