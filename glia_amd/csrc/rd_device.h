// rd_device.h -- device-side data structures shared by the kernels and the host solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace glia {

// complex work slots of one ensemble member (each n0*n1*(n2/2) complex, the packed half spectrum)
enum CSlot {
  CS_X = 0,   // forward state c (spectral)
  CS_A,       // adjoint state alpha (spectral)
  CS_H,       // half-step recomputation buffer (spectral)
  CS_R,       // PCG residual
  CS_P,       // PCG search direction
  CS_W,       // PCG A*p
  CS_T0,      // transform intermediates
  CS_T1,
  CS_T2,
  CS_T3,
  CS_T4,
  CS_T5,
  CS_T6,
  CS_T7,
  CS_BK,      // copy of the adjoint state at the start of a captured adjoint step (re-run source)
  CS_NUM
};

// real fields of one member (each n0*n1*n2 floats)
enum RSlot {
  RS_K = 0,   // isotropic diffusion coefficient k(x)
  RS_RHO,     // reaction coefficient rho(x)
  RS_WM,
  RS_GM,
  RS_CSF,
  RS_VT,
  RS_FILTER,  // brain mask of MatProp
  RS_OBS,     // observation mask O
  RS_D1,      // data d1
  RS_C0,      // c(0) = Phi p
  RS_CT,      // c(T)
  RS_A0,      // alpha(0)
  RS_CH,      // half-step state c_half (real) for the adjoint reaction
  RS_KS,      // secondary diffusion coefficient k~ (wm + k_gm gm + k_glm csf) of the Hessian
  RS_NUM
};

// PETSc KSPConvergedReason values used by the CG state machine
enum {
  KSPR_CONVERGED_RTOL = 2,
  KSPR_CONVERGED_ATOL = 3,
  KSPR_DIVERGED_ITS = -3,
  KSPR_DIVERGED_DTOL = -4,
  KSPR_DIVERGED_INDEFINITE_PC = -8,
  KSPR_DIVERGED_NANORINF = -9,
  KSPR_DIVERGED_INDEFINITE_MAT = -10
};

// scalar state of one member, in device memory
struct PcgState {
  // CN solve parameters
  double tau;      // solve argument (A = I - tau/2 D)
  double pc;       // preconditioner coefficient 0.25 * prec_dt * kbar
  double rtol, atol, dtol;
  int maxit;
  int active;      // 0 when k_scale == 0 (solve is a no-op)
  // CG iteration state
  double beta, betaold, dpi, dpiold, alpha, bcoef, dp, rnorm0, ttol;
  int i, its, reason, done;
  int its_total;   // accumulated CG iterations since the last reset
  int solves;      // number of CN solves since the last reset
  int worst_reason;  // most negative reason seen since the last reset
  int pad0;
  // coefficient scalars
  double ksum, kavg, filter_sum;
  // objective / gradient accumulators (FP64)
  double m1;       // ||O c(T) - d1||^2
  double c0sq;     // ||c0||^2
  double gk, gr;   // trapezoid sums of <m_k, grad c . grad alpha> and <m_rho, (c^2 - c) alpha>
  // accumulators and counters at the start of a captured time step, restored when the step runs
  // again (one of its solves needed more iterations than the capture allowed)
  double gk_snap, gr_snap;
  int its_snap, solves_snap, worst_snap, pad1;
};

struct Brick {
  int o[3];     // origin (grid indices)
  int b[3];     // extent
  int64_t off;  // offset into the member's brick storage
};

// per-member pointers, in device memory (one array entry per member of a launch group)
struct MemberDev {
  float2* c[CS_NUM];
  float* r[RS_NUM];
  PcgState* st;
  double* part;        // per-workgroup partial sums (FP64)
  float2** hist;       // spectral states kept for the adjoint (store-all or checkpoints)
  float** chalf;       // real half-step states (adjoint_store), may be null
  // Gaussian basis
  const Brick* bricks;
  const float* brick_data;
  const double* phi_scale;  // per-basis scale used by apply (p_i / phi_max)
  double* phit_out;    // Phi^T results (np doubles)
  int np;
  int pad1;
  // coefficient scalars
  float rho_dt;        // unused placeholder for alignment
  float kr_gm;         // k_gm_wm ratio (gradient weight)
  float rr_gm;         // r_gm_wm ratio (gradient weight)
  float pad2;
};

// one brick of the Gaussian basis to build: the truncated, smoothed, masked Gaussian is computed
// on the brick [o, o+b) (within the grid) from an extended source box of kr more voxels per side
struct BrickJob {
  int o[3];
  int b[3];
  int kr[3];    // smoothing radius per axis (the brick is extended by kr[a] on both sides)
  int member;
  double ctr[3];
  int64_t ext_off;  // offset of this job's extended scratch (doubles)
  int64_t out_off;  // offset into the member's brick storage (floats)
};

struct Dims {
  int n0, n1, n2, m2;  // m2 = n2/2 complex entries per packed line
  int64_t nspec;       // n0*n1*m2
  int64_t ng;          // n0*n1*n2
  double inv_ng;       // 1/ng
};

}  // namespace glia
