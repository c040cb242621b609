// sl_capi.cpp -- extern "C" boundary of the semi-Lagrangian solver (include/glia_adv.h).
#include <algorithm>
#include <cmath>
#include <memory>
#include <new>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/glia_adv.h"
#include "rd_solver.h"
#include "sl_launch.h"

using glia::sl::SLDims;
using glia::sl::SLFields;

struct glia_sl_s {
  // spectral engine of one member: its real slots RS_WM, RS_GM, RS_CSF hold the velocity and RS_VT
  // holds div v; its stream is the handle's stream
  std::unique_ptr<glia::RDEnsemble> e;
  SLDims d{};
  hipStream_t st = nullptr;
  float* q[3] = {nullptr, nullptr, nullptr};  // departure offsets
  float* dvi = nullptr;                       // div v at the departure points
  float* work[2] = {nullptr, nullptr};
  bool have_v = false, traj = false, div_ok = false;
  std::string err;
  ~glia_sl_s() {
    for (float* p : {q[0], q[1], q[2], dvi, work[0], work[1]})
      if (p) (void)hipFree(p);
  }
};

namespace {

int fail(glia_sl_s* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}

int hip_fail(glia_sl_s* h, hipError_t e, const char* what) {
  h->err = std::string(what) + ": " + hipGetErrorString(e);
  return (int)e;
}

#define SL_HIP(h, x)                               \
  do {                                             \
    hipError_t e_ = (x);                           \
    if (e_ != hipSuccess) return hip_fail(h, e_, #x); \
  } while (0)

float* vel(glia_sl_s* h, int a) {
  static const int slots[3] = {glia::RS_WM, glia::RS_GM, glia::RS_CSF};
  return h->e->field_ptr(0, slots[a]);
}

int copy_in(glia_sl_s* h, float* dst, const float* src, int kind) {
  const size_t b = sizeof(float) * h->d.ng;
  SL_HIP(h, hipMemcpyAsync(dst, src, b, kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           h->st));
  return 0;
}

int copy_out(glia_sl_s* h, float* dst, const float* src, int kind) {
  const size_t b = sizeof(float) * h->d.ng;
  SL_HIP(h, hipMemcpyAsync(dst, src, b, kind == GLIA_RD_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                           h->st));
  return 0;
}

int finish(glia_sl_s* h, int kind) {
  if (kind == GLIA_RD_HOST) SL_HIP(h, hipStreamSynchronize(h->st));
  return 0;
}

int trajectories(glia_sl_s* h, double dt) {
  if (!h->have_v) return fail(h, GLIA_RD_ERR_STATE, "no velocity set");
  const double s0 = dt * h->d.n0 / (2.0 * M_PI), s1 = dt * h->d.n1 / (2.0 * M_PI), s2 = dt * h->d.n2 / (2.0 * M_PI);
  const int rc = glia::sl::launch_sl_trajectories(h->d, h->st, vel(h, 0), vel(h, 1), vel(h, 2), (float)s0,
                                                  (float)s1, (float)s2, h->q[0], h->q[1], h->q[2]);
  if (rc) return hip_fail(h, (hipError_t)rc, "trajectories launch");
  h->traj = true;
  h->div_ok = false;  // div v at the departure points follows the trajectories
  return 0;
}

int apply(glia_sl_s* h, const SLFields& F, double dt) {
  const int rc = glia::sl::launch_sl_apply(h->d, h->st, F, h->q[0], h->q[1], h->q[2],
                                           h->e->field_ptr(0, glia::RS_VT), h->dvi, (float)dt);
  if (rc) return hip_fail(h, (hipError_t)rc, "advection launch");
  return 0;
}

SLFields one_field(const float* in, float* out, int mode) {
  SLFields F{};
  F.in[0] = in;
  F.out[0] = out;
  F.mode[0] = mode;
  F.n = 1;
  F.any_mass = mode == GLIA_SL_MASS_CONSERVATION;
  return F;
}

// div v on the grid (spectral, computeDivergence) and at the departure points
int ensure_div(glia_sl_s* h) {
  if (h->div_ok) return 0;
  int rc = h->e->divergence(glia::RS_WM, glia::RS_GM, glia::RS_CSF, glia::RS_VT);
  if (rc) return fail(h, rc, h->e->error());
  if ((rc = apply(h, one_field(h->e->field_ptr(0, glia::RS_VT), h->dvi, GLIA_SL_PURE_ADVECTION), 0.0))) return rc;
  h->div_ok = true;
  return 0;
}

bool valid_kind(int kind) { return kind == GLIA_RD_HOST || kind == GLIA_RD_DEVICE; }

}  // namespace

extern "C" {

int glia_sl_create(const int n[3], void* stream, glia_sl_t* out) {
  if (!n || !out) return GLIA_RD_ERR_ARG;
  *out = nullptr;
  glia::SolverConfig sc;
  for (int i = 0; i < 3; i++) sc.n[i] = n[i];
  sc.nmembers = 1;
  sc.nt = 1;
  sc.stream = (hipStream_t)stream;
  glia_sl_s* h = new (std::nothrow) glia_sl_s();
  if (!h) return GLIA_RD_ERR_ARG;
  h->e.reset(new (std::nothrow) glia::RDEnsemble(sc));
  if (!h->e) {
    delete h;
    return GLIA_RD_ERR_ARG;
  }
  int rc = h->e->init();
  if (rc) {
    delete h;
    return rc;
  }
  h->d.n0 = n[0];
  h->d.n1 = n[1];
  h->d.n2 = n[2];
  h->d.ng = (int64_t)n[0] * n[1] * n[2];
  if (h->d.ng * 4 >= ((int64_t)1 << 31)) {  // buffer offsets of the interpolation are 32-bit signed
    delete h;
    return GLIA_RD_ERR_UNSUPPORTED;
  }
  h->st = h->e->stream();
  const size_t b = sizeof(float) * h->d.ng;
  for (float** p : {&h->q[0], &h->q[1], &h->q[2], &h->dvi, &h->work[0], &h->work[1]}) {
    const hipError_t e = hipMalloc((void**)p, b);
    if (e != hipSuccess) {
      delete h;
      return (int)e;
    }
  }
  *out = h;
  return 0;
}

int glia_sl_destroy(glia_sl_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  (void)hipStreamSynchronize(h->st);
  delete h;
  return 0;
}

const char* glia_sl_last_error(glia_sl_t h) { return h ? h->err.c_str() : "null handle"; }

int glia_sl_sync(glia_sl_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  SL_HIP(h, hipStreamSynchronize(h->st));
  return 0;
}

int glia_sl_set_velocity(glia_sl_t h, const float* vx, const float* vy, const float* vz, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!vx || !vy || !vz || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null velocity or bad mem_kind");
  const float* v[3] = {vx, vy, vz};
  for (int a = 0; a < 3; a++) {
    const int rc = copy_in(h, vel(h, a), v[a], mem_kind);
    if (rc) return rc;
  }
  h->have_v = true;
  h->traj = false;
  h->div_ok = false;
  return finish(h, mem_kind);
}

int glia_sl_invalidate(glia_sl_t h) {
  if (!h) return GLIA_RD_ERR_ARG;
  h->traj = false;
  return 0;
}

int glia_sl_compute_trajectories(glia_sl_t h, double dt) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!std::isfinite(dt)) return fail(h, GLIA_RD_ERR_ARG, "dt must be finite");
  return trajectories(h, dt);
}

int glia_sl_get_offsets(glia_sl_t h, float* qx, float* qy, float* qz, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!qx || !qy || !qz || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null output or bad mem_kind");
  if (!h->traj) return fail(h, GLIA_RD_ERR_STATE, "trajectories not computed");
  float* o[3] = {qx, qy, qz};
  for (int a = 0; a < 3; a++) {
    const int rc = copy_out(h, o[a], h->q[a], mem_kind);
    if (rc) return rc;
  }
  return finish(h, mem_kind);
}

int glia_sl_interpolate(glia_sl_t h, const float* in, float* out, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!in || !out || in == out || !valid_kind(mem_kind))
    return fail(h, GLIA_RD_ERR_ARG, "null or aliasing buffers, or bad mem_kind");
  if (!h->traj) return fail(h, GLIA_RD_ERR_STATE, "trajectories not computed");
  int rc;
  if (mem_kind == GLIA_RD_DEVICE) return apply(h, one_field(in, out, GLIA_SL_PURE_ADVECTION), 0.0);
  if ((rc = copy_in(h, h->work[0], in, mem_kind))) return rc;
  if ((rc = apply(h, one_field(h->work[0], h->work[1], GLIA_SL_PURE_ADVECTION), 0.0))) return rc;
  if ((rc = copy_out(h, out, h->work[1], mem_kind))) return rc;
  return finish(h, mem_kind);
}

int glia_sl_solve(glia_sl_t h, float* scalar, double dt, int mode, int mem_kind) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (!scalar || !valid_kind(mem_kind)) return fail(h, GLIA_RD_ERR_ARG, "null field or bad mem_kind");
  if (mode != GLIA_SL_MASS_CONSERVATION && mode != GLIA_SL_PURE_ADVECTION)
    return fail(h, GLIA_RD_ERR_ARG, "advection mode must be 1 (mass conservation) or 2 (pure advection)");
  if (!std::isfinite(dt)) return fail(h, GLIA_RD_ERR_ARG, "dt must be finite");
  int rc;
  if (!h->traj && (rc = trajectories(h, dt))) return rc;
  if (mode == GLIA_SL_MASS_CONSERVATION && (rc = ensure_div(h))) return rc;
  const float* src = scalar;
  if (mem_kind == GLIA_RD_HOST) {
    if ((rc = copy_in(h, h->work[0], scalar, mem_kind))) return rc;
    src = h->work[0];
  }
  if ((rc = apply(h, one_field(src, h->work[1], mode), dt))) return rc;
  if ((rc = copy_out(h, scalar, h->work[1], mem_kind))) return rc;
  return finish(h, mem_kind);
}

int glia_sl_advect(glia_sl_t h, int nfields, const float* const* in, float* const* out, const int* modes, double dt) {
  if (!h) return GLIA_RD_ERR_ARG;
  if (nfields < 0 || (nfields > 0 && (!in || !out || !modes))) return fail(h, GLIA_RD_ERR_ARG, "bad field lists");
  if (!std::isfinite(dt)) return fail(h, GLIA_RD_ERR_ARG, "dt must be finite");
  bool mass = false;
  for (int f = 0; f < nfields; f++) {
    if (!in[f] || !out[f]) return fail(h, GLIA_RD_ERR_ARG, "null field");
    if (modes[f] != GLIA_SL_MASS_CONSERVATION && modes[f] != GLIA_SL_PURE_ADVECTION)
      return fail(h, GLIA_RD_ERR_ARG, "advection mode must be 1 or 2");
    for (int g = 0; g < nfields; g++)
      if ((const float*)out[f] == in[g]) return fail(h, GLIA_RD_ERR_ARG, "an output aliases an input");
    mass = mass || modes[f] == GLIA_SL_MASS_CONSERVATION;
  }
  if (nfields == 0) return 0;
  int rc;
  if (!h->traj && (rc = trajectories(h, dt))) return rc;
  if (mass && (rc = ensure_div(h))) return rc;
  for (int f0 = 0; f0 < nfields; f0 += glia::sl::SL_MAX_FIELDS) {
    SLFields F{};
    F.n = std::min(glia::sl::SL_MAX_FIELDS, nfields - f0);
    for (int f = 0; f < F.n; f++) {
      F.in[f] = in[f0 + f];
      F.out[f] = out[f0 + f];
      F.mode[f] = modes[f0 + f];
      F.any_mass |= F.mode[f] == GLIA_SL_MASS_CONSERVATION;
    }
    if ((rc = apply(h, F, dt))) return rc;
  }
  return 0;
}

}  // extern "C"
