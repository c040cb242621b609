// rd_launch.h -- host-side launchers of the kernels in rd_kernels.hip (size dispatch included).
// All launchers enqueue on c.stream and return 0 or a hipError_t / -1 (unsupported size).
#pragma once
#include <hip/hip_runtime.h>

#include "rd_device.h"

namespace glia {

struct LaunchCtx {
  Dims d;
  const MemberDev* mem;  // device array of nm member descriptors
  int nm;
  hipStream_t stream;
  const float2* tw0;  // exp(-2 pi i m / n0), m < n0
  const float2* tw1;  // exp(-2 pi i m / n1)
  const float2* twm;  // exp(-2 pi i m / (n2/2))
  const float2* tw2;  // exp(-2 pi i m / n2), m < n2/2
  // variant of the forward divergence passes: split-group kernels (one field per thread group)
  // instead of one group holding every field; chosen per handle (autotuned or by default)
  int yf_split = 0, xf_split = 0;
};

// modes of launch_xf_div and launch_z_point (values match the kernel enums)
enum { XFM_MATVEC = 0, XFM_SETUP = 1, XFM_D = 2 };
enum { ZPM_REACT = 0, ZPM_REACT_ADJ = 1, ZPM_C2R = 2, ZPM_R2C = 3, ZPM_MISFIT = 4, ZPM_OBS2 = 5, ZPM_COEF = 6,
       ZPM_SCREEN = 7, ZPM_C2R_DOT = 8 };
enum { FINK_FILTER = 0, FINK_KSUM = 1, FINK_MISFIT = 2, FINK_C0SQ = 3, FINK_QUAD = 4 };

bool supported_line(int L);
int xf_div_blocks(const Dims& d);
int z_blocks(const Dims& d);
int z_lines_per_block(int m2);
int launch_smooth_axis(const LaunchCtx& c, const float* in, float* out, int axis, const double* taps, int kr);
// per lattice candidate of the data-driven basis support: voxels above thr in its ball (k_support_count)
int launch_support_count(const LaunchCtx& c, const float* f, int step, int l0, int l1, int l2, int R, int d2max,
                         float thr, int* cnt);
int pcg_update_blocks();
int ew_blocks();

int launch_xi_grad(const LaunchCtx& c, int s_in, int s_r, int s_u, int s_v, bool pupd);
int launch_yi_grad(const LaunchCtx& c, int s_u, int s_v, int s_gx, int s_gy, int s_gz, int s_g0);
int launch_yf_div(const LaunchCtx& c, int s_hx, int s_hy, int s_q, int s_yx, int s_rr);
int launch_xf_div(const LaunchCtx& c, int mode, int s_yx, int s_rr, int s_p, int s_out);
int launch_col_plain(const LaunchCtx& c, int dir, int sign, int s_in, int s_out, bool honor_done);
int launch_z_diff(const LaunchCtx& c, int s_gx, int s_gy, int s_gz, int k_slot);
int launch_z_point(const LaunchCtx& c, int mode, int s_in, int s_out, int r_slot, int ch_idx, float dt, bool honor_done);
// Z pass of a forward transform with the elasticity CG's vector update folded in (k_z_upd_r2c; the
// scalars from lane 0's PcgState and sc): ZUM_AXPBY: p = z + (beta / betaold) p, out = R2C(p)
// [ra p, rb z]; ZUM_CGUPD: x += t p, r -= t w with t = beta / sc[0], out = R2C(r) [ra x, rb r, rc p, rd w]
enum { ZUM_AXPBY = 0, ZUM_CGUPD = 1 };
int launch_z_upd_r2c(const LaunchCtx& c, int mode, int ra, int rb, int rc, int rd, const double* sc, int s_out,
                     bool honor_done);
// the elasticity CG's state machine in lane 0's PcgState (k_el_cg_begin / k_el_cg_fin)
int launch_el_cg_begin(const LaunchCtx& c, double beta0, double dp0, double rnorm0, double ttol, double atol,
                       double dtol, int maxit);
int launch_el_cg_fin(const LaunchCtx& c, const double* sc, int* hflag);
int launch_z_gradq(const LaunchCtx& c, int4 sc, int4 sa);

int launch_pcg_begin(const LaunchCtx& c, double tau);
int launch_pcg_update(const LaunchCtx& c, int s_x);
int launch_pcg_fin_setup(const LaunchCtx& c, int nb, int* hflag, int hslot);
int launch_pcg_fin_dpi(const LaunchCtx& c, int nb);
int launch_pcg_fin_update(const LaunchCtx& c, int nb, int* hflag, int hslot);

int launch_copy(const LaunchCtx& c, int s_in, int s_out);  // slots >= CS_NUM address the state pool
int launch_tissue(const LaunchCtx& c);
int launch_coef(const LaunchCtx& c, int r_out, const double* coefs_dev);
int launch_obs_default(const LaunchCtx& c, float thr);
int launch_fin_state(const LaunchCtx& c, int nb, int kind, double w);
int launch_prec_factor(const LaunchCtx& c, double prec_dt);
int launch_state_reset(const LaunchCtx& c, int what);
int launch_phi_apply(const LaunchCtx& c);
int launch_phi_applyT(const LaunchCtx& c, int np_max, int r_in, double scale);
int launch_dot_real(const LaunchCtx& c, int r_x, int r_y);
int launch_set_done(const LaunchCtx& c, int v);
int launch_axpby_real(const LaunchCtx& c, int r_x, double a, int r_y, double b, int r_out);
int launch_axpby_c(const LaunchCtx& c, int s_x, double a, int s_y, double b, int s_out);
int launch_spec_div(const LaunchCtx& c, int s_x, int s_y, int s_z, int s_out);
int launch_spec_lin(const LaunchCtx& c, int nterms, const int* slots, const int* dirs, int s_out);
int launch_el_prec(const LaunchCtx& c, int4 sf, int4 su, double mu, double lam, double scr);
// members as lanes (at most 3, c.nm of them written): output lane l = i sum_t w_{dir[l][t]} X with X the
// slot slot[l][t] of lane lane[l][t] (1 to 4 terms, the same count for every lane of a launch)
int launch_lane_lin(const LaunchCtx& c, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                    const int (*dir)[4], int s_out);
// the inverse X pass of the same combination, formed as the columns load, into complex slot s_out
// (honor_done: lanes whose CG has finished return at once)
int launch_col_lin(const LaunchCtx& c, const int* nterms, const int (*lane)[4], const int (*slot)[4],
                   const int (*dir)[4], int s_out, bool honor_done);
// out[q], q < nq: the Z-pass block partials (nq per block) of the c.nm lanes, summed on the device
int launch_lane_part_sum(const LaunchCtx& c, int nq, double* out);
// the elasticity preconditioner with F_d and U_d in lane d (members 0..2)
int launch_el_prec_lanes(const LaunchCtx& c, int sf, int su, double mu, double lam, double scr, bool honor_done);
int launch_bricks(const LaunchCtx& c, const BrickJob* jobs_host, int njobs, const double* taps_host, int tap_stride,
                  double h0, double h1, double h2, double sigma, float* store_dev, double* max_out_host);

}  // namespace glia
