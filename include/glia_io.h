/*
 * glia_io.h -- host-side input of the reaction-diffusion path: NetCDF classic fields, segmentation
 * labels, config files.
 *
 * These replace the reference's IO layer for the RD path (src/utils/IO.cpp) with plain C entry
 * points in libglia_rd.so.  They use no GPU and no external library.  Return value: 0 on
 * success, negative GLIA_RD_ERR_* codes otherwise (glia_rd.h).
 */
#ifndef GLIA_IO_H_
#define GLIA_IO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NetCDF classic formats (CDF-1 and CDF-2; big-endian, non-record variables).
   glia_nc_inquire: dimensions of variable 'var' (the reference reads "data"); ndims <= 8;
   type is the NetCDF external type (1 byte, 3 short, 4 int, 5 float, 6 double). */
int glia_nc_inquire(const char* path, const char* var, int* ndims, int64_t* dims, int* type);
/* dataIn (src/utils/IO.cpp:511-539): the whole variable, row-major (last dimension fastest), as
   FP32 or FP64; count must equal the product of the dimensions */
int glia_nc_read_f32(const char* path, const char* var, float* out, int64_t count);
int glia_nc_read_f64(const char* path, const char* var, double* out, int64_t count);
/* dataOut (src/utils/IO.cpp:541-601): CDF-2 (64-bit offsets) file with dimensions "x", "y", "z",
   one variable 'var' (NC_FLOAT, or NC_DOUBLE when as_double != 0) and the global attribute
   "CDF-5 mode" = 0 the reference writes */
int glia_nc_write3_f32(const char* path, const char* var, const int* n, const float* data, int as_double);

/* splitSegmentation (src/utils/Utils.cpp:592-660): labels[8] = {wm, gm, vt, csf, tc, nec, en, ed}
   (values <= 0: absent).  wm, gm, vt are required; csf, tc, ed may be NULL.  With a tumor map
   (tc non-NULL), the tumor core (tc, or nec together with en) and edema also count as white
   matter and ed receives the edema label; without one, ed is zero -- as in the reference. */
int glia_split_segmentation(const float* seg, int64_t count, const int* labels, float* wm, float* gm, float* vt,
                            float* csf, float* tc, float* ed);

/* The RD-relevant part of a reference config file (parseConfig / setParameter,
   src/utils/IO.cpp:6-257): "key=value" lines, whitespace removed, '#' comments. */
typedef struct glia_rd_file_config {
  int n;                    /* n (cubic grid) */
  int model;                /* model (1 = reaction-diffusion) */
  int nt;                   /* nt_inv */
  double dt;                /* dt_inv */
  double k;                 /* init_k */
  double rho;               /* init_rho */
  double k_gm_wm, r_gm_wm;  /* k_gm_wm, r_gm_wm */
  double beta;              /* beta_p */
  double smoothing_factor;  /* smoothing_factor */
  double obs_threshold_1;   /* obs_threshold_1 */
  int np;                   /* number_gaussians */
  double sigma_factor;      /* sigma_factor (sigma = factor * 2 pi / n) */
  double sigma_spacing;     /* sigma_spacing */
  int phi_store, adjoint_store; /* store_phi, store_adjoint */
  int atlas_labels[8];      /* atlas_labels [wm, gm, vt, csf, tc, nec, en, ed]; -1 absent */
  char a_seg_path[512], a_wm_path[512], a_gm_path[512], a_csf_path[512], a_vt_path[512];
  char d1_path[512], obs_mask_path[512];
  double obs_lambda;            /* obs_lambda (< 0: no observation mask from the segmentation) */
  double smoothing_factor_data; /* smoothing_factor_data (d1 smoothed with factor * 2 pi / n) */
  int obs_threshold_rel;        /* obs_threshold_rel (obs_threshold_1 relative to max d1) */
  int patient_labels[8];        /* patient_labels, as atlas_labels; -1 absent */
  char p_seg_path[512];         /* p_seg_path */
  double threshold_data_driven; /* threshold_data_driven (data-driven basis support threshold) */
  double gaussian_volume_fraction; /* gaussian_volume_fraction */
  double sigma_data_driven;     /* phi_sigma_data_driven: 2 pi / 256, or sigma_factor * 2 pi / n */
  char support_data_path[512];  /* support_data_path (NetCDF map the support is taken from) */
  char gaussian_cm_path[512];   /* gaussian_cm_path (phi-mesh file of centers) */
  char data_comp_data_path[512]; /* data_comp_data_path (connected-component file, readConCompDat) */
  char d0_path[512];            /* d0_path (tumor concentration c(0) of the two-time-point data) */
  double smoothing_factor_data_t0; /* smoothing_factor_data_t0 (c(0) smoothed with factor * 2 pi / n) */
  double obs_threshold_0;       /* obs_threshold_0 (relative to max c(0) under obs_threshold_rel) */
} glia_rd_file_config;

/* Defaults of the reference parameter structs for the fields above, then the file's values.
   Unknown keys are ignored; a missing file is GLIA_RD_ERR_ARG. */
int glia_rd_parse_config(const char* path, glia_rd_file_config* out);

/* createEdemaBasedObservationMask (src/utils/Utils.cpp:567-590): the observation operator
   OBS = 1[TC] + lambda 1[B/WT] from the tumor-core and edema maps of a segmentation:
   mask = 1 where tc > 0.99, else 0 where ed > 0.99, else lambda. */
int glia_edema_obs_mask(const float* tc, const float* ed, int64_t count, double lambda, float* mask);

/* Phi::phiMesh (src/mat/Phi.cpp:148-234): the bounding-box basis of Phi::setGaussians(cm, sigma,
   spacing_factor, np) (:33-73).  np = 1 gives cm itself; odd np a cube of h = round(np^(1/3))
   centers per axis spaced spacing_factor * sigma around cm; even np the 8 octants' cubes offset by
   half a spacing (h/2 per side).  Writes the centers (x, y, z, first axis fastest) into centers
   (room for 3 np values) and returns their number. */
int glia_phi_mesh(int np, const double cm[3], double sigma, double spacing_factor, double* centers);
/* readPhiMesh (src/utils/IO.cpp:717-790): a phi-mesh text file ("sigma = s, spacing = d", one line
   skipped, then "x, y, z[, component]" per center up to a line holding "]").  Writes at most
   max_np centers (3 values each) and, when comps is not NULL, their component labels; *np receives
   the number of centers in the file.  sigma and spacing (may be NULL) receive the header values.
   Returns 0, -1 (file not readable) or -2 (more than max_np centers). */
int glia_read_phi_mesh(const char* path, int max_np, double* centers, int* comps, int* np, double* sigma,
                       double* spacing);
/* readConCompDat (src/utils/IO.cpp:929-990): a connected-component file ("#components" line, the
   count, a header line, one "x,y,z" center of mass per component, a header line, one relative mass
   per component).  Writes at most max_comp weights and centers (3 values each; centers may be
   NULL); *ncomp receives the count in the file.  Returns 0, -1 (file not readable) or -2 (more
   than max_comp components). */
int glia_read_concomp(const char* path, int max_comp, double* weights, double* centers, int* ncomp);

#ifdef __cplusplus
}
#endif

#endif
