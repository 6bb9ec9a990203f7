/*
 * ntt_oracle.h -- C ABI of the CPU oracle (libntt_oracle.so).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain, sequential CPU restatement of
 * Entity's per-timestep SR PIC hot path and of its GR pusher, written from the
 * reference sources cited next to every function in ntt_oracle.cpp, in the
 * reference's own operation order (built with -ffp-contract=off).  It exists
 * to check the HIP library and to time a CPU baseline; nothing in the product
 * path may call it.
 *
 * Every function exists in a single- and a double-precision build (suffix
 * _f32 / _f64), mirroring the reference's `precision` CMake option.  Real
 * arrays are passed as void* and hold float (f32) or double (f64); in-cell
 * offsets dx* are float in both builds (prtldx_t, src/global/global.h:347).
 * All pointers are host pointers.
 */
#ifndef NTT_ORACLE_H
#define NTT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ort_particles {
  int32_t *i1, *i2, *i3;
  int32_t *i1_prev, *i2_prev, *i3_prev;
  float   *dx1, *dx2, *dx3;
  float   *dx1_prev, *dx2_prev, *dx3_prev;
  void    *ux1, *ux2, *ux3; /* real */
  void    *weight;          /* real */
  void    *phi;             /* real, GR 2D only */
  int16_t *tag;
  int64_t  npart;
} ort_particles_t;

typedef struct ort_field {
  void*   data; /* real; LayoutLeft, component slowest */
  int32_t dim;
  int32_t n[3];
  int32_t ngh;
  int32_t ncomp;
} ort_field_t;

typedef struct ort_sr_params {
  float   mass, charge;      /* species (PusherContext, context.h:82)        */
  double  dt, omegaB0;       /* global constants (context.h:88)              */
  double  dx;                /* Minkowski cell size                          */
  int32_t ni1, ni2, ni3;
  uint8_t pusher_flags;      /* NTT_PUSHER_* values of include/ntt.h         */
  uint8_t prtl_bc[3][2];     /* NTT_PRTL_BC_* values                         */
  int32_t shape_order;       /* 0, 1, 2                                      */
} ort_sr_params_t;

typedef struct ort_fd_coeffs {
  double delta_x, delta_y, delta_z;
  double beta_xy, beta_yx, beta_xz, beta_zx, beta_yz, beta_zy;
} ort_fd_coeffs_t;

typedef struct ort_species {
  ort_particles_t prtls;
  float           mass, charge;
  uint8_t         pusher_flags;
} ort_species_t;

typedef struct ort_step_params {
  double  dt, correction, omegaB0;
  double  q0, B0, V0, ppc0;
  double  dx;
  int32_t shape_order;
  int32_t n_filters;
  int32_t fieldsolver_enable, deposit_enable;
  int64_t step;            /* step counter: step 0 fills ghosts first        */
  ort_fd_coeffs_t fd;
} ort_step_params_t;

/* GR pusher parameters (kernel::gr::PusherContext, context.h:182-211, plus the
 * QKerrSchild metric parameters) */
typedef struct ort_gr_params {
  float   mass, charge;
  double  dt, omegaB0;
  int32_t niter;
  int32_t ni1, ni2;
  int32_t massless;         /* 1: photon pusher (Massless_t)                  */
  int32_t absorb_i1min, absorb_i1max, axis_i2min, axis_i2max;
  int32_t metric_kind;      /* 0 = KerrSchild, 1 = QKerrSchild, 2 = KerrSchild0 */
  double  a;                /* spin                                          */
  double  r0, h;            /* QKS parameters                                */
  double  x1_min, x1_max, x2_min, x2_max; /* physical extent (r, theta)     */
  int32_t nx1, nx2;
} ort_gr_params_t;

/* GR field boundaries and the spherical current filter
 * (ntt_oracle_gr_bc_*): arguments of one boundary operation */
typedef struct ort_gr_bc_args {
  double   xg_edge, ds;     /* match / absorb: edge radius and layer width   */
  int32_t  i1_lo, i1_hi;    /* match / absorb: x1 cell range [lo, hi), ghost-
                               extended indices                              */
  uint32_t tags;            /* bits 0-2: D (E), bits 3-5: B (H)              */
  uint32_t comps;           /* match: components the target defines, bit c
                               for em component c                            */
  int32_t  nfilter;         /* horizon: number of current filter passes      */
  int32_t  upper;           /* axis: 1 for the theta = pi side               */
  int32_t  axis_min, axis_max;
} ort_gr_bc_args_t;

#define ORT_DECL(SUF, R)                                                        \
  int ntt_oracle_sr_push_##SUF(const ort_sr_params_t* prm,                       \
                               ort_particles_t* prtls, const ort_field_t* em);   \
  int ntt_oracle_interp_##SUF(const ort_sr_params_t* prm,                        \
                              const ort_particles_t* prtls,                      \
                              const ort_field_t* em, R* e_out, R* b_out);        \
  int ntt_oracle_deposit_##SUF(const ort_particles_t* prtls, ort_field_t* cur,   \
                               double charge, double dt, double dx,              \
                               int32_t shape_order);                             \
  int ntt_oracle_faraday_##SUF(ort_field_t* em, double coeff1, double coeff2,    \
                               const ort_fd_coeffs_t* fd);                       \
  int ntt_oracle_ampere_##SUF(ort_field_t* em, double coeff1, double coeff2);    \
  int ntt_oracle_currents_ampere_##SUF(ort_field_t* em, ort_field_t* cur,        \
                                       double coeff, double ppc0);               \
  int ntt_oracle_filter_##SUF(ort_field_t* array, const ort_field_t* buffer);    \
  int ntt_oracle_communicate_##SUF(ort_field_t* fld, int32_t c0, int32_t c1);    \
  int ntt_oracle_synchronize_##SUF(ort_field_t* cur, ort_field_t* buff);         \
  int ntt_oracle_sr_step_##SUF(ort_field_t* em, ort_field_t* cur,                \
                               ort_field_t* buff, ort_species_t* species,        \
                               int32_t nspecies, const ort_step_params_t* prm,   \
                               const uint8_t prtl_bc[3][2]);                     \
  int ntt_oracle_gr_push_##SUF(const ort_gr_params_t* prm, ort_particles_t* prtls, \
                               const ort_field_t* em, const ort_field_t* em0,    \
                               int32_t shape_order);                             \
  int ntt_oracle_gr_metric_##SUF(const ort_gr_params_t* prm, int64_t n,          \
                                 const double* x1, const double* x2,             \
                                 double* out);                                   \
  int ntt_oracle_gr_transform_##SUF(const ort_gr_params_t* prm, int32_t which,   \
                                    int64_t n, const double* x1,                 \
                                    const double* x2, const double* vin,         \
                                    double* vout);

ORT_DECL(f32, float)
ORT_DECL(f64, double)

#undef ORT_DECL

#ifdef __cplusplus
}
#endif

#endif /* NTT_ORACLE_H */
