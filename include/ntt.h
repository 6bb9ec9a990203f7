/*
 * ntt.h -- C ABI of the MI355X-native Entity PIC hot path (libntt_hip.so).
 *
 * Every entry point is stream-ordered and asynchronous: it validates its
 * arguments on the host, enqueues HIP kernels on `stream` (a hipStream_t, NULL
 * for the null stream) and returns.  Device-side faults that the reference
 * turns into `raise::KernelError` (invalid particle tag, a stencil that leaves
 * the ghost-extended array) set bits of a device error word instead of
 * aborting; read and clear it with ntt_get_device_errors().
 *
 * Pointers inside the structs below are caller-owned DEVICE pointers.  The
 * library never allocates: scratch memory is passed in explicitly.
 *
 * Memory image (the reference's GPU layout, Kokkos LayoutLeft):
 *   field(i1, i2, i3, c) = data[i1 + N1 * (i2 + N2 * (i3 + N3 * c))]
 *   with N_d = n[d] + 2 * ngh for d < dim and N_d = 1 otherwise; active cells
 *   are [ngh, ngh + n[d]) in every simulated direction.
 *   em components: 0..2 = ex1..ex3 (Yee edges), 3..5 = bx1..bx3 (Yee faces);
 *   cur components: 0..2 = jx1..jx3.
 * Particles are a structure of arrays with an active prefix [0, npart).
 *
 * Reference interfaces replaced (paths relative to the Entity source tree):
 *   ntt_particles_t         ntt::ParticleArrays / Particles<D,C>
 *                           src/framework/containers/particles.h:47-71,78-320
 *   ntt_field_t             ndfield_t<D,N> (Kokkos::View<real_t*** [N]>)
 *                           src/global/arch/kokkos_aliases.h:85-109,
 *                           src/framework/containers/fields.h:54-82
 *   ntt_sr_pusher_ctx_t     kernel::sr::PusherContext
 *                           src/kernels/pushers/context.h:74-122
 *   ntt_prtl_bc_t           kernel::sr::PusherBoundaries<D>
 *                           src/kernels/pushers/context.h:124-178
 *   ntt_minkowski_t         metric::Minkowski<D> (dx, dx_inv)
 *                           src/metrics/minkowski.h:30-84
 */
#ifndef NTT_H
#define NTT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NTT_ABI_VERSION 1

#if defined(__GNUC__) || defined(__clang__)
#define NTT_API __attribute__((visibility("default")))
#else
#define NTT_API
#endif

typedef enum ntt_status {
  NTT_OK              = 0,
  NTT_ERR_ARG         = 1, /* invalid argument (null pointer, bad size, ...)  */
  NTT_ERR_HIP         = 2, /* a HIP runtime call failed                       */
  NTT_ERR_DEVICE      = 3, /* device error word is non-zero                   */
  NTT_ERR_UNSUPPORTED = 4, /* valid request outside what the library supports */
  NTT_ERR_CAPACITY    = 5  /* workspace or capacity too small                 */
} ntt_status;

/* device error word bits (ntt_get_device_errors) */
#define NTT_DEVERR_INVALID_TAG   0x1u /* tag neither alive nor dead (sr.hpp:118-122) */
#define NTT_DEVERR_OUT_OF_BOUNDS 0x2u /* gather/deposit stencil outside the array     */
#define NTT_DEVERR_PUSHER_FLAGS  0x4u /* invalid pusher algorithm flags               */
#define NTT_DEVERR_CURRENT_RANGE 0x8u /* a particle's current exceeds the fixed-point
                                         range (|weight| far above weight_max)     */

/* ParticleTag (src/global/global.h:168-171) */
#define NTT_TAG_DEAD  0
#define NTT_TAG_ALIVE 1

/* ParticlePusher flags (src/global/enums.h:317-324) */
#define NTT_PUSHER_NONE   0u
#define NTT_PUSHER_PHOTON (1u << 0)
#define NTT_PUSHER_BORIS  (1u << 1)
#define NTT_PUSHER_VAY    (1u << 2)
#define NTT_PUSHER_GCA    (1u << 3)

/* PrtlBC values (src/global/enums.h:185-195) */
#define NTT_PRTL_BC_PERIODIC   1u
#define NTT_PRTL_BC_ABSORB     2u
#define NTT_PRTL_BC_ATMOSPHERE 3u
#define NTT_PRTL_BC_REFLECT    5u
#define NTT_PRTL_BC_HORIZON    6u
#define NTT_PRTL_BC_AXIS       7u

/* FldsBC values (src/global/enums.h:215-226) */
#define NTT_FLDS_BC_PERIODIC  1u
#define NTT_FLDS_BC_MATCH     2u
#define NTT_FLDS_BC_FIXED     3u
#define NTT_FLDS_BC_HORIZON   6u
#define NTT_FLDS_BC_AXIS      7u
#define NTT_FLDS_BC_CONDUCTOR 8u

typedef struct ntt_particles {
  int32_t* i1;
  int32_t* i2;
  int32_t* i3;
  int32_t* i1_prev;
  int32_t* i2_prev;
  int32_t* i3_prev;
  float*   dx1;
  float*   dx2;
  float*   dx3;
  float*   dx1_prev;
  float*   dx2_prev;
  float*   dx3_prev;
  float*   ux1;
  float*   ux2;
  float*   ux3;
  float*   weight;
  float*   phi;   /* 2D axisymmetric (GR) only; may be NULL otherwise       */
  int16_t* tag;
  int64_t* pld_i; /* optional integer payload carried through sorts, or NULL */
  int64_t  npart;
  int64_t  maxnpart;
} ntt_particles_t;

typedef struct ntt_field {
  float*  data;
  int32_t dim;   /* 2 or 3                                         */
  int32_t n[3];  /* active cells; n[d] = 1 for d >= dim            */
  int32_t ngh;   /* ghost layers per side (2 for shape order <= 2) */
  int32_t ncomp; /* 6 (em) or 3 (cur)                              */
} ntt_field_t;

typedef struct ntt_sr_pusher_ctx {
  float   mass;
  float   charge;
  float   dt;
  float   omegaB0;
  int32_t ni1, ni2, ni3;
  uint8_t pusher_flags; /* NTT_PUSHER_* */
} ntt_sr_pusher_ctx_t;

typedef struct ntt_prtl_bc {
  uint8_t bc[3][2]; /* [direction][0 = min, 1 = max], NTT_PRTL_BC_* */
} ntt_prtl_bc_t;

typedef struct ntt_flds_bc {
  uint8_t bc[3][2]; /* NTT_FLDS_BC_*; the library filters the periodic interior */
} ntt_flds_bc_t;

typedef struct ntt_minkowski {
  int32_t dim;
  float   dx; /* uniform cell size: (x1_max - x1_min) / nx1 */
} ntt_minkowski_t;

/* Extended-stencil Faraday coefficients (algorithms.fieldsolver.*,
 * src/global/defaults.h:29-41; all zero = standard Yee). */
typedef struct ntt_fd_coeffs {
  float delta_x, delta_y, delta_z;
  float beta_xy, beta_yx, beta_xz, beta_zx, beta_yz, beta_zy;
} ntt_fd_coeffs_t;

/* ------------------------------------------------------------------ misc */

/* library ABI version (NTT_ABI_VERSION) */
NTT_API int32_t ntt_abi_version(void);

/* human-readable name of a status code */
NTT_API const char* ntt_status_string(ntt_status status);

/* Waits for the device, returns the error word (NTT_DEVERR_*) of the current
 * device in *flags and clears it.  No reference counterpart: it replaces the
 * Kokkos::abort of raise::KernelError (src/global/utils/error.h:101-107). */
NTT_API ntt_status ntt_get_device_errors(uint32_t* flags);

/* --------------------------------------------------- kernel-level entries */

/* One launch of the SR pusher over [0, npart).
 * Replaces kernel::sr::Pusher_kernel<Minkowski<D>, PusherPolicy<M>> with the
 * no-policy path (src/kernels/pushers/sr.hpp:63-332; launch site
 * src/engines/srpic/particle_pusher.h:152-160).  shape_order is the
 * compile-time SHAPE_ORDER of the reference (0 = linear "direct"
 * interpolation, 1, 2 = B-spline shapes). */
NTT_API ntt_status ntt_sr_push(const ntt_sr_pusher_ctx_t* ctx,
                       const ntt_prtl_bc_t*       bc,
                       const ntt_particles_t*     prtls,
                       const ntt_field_t*         em,
                       const ntt_minkowski_t*     metric,
                       int32_t                    shape_order,
                       void*                      stream);

/* ------------------------------------------------------------ GR pusher */

/* Metrics of the Kerr-Schild family on a 2D axisymmetric (r, theta) grid:
 * metric::KerrSchild (src/metrics/kerr_schild.h), metric::QKerrSchild
 * (qkerr_schild.h: r = r0 + exp(chi), theta(eta) stretched by h) and
 * metric::KerrSchild0 (kerr_schild_0.h: a = 0, flat). */
#define NTT_METRIC_KERR_SCHILD   0
#define NTT_METRIC_QKERR_SCHILD  1
#define NTT_METRIC_KERR_SCHILD_0 2

typedef struct ntt_ks_metric {
  int32_t kind;           /* NTT_METRIC_*                                   */
  int32_t nx1, nx2;       /* active cells in r and theta                    */
  double  x1_min, x1_max; /* r extent [r_g]                                 */
  double  x2_min, x2_max; /* theta extent [rad]                             */
  double  a;              /* spin (ignored by KerrSchild0)                  */
  double  r0, h;          /* quasi-spherical parameters (QKerrSchild)       */
} ntt_ks_metric_t;

/* kernel::gr::PusherContext (src/kernels/pushers/context.h:182-211). */
typedef struct ntt_gr_pusher_ctx {
  float   mass, charge;
  float   dt, omegaB0;
  float   epsilon;        /* algorithms.gr.pusher_eps (unused by the pusher) */
  int32_t niter;          /* algorithms.gr.pusher_niter (default 10)        */
  int32_t ni1, ni2, ni3;
  uint8_t pusher_flags;   /* NTT_PUSHER_PHOTON (massless) or NTT_PUSHER_BORIS */
  uint8_t precision;      /* 64 (or 0): metric and pusher in FP64 on FP32
                             storage; 32: the reference's FP32 arithmetic   */
} ntt_gr_pusher_ctx_t;

/* GR particle push of one species (2D): photons along geodesics, massive
 * particles by EM half push -> geodesic momentum push -> EM half push ->
 * geodesic coordinate push, with niter fixed-point iterations each.  D is
 * read from em, B from em0 (contravariant components, ghost-extended; both
 * may be NULL for photons).  Boundaries: ABSORB or HORIZON at r kills the
 * particle, AXIS at theta reflects it.
 * Replaces kernel::gr::Pusher_kernel<M> (src/kernels/pushers/gr.hpp:71-865)
 * as launched by grpic::ParticlePush (src/engines/grpic/particle_pusher.h:26-89). */
NTT_API ntt_status ntt_gr_push(const ntt_gr_pusher_ctx_t* ctx,
                               const ntt_prtl_bc_t*       bc,
                               ntt_particles_t*           prtls,
                               const ntt_field_t*         em,
                               const ntt_field_t*         em0,
                               const ntt_ks_metric_t*     metric,
                               int32_t                    shape_order,
                               void*                      stream);

/* One species of one run in a batched GR push: the arguments ntt_gr_push
 * takes, by value (pointers inside are device pointers). */
typedef struct ntt_gr_batch_entry {
  ntt_gr_pusher_ctx_t ctx;
  ntt_prtl_bc_t       bc;
  ntt_particles_t     prtls;
  ntt_field_t         em;   /* D (ignored for photons)        */
  ntt_field_t         em0;  /* B (ignored for photons)        */
  ntt_ks_metric_t     metric;
} ntt_gr_batch_entry_t;

typedef struct ntt_gr_batch* ntt_gr_batch_t;

/* Plans the GR push of many independent runs (each a box with its own metric,
 * spin, grid, fields and species) as one launch per (precision, metric kind,
 * massless/massive) group -- a single launch when all entries share them.
 * Every entry is validated as ntt_gr_push validates it; the results are
 * bit-identical to ntt_gr_push of each entry.  The plan records the entries'
 * device pointers and particle counts (which bound the launch), and
 * allocates device memory for its tables.  No reference counterpart: the
 * reference runs such scans as separate simulations, each launching
 * kernel::gr::Pusher_kernel per species (src/engines/grpic/particle_pusher.h:26-89). */
NTT_API ntt_status ntt_gr_batch_create(const ntt_gr_batch_entry_t* entries, int32_t nentries,
                                       int32_t shape_order, ntt_gr_batch_t* batch);

/* One push of every entry of the plan.  Enqueues kernels only (no
 * allocation, copy or synchronisation): graph-capturable. */
NTT_API ntt_status ntt_gr_batch_push(ntt_gr_batch_t batch, void* stream);

/* Sets the particle count of one entry (e.g. after dead particles were
 * compacted away), at most its count at planning; stream-ordered. */
NTT_API ntt_status ntt_gr_batch_set_npart(ntt_gr_batch_t batch, int32_t entry, int64_t npart,
                                          void* stream);

NTT_API ntt_status ntt_gr_batch_destroy(ntt_gr_batch_t batch);

/* ------------------------------------------------ GR field solver (2D) */
/* Fields of the GR step on the (r, theta) grid of a Kerr-Schild metric, each
 * 6 components with ghosts: em/em0 hold D (0..2) and B (3..5), aux holds E
 * (0..2) and H (3..5), all contravariant code components except the
 * covariant auxiliary E and H; cur/cur0 hold J (3 components).  bc gives the
 * field boundaries (NTT_FLDS_BC_AXIS at theta flags the polar rows); it may
 * be NULL (no axis).  The metric is evaluated in FP64, the fields are FP32.
 * The launch ranges are the reference drivers'
 * (src/engines/grpic/fieldsolvers.h:53-266). */

/* E = alpha D_cov + beta x B over active cells plus one cell below in r (and
 * one row past the last theta row on an axis).  Replaces
 * kernel::gr::ComputeAuxE_kernel (src/kernels/aux_fields_gr.hpp:31-125). */
NTT_API ntt_status ntt_gr_aux_e(const ntt_field_t* D, const ntt_field_t* B, ntt_field_t* E,
                                const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc,
                                void* stream);

/* H = alpha B_cov - beta x D over the same range as ntt_gr_aux_e (D and B
 * may be the same array).  Replaces kernel::gr::ComputeAuxH_kernel
 * (src/kernels/aux_fields_gr.hpp:136-234). */
NTT_API ntt_status ntt_gr_aux_h(const ntt_field_t* D, const ntt_field_t* B, ntt_field_t* H,
                                const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc,
                                void* stream);

/* Bout = Bin - coeff curl E / sqrt(det h) over active cells (Bin and Bout
 * may be the same array; B^theta = 0 on an axis row).  Replaces
 * kernel::gr::Faraday_kernel (src/kernels/faraday_gr.hpp:32-106). */
NTT_API ntt_status ntt_gr_faraday(const ntt_field_t* Bin, ntt_field_t* Bout, const ntt_field_t* E,
                                  const ntt_ks_metric_t* metric, float coeff,
                                  const ntt_flds_bc_t* bc, void* stream);

/* Dout = Din + coeff curl H / sqrt(det h) over active cells plus one row
 * past the last theta row (Din and Dout may be the same array; on an axis
 * row D^r takes the value of the row next to it, read before the update).
 * Replaces kernel::gr::Ampere_kernel (src/kernels/ampere_gr.hpp:33-100). */
NTT_API ntt_status ntt_gr_ampere(const ntt_field_t* Din, ntt_field_t* Dout, const ntt_field_t* H,
                                 const ntt_ks_metric_t* metric, float coeff,
                                 const ntt_flds_bc_t* bc, void* stream);

/* D += coeff J / sqrt(det h) (the polar area on an axis row), over the
 * range of ntt_gr_ampere.  Replaces kernel::gr::CurrentsAmpere_kernel
 * (src/kernels/ampere_gr.hpp:105-178). */
NTT_API ntt_status ntt_gr_currents_ampere(ntt_field_t* D, const ntt_field_t* J,
                                          const ntt_ks_metric_t* metric, float coeff,
                                          const ntt_flds_bc_t* bc, void* stream);

/* BD0 = (BD0 + BD) / 2 (6 components) and J = (J + J0) / 2 (3 components)
 * over active cells.  Replace kernel::gr::TimeAverageDB_kernel and
 * TimeAverageJ_kernel (src/kernels/aux_fields_gr.hpp:240-303). */
NTT_API ntt_status ntt_gr_time_average_db(const ntt_field_t* BD, ntt_field_t* BD0, void* stream);
NTT_API ntt_status ntt_gr_time_average_j(ntt_field_t* J, const ntt_field_t* J0, void* stream);

/* ---- GR field boundaries, current filter and deposit (2D r-theta) -------- */

/* BC tags (src/global/global.h:306-329): D shares the bits of E, H those of B. */
#define NTT_BC_D 0x7u
#define NTT_BC_B 0x38u

/* One match layer at an r boundary: its cell range, as
 * grpic::MatchFieldsIn computes it with Mesh::ExtentToRange
 * (src/engines/grpic/fields_bcs.h:41-94, src/framework/domain/mesh.h:133-203),
 * the edge radius and the layer width (grid.boundaries.match.ds). */
typedef struct ntt_gr_match {
  double   r_edge, ds;
  int32_t  i1_lo, i1_hi;  /* x1 cells [lo, hi), ghost-extended indices          */
  uint32_t tags;          /* NTT_BC_D and/or NTT_BC_B                           */
  uint32_t comps;         /* components `target` defines: bit c for em component c */
} ntt_gr_match_t;

/* Blends em towards `target` over the layer, every x2 row (ghosts
 * included): F = s F + (1 - s) T, s = tanh(4 |r - r_edge| / ds) at each
 * component's own x1 position; D^3 on and past an axis row is only damped.
 * `target` holds the problem's field at every component's staggered point
 * (the values the reference's field setter returns there).  Replaces
 * kernel::bc::MatchBoundaries_kernel<GRPIC, M, FS, in::x1>
 * (src/kernels/fields_bcs.hpp:176-345). */
NTT_API ntt_status ntt_gr_match_boundaries(ntt_field_t* em, const ntt_field_t* target,
                                           const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc,
                                           const ntt_gr_match_t* match, void* stream);

/* Inner (horizon) boundary: x1 rows 0 .. 2 + nfilter take the values of row
 * 3 + nfilter, for x2 rows [2, 2 + n2] and the tagged components.  Replaces
 * kernel::bc::gr::HorizonBoundaries_kernel (src/kernels/fields_bcs.hpp:
 * 1188-1240). */
NTT_API ntt_status ntt_gr_horizon_boundaries(ntt_field_t* em, uint32_t tags, int32_t nfilter,
                                             void* stream);

/* Axis rows and their ghosts for every x1 column (upper = 0: theta = 0,
 * 1: theta = pi).  Replaces kernel::bc::AxisBoundaries_kernel
 * (src/kernels/fields_bcs.hpp:825-867). */
NTT_API ntt_status ntt_gr_axis_boundaries(ntt_field_t* em, int32_t upper, uint32_t tags,
                                          void* stream);

/* J *= tanh(|r(x1) - r_edge| / (ds / 4)) over x1 cells [i1_lo, i1_hi), every
 * x2 row.  Replaces kernel::bc::gr::AbsorbCurrents_kernel
 * (src/kernels/fields_bcs.hpp:1243-1280). */
NTT_API ntt_status ntt_gr_absorb_currents(ntt_field_t* cur, const ntt_ks_metric_t* metric,
                                          double r_edge, double ds, int32_t i1_lo, int32_t i1_hi,
                                          void* stream);

/* One pass of the current filter of a spherical grid, buff -> cur, over x1
 * [2, 2 + n1) and x2 [2, 2 + n2] (axis rows per bc).  Replaces
 * kernel::DigitalFilter_kernel<2D, non-Cartesian> (src/kernels/
 * digital_filter.hpp:196-266) as grpic::CurrentsFilter launches it
 * (src/engines/grpic/currents.h:80-110). */
NTT_API ntt_status ntt_gr_digital_filter(ntt_field_t* cur, const ntt_field_t* buff,
                                         const ntt_flds_bc_t* bc, void* stream);

/* The GR current deposit (order 1, whatever the pusher's shape order), added
 * atomically into `cur` with its ghosts (summed per workgroup in an LDS
 * window where its particles' stencils fit one); precision 64 (or 0)
 * evaluates the metric in FP64, 32 in FP32.  Replaces kernel::DepositCurrents_kernel<GRPIC,
 * M> (src/kernels/currents_deposit.hpp:108-555) as grpic::CurrentsDeposit
 * launches it (src/engines/grpic/currents.h:34-77). */
NTT_API ntt_status ntt_gr_deposit_currents(const ntt_particles_t* prtls, ntt_field_t* cur,
                                           const ntt_ks_metric_t* metric, float charge, float dt,
                                           int32_t precision, void* stream);

/* Charge-conserving current deposit, accumulated (atomically) into `cur`
 * including its ghost cells, exactly where the reference writes (summed per
 * workgroup of 1,024 consecutive particles in an LDS window when their
 * stencils fit one, as spatially sorted particles do).
 * Replaces kernel::DepositCurrents_kernel<SRPIC, Minkowski<D>, O>
 * (src/kernels/currents_deposit.hpp:33-762): O = 0 zig-zag, 1..2 Esirkepov.
 * The caller zeroes `cur` first, as srpic::CurrentsDeposit does
 * (src/engines/srpic/currents.h:68). */
NTT_API ntt_status ntt_deposit_currents(const ntt_particles_t* prtls,
                                ntt_field_t*           cur,
                                const ntt_minkowski_t* metric,
                                float                  charge,
                                float                  dt,
                                int32_t                shape_order,
                                void*                  stream);

/* B += coeff * curl E over active cells (in place on em).
 * Replaces kernel::mink::Faraday_kernel<D> (src/kernels/faraday_mink.hpp:25-167).
 * fd may be NULL (standard Yee). */
NTT_API ntt_status ntt_faraday_mink(ntt_field_t*           em,
                            float                  coeff1,
                            float                  coeff2,
                            const ntt_fd_coeffs_t* fd,
                            void*                  stream);

/* E += coeff * curl B over active cells.
 * Replaces kernel::mink::Ampere_kernel<D> (src/kernels/ampere_mink.hpp:31-89). */
NTT_API ntt_status ntt_ampere_mink(ntt_field_t* em, float coeff1, float coeff2, void* stream);

/* E += coeff * J, then J /= ppc0, over active cells.
 * Replaces kernel::mink::CurrentsAmpere_kernel<D> (no external current)
 * (src/kernels/ampere_mink.hpp:96-215). */
NTT_API ntt_status ntt_currents_ampere_mink(ntt_field_t* em,
                                    ntt_field_t* cur,
                                    float        coeff,
                                    float        ppc0,
                                    void*        stream);

/* One binomial pass array = K * buffer over active cells (periodic interior
 * stencil; 3D keeps the reference's stencil as written).
 * Replaces kernel::DigitalFilter_kernel<D, Cartesian>
 * (src/kernels/digital_filter.hpp:58-389). bc may be NULL (periodic). */
NTT_API ntt_status ntt_digital_filter(ntt_field_t*         array,
                              const ntt_field_t*   buffer,
                              const ntt_flds_bc_t* bc,
                              void*                stream);

/* Periodic ghost fill (copy the active edges into the ghost layers) of
 * components [comp_begin, comp_end).  Replaces the single-domain periodic path
 * of Metadomain::CommunicateFields (src/framework/domain/metadomain_comm.cpp:204-367,
 * comm_nompi.hpp:28-63). */
NTT_API ntt_status ntt_communicate_fields(ntt_field_t* fld,
                                  int32_t      comp_begin,
                                  int32_t      comp_end,
                                  void*        stream);

/* Periodic fold of ghost-deposited currents into their images (additive),
 * using `buff` as the receive buffer.  Replaces Metadomain::SynchronizeFields
 * for Comm::J (src/framework/domain/metadomain_comm.cpp:369-562). */
NTT_API ntt_status ntt_synchronize_currents(ntt_field_t* cur, ntt_field_t* buff, void* stream);

/* ------------------------------------------------------ engine-level entries */

/* One species of a domain (ntt::Particles + ParticleSpecies mass/charge/pusher,
 * src/framework/containers/species.h:29-32). */
typedef struct ntt_species {
  ntt_particles_t prtls;
  float           mass;
  float           charge;
  uint8_t         pusher_flags; /* NTT_PUSHER_*; NONE skips the species             */
  const uint32_t* tile_offsets; /* device: first particle of every tile (ntiles + 1
                                   entries) from the last ntt_sort_particles, or NULL */
  int64_t         ntiles;
} ntt_species_t;

/* Scalars of one SRPIC step (engine params and SimulationParams the drivers
 * read: src/engines/srpic/fieldsolvers.h:36-200, currents.h:64-119,
 * particle_pusher.h:93-105). */
typedef struct ntt_sr_step_params {
  float           dt;                 /* CFL * dx / sqrt(D) (algorithms.cpp:21-22) */
  float           correction;         /* algorithms.timestep.correction            */
  float           omegaB0;            /* scales.omegaB0 = 1 / larmor0               */
  float           q0, B0, V0, ppc0;   /* scales.q0, scales.B0, scales.V0, ppc0      */
  int32_t         shape_order;        /* SHAPE_ORDER: 0, 1, 2                       */
  int32_t         n_filters;          /* algorithms.current_filters                 */
  int32_t         fieldsolver_enable; /* algorithms.fieldsolver.enable              */
  int32_t         deposit_enable;     /* algorithms.deposit.enable                  */
  int64_t         step;               /* step 0 fills the E/B ghosts first          */
  ntt_fd_coeffs_t fd;                 /* extended Faraday stencil (zero = Yee)      */
  ntt_prtl_bc_t   prtl_bc;            /* must be periodic in every direction        */
  int32_t         particles_per_thread; /* chunk = 256 * this when a species has no
                                           tile offsets (0: 4)                        */
  float           weight_max;         /* bound on |weight| of every particle (0: 1);
                                           sets the fixed-point scale of the current;
                                           heavier particles stay exact, only slower  */
  float*          em_half_out;        /* device, or NULL: receives em (6 components,
                                           ghosts included) as the push reads it,
                                           after the first Faraday half.  No reference
                                           counterpart: for adjoint recording (read by
                                           ntt_sr_step only)                          */
} ntt_sr_step_params_t;

/* One full SRPIC step of a periodic box, in the reference's phase order:
 * Faraday(1/2) -> push + deposit of every species -> current fold -> n
 * filter passes -> Faraday(1/2) -> Ampere -> current source (E += coeff J,
 * J /= ppc0), leaving E, B and J with consistent periodic ghosts.
 * Replaces SRPICEngine::step_forward (src/engines/srpic/srpic.hpp:65-188) for
 * a single periodic domain.  Push and deposit run LDS-windowed (the species
 * arrays should be sorted by spatial tile, see ntt_sort_particles; unsorted
 * input is handled, only slower): fused in one kernel in 2D, and in 3D (shape
 * orders 1 and 2, Boris and Vay, sorted species) as a push launch and a
 * deposit launch when the workspace holds their chunk flags and box records.
 * `buff` is scratch (its contents afterwards are unspecified).  The deposit
 * accumulates the current as 64-bit fixed point at the start of `workspace`
 * (ntt_sr_step_workspace_size bytes; the 3D accumulator alone, 24 bytes per
 * ghost-extended cell, is enough, the fused kernel then runs), in units of
 * 2^-k with 2^k the largest power of two for which every contribution of a
 * particle with |charge * weight| <= max|charge| * weight_max stays below
 * 2^30 units (each contribution is truncated to a unit), so J does not depend
 * on the order of the atomic additions, nor on the path. */
NTT_API ntt_status ntt_sr_step(ntt_field_t*                em,
                               ntt_field_t*                cur,
                               ntt_field_t*                buff,
                               ntt_species_t*              species,
                               int32_t                     nspecies,
                               const ntt_sr_step_params_t* prm,
                               const ntt_minkowski_t*      metric,
                               void*                       workspace,
                               uint64_t                    workspace_bytes,
                               void*                       stream);

/* bytes of the workspace ntt_sr_step uses for a current field of this shape:
 * the accumulator (24 bytes per ghost-extended cell), and in 3D the split
 * launches' chunk flags (one bit pair per cell) and one set of box records
 * (16 bytes per 16 cells), shared by the species: every push then finds its
 * chunks' boxes with a pass over the particles */
NTT_API ntt_status ntt_sr_step_workspace_size(const ntt_field_t* cur, uint64_t* bytes);

/* the same with one set of box records per species (3D): a workspace of this
 * size, kept from step to step, lets each species' push start from the boxes
 * its previous push recorded (a superset of the cells, no pass over the
 * particles).  A record that no longer covers a particle (the particles were
 * changed, or the workspace is new) only costs time: that particle is pushed
 * and deposited by the fallback launch.  Layout: accumulator, flags, then the
 * record sets in species order.  No reference counterpart. */
NTT_API ntt_status ntt_sr_step_workspace_size_species(const ntt_field_t* cur, int32_t nspecies,
                                                      uint64_t* bytes);

/* Size of the device array ntt_sr_step_static works in: 32 MiB.  The
 * current accumulator must fit (1,398,101 ghost-extended cells: 2D up to
 * 1178 x 1178 active cells, 3D up to 107^3); the 3D split launches' flags and
 * records are used when they fit after it (3D up to 106^3), else the fused
 * kernel runs. */
#define NTT_SR_STEP_STATIC_MAX_BYTES (32ull << 20)

/* ntt_sr_step without the workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the fixed-point current accumulator is a
 * statically allocated device array of NTT_SR_STEP_STATIC_MAX_BYTES that the
 * library holds on every GPU it runs on until the process ends.  The call
 * allocates nothing (no hipMalloc/hipFree); fields and particles are the
 * caller's, as for ntt_sr_step, and results are identical to it.  A box whose
 * workspace would exceed the maximum is rejected with NTT_ERR_CAPACITY.  All
 * calls on one GPU share the array (the step clears it before depositing),
 * so calls that overlap on one GPU (concurrent streams or host threads) race.
 * No counterpart in the reference. */
NTT_API ntt_status ntt_sr_step_static(ntt_field_t*                em,
                                      ntt_field_t*                cur,
                                      ntt_field_t*                buff,
                                      ntt_species_t*              species,
                                      int32_t                     nspecies,
                                      const ntt_sr_step_params_t* prm,
                                      const ntt_minkowski_t*      metric,
                                      void*                       stream);

/* Cells per dimension (w[2] = 1 in 2D) of the LDS window of ntt_sr_step's
 * fused push + deposit: a chunk whose live particles span a box of b cells
 * runs from LDS when b + 3 <= w in every dimension (the [-1, +2] reach of
 * the order-2 stencil), else from HBM.  A sort tile of t cells therefore
 * keeps the fast path while t + 3 (+ the drift since the last sort) <= w.
 * No reference counterpart (a launch-shape query for tuners). */
NTT_API ntt_status ntt_fused_window(int32_t dim, int32_t w[3]);

/* ---------------------------------------------------------------- adjoint */

/* Per-particle adjoints (d loss / d state) of a 2D species: offsets dx1,
 * dx2 (the cell index is held fixed) and momenta u1..3; float or double
 * arrays of `size` entries, as the precision argument says.
 * Indexing: when the particles carry an ID payload (pld_i != NULL), particle
 * p's adjoint is entry pld_i[p], so the adjoint follows a particle through
 * re-sorts; IDs outside [0, size) skip the particle and raise
 * NTT_DEVERR_OUT_OF_BOUNDS.  Without a payload it is entry p, and size must
 * be at least npart. */
typedef struct ntt_particle_adjoint {
  void *lx1, *lx2;
  void *lu1, *lu2, *lu3;
  int64_t size;
} ntt_particle_adjoint_t;

#define NTT_VJP_FARADAY 0
#define NTT_VJP_AMPERE  1
#define NTT_VJP_FILTER  2

/* Vector-Jacobian product of the 2D Boris push of one species (the forward
 * of ntt_sr_push / ntt_sr_step: gather of E and B from em at the pre-push
 * position, Boris, position update), shape order 1 or 2.  `pre` holds the
 * particles before the push, `post_adj` the adjoint of their state after it.
 * Adds the position adjoint to pre_adj->lx*, writes pre_adj->lu*, and adds
 * the adjoint of the gathered fields to lam_em (6 components, active cells,
 * layout [comp][n2][n1]).  precision 32 or 64 selects float or double
 * adjoint arrays and arithmetic.  No reference counterpart. */
NTT_API ntt_status ntt_sr_push_vjp(const ntt_sr_pusher_ctx_t*   ctx,
                                   const ntt_particles_t*       pre,
                                   const ntt_field_t*           em,
                                   const ntt_minkowski_t*       metric,
                                   int32_t                      shape_order,
                                   int32_t                      precision,
                                   const ntt_particle_adjoint_t* post_adj,
                                   ntt_particle_adjoint_t*       pre_adj,
                                   void*                        lam_em,
                                   void*                        stream);

/* Vector-Jacobian product of the 2D Esirkepov deposit (order 1 or 2) of one
 * species: `post` holds the particles after the push (i_prev, dx_prev, i, dx,
 * u, weight), lam_j the adjoint of the folded current (3 components, active
 * cells, [comp][n2][n1]).  Adds to post_adj (positions and momenta after
 * the push) and writes the adjoint of the pre-push offsets to pre_adj->lx1/2
 * (pre_adj->lu* are not used).  No reference counterpart. */
NTT_API ntt_status ntt_deposit_vjp(const ntt_particles_t* post, const int32_t n[2],
                                   float charge, float dt, int32_t shape_order,
                                   int32_t precision, const void* lam_j,
                                   ntt_particle_adjoint_t* post_adj,
                                   ntt_particle_adjoint_t* pre_adj, void* stream);

/* The deposit VJP and the push VJP of one species and one step in one pass:
 * `pre` and `post` are the particles before and after that step's push (same
 * order, same count).  Equivalent to ntt_deposit_vjp(post, ..., post_adj,
 * pre_adj) followed by ntt_sr_push_vjp(pre, ..., post_adj, pre_adj), except
 * that post_adj is only read (the deposit's contribution is carried in
 * registers) and pre_adj->lx* are written, not accumulated.  The grid size
 * and charge come from em and ctx.  No reference counterpart. */
NTT_API ntt_status ntt_sr_particle_vjp(const ntt_sr_pusher_ctx_t*    ctx,
                                       const ntt_particles_t*        pre,
                                       const ntt_particles_t*        post,
                                       const ntt_field_t*            em,
                                       const ntt_minkowski_t*        metric,
                                       int32_t                       shape_order,
                                       int32_t                       precision,
                                       const void*                   lam_j,
                                       const ntt_particle_adjoint_t* post_adj,
                                       ntt_particle_adjoint_t*       pre_adj,
                                       void*                         lam_em,
                                       void*                         stream);

/* Transposed 2D periodic field operators on active-cell adjoints
 * ([comp][n2][n1], 3 components each):
 *   NTT_VJP_FARADAY: lam_out(E) += M^T lam_in(B), M the Faraday update with
 *                    coefficients c1, c2 (ntt_faraday_mink);
 *   NTT_VJP_AMPERE:  lam_out(B) += N^T lam_in(E), N the Ampere update;
 *   NTT_VJP_FILTER:  lam_out = F lam_in, the (symmetric) binomial filter.
 * No reference counterpart. */
NTT_API ntt_status ntt_yee_vjp(int32_t op, const int32_t n[2], float c1, float c2,
                               int32_t precision, const void* lam_in, void* lam_out,
                               void* stream);

/* ------------------------------------------------------- batched ensembles */

/* One box of an ensemble: the arguments ntt_sr_step takes for that box. */
typedef struct ntt_box {
  ntt_field_t          em, cur, buff;
  ntt_species_t*       species;          /* host array of nspecies            */
  int32_t              nspecies;
  ntt_sr_step_params_t prm;              /* prm.step is ignored (see step)    */
  ntt_minkowski_t      metric;
  void*                workspace;        /* ntt_sr_step_workspace_size bytes  */
  uint64_t             workspace_bytes;
  int32_t              tile[3];          /* tile of ntt_ensemble_sort; tile[0]
                                            = 0: the ensemble does not sort
                                            this box                          */
} ntt_box_t;

typedef struct ntt_ensemble* ntt_ensemble_t;

/* Plans the batched step of independent periodic boxes of any sizes, shape
 * orders, pushers and filter counts (parameter scans).  The plan records the
 * device pointers of every box (fields, particle arrays, tile offsets,
 * workspaces), which must stay valid and in place while it is used; particle
 * counts are those at creation.  Allocates device memory for its tables.
 * No reference counterpart: the reference runs one simulation per process. */
NTT_API ntt_status ntt_ensemble_create(const ntt_box_t* boxes, int32_t nboxes,
                                       ntt_ensemble_t* ens);

/* One SRPIC step of every box of the plan: each phase of ntt_sr_step is one
 * launch over all boxes (the push + deposit one launch per dimension, shape
 * order and pusher), so every box's fields and particles are bit-identical
 * to ntt_sr_step on that box alone.  step == 0 fills the E/B ghosts first.
 * Enqueues kernels only (no allocation, copy or synchronisation), so the
 * call can be captured in a HIP graph. */
NTT_API ntt_status ntt_ensemble_step(ntt_ensemble_t ens, int64_t step, void* stream);

/* The tile re-sort (ntt_sort_particles, without compaction) of every species
 * of every box whose ntt_box_t.tile was set, as one counting sort over all of
 * them: a fixed number of launches whatever the number of boxes.  Writes
 * each species' tile offsets.  Particle order inside a tile may differ from
 * ntt_sort_particles; results of later steps do not depend on it.  The prev
 * arrays (i*_prev, dx*_prev) of species with a pusher are left unpermuted:
 * the next ntt_ensemble_step rewrites them before reading them.  Enqueues
 * kernels and device-to-device copies only (graph-capturable); the sort
 * workspace is allocated by ntt_ensemble_create.  No reference counterpart. */
NTT_API ntt_status ntt_ensemble_sort(ntt_ensemble_t ens, void* stream);

NTT_API ntt_status ntt_ensemble_destroy(ntt_ensemble_t ens);

/* Launch timing of the fused push+deposit kernels that ntt_sr_step enqueues:
 * while enabled, every such launch is bracketed by HIP events recorded on its
 * stream; ntt_profile_read waits for them and returns the summed elapsed
 * milliseconds and the number of launches (reset != 0 clears the record).
 * No reference counterpart (the reference wraps whole phases in wall-clock
 * timers, src/global/utils/timer.h:51-148). */
NTT_API ntt_status ntt_profile_enable(int32_t on);
NTT_API ntt_status ntt_profile_read(double* ms_total, int64_t* launches, int32_t reset);

/* Spatial sort of one species by tile, dead particles last.
 * Replaces Particles::SortSpatially + RemoveDead
 * (src/framework/containers/particles_sort.cpp:103-253) with the tile key of
 * sort::PositionToTileIndex (src/global/utils/sorting.h:65-137): tiles of
 * tile[0] x tile[1] (x tile[2]) cells numbered x fastest.  Every non-NULL
 * SoA array is permuted; only the cell indices and tag are required.  A
 * caller whose next push rewrites the prev arrays (i*_prev, dx*_prev) may pass
 * them as NULL: they are then left unpermuted, which saves their traffic (no
 * reference counterpart: SortSpatially permutes every array).  The order
 * inside a tile is unspecified.  `workspace` must hold
 * ntt_sort_workspace_size() bytes.  When npart_alive is non-NULL the call
 * waits for the stream and returns the number of live particles (which lead
 * the arrays afterwards); the caller may then shrink npart to it.  When
 * tile_offsets (device, ntiles + 1 entries) is non-NULL it receives the index
 * of the first particle of every tile, for ntt_species_t.tile_offsets. */
NTT_API ntt_status ntt_sort_workspace_size(int64_t       maxnpart,
                                           int32_t       dim,
                                           const int32_t n[3],
                                           const int32_t tile[3],
                                           uint64_t*     bytes);
NTT_API ntt_status ntt_sort_particles(ntt_particles_t* prtls,
                                      int32_t          dim,
                                      const int32_t    n[3],
                                      const int32_t    tile[3],
                                      void*            workspace,
                                      uint64_t         workspace_bytes,
                                      uint32_t*        tile_offsets,
                                      int64_t*         npart_alive,
                                      void*            stream);

/* The same sort out of place: every array of src (npart entries) is gathered
 * in tile order straight into the matching array of dst, which must not
 * alias any array of src and must hold npart entries; dst->npart is set.
 * It moves each array once instead of staging it, about half the memory
 * traffic of ntt_sort_particles, at the price of a second set of arrays (the
 * caller swaps the two sets).  `workspace` must hold
 * ntt_sort_workspace_size_to() bytes (no staging buffer).  Same key, tile
 * offsets and npart_alive as ntt_sort_particles.  No reference counterpart
 * as an interface (the reference sorts in place). */
NTT_API ntt_status ntt_sort_particles_to(const ntt_particles_t* src,
                                         ntt_particles_t*       dst,
                                         int32_t                dim,
                                         const int32_t          n[3],
                                         const int32_t          tile[3],
                                         void*                  workspace,
                                         uint64_t               workspace_bytes,
                                         uint32_t*              tile_offsets,
                                         int64_t*               npart_alive,
                                         void*                  stream);
NTT_API ntt_status ntt_sort_workspace_size_to(int64_t       maxnpart,
                                              int32_t       dim,
                                              const int32_t n[3],
                                              const int32_t tile[3],
                                              uint64_t*     bytes);

/* number of tiles of the key used by ntt_sort_particles (tile_offsets of the
 * sort then holds ntiles + 1 entries, the last one counting the live particles) */
NTT_API ntt_status ntt_sort_ntiles(int32_t dim, const int32_t n[3], const int32_t tile[3],
                                   int64_t* ntiles);

/* Seeded synthetic plasma: ppc particles per cell over the n[] box, thermal
 * 4-velocities (Gaussian in u for temperature < 0.5, Sobol rejection above,
 * as archetypes' JuttnerSinge, src/archetypes/energy_dist.h:79-135) boosted
 * along one principal axis by the flipping method (:136-170).  Particles are
 * laid out tile by tile (tiles of tile[] cells, cells visited round-robin
 * inside a tile), i.e. already sorted for ntt_sort_particles with the same
 * tile.  Species sharing position_seed are co-located particle by particle.
 * Sets prtls->npart.  No reference counterpart as an interface: it stands in
 * for the reference's injectors (src/kernels/injectors.hpp) to produce the
 * benchmark inputs. */
typedef struct ntt_plasma_init {
  uint64_t seed;          /* velocity stream seed                   */
  uint64_t position_seed; /* position stream seed                   */
  int32_t  species_id;    /* separates the velocity streams          */
  int32_t  ppc;           /* particles per cell                      */
  float    temperature;   /* theta, in units of m c^2                */
  float    drift_u[3];    /* drift 4-velocity (one non-zero entry)   */
  float    weight;
  int32_t  tile[3];       /* generation tile; must divide n          */
} ntt_plasma_init_t;

NTT_API ntt_status ntt_init_uniform_plasma(ntt_particles_t*         prtls,
                                           int32_t                  dim,
                                           const int32_t            n[3],
                                           const ntt_plasma_init_t* init,
                                           void*                    stream);

#ifdef __cplusplus
}
#endif

#endif /* NTT_H */
