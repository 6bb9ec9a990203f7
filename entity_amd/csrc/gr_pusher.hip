// gr_pusher.hip -- GR particle pusher of a 2D axisymmetric (r, theta) box over
// the Kerr-Schild family of metrics, on gfx950.
//
// Reference: kernel::gr::Pusher_kernel<M> (src/kernels/pushers/gr.hpp:71-865)
// with M = KerrSchild / QKerrSchild / KerrSchild0 (src/metrics/kerr_schild.h,
// qkerr_schild.h, kerr_schild_0.h), launched per species by
// grpic::ParticlePush (src/engines/grpic/particle_pusher.h:26-89).
//
// Design:
// * One thread per particle; the work is FP64 VALU (metric evaluation inside
//   two fixed-point loops of niter iterations), not memory.  The real type R
//   is double by default (FP32 storage, FP64 metric and pusher) or float
//   (the reference's single-precision arithmetic).
// * The geodesic momentum push evaluates every metric quantity at one fixed
//   position: they are computed once per particle and reused by all niter
//   iterations (the reference re-evaluates them per iteration; the values
//   are the same).  The EM half pushes share the tetrad at the particle.
//   The coordinate push moves its midpoint every iteration and evaluates the
//   inverse metric, lapse and shift there, with one sincos per evaluation.
// * The theta clamp of the midpoint (SMALL_ANGLE_GR) uses code-unit bounds
//   computed on the host.
// * UpdatePhi of the reference takes phi by value and has no effect; it is
//   not evaluated.
#include "gr_metric.hpp"

#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace ntt {
namespace {

// tetrad transformations at a point (kerr_schild.h:599-630)
template <typename R, int KIND>
__device__ __forceinline__ void cov_to_hat(const MPt<R>& m, const R* vi, R* vo) {
  if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
    vo[0] = vi[0] * sqrt(m.h11);
    vo[1] = vi[1] / sqrt(m.h_22);
    vo[2] = vi[2] / sqrt(m.h_33);
  } else {
    const R A0 = sqrt(m.h11);
    vo[0] = vi[0] * A0 - vi[2] * A0 * m.h_13 / m.h_33;
    vo[1] = vi[1] / sqrt(m.h_22);
    vo[2] = vi[2] / sqrt(m.h_33);
  }
}
template <typename R, int KIND>
__device__ __forceinline__ void hat_to_cov(const MPt<R>& m, const R* vi, R* vo) {
  if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
    vo[0] = vi[0] / sqrt(m.h11);
    vo[1] = vi[1] * sqrt(m.h_22);
    vo[2] = vi[2] * sqrt(m.h_33);
  } else {
    vo[0] = vi[0] / sqrt(m.h11) + vi[2] * m.h_13 / sqrt(m.h_33);
    vo[1] = vi[1] * sqrt(m.h_22);
    vo[2] = vi[2] * sqrt(m.h_33);
  }
}
template <typename R, int KIND>
__device__ __forceinline__ void cntrv_to_hat(const MPt<R>& m, const R* vi, R* vo) {
  if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
    vo[0] = vi[0] / sqrt(m.h11);
    vo[1] = vi[1] * sqrt(m.h_22);
    vo[2] = vi[2] * sqrt(m.h_33);
  } else {
    vo[0] = vi[0] / sqrt(m.h11);
    vo[1] = vi[1] * sqrt(m.h_22);
    vo[2] = vi[2] * sqrt(m.h_33) + vi[0] * m.h_13 / sqrt(m.h_33);
  }
}
template <typename R>
__device__ __forceinline__ void cov_to_cntrv(const MPt<R>& m, const R* vi, R* vo) {
  vo[0] = vi[0] * m.h11 + vi[2] * m.h13;
  vo[1] = vi[1] * m.h22;
  vo[2] = vi[0] * m.h13 + vi[2] * m.h33;
}

struct GRArgs {
  ParticleView P;
  FieldView    DB, DB0;    // D from em, B from em0 (ghost-extended, 6 components)
  int32_t      ni1, ni2;
  int32_t      niter;
  bool         absorb_i1min, absorb_i1max, axis_i2min, axis_i2max;
  uint32_t*    err;
};

// field gather of one component: O = 0 "direct" (gr.hpp:488-578), O >= 1
// shape functions (gr.hpp:579-662)
template <typename R, int O>
struct Gather2D {
  int i, j;
  R   dx1, dx2;
  // O = 0
  int indx, indy;
  R   ponpmx, ponppx, ponpmy, ponppy, pondmx, pondpx, pondmy, pondpy;
  // O >= 1
  R   S1p[O + 1 > 1 ? O + 1 : 2], S1d[O + 1 > 1 ? O + 1 : 2];
  R   S2p[O + 1 > 1 ? O + 1 : 2], S2d[O + 1 > 1 ? O + 1 : 2];
  int ip, id, jp, jd;

  __device__ __forceinline__ void shapes(int i_, int j_, R d1, R d2) {
    i = i_;
    j = j_;
    dx1 = d1;
    dx2 = d2;
    if constexpr (O == 0) {
      indx   = static_cast<int>(dx1 + R(0.5));
      indy   = static_cast<int>(dx2 + R(0.5));
      ponpmx = R(1) - dx1;
      ponppx = dx1;
      ponpmy = R(1) - dx2;
      ponppy = dx2;
      pondmx = static_cast<R>(indx + 1) - (dx1 + R(0.5));
      pondpx = R(1) - pondmx;
      pondmy = static_cast<R>(indy + 1) - (dx2 + R(0.5));
      pondpy = R(1) - pondmy;
    } else {
      shape<false>(i, dx1, ip, S1p);
      shape<false>(j, dx2, jp, S2p);
      shape<true>(i, dx1, id, S1d);
      shape<true>(j, dx2, jd, S2d);
    }
  }
  // prtl_shape::order<STAG, O> (particle_shapes.hpp:544-588)
  template <bool STAG>
  __device__ __forceinline__ static void shape(int ii, R di, int& imin, R* S) {
    if constexpr (O == 1) {
      if constexpr (!STAG) {
        imin = ii;
        S[0] = R(1) - di;
        S[1] = di;
      } else if (di < R(0.5)) {
        imin = ii - 1;
        S[0] = R(0.5) - di;
        S[1] = R(1) - S[0];
      } else {
        imin = ii;
        S[0] = R(1.5) - di;
        S[1] = R(1) - S[0];
      }
    } else if constexpr (O == 2) {
      if constexpr (!STAG) {
        if (di < R(0.5)) {
          imin = ii - 1;
          S[0] = R(0.5) * ((R(0.5) - di) * (R(0.5) - di));
          S[1] = R(0.75) - di * di;
          S[2] = R(1) - S[0] - S[1];
        } else {
          imin = ii;
          S[0] = R(0.5) * ((R(1.5) - di) * (R(1.5) - di));
          S[1] = R(0.75) - (R(1) - di) * (R(1) - di);
          S[2] = R(1) - S[0] - S[1];
        }
      } else {
        imin = ii - 1;
        S[0] = R(0.5) * ((R(1) - di) * (R(1) - di));
        S[2] = R(0.5) * (di * di);
        S[1] = R(1) - S[0] - S[2];
      }
    }
  }
  __device__ __forceinline__ R F(const FieldView& f, int a, int b, int c) const {
    return static_cast<R>(f(a, b, 0, c));
  }
  template <int DUAL1, int DUAL2>
  __device__ __forceinline__ R get(const FieldView& f, int c) const {
    if constexpr (O == 0) {
      const int a0 = DUAL1 ? i - 1 + indx : i, b0 = DUAL2 ? j - 1 + indy : j;
      const R   w0 = DUAL1 ? pondmx : ponpmx, w1 = DUAL1 ? pondpx : ponppx;
      const R   v0 = DUAL2 ? pondmy : ponpmy, v1 = DUAL2 ? pondpy : ponppy;
      const R c00 = F(f, a0, b0, c) * w0 + F(f, a0 + 1, b0, c) * w1;
      const R c10 = F(f, a0, b0 + 1, c) * w0 + F(f, a0 + 1, b0 + 1, c) * w1;
      return c00 * v0 + c10 * v1;
    } else {
      const R*  S1 = DUAL1 ? S1d : S1p;
      const R*  S2 = DUAL2 ? S2d : S2p;
      const int im = DUAL1 ? id : ip, jm = DUAL2 ? jd : jp;
      R e = R(0);
#pragma unroll
      for (int b = 0; b < O + 1; ++b) {
        R c0 = R(0);
#pragma unroll
        for (int a = 0; a < O + 1; ++a) c0 += S1[a] * F(f, im + a, jm + b, c);
        e += c0 * S2[b];
      }
      return e;
    }
  }
};

// The push of particle p of one species (the body of Pusher_kernel::operator()).
template <typename R, int KIND, int O, bool MASSLESS>
__device__ __forceinline__ void gr_push_one(const GRArgs& g, const KSMetric<R, KIND>& M, R dt,
                                            R ndth, int64_t p) {
  const int16_t tag = g.P.tag[p];
  if (tag != NTT_TAG_ALIVE) {
    if (tag != NTT_TAG_DEAD) raise_device_error(g.err, NTT_DEVERR_INVALID_TAG);
    return;
  }
  const int32_t i1 = g.P.i1[p], i2 = g.P.i2[p];
  const float   fdx1 = g.P.dx1[p], fdx2 = g.P.dx2[p];
  if (!MASSLESS) {
    // gather stencil [i - 1, i + 2] (ghost-extended) must lie in the arrays
    const int G = g.DB.ngh;
    if (i1 + G - 1 < 0 || i1 + G + 2 >= g.DB.N1 || i2 + G - 1 < 0 || i2 + G + 2 >= g.DB.N2 ||
        !(fdx1 >= 0.0f && fdx1 <= 1.0f) || !(fdx2 >= 0.0f && fdx2 <= 1.0f)) {
      raise_device_error(g.err, NTT_DEVERR_OUT_OF_BOUNDS);
      return;
    }
  }
  g.P.i1p[p]  = i1;
  g.P.i2p[p]  = i2;
  g.P.dx1p[p] = fdx1;
  g.P.dx2p[p] = fdx2;
  const R x0 = static_cast<R>(i1) + static_cast<R>(fdx1);
  const R x1 = static_cast<R>(i2) + static_cast<R>(fdx2);
  R vp[3] = { static_cast<R>(g.P.ux1[p]), static_cast<R>(g.P.ux2[p]),
              static_cast<R>(g.P.ux3[p]) };
  auto clamp_theta = [&](R tcd) -> R {
    const R tph = M.theta_ph(tcd);
    if (tph < M.k.small_angle) return M.k.th_small_cd;
    if (static_cast<double>(tph) >= M.k.large_angle) return M.k.th_large_cd;
    return tcd;
  };
  auto gamma_of = [&](const R* ucov, const R* ucnt) -> R {
    const R s = ucov[0] * ucnt[0] + ucov[1] * ucnt[1] + ucov[2] * ucnt[2];
    return MASSLESS ? sqrt(s) : sqrt(R(1) + s);
  };

  // GeodesicMomentumPush at a fixed position (gr.hpp:260-331)
  auto momentum_push = [&](const MPt<R>& m, const R* v, R* vu) {
    typename KSMetric<R, KIND>::Deriv d;
    M.derivs(m, d);
    vu[0] = v[0];
    vu[1] = v[1];
    vu[2] = v[2];
    R vm[3], vmc[3];
    // the iteration is a fixed map of (vu[0], vu[1]): once an iteration
    // leaves them unchanged bit for bit, every later one does too, so
    // stopping there gives exactly the niter-iteration result
    for (int it = 0; it < g.niter; ++it) {
      const R prev0 = vu[0], prev1 = vu[1];
      vm[0] = R(0.5) * (v[0] + vu[0]);
      vm[1] = R(0.5) * (v[1] + vu[1]);
      vm[2] = v[2];
      cov_to_cntrv(m, vm, vmc);
      const R u0 = gamma_of(vm, vmc) / m.alpha;
      vu[0] = v[0] + dt * (-m.alpha * u0 * d.dr_alpha + vm[0] * d.dr_beta1 -
                           (R(0.5) / u0) *
                             (d.dr_h11 * sqr(vm[0]) + d.dr_h22 * sqr(vm[1]) + d.dr_h33 * sqr(vm[2]) +
                              R(2) * d.dr_h13 * vm[0] * vm[2]));
      vu[1] = v[1] + dt * (-m.alpha * u0 * d.dt_alpha + vm[0] * d.dt_beta1 -
                           (R(0.5) / u0) *
                             (d.dt_h11 * sqr(vm[0]) + d.dt_h22 * sqr(vm[1]) + d.dt_h33 * sqr(vm[2]) +
                              R(2) * d.dt_h13 * vm[0] * vm[2]));
      if (vu[0] == prev0 && vu[1] == prev1) break;
    }
  };

  MPt<R> mx;
  M.eval(x0, x1, mx);
  R vp_upd[3];
  if constexpr (MASSLESS) {
    momentum_push(mx, vp, vp_upd);
  } else {
    // fields at the particle (contravariant), to the tetrad at x
    Gather2D<R, O> gt;
    gt.shapes(i1 + g.DB.ngh, i2 + g.DB.ngh, static_cast<R>(fdx1), static_cast<R>(fdx2));
    R Dc[3], Bc[3], Dh[3], Bh[3];
    Dc[0] = gt.template get<1, 0>(g.DB, 0);
    Dc[1] = gt.template get<0, 1>(g.DB, 1);
    Dc[2] = gt.template get<0, 0>(g.DB, 2);
    Bc[0] = gt.template get<0, 1>(g.DB0, 3);
    Bc[1] = gt.template get<1, 0>(g.DB0, 4);
    Bc[2] = gt.template get<1, 1>(g.DB0, 5);
    cntrv_to_hat<R, KIND>(mx, Dc, Dh);
    cntrv_to_hat<R, KIND>(mx, Bc, Bh);
    // EMHalfPush (gr.hpp:177-217), twice, with the same D, B and tetrad
    auto em_half = [&](const R* v, R* vout) {
      R D0[3] = { Dh[0], Dh[1], Dh[2] }, B0[3] = { Bh[0], Bh[1], Bh[2] };
      R vuh[3], vh[3];
      cov_to_hat<R, KIND>(mx, v, vuh);
      R COEFF = ndth * R(0.5) * mx.alpha;
      D0[0] *= COEFF; D0[1] *= COEFF; D0[2] *= COEFF;
      vuh[0] += D0[0]; vuh[1] += D0[1]; vuh[2] += D0[2];
      COEFF *= R(1) / sqrt(R(1) + sqr(vuh[0]) + sqr(vuh[1]) + sqr(vuh[2]));
      B0[0] *= COEFF; B0[1] *= COEFF; B0[2] *= COEFF;
      COEFF = R(2) / (R(1) + sqr(B0[0]) + sqr(B0[1]) + sqr(B0[2]));
      vh[0] = (vuh[0] + vuh[1] * B0[2] - vuh[2] * B0[1]) * COEFF;
      vh[1] = (vuh[1] + vuh[2] * B0[0] - vuh[0] * B0[2]) * COEFF;
      vh[2] = (vuh[2] + vuh[0] * B0[1] - vuh[1] * B0[0]) * COEFF;
      vuh[0] += vh[1] * B0[2] - vh[2] * B0[1] + D0[0];
      vuh[1] += vh[2] * B0[0] - vh[0] * B0[2] + D0[1];
      vuh[2] += vh[0] * B0[1] - vh[1] * B0[0] + D0[2];
      hat_to_cov<R, KIND>(mx, vuh, vout);
    };
    R vtmp[3];
    em_half(vp, vtmp);
    // the momentum push runs at the theta-clamped position (gr.hpp:742-756)
    const R x1c = clamp_theta(x1);
    if (x1c == x1) {
      momentum_push(mx, vtmp, vp_upd);
    } else {
      MPt<R> mc;
      M.eval(x0, x1c, mc);
      momentum_push(mc, vtmp, vp_upd);
    }
    em_half(vp_upd, vtmp);
    vp_upd[0] = vtmp[0];
    vp_upd[1] = vtmp[1];
    vp_upd[2] = vtmp[2];
  }

  // GeodesicCoordinatePush (gr.hpp:333-383)
  R xu0 = x0, xu1 = x1;
  for (int it = 0; it < g.niter; ++it) {   // exact early stop as in momentum_push
    const R prev0 = xu0, prev1 = xu1;
    const R xm0 = R(0.5) * (x0 + xu0);
    const R xm1 = clamp_theta(R(0.5) * (x1 + xu1));
    MPt<R> mm;
    M.eval(xm0, xm1, mm);
    R vc[3];
    cov_to_cntrv(mm, vp_upd, vc);
    const R u0 = gamma_of(vp_upd, vc) / mm.alpha;
    xu0 = x0 + dt * (vc[0] / u0 - mm.beta1);
    xu1 = x1 + dt * (vc[1] / u0);
    if (xu0 == prev0 && xu1 == prev1) break;
  }

  // from_Xi_to_i_di (gr.hpp:36-45) and boundaryConditions (gr.hpp:818-865)
  int32_t ni1_ = static_cast<int32_t>(xu0 + R(1)) - 1;
  int32_t ni2_ = static_cast<int32_t>(xu1 + R(1)) - 1;
  float   ndx1 = static_cast<float>(xu0) - static_cast<float>(ni1_);
  float   ndx2 = static_cast<float>(xu1) - static_cast<float>(ni2_);
  float   u2   = static_cast<float>(vp_upd[1]);
  int16_t ntag = NTT_TAG_ALIVE;
  if ((ni1_ < 0 && g.absorb_i1min) || (ni1_ >= g.ni1 && g.absorb_i1max)) ntag = NTT_TAG_DEAD;
  if (ni2_ < 0) {
    if (g.axis_i2min) {
      ni2_ = 0;
      ndx2 = 1.0f - ndx2;
      u2   = -u2;
    }
  } else if (ni2_ >= g.ni2) {
    if (g.axis_i2max) {
      ni2_ = g.ni2 - 1;
      ndx2 = 1.0f - ndx2;
      u2   = -u2;
    }
  }
  g.P.i1[p]  = ni1_;
  g.P.dx1[p] = ndx1;
  g.P.i2[p]  = ni2_;
  g.P.dx2[p] = ndx2;
  g.P.ux1[p] = static_cast<float>(vp_upd[0]);
  g.P.ux2[p] = u2;
  g.P.ux3[p] = static_cast<float>(vp_upd[2]);
  if (ntag != NTT_TAG_ALIVE) g.P.tag[p] = ntag;
}

template <typename R, int KIND, int O, bool MASSLESS>
__global__ __launch_bounds__(256) void k_gr_push(GRArgs g, KSMetric<R, KIND> M, R dt, R ndth) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= g.P.n) return;
  gr_push_one<R, KIND, O, MASSLESS>(g, M, dt, ndth, p);
}

// One species of one run in a batched launch: the push's arguments and its
// metric constants, as the single-species launch derives them.
template <typename R>
struct GREntry {
  GRArgs     g;
  KSConst<R> k;
  R          dt, ndth;
};

// Every (run, species) entry of a batch in one launch: entry e owns the
// blocks [blk0[e], blk0[e + 1]), one particle per thread.  The entry index is
// uniform across the block, so its arguments are read in place from the
// table (scalar loads) rather than copied into registers.
template <typename R, int KIND, int O, bool MASSLESS>
__global__ __launch_bounds__(256) void k_gr_push_batched(const GREntry<R>* __restrict__ table,
                                                         const int64_t* __restrict__ blk0,
                                                         int nent) {
  const int64_t b = blockIdx.x;
  int lo = 0, hi = nent - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const int        e = __builtin_amdgcn_readfirstlane(lo);
  const GREntry<R>& E = table[e];
  const int64_t    p = (b - blk0[e]) * 256 + threadIdx.x;
  if (p >= E.g.P.n) return;
  // KSMetric is KSConst plus methods (same layout): read the constants in place
  static_assert(sizeof(KSMetric<R, KIND>) == sizeof(KSConst<R>), "KSMetric layout");
  const KSMetric<R, KIND>& M = *reinterpret_cast<const KSMetric<R, KIND>*>(&E.k);
  gr_push_one<R, KIND, O, MASSLESS>(E.g, M, E.dt, E.ndth, p);
}

template <typename R, int KIND, int O>
void launch_gr(const GRArgs& g, const KSConst<R>& k, R dt, R ndth, bool massless,
               hipStream_t st) {
  KSMetric<R, KIND> M{ k };
  const unsigned grid = grid_for(g.P.n, 256);
  if (massless) k_gr_push<R, KIND, O, true><<<grid, 256, 0, st>>>(g, M, dt, ndth);
  else k_gr_push<R, KIND, O, false><<<grid, 256, 0, st>>>(g, M, dt, ndth);
}

template <typename R, int KIND>
void launch_gr_o(int order, const GRArgs& g, const KSConst<R>& k, R dt, R ndth, bool massless,
                 hipStream_t st) {
  if (order == 0) launch_gr<R, KIND, 0>(g, k, dt, ndth, massless, st);
  else if (order == 1) launch_gr<R, KIND, 1>(g, k, dt, ndth, massless, st);
  else launch_gr<R, KIND, 2>(g, k, dt, ndth, massless, st);
}

template <typename R>
ntt_status run_gr(const ntt_gr_pusher_ctx_t* ctx, const GRArgs& g, const ntt_ks_metric_t* metric,
                  int order, bool massless, hipStream_t st) {
  KSConst<R> k;
  ntt_status s = make_const<R>(metric, &k);
  if (s != NTT_OK) return s;
  const R dt   = static_cast<R>(ctx->dt);
  // gr.hpp:108-109: normalized_dt_half = 1/2 (q/m) omegaB0 dt
  const R ndth = R(0.5) * static_cast<R>(ctx->charge / ctx->mass) *
                 static_cast<R>(ctx->omegaB0) * dt;
  switch (metric->kind) {
    case NTT_METRIC_KERR_SCHILD:
      launch_gr_o<R, NTT_METRIC_KERR_SCHILD>(order, g, k, dt, ndth, massless, st);
      break;
    case NTT_METRIC_QKERR_SCHILD:
      launch_gr_o<R, NTT_METRIC_QKERR_SCHILD>(order, g, k, dt, ndth, massless, st);
      break;
    default:
      launch_gr_o<R, NTT_METRIC_KERR_SCHILD_0>(order, g, k, dt, ndth, massless, st);
      break;
  }
  return check_launch();
}

// Validates one species' push as ntt_gr_push does and fills its GRArgs;
// *skip is set for a species with nothing to push.
ntt_status prepare_gr(const ntt_gr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                      const ntt_particles_t* prtls, const ntt_field_t* em,
                      const ntt_field_t* em0, const ntt_ks_metric_t* metric, int32_t shape_order,
                      GRArgs* g, bool* massless, bool* skip) {
  *skip = false;
  if (ctx == nullptr || bc == nullptr || prtls == nullptr || metric == nullptr)
    return NTT_ERR_ARG;
  if (shape_order < 0 || shape_order > 2) return NTT_ERR_UNSUPPORTED;
  if (metric->kind != NTT_METRIC_KERR_SCHILD && metric->kind != NTT_METRIC_QKERR_SCHILD &&
      metric->kind != NTT_METRIC_KERR_SCHILD_0)
    return NTT_ERR_UNSUPPORTED;
  // grpic::ParticlePush (particle_pusher.h:53-87): PHOTON -> massless,
  // BORIS -> massive, anything else is not implemented
  if (ctx->pusher_flags == NTT_PUSHER_PHOTON) *massless = true;
  else if (ctx->pusher_flags == NTT_PUSHER_BORIS) *massless = false;
  else if (ctx->pusher_flags == NTT_PUSHER_NONE) {
    *skip = true;
    return NTT_OK;
  } else return NTT_ERR_UNSUPPORTED;
  if (ctx->niter < 0) return NTT_ERR_ARG;
  if (!*massless && !(ctx->mass != 0.0f)) return NTT_ERR_ARG;
  if (ctx->precision != 0 && ctx->precision != 32 && ctx->precision != 64)
    return NTT_ERR_ARG;
  *g = GRArgs{};
  ntt_status s = make_particle_view(prtls, 2, true, &g->P);
  if (s != NTT_OK) return s;
  if (!*massless) {
    if (em == nullptr || em0 == nullptr) return NTT_ERR_ARG;
    s = make_field_view(em, 6, &g->DB);
    if (s != NTT_OK) return s;
    s = make_field_view(em0, 6, &g->DB0);
    if (s != NTT_OK) return s;
    if (em->dim != 2 || em0->dim != 2) return NTT_ERR_UNSUPPORTED;
    if (g->DB.N1 != g->DB0.N1 || g->DB.N2 != g->DB0.N2 || g->DB.ngh != g->DB0.ngh)
      return NTT_ERR_ARG;
    if (g->DB.n1 != ctx->ni1 || g->DB.n2 != ctx->ni2) return NTT_ERR_ARG;
  }
  g->ni1 = ctx->ni1;
  g->ni2 = ctx->ni2;
  g->niter = ctx->niter;
  // kernel::gr::PusherBoundaries (context.h:213-228)
  g->absorb_i1min = bc->bc[0][0] == NTT_PRTL_BC_ABSORB || bc->bc[0][0] == NTT_PRTL_BC_HORIZON;
  g->absorb_i1max = bc->bc[0][1] == NTT_PRTL_BC_ABSORB || bc->bc[0][1] == NTT_PRTL_BC_HORIZON;
  g->axis_i2min   = bc->bc[1][0] == NTT_PRTL_BC_AXIS;
  g->axis_i2max   = bc->bc[1][1] == NTT_PRTL_BC_AXIS;
  g->err = device_error_word();
  if (g->err == nullptr) return NTT_ERR_HIP;
  if (g->P.n == 0) *skip = true;
  return NTT_OK;
}

template <typename R>
ntt_status make_entry(const ntt_gr_pusher_ctx_t* ctx, const GRArgs& g,
                      const ntt_ks_metric_t* metric, GREntry<R>* e) {
  e->g = g;
  ntt_status s = make_const<R>(metric, &e->k);
  if (s != NTT_OK) return s;
  e->dt   = static_cast<R>(ctx->dt);
  e->ndth = R(0.5) * static_cast<R>(ctx->charge / ctx->mass) * static_cast<R>(ctx->omegaB0) *
            e->dt;
  return NTT_OK;
}

}  // namespace
}  // namespace ntt

using namespace ntt;

// ---- batches of (run, species) entries ---------------------------------------
struct ntt_gr_batch {
  struct Group {
    int      precision, kind;       // 32 / 64, NTT_METRIC_*
    bool     massless;
    std::vector<char>    host;      // GREntry<R> table
    std::vector<int64_t> blk0;      // nent + 1 block offsets
    std::vector<int32_t> entry_of;  // batch entry index of each table entry
    void*    d_table = nullptr;
    int64_t* d_blk0  = nullptr;
  };
  int32_t                 order = 1;
  std::vector<Group>      groups;
  std::vector<std::pair<int, int>> where;  // batch entry -> (group, table index) or (-1, -1)
  std::vector<int64_t>    planned;          // particles each entry was planned for
};

namespace {

template <typename R>
size_t entry_size() { return sizeof(GREntry<R>); }

template <typename R, int KIND, bool MASSLESS>
void launch_batched_k(int order, const void* table, const int64_t* blk0, int nent, int64_t nb,
                      hipStream_t st) {
  const auto* t = static_cast<const GREntry<R>*>(table);
  if (order == 0) k_gr_push_batched<R, KIND, 0, MASSLESS><<<nb, 256, 0, st>>>(t, blk0, nent);
  else if (order == 1) k_gr_push_batched<R, KIND, 1, MASSLESS><<<nb, 256, 0, st>>>(t, blk0, nent);
  else k_gr_push_batched<R, KIND, 2, MASSLESS><<<nb, 256, 0, st>>>(t, blk0, nent);
}

template <typename R, bool MASSLESS>
void launch_batched_kind(int kind, int order, const void* table, const int64_t* blk0, int nent,
                         int64_t nb, hipStream_t st) {
  if (kind == NTT_METRIC_KERR_SCHILD)
    launch_batched_k<R, NTT_METRIC_KERR_SCHILD, MASSLESS>(order, table, blk0, nent, nb, st);
  else if (kind == NTT_METRIC_QKERR_SCHILD)
    launch_batched_k<R, NTT_METRIC_QKERR_SCHILD, MASSLESS>(order, table, blk0, nent, nb, st);
  else
    launch_batched_k<R, NTT_METRIC_KERR_SCHILD_0, MASSLESS>(order, table, blk0, nent, nb, st);
}

void free_batch(ntt_gr_batch* b) {
  for (auto& g : b->groups) {
    if (g.d_table) (void)hipFree(g.d_table);
    if (g.d_blk0) (void)hipFree(g.d_blk0);
  }
  delete b;
}

}  // namespace

extern "C" ntt_status ntt_gr_batch_create(const ntt_gr_batch_entry_t* entries, int32_t nentries,
                                          int32_t shape_order, ntt_gr_batch_t* out) {
  if (out == nullptr || nentries < 0 || (nentries > 0 && entries == nullptr)) return NTT_ERR_ARG;
  *out = nullptr;
  if (shape_order < 0 || shape_order > 2) return NTT_ERR_UNSUPPORTED;
  auto* b = new ntt_gr_batch;
  b->order = shape_order;
  b->where.assign(nentries, { -1, -1 });
  b->planned.assign(nentries, 0);
  ntt_status s = NTT_OK;
  for (int32_t i = 0; i < nentries && s == NTT_OK; ++i) {
    const ntt_gr_batch_entry_t& E = entries[i];
    GRArgs g;
    bool   massless = false, skip = false;
    s = prepare_gr(&E.ctx, &E.bc, &E.prtls, &E.em, &E.em0, &E.metric, shape_order, &g,
                   &massless, &skip);
    if (s != NTT_OK || skip) continue;
    const int prec = E.ctx.precision == 32 ? 32 : 64;
    int gi = -1;
    for (size_t k = 0; k < b->groups.size(); ++k) {
      const auto& G = b->groups[k];
      if (G.precision == prec && G.kind == E.metric.kind && G.massless == massless)
        gi = static_cast<int>(k);
    }
    if (gi < 0) {
      ntt_gr_batch::Group G;
      G.precision = prec;
      G.kind      = E.metric.kind;
      G.massless  = massless;
      G.blk0      = { 0 };
      b->groups.push_back(std::move(G));
      gi = static_cast<int>(b->groups.size()) - 1;
    }
    auto&        G  = b->groups[gi];
    const size_t sz = prec == 32 ? entry_size<float>() : entry_size<double>();
    const size_t at = G.host.size();
    G.host.resize(at + sz);
    if (prec == 32) {
      GREntry<float> e;
      s = make_entry<float>(&E.ctx, g, &E.metric, &e);
      std::memcpy(G.host.data() + at, &e, sz);
    } else {
      GREntry<double> e;
      s = make_entry<double>(&E.ctx, g, &E.metric, &e);
      std::memcpy(G.host.data() + at, &e, sz);
    }
    const int64_t nb = (g.P.n + 255) / 256;
    b->where[i]   = { gi, static_cast<int>(G.entry_of.size()) };
    b->planned[i] = g.P.n;
    G.entry_of.push_back(i);
    G.blk0.push_back(G.blk0.back() + nb);
    if (G.blk0.back() > 0x7FFFFFFF) s = NTT_ERR_UNSUPPORTED;
  }
  for (auto& G : b->groups) {
    if (s != NTT_OK) break;
    if (hipMalloc(&G.d_table, G.host.size()) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&G.d_blk0), G.blk0.size() * sizeof(int64_t)) !=
          hipSuccess ||
        hipMemcpy(G.d_table, G.host.data(), G.host.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(G.d_blk0, G.blk0.data(), G.blk0.size() * sizeof(int64_t),
                  hipMemcpyHostToDevice) != hipSuccess)
      s = NTT_ERR_HIP;
  }
  if (s != NTT_OK) {
    free_batch(b);
    return s;
  }
  *out = b;
  return NTT_OK;
}

extern "C" ntt_status ntt_gr_batch_push(ntt_gr_batch_t b, void* stream) {
  if (b == nullptr) return NTT_ERR_ARG;
  hipStream_t st = as_stream(stream);
  for (const auto& G : b->groups) {
    const int     nent = static_cast<int>(G.entry_of.size());
    const int64_t nb   = G.blk0.back();
    if (nent == 0 || nb == 0) continue;
    if (G.precision == 32) {
      if (G.massless) launch_batched_kind<float, true>(G.kind, b->order, G.d_table, G.d_blk0, nent, nb, st);
      else launch_batched_kind<float, false>(G.kind, b->order, G.d_table, G.d_blk0, nent, nb, st);
    } else {
      if (G.massless) launch_batched_kind<double, true>(G.kind, b->order, G.d_table, G.d_blk0, nent, nb, st);
      else launch_batched_kind<double, false>(G.kind, b->order, G.d_table, G.d_blk0, nent, nb, st);
    }
  }
  return check_launch();
}

extern "C" ntt_status ntt_gr_batch_set_npart(ntt_gr_batch_t b, int32_t entry, int64_t npart,
                                             void* stream) {
  if (b == nullptr || entry < 0 || entry >= static_cast<int32_t>(b->where.size()) || npart < 0)
    return NTT_ERR_ARG;
  const auto [gi, ti] = b->where[entry];
  if (gi < 0) return npart == 0 ? NTT_OK : NTT_ERR_ARG;
  if (npart > b->planned[entry]) return NTT_ERR_CAPACITY;
  auto&        G  = b->groups[gi];
  const size_t sz = G.precision == 32 ? entry_size<float>() : entry_size<double>();
  const size_t off = static_cast<size_t>(ti) * sz +
                     (G.precision == 32 ? offsetof(GREntry<float>, g) : offsetof(GREntry<double>, g)) +
                     offsetof(GRArgs, P) + offsetof(ParticleView, n);
  // the host copy stays current, so the copy source outlives the stream
  std::memcpy(G.host.data() + off, &npart, sizeof(npart));
  if (hipMemcpyAsync(static_cast<char*>(G.d_table) + off, G.host.data() + off, sizeof(npart),
                     hipMemcpyHostToDevice, as_stream(stream)) != hipSuccess)
    return NTT_ERR_HIP;
  return NTT_OK;
}

extern "C" ntt_status ntt_gr_batch_destroy(ntt_gr_batch_t b) {
  if (b == nullptr) return NTT_OK;
  (void)hipDeviceSynchronize();
  free_batch(b);
  return NTT_OK;
}

extern "C" ntt_status ntt_gr_push(const ntt_gr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                                  ntt_particles_t* prtls, const ntt_field_t* em,
                                  const ntt_field_t* em0, const ntt_ks_metric_t* metric,
                                  int32_t shape_order, void* stream) {
  GRArgs     g;
  bool       massless = false, skip = false;
  ntt_status s = prepare_gr(ctx, bc, prtls, em, em0, metric, shape_order, &g, &massless, &skip);
  if (s != NTT_OK || skip) return s;
  hipStream_t st = as_stream(stream);
  if (ctx->precision == 32) return run_gr<float>(ctx, g, metric, shape_order, massless, st);
  return run_gr<double>(ctx, g, metric, shape_order, massless, st);
}
