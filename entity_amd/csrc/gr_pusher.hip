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
#include "internal.hpp"

#include <cmath>

namespace ntt {
namespace {

constexpr double kPi          = 3.14159265358979323846;
constexpr double kInvPi       = 0.31830988618379067154;
constexpr double kPiSqr       = 9.86960440108935861882;
constexpr double kInvPiSqr    = 0.10132118364233777144;
constexpr double kSmallAngleGR = 1e-5;

template <typename R> __host__ __device__ __forceinline__ R sqr(R x) { return x * x; }
template <typename R> __host__ __device__ __forceinline__ R cube(R x) { return x * x * x; }

template <typename R>
__host__ __device__ __forceinline__ bool almost_zero(R a) {
  const R eps = sizeof(R) == 4 ? static_cast<R>(1.1920928955078125e-7)
                               : static_cast<R>(2.220446049250313e-16);
  return (a < 0 ? -a : a) <= eps;
}

__device__ __forceinline__ void sin_cos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void sin_cos(double x, double* s, double* c) { sincos(x, s, c); }

// Metric constants (code-unit spacings) computed on the host in R
template <typename R>
struct KSConst {
  R a, r0, h0, x1_min, x2_min;
  R dr, dtheta, dr_inv, dtheta_inv;
  R chi_min, eta_min, dchi, deta, dchi_inv, deta_inv;
  R th_small_cd, th_large_cd;  // clamp targets: convert<2, Ph, Cd>(small / large)
  double large_angle;          // PI - SMALL_ANGLE_GR, compared in double
  R      small_angle;
};

// Metric quantities at one point.  Lower components h_ij carry an underscore.
template <typename R>
struct MPt {
  R r, th, s, c;       // r, theta, sin, cos
  R S, Dl, Av, z;      // Sigma, Delta, A, z = 2r/Sigma
  R e, eta, dtde;      // QKS: exp(chi), eta, dtheta/deta
  R h11, h22, h33, h13;
  R h_22, h_33, h_13;
  R alpha, beta1;
};

template <typename R, int KIND>
struct KSMetric {
  KSConst<R> k;

  __device__ __forceinline__ R eta2theta(R eta) const {
    if (almost_zero(k.h0)) return eta;
    return eta + R(2) * k.h0 * eta * (static_cast<R>(kPi) - R(2) * eta) *
                   (static_cast<R>(kPi) - eta) * static_cast<R>(kInvPiSqr);
  }
  __device__ __forceinline__ R dtheta_deta(R eta) const {
    if (almost_zero(k.h0)) return R(1);
    return (R(1) + R(2) * k.h0 +
            static_cast<R>(12.0) * k.h0 * (eta * static_cast<R>(kInvPi)) *
              ((eta * static_cast<R>(kInvPi)) - R(1)));
  }
  __device__ __forceinline__ R dx_dt(R eta) const {
    if (almost_zero(k.h0)) return k.deta;
    return k.deta * (R(1) + R(2) * k.h0 * static_cast<R>(kInvPiSqr) *
                              (R(2) * R(3) * sqr(eta) - R(2) * R(3) * static_cast<R>(kPi) * eta +
                               static_cast<R>(kPiSqr)));
  }
  // convert<2, Cd, Ph>
  __device__ __forceinline__ R theta_ph(R x2) const {
    if constexpr (KIND == NTT_METRIC_QKERR_SCHILD) return eta2theta(x2 * k.deta + k.eta_min);
    else return x2 * k.dtheta + k.x2_min;
  }

  // r, theta, Sigma, ..., lapse, shift, inverse metric and the lower
  // components the transformations use (kerr_schild.h:143-330,
  // qkerr_schild.h:145-385, kerr_schild_0.h)
  __device__ __forceinline__ void eval(R x1, R x2, MPt<R>& m) const {
    const R a2 = sqr(k.a);
    if constexpr (KIND == NTT_METRIC_QKERR_SCHILD) {
      const R chi = x1 * k.dchi + k.chi_min;
      m.e   = exp(chi);
      m.r   = k.r0 + m.e;
      m.eta = x2 * k.deta + k.eta_min;
      m.th  = eta2theta(m.eta);
      m.dtde = dtheta_deta(m.eta);
    } else {
      m.e = R(0);
      m.eta = R(0);
      m.dtde = R(1);
      m.r  = x1 * k.dr + k.x1_min;
      m.th = x2 * k.dtheta + k.x2_min;
    }
    sin_cos(m.th, &m.s, &m.c);
    const R s2 = sqr(m.s);
    if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
      m.S = sqr(m.r);
      m.Dl = m.Av = m.z = R(0);
      m.h11 = sqr(k.dr_inv);
      m.h22 = sqr(k.dtheta_inv / m.r);
      m.h33 = R(1) / sqr(m.r * m.s);
      m.h13 = R(0);
      m.h_22 = sqr(k.dtheta) * sqr(m.r);
      m.h_33 = sqr(m.r * m.s);
      m.h_13 = R(0);
      m.alpha = R(1);
      m.beta1 = R(0);
    } else {
      m.S  = sqr(m.r) + a2 * sqr(m.c);
      m.Dl = sqr(m.r) - R(2) * m.r + a2;
      m.Av = sqr(sqr(m.r) + a2) - a2 * m.Dl * s2;
      m.z  = R(2) * m.r / m.S;
      const R opz = R(1) + m.z;
      m.h33  = R(1) / (m.S * s2);
      m.h_33 = m.Av * s2 / m.S;
      m.alpha = R(1) / sqrt(opz);
      if constexpr (KIND == NTT_METRIC_KERR_SCHILD) {
        m.h11  = sqr(k.dr_inv) * m.Av / (m.S * (m.S + R(2) * m.r));
        m.h22  = sqr(k.dtheta_inv) / m.S;
        m.h13  = k.dr_inv * k.a / m.S;
        m.h_22 = sqr(k.dtheta) * m.S;
        m.h_13 = -k.dr * k.a * opz * s2;
        m.beta1 = k.dr_inv * m.z / opz;
      } else {
        const R einv = R(1) / m.e;
        m.h11  = (sqr(einv) / sqr(k.dchi)) * m.Av / (m.S * (m.S + R(2) * m.r));
        m.h22  = R(1) / (m.S * sqr(m.dtde) * sqr(k.deta));
        m.h13  = (einv * k.dchi_inv) * k.a / m.S;
        m.h_22 = sqr(k.deta) * sqr(m.dtde) * m.S;
        m.h_13 = -k.dchi * m.e * k.a * opz * s2;
        m.beta1 = einv * k.dchi_inv * m.z / opz;
      }
    }
  }

  // d/dx1 and d/dx2 of alpha, beta1, h^11, h^22, h^33, h^13 at m (the
  // metrics' dr_* / dt_* functions, same quirks: the quasi-spherical h^22
  // derivatives omit the angle-stretching factor)
  struct Deriv {
    R dr_alpha, dt_alpha, dr_beta1, dt_beta1;
    R dr_h11, dr_h22, dr_h33, dr_h13, dt_h11, dt_h22, dt_h33, dt_h13;
  };
  __device__ __forceinline__ void derivs(const MPt<R>& m, Deriv& d) const {
    const R a2 = sqr(k.a);
    const R s2 = sqr(m.s);
    if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
      d.dr_alpha = d.dt_alpha = d.dr_beta1 = d.dt_beta1 = R(0);
      d.dr_h11 = d.dr_h13 = d.dt_h11 = d.dt_h22 = d.dt_h13 = R(0);
      d.dr_h22 = -R(2) / cube(m.r) * sqr(k.dtheta_inv) * k.dr;
      d.dr_h33 = -R(2) / cube(m.r) / s2 * k.dr;
      d.dt_h33 = -R(2) * m.c / sqr(m.r) / cube(m.s) * k.dtheta;
      return;
    }
    const R S = m.S, r = m.r;
    const R opz = R(1) + m.z;
    const R al3 = cube(m.alpha);
    if constexpr (KIND == NTT_METRIC_KERR_SCHILD) {
      const R dr_S = R(2) * r * k.dr;
      R dt_S = -R(2) * a2 * m.s * m.c * k.dtheta;
      if (almost_zero(dt_S)) dt_S = R(0);
      R dt_A = -R(2) * a2 * m.s * m.c * m.Dl * k.dtheta;
      if (almost_zero(dt_A)) dt_A = R(0);
      const R dr_D = R(2) * k.dr * (r - R(1));
      const R dr_A = R(4) * r * k.dr * (sqr(r) + a2) - a2 * s2 * dr_D;
      d.dr_alpha = -(k.dr * S - r * dr_S) * al3 / sqr(S);
      d.dt_alpha = al3 * r * dt_S / sqr(S);
      d.dr_beta1 = k.dr_inv * R(2) * (k.dr * S - r * dr_S) / sqr(S + R(2) * r);
      d.dt_beta1 = -k.dr_inv * R(2) * r * dt_S / sqr(S * opz);
      const R SS = S * (S + R(2) * r);
      d.dr_h11 = (SS * dr_A - R(2) * m.Av * (r * dr_S + S * (dr_S + k.dr))) / sqr(SS) *
                 sqr(k.dr_inv);
      d.dr_h22 = -dr_S / sqr(S) * sqr(k.dtheta_inv);
      d.dr_h33 = -dr_S / sqr(S) / s2;
      d.dr_h13 = -k.a * dr_S / sqr(S) * k.dr_inv;
      d.dt_h11 = (SS * dt_A - R(2) * m.Av * dt_S * (r + S)) / sqr(SS) * sqr(k.dr_inv);
      d.dt_h22 = -dt_S / sqr(S) * sqr(k.dtheta_inv);
      d.dt_h33 = -R(2) * k.dtheta * m.c * (S - a2 * s2) / cube(m.s) / sqr(S);
      d.dt_h13 = -k.a * dt_S / sqr(S) * k.dr_inv;
    } else {
      const R eta  = m.eta;
      const R einv = R(1) / m.e;
      const R dx_r = k.dchi * m.e;
      const R dr_S = R(2) * r * dx_r;
      const R dxdt = dx_dt(eta);
      // dt_alpha uses the stretching derivative without the AlmostZero cuts
      const R dxdt_inline = k.deta * (R(1) + R(2) * k.h0 * static_cast<R>(kInvPiSqr) *
                                                  (R(2) * R(3) * sqr(eta) -
                                                   R(2) * R(3) * static_cast<R>(kPi) * eta +
                                                   static_cast<R>(kPiSqr)));
      R dt_S = -R(2) * a2 * m.s * m.c * dxdt;
      if (almost_zero(dt_S)) dt_S = R(0);
      R dt_A = -R(2) * a2 * m.s * m.c * m.Dl * dxdt;
      if (almost_zero(dt_A)) dt_A = R(0);
      const R dr_D = R(2) * dx_r * (r - R(1));
      const R dr_A = R(4) * r * dx_r * (sqr(r) + a2) - a2 * s2 * dr_D;
      const R e2inv = sqr(einv) / sqr(k.dchi);
      d.dr_alpha = -(dx_r * S - r * dr_S) * al3 / sqr(S);
      d.dt_alpha = r * (-R(2) * a2 * m.s * m.c * dxdt_inline) * al3 / sqr(S);
      d.dr_beta1 = einv * k.dchi_inv * R(2) * (dx_r * S - r * dr_S) / sqr(S + R(2) * r) -
                   k.dchi * einv * k.dchi_inv * m.z / opz;
      d.dt_beta1 = -einv * k.dchi_inv * R(2) * r * dt_S / sqr(S * opz);
      const R SS = S * (S + R(2) * r);
      d.dr_h11 = (e2inv * (SS * dr_A - R(2) * m.Av * (r * dr_S + S * (dr_S + dx_r))) / sqr(SS)) -
                 R(2) * k.dchi * e2inv * m.Av / SS;
      d.dr_h22 = -dr_S / sqr(S) / sqr(k.deta);
      d.dr_h33 = -dr_S / sqr(S) / s2;
      d.dr_h13 = -k.a * dr_S / sqr(S) * (einv * k.dchi_inv) -
                 k.dchi * (einv * k.dchi_inv) * k.a / S;
      d.dt_h11 = e2inv * (SS * dt_A - R(2) * m.Av * dt_S * (r + S)) / sqr(SS);
      d.dt_h22 = -dt_S / sqr(S) / sqr(k.deta);
      d.dt_h33 = -(dt_S + R(2) * m.c / m.s * S * dxdt) / sqr(S * m.s);
      d.dt_h13 = -k.a * dt_S / sqr(S) * (einv * k.dchi_inv);
    }
  }
};

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

template <typename R, int KIND, int O, bool MASSLESS>
__global__ __launch_bounds__(256) void k_gr_push(GRArgs g, KSMetric<R, KIND> M, R dt, R ndth) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= g.P.n) return;
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

// host: metric constants in R, as the reference's metric constructors
// compute them (kerr_schild.h:79-90, qkerr_schild.h:79-94)
template <typename R>
R theta2eta_host(R theta, R h0) {
  if (almost_zero(h0)) return theta;
  const double hd = static_cast<double>(h0), th = static_cast<double>(theta);
  const double two_pi = 6.28318530717958647692, sqrt3 = 1.73205080756887729352;
  const double Rv = std::pow(
    -9.0 * hd * hd * (kPi - 2.0 * th) +
      sqrt3 * std::sqrt(hd * hd * hd *
                        ((4.0 - hd) * (kPi + hd * two_pi) * (kPi + hd * two_pi) -
                         108.0 * hd * kPi * th + 108.0 * hd * th * th)),
    1.0 / 3.0);
  const double PI_23 = 2.14502939711102560008, PI_13 = 1.46459188756152326302;
  const double TWO_23 = 1.58740105196819947475, THREE_13 = 1.442249570307408382321;
  const double TWO_13 = 1.2599210498948731647672, THREE_PI_23 = 4.46184094890142313715794;
  return static_cast<R>(PI_23 *
                        (6.0 * PI_13 + 2.0 * TWO_13 * (hd - 1.0) * THREE_PI_23 / Rv +
                         TWO_23 * THREE_13 * Rv / hd) /
                        12.0);
}

template <typename R>
ntt_status make_const(const ntt_ks_metric_t* m, KSConst<R>* k) {
  if (m->nx1 < 1 || m->nx2 < 1) return NTT_ERR_ARG;
  if (!(m->x1_max > m->x1_min) || !(m->x2_max > m->x2_min)) return NTT_ERR_ARG;
  k->a      = (m->kind == NTT_METRIC_KERR_SCHILD_0) ? R(0) : static_cast<R>(m->a);
  k->r0     = static_cast<R>(m->r0);
  k->h0     = static_cast<R>(m->h);
  k->x1_min = static_cast<R>(m->x1_min);
  k->x2_min = static_cast<R>(m->x2_min);
  const R x1_max = static_cast<R>(m->x1_max), x2_max = static_cast<R>(m->x2_max);
  k->dr         = (x1_max - k->x1_min) / static_cast<R>(m->nx1);
  k->dtheta     = (x2_max - k->x2_min) / static_cast<R>(m->nx2);
  k->dr_inv     = R(1) / k->dr;
  k->dtheta_inv = R(1) / k->dtheta;
  k->chi_min = k->eta_min = k->dchi = k->deta = k->dchi_inv = k->deta_inv = R(0);
  if (m->kind == NTT_METRIC_QKERR_SCHILD) {
    if (!(k->x1_min > k->r0)) return NTT_ERR_ARG;
    k->chi_min  = std::log(k->x1_min - k->r0);
    k->eta_min  = theta2eta_host(k->x2_min, k->h0);
    k->dchi     = (std::log(x1_max - k->r0) - k->chi_min) / static_cast<R>(m->nx1);
    k->deta     = (theta2eta_host(x2_max, k->h0) - k->eta_min) / static_cast<R>(m->nx2);
    k->dchi_inv = R(1) / k->dchi;
    k->deta_inv = R(1) / k->deta;
  }
  k->small_angle = static_cast<R>(kSmallAngleGR);
  k->large_angle = kPi - static_cast<double>(k->small_angle);
  const R large  = static_cast<R>(k->large_angle);
  if (m->kind == NTT_METRIC_QKERR_SCHILD) {
    k->th_small_cd = (theta2eta_host(k->small_angle, k->h0) - k->eta_min) * k->deta_inv;
    k->th_large_cd = (theta2eta_host(large, k->h0) - k->eta_min) * k->deta_inv;
  } else {
    k->th_small_cd = (k->small_angle - k->x2_min) * k->dtheta_inv;
    k->th_large_cd = (large - k->x2_min) * k->dtheta_inv;
  }
  return NTT_OK;
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

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_gr_push(const ntt_gr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                                  ntt_particles_t* prtls, const ntt_field_t* em,
                                  const ntt_field_t* em0, const ntt_ks_metric_t* metric,
                                  int32_t shape_order, void* stream) {
  if (ctx == nullptr || bc == nullptr || prtls == nullptr || metric == nullptr)
    return NTT_ERR_ARG;
  if (shape_order < 0 || shape_order > 2) return NTT_ERR_UNSUPPORTED;
  if (metric->kind != NTT_METRIC_KERR_SCHILD && metric->kind != NTT_METRIC_QKERR_SCHILD &&
      metric->kind != NTT_METRIC_KERR_SCHILD_0)
    return NTT_ERR_UNSUPPORTED;
  // grpic::ParticlePush (particle_pusher.h:53-87): PHOTON -> massless,
  // BORIS -> massive, anything else is not implemented
  bool massless;
  if (ctx->pusher_flags == NTT_PUSHER_PHOTON) massless = true;
  else if (ctx->pusher_flags == NTT_PUSHER_BORIS) massless = false;
  else if (ctx->pusher_flags == NTT_PUSHER_NONE) return NTT_OK;
  else return NTT_ERR_UNSUPPORTED;
  if (ctx->niter < 0) return NTT_ERR_ARG;
  if (!massless && !(ctx->mass != 0.0f)) return NTT_ERR_ARG;
  if (ctx->precision != 0 && ctx->precision != 32 && ctx->precision != 64)
    return NTT_ERR_ARG;
  GRArgs g{};
  ntt_status s = make_particle_view(prtls, 2, true, &g.P);
  if (s != NTT_OK) return s;
  if (g.P.n == 0) return NTT_OK;
  if (!massless) {
    if (em == nullptr || em0 == nullptr) return NTT_ERR_ARG;
    s = make_field_view(em, 6, &g.DB);
    if (s != NTT_OK) return s;
    s = make_field_view(em0, 6, &g.DB0);
    if (s != NTT_OK) return s;
    if (em->dim != 2 || em0->dim != 2) return NTT_ERR_UNSUPPORTED;
    if (g.DB.N1 != g.DB0.N1 || g.DB.N2 != g.DB0.N2 || g.DB.ngh != g.DB0.ngh)
      return NTT_ERR_ARG;
    if (g.DB.n1 != ctx->ni1 || g.DB.n2 != ctx->ni2) return NTT_ERR_ARG;
  }
  g.ni1 = ctx->ni1;
  g.ni2 = ctx->ni2;
  g.niter = ctx->niter;
  // kernel::gr::PusherBoundaries (context.h:213-228)
  g.absorb_i1min = bc->bc[0][0] == NTT_PRTL_BC_ABSORB || bc->bc[0][0] == NTT_PRTL_BC_HORIZON;
  g.absorb_i1max = bc->bc[0][1] == NTT_PRTL_BC_ABSORB || bc->bc[0][1] == NTT_PRTL_BC_HORIZON;
  g.axis_i2min   = bc->bc[1][0] == NTT_PRTL_BC_AXIS;
  g.axis_i2max   = bc->bc[1][1] == NTT_PRTL_BC_AXIS;
  g.err = device_error_word();
  if (g.err == nullptr) return NTT_ERR_HIP;
  hipStream_t st = as_stream(stream);
  if (ctx->precision == 32) return run_gr<float>(ctx, g, metric, shape_order, massless, st);
  return run_gr<double>(ctx, g, metric, shape_order, massless, st);
}
