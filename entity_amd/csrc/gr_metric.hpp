// gr_metric.hpp -- the Kerr-Schild family of metrics on a 2D axisymmetric
// (r, theta) grid, evaluated on the device in the real type R: the pusher's
// quantities (lapse, shift, inverse metric, their derivatives) and the field
// solver's (lower metric components, sqrt(det h), the polar area).
//
// Reference: metric::KerrSchild (src/metrics/kerr_schild.h:30-759),
// metric::QKerrSchild (src/metrics/qkerr_schild.h:28-914), metric::KerrSchild0
// (src/metrics/kerr_schild_0.h).
#pragma once

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
  // polar_area: the angle in its factor and whether the small-angle form is
  // used (kerr_schild.h:519-532, qkerr_schild.h:590-609)
  R      polar_delta;
  int    polar_small;
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
  // convert<1, Cd, Ph>: the radius of code x1
  __device__ __forceinline__ R r_of(R x1) const {
    if constexpr (KIND == NTT_METRIC_QKERR_SCHILD) return k.r0 + exp(x1 * k.dchi + k.chi_min);
    else return x1 * k.dr + k.x1_min;
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

  // The field solver's quantities at a point: lapse, shift, lower metric
  // components, sqrt(det h) and sqrt(det h) / sin(theta)
  // (kerr_schild.h:143-177,294-330,486-512; qkerr_schild.h:145-190,
  // 553-582; kerr_schild_0.h:127-140,346-365)
  struct FPt {
    R alpha, beta1, h_11, h_22, h_33, h_13, sdh, sdht;
  };
  __device__ __forceinline__ FPt fpoint(R x1, R x2) const {
    MPt<R> m;
    eval(x1, x2, m);
    FPt f;
    f.alpha = m.alpha;
    f.beta1 = m.beta1;
    f.h_22  = m.h_22;
    f.h_33  = m.h_33;
    f.h_13  = m.h_13;
    if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
      f.h_11 = sqr(k.dr);
      f.sdht = k.dr * k.dtheta * sqr(m.r);
    } else if constexpr (KIND == NTT_METRIC_KERR_SCHILD) {
      f.h_11 = sqr(k.dr) * (R(1) + m.z);
      f.sdht = k.dr * k.dtheta * m.S * sqrt(R(1) + m.z);
    } else {
      f.h_11 = sqr(k.dchi) * sqr(m.e) * (R(1) + m.z);
      f.sdht = k.dchi * m.e * m.dtde * k.deta * m.S * sqrt(R(1) + m.z);
    }
    f.sdh = f.sdht * m.s;
    return f;
  }
  // differential area at the pole, at radial code coordinate x1
  __device__ __forceinline__ R polar_area(R x1) const {
    const R dth = k.polar_delta;
    if constexpr (KIND == NTT_METRIC_KERR_SCHILD_0) {
      const R r = x1 * k.dr + k.x1_min;
      return k.dr * sqr(r) * (R(1) - cos(R(0.5) * dth));
    } else {
      R dx, r;
      if constexpr (KIND == NTT_METRIC_QKERR_SCHILD) {
        const R e = exp(x1 * k.dchi + k.chi_min);
        dx = k.dchi * e;
        r  = k.r0 + e;
      } else {
        dx = k.dr;
        r  = x1 * k.dr + k.x1_min;
      }
      const R ra = sqr(r) + sqr(k.a);
      const R ang = k.polar_small ? (R(48) - sqr(dth)) * sqr(dth) / R(384)
                                  : (R(1) - cos(KIND == NTT_METRIC_QKERR_SCHILD ? dth
                                                                               : R(0.5) * dth));
      return dx * ra * sqrt(R(1) + R(2) * r / ra) * ang;
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
  // SMALL_ANGLE = 1e-3 (numeric.h:102) decides the polar area's form:
  // KerrSchild compares dtheta / 2 and uses dtheta in the expansion,
  // QKerrSchild compares and uses theta(deta / 2)
  constexpr double kSmallAnglePolar = 1e-3;
  if (m->kind == NTT_METRIC_QKERR_SCHILD) {
    const R half = R(0.5) * k->deta;
    // eta2theta (qkerr_schild.h:845-853) of half a cell
    const R th = almost_zero(k->h0)
                   ? half
                   : half + R(2) * k->h0 * half * (static_cast<R>(kPi) - R(2) * half) *
                              (static_cast<R>(kPi) - half) * static_cast<R>(kInvPiSqr);
    k->polar_delta = th;
    k->polar_small = static_cast<double>(th) < kSmallAnglePolar;
  } else if (m->kind == NTT_METRIC_KERR_SCHILD) {
    k->polar_delta = k->dtheta;
    k->polar_small = static_cast<double>(R(0.5) * k->dtheta) < kSmallAnglePolar;
  } else {
    k->polar_delta = k->dtheta;
    k->polar_small = 0;
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

}  // namespace
}  // namespace ntt
