/*
 * ntt_oracle_gr.cpp -- CPU restatement of Entity's GR particle pusher over
 * the Kerr-Schild family of metrics (2D axisymmetric, r-theta).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP GR pusher.  Product code
 * never links or calls this file.
 *
 * Metric functions follow src/metrics/kerr_schild.h, qkerr_schild.h and
 * kerr_schild_0.h, function by function, in their operation order; the
 * pusher follows src/kernels/pushers/gr.hpp:177-865.  Every function is a
 * template on the real type (float: the reference's SINGLE_PRECISION build;
 * double: its double build).
 */
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "ntt_oracle_common.hpp"

namespace {

using namespace ort;

constexpr double PI_D         = 3.14159265358979323846;
constexpr double INV_PI_D     = 0.31830988618379067154;
constexpr double PI_SQR_D     = 9.86960440108935861882;
constexpr double INV_PI_SQR_D = 0.10132118364233777144;
constexpr double SQRT3_D      = 1.73205080756887729352;
constexpr double SMALL_ANGLE_GR_D = 1e-5;

template <typename R> inline R sqr(R x) { return x * x; }
template <typename R> inline R cube(R x) { return x * x * x; }

// cmp::AlmostZero with the machine epsilon of R (utils/comparators.h:38-42)
template <typename R> inline bool almost_zero(R a) {
  return std::fabs(a) <= std::numeric_limits<R>::epsilon();
}

/* ---- metrics --------------------------------------------------------------
 * kind 0: KerrSchild (kerr_schild.h), 1: QKerrSchild (qkerr_schild.h),
 * 2: KerrSchild0 (kerr_schild_0.h).  Coordinates x are code units. */
template <typename R>
struct GRMetric {
  int kind;
  R   a, r0, h0;
  R   x1_min, x2_min;
  // KS / KS0
  R dr, dtheta, dr_inv, dtheta_inv;
  // QKS
  R chi_min, eta_min, dchi, deta, dchi_inv, deta_inv;

  explicit GRMetric(const ort_gr_params_t* p)
    : kind(p->metric_kind), a(static_cast<R>(p->a)), r0(static_cast<R>(p->r0)),
      h0(static_cast<R>(p->h)), x1_min(static_cast<R>(p->x1_min)),
      x2_min(static_cast<R>(p->x2_min)) {
    const R x1_max = static_cast<R>(p->x1_max), x2_max = static_cast<R>(p->x2_max);
    if (kind == 2) a = K<R>::ZERO;
    // kerr_schild.h:79-90, kerr_schild_0.h:47-56
    dr         = (x1_max - x1_min) / static_cast<R>(p->nx1);
    dtheta     = (x2_max - x2_min) / static_cast<R>(p->nx2);
    dr_inv     = K<R>::ONE / dr;
    dtheta_inv = K<R>::ONE / dtheta;
    // qkerr_schild.h:79-94
    chi_min = eta_min = dchi = deta = dchi_inv = deta_inv = K<R>::ZERO;
    if (kind == 1) {
      chi_min  = std::log(x1_min - r0);
      eta_min  = theta2eta(x2_min);
      dchi     = (std::log(x1_max - r0) - chi_min) / static_cast<R>(p->nx1);
      deta     = (theta2eta(x2_max) - eta_min) / static_cast<R>(p->nx2);
      dchi_inv = K<R>::ONE / dchi;
      deta_inv = K<R>::ONE / deta;
    }
  }

  // Kerr functions (kerr_schild.h:44-58)
  R Delta(R r) const { return sqr(r) - K<R>::TWO * r + sqr(a); }
  R Sigma(R r, R th) const { return sqr(r) + sqr(a) * sqr(std::cos(th)); }
  R A(R r, R th) const {
    return sqr(sqr(r) + sqr(a)) - sqr(a) * Delta(r) * sqr(std::sin(th));
  }
  R z(R r, R th) const { return K<R>::TWO * r / Sigma(r, th); }

  // quasi-spherical angle stretching (qkerr_schild.h:857-914)
  R dtheta_deta(R eta) const {
    if (almost_zero(h0)) return K<R>::ONE;
    return (K<R>::ONE + K<R>::TWO * h0 +
            static_cast<R>(12.0) * h0 * (eta * static_cast<R>(INV_PI_D)) *
              ((eta * static_cast<R>(INV_PI_D)) - K<R>::ONE));
  }
  R eta2theta(R eta) const {
    if (almost_zero(h0)) return eta;
    return eta + K<R>::TWO * h0 * eta * (static_cast<R>(PI_D) - K<R>::TWO * eta) *
                   (static_cast<R>(PI_D) - eta) * static_cast<R>(INV_PI_SQR_D);
  }
  R dx_dt(R eta) const {
    if (almost_zero(h0)) return deta;
    return deta * (K<R>::ONE + K<R>::TWO * h0 * static_cast<R>(INV_PI_SQR_D) *
                                 (K<R>::TWO * K<R>::THREE * sqr(eta) -
                                  K<R>::TWO * K<R>::THREE * static_cast<R>(PI_D) * eta +
                                  static_cast<R>(PI_SQR_D)));
  }
  // closed-form inverse of eta2theta, evaluated in double (qkerr_schild.h:889-913)
  R theta2eta(R theta) const {
    if (almost_zero(h0)) return theta;
    const double hd = static_cast<double>(h0), th = static_cast<double>(theta);
    const double TWO_PI_D = 6.28318530717958647692;
    const double Rv = std::pow(
      -9.0 * hd * hd * (PI_D - 2.0 * th) +
        SQRT3_D * std::sqrt((hd * hd * hd *
                             ((4.0 - hd) * (PI_D + hd * TWO_PI_D) * (PI_D + hd * TWO_PI_D) -
                              108.0 * hd * PI_D * th + 108.0 * hd * th * th))),
      1.0 / 3.0);
    const double PI_23 = 2.14502939711102560008, PI_13 = 1.46459188756152326302;
    const double TWO_23 = 1.58740105196819947475, THREE_13 = 1.442249570307408382321;
    const double TWO_13 = 1.2599210498948731647672, THREE_PI_23 = 4.46184094890142313715794;
    return static_cast<R>(PI_23 *
                          (6.0 * PI_13 + 2.0 * TWO_13 * (hd - 1.0) * THREE_PI_23 / Rv +
                           TWO_23 * THREE_13 * Rv / hd) /
                          12.0);
  }

  // physical r, theta of code coordinates
  R r_of(const R* x) const {
    if (kind == 1) return r0 + std::exp(x[0] * dchi + chi_min);
    return x[0] * dr + x1_min;
  }
  R th_of(const R* x) const {
    if (kind == 1) return eta2theta(x[1] * deta + eta_min);
    return x[1] * dtheta + x2_min;
  }
  // convert<2, Cd, Ph> and convert<2, Ph, Cd>
  R theta_cd2ph(R x) const {
    if (kind == 1) return eta2theta(x * deta + eta_min);
    return x * dtheta + x2_min;
  }
  R theta_ph2cd(R th) const {
    if (kind == 1) return (theta2eta(th) - eta_min) * deta_inv;
    return (th - x2_min) * dtheta_inv;
  }

  // ---- h_ij (lower) ----
  R h_11(const R* x) const {
    if (kind == 0) return sqr(dr) * (K<R>::ONE + z(x[0] * dr + x1_min, x[1] * dtheta + x2_min));
    if (kind == 2) return sqr(dr);
    return sqr(dchi) * std::exp(K<R>::TWO * (x[0] * dchi + chi_min)) *
           (K<R>::ONE + z(r0 + std::exp(x[0] * dchi + chi_min), eta2theta(x[1] * deta + eta_min)));
  }
  R h_22(const R* x) const {
    if (kind == 0) return sqr(dtheta) * Sigma(x[0] * dr + x1_min, x[1] * dtheta + x2_min);
    if (kind == 2) return sqr(dtheta) * sqr(x[0] * dr + x1_min);
    return sqr(deta) * sqr(dtheta_deta(x[1] * deta + eta_min)) *
           Sigma(r0 + std::exp(x[0] * dchi + chi_min), eta2theta(x[1] * deta + eta_min));
  }
  R h_33(const R* x) const {
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return A(r, th) * sqr(std::sin(th)) / Sigma(r, th);
    }
    if (kind == 2) return sqr((x[0] * dr + x1_min) * std::sin(x[1] * dtheta + x2_min));
    const R th = eta2theta(x[1] * deta + eta_min);
    return A(r0 + std::exp(x[0] * dchi + chi_min), th) * sqr(std::sin(th)) /
           Sigma(r0 + std::exp(x[0] * dchi + chi_min), th);
  }
  R h_13(const R* x) const {
    if (kind == 0)
      return -dr * a * (K<R>::ONE + z(x[0] * dr + x1_min, x[1] * dtheta + x2_min)) *
             sqr(std::sin(x[1] * dtheta + x2_min));
    if (kind == 2) return K<R>::ZERO;
    const R th = eta2theta(x[1] * deta + eta_min);
    return -dchi * std::exp(x[0] * dchi + chi_min) * a *
           (K<R>::ONE + z(r0 + std::exp(x[0] * dchi + chi_min), th)) * sqr(std::sin(th));
  }
  // ---- h^ij (upper) ----
  R h11(const R* x) const {
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      const R S = Sigma(r, th);
      return sqr(dr_inv) * A(r, th) / (S * (S + K<R>::TWO * r));
    }
    if (kind == 2) return sqr(dr_inv);
    const R th = eta2theta(x[1] * deta + eta_min);
    const R r  = r0 + std::exp(x[0] * dchi + chi_min);
    const R S  = Sigma(r, th);
    return (std::exp(-K<R>::TWO * (x[0] * dchi + chi_min)) / sqr(dchi)) * A(r, th) /
           (S * (S + K<R>::TWO * r));
  }
  R h22(const R* x) const {
    if (kind == 0) return sqr(dtheta_inv) / Sigma(x[0] * dr + x1_min, x[1] * dtheta + x2_min);
    if (kind == 2) return sqr(dtheta_inv / (x[0] * dr + x1_min));
    const R th = eta2theta(x[1] * deta + eta_min);
    return K<R>::ONE / (Sigma(r0 + std::exp(x[0] * dchi + chi_min), th) *
                        sqr(dtheta_deta(x[1] * deta + eta_min)) * sqr(deta));
  }
  R h33(const R* x) const {
    if (kind == 0) {
      const R th = x[1] * dtheta + x2_min;
      return K<R>::ONE / (Sigma(x[0] * dr + x1_min, th) * sqr(std::sin(th)));
    }
    if (kind == 2) return K<R>::ONE / sqr((x[0] * dr + x1_min) * std::sin(x[1] * dtheta + x2_min));
    const R th = eta2theta(x[1] * deta + eta_min);
    return K<R>::ONE / (Sigma(r0 + std::exp(x[0] * dchi + chi_min), th) * sqr(std::sin(th)));
  }
  R h13(const R* x) const {
    if (kind == 0) return dr_inv * a / Sigma(x[0] * dr + x1_min, x[1] * dtheta + x2_min);
    if (kind == 2) return K<R>::ZERO;
    const R th = eta2theta(x[1] * deta + eta_min);
    return (std::exp(-(x[0] * dchi + chi_min)) * dchi_inv) * a /
           Sigma(r0 + std::exp(x[0] * dchi + chi_min), th);
  }
  // ---- sqrt(det h), sqrt(det h) / sin(theta), polar area ----
  // kerr_schild.h:486-532, qkerr_schild.h:553-609, kerr_schild_0.h:346-372
  R sqrt_det_h_tilde(const R* x) const {
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return dr * dtheta * Sigma(r, th) * std::sqrt(K<R>::ONE + z(r, th));
    }
    if (kind == 2) return dr * dtheta * sqr(x[0] * dr + x1_min);
    const R expchi = std::exp(x[0] * dchi + chi_min);
    const R th     = eta2theta(x[1] * deta + eta_min);
    return dchi * expchi * dtheta_deta(x[1] * deta + eta_min) * deta * Sigma(r0 + expchi, th) *
           std::sqrt(K<R>::ONE + z(r0 + expchi, th));
  }
  R sqrt_det_h(const R* x) const {
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return dr * dtheta * Sigma(r, th) * std::sin(th) * std::sqrt(K<R>::ONE + z(r, th));
    }
    if (kind == 2) return dr * dtheta * sqr(x[0] * dr + x1_min) * std::sin(x[1] * dtheta + x2_min);
    const R expchi = std::exp(x[0] * dchi + chi_min);
    const R th     = eta2theta(x[1] * deta + eta_min);
    return dchi * expchi * dtheta_deta(x[1] * deta + eta_min) * deta * Sigma(r0 + expchi, th) *
           std::sin(th) * std::sqrt(K<R>::ONE + z(r0 + expchi, th));
  }
  R polar_area(R x1) const {
    constexpr double SMALL_ANGLE_D = 1e-3;  // numeric.h:102
    if (kind == 2) return dr * sqr(x1 * dr + x1_min) * (K<R>::ONE - std::cos(K<R>::HALF * dtheta));
    R dx, r, ang;
    if (kind == 0) {
      dx = dr;
      r  = x1 * dr + x1_min;
      const bool small = static_cast<double>(K<R>::HALF * dtheta) < SMALL_ANGLE_D;
      ang = small ? (static_cast<R>(48) - sqr(dtheta)) * sqr(dtheta) / static_cast<R>(384)
                  : (K<R>::ONE - std::cos(K<R>::HALF * dtheta));
    } else {
      const R e = std::exp(x1 * dchi + chi_min);
      dx = dchi * e;
      r  = r0 + e;
      const R dth = eta2theta(K<R>::HALF * deta);
      const bool small = static_cast<double>(dth) < SMALL_ANGLE_D;
      ang = small ? (static_cast<R>(48) - sqr(dth)) * sqr(dth) / static_cast<R>(384)
                  : (K<R>::ONE - std::cos(dth));
    }
    const R ra = sqr(r) + sqr(a);
    return dx * ra * std::sqrt(K<R>::ONE + K<R>::TWO * r / ra) * ang;
  }
  // ---- lapse, shift ----
  R alpha(const R* x) const {
    if (kind == 2) return K<R>::ONE;
    return K<R>::ONE / std::sqrt(K<R>::ONE + z(r_of(x), th_of(x)));
  }
  R beta1(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R z_ = z(x[0] * dr + x1_min, x[1] * dtheta + x2_min);
      return dr_inv * z_ / (K<R>::ONE + z_);
    }
    const R chi = x[0] * dchi + chi_min;
    const R z_  = z(r0 + std::exp(chi), eta2theta(x[1] * deta + eta_min));
    return std::exp(-chi) * dchi_inv * z_ / (K<R>::ONE + z_);
  }
  // d/d(theta) of Sigma and A (with the AlmostZero cut of the reference)
  R dt_Sigma_ks(R th) const {
    const R v = -K<R>::TWO * sqr(a) * std::sin(th) * std::cos(th) * dtheta;
    return almost_zero(v) ? K<R>::ZERO : v;
  }
  R dt_A_ks(R r, R th) const {
    const R v = -K<R>::TWO * sqr(a) * std::sin(th) * std::cos(th) * Delta(r) * dtheta;
    return almost_zero(v) ? K<R>::ZERO : v;
  }
  R dt_Sigma_q(R eta) const {
    const R th = eta2theta(eta);
    const R v  = -K<R>::TWO * sqr(a) * std::sin(th) * std::cos(th) * dx_dt(eta);
    return almost_zero(v) ? K<R>::ZERO : v;
  }
  R dt_A_q(R r, R eta) const {
    const R th = eta2theta(eta);
    const R v  = -K<R>::TWO * sqr(a) * std::sin(th) * std::cos(th) * Delta(r) * dx_dt(eta);
    return almost_zero(v) ? K<R>::ZERO : v;
  }

  R dr_alpha(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    const R r = r_of(x), th = th_of(x);
    const R dx_r = (kind == 0) ? dr : dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S = K<R>::TWO * r * dx_r;
    return -(dx_r * Sigma(r, th) - r * dr_S) * cube(alpha(x)) / sqr(Sigma(r, th));
  }
  R dt_alpha(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return cube(alpha(x)) * r * dt_Sigma_ks(th) / sqr(Sigma(r, th));
    }
    const R r   = r0 + std::exp(x[0] * dchi + chi_min);
    const R eta = x[1] * deta + eta_min;
    const R th  = eta2theta(eta);
    const R dxdt = deta * (K<R>::ONE + K<R>::TWO * h0 * static_cast<R>(INV_PI_SQR_D) *
                                         (K<R>::TWO * K<R>::THREE * sqr(eta) -
                                          K<R>::TWO * K<R>::THREE * static_cast<R>(PI_D) * eta +
                                          static_cast<R>(PI_SQR_D)));
    const R dt_S = -K<R>::TWO * sqr(a) * std::sin(th) * std::cos(th) * dxdt;
    return r * dt_S * cube(alpha(x)) / sqr(Sigma(r, th));
  }
  R dr_beta1(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      const R dr_S = K<R>::TWO * r * dr;
      return dr_inv * K<R>::TWO * (dr * Sigma(r, th) - r * dr_S) /
             sqr(Sigma(r, th) + K<R>::TWO * r);
    }
    const R chi  = x[0] * dchi + chi_min;
    const R r    = r0 + std::exp(chi);
    const R th   = eta2theta(x[1] * deta + eta_min);
    const R z_   = z(r, th);
    const R dx_r = dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S = K<R>::TWO * r * dx_r;
    return std::exp(-chi) * dchi_inv * K<R>::TWO * (dx_r * Sigma(r, th) - r * dr_S) /
             sqr(Sigma(r, th) + K<R>::TWO * r) -
           dchi * std::exp(-chi) * dchi_inv * z_ / (K<R>::ONE + z_);
  }
  R dt_beta1(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -dr_inv * K<R>::TWO * r * dt_Sigma_ks(th) /
             sqr(Sigma(r, th) * (K<R>::ONE + z(r, th)));
    }
    const R chi = x[0] * dchi + chi_min;
    const R r   = r0 + std::exp(chi);
    const R eta = x[1] * deta + eta_min;
    const R th  = eta2theta(eta);
    return -std::exp(-chi) * dchi_inv * K<R>::TWO * r * dt_Sigma_q(eta) /
           sqr(Sigma(r, th) * (K<R>::ONE + z(r, th)));
  }
  R dr_h11(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    const R r = r_of(x), th = th_of(x);
    const R dx_r  = (kind == 0) ? dr : dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S  = K<R>::TWO * r * dx_r;
    const R dr_D  = K<R>::TWO * dx_r * (r - K<R>::ONE);
    const R dr_A  = K<R>::FOUR * r * dx_r * (sqr(r) + sqr(a)) - sqr(a) * sqr(std::sin(th)) * dr_D;
    const R S = Sigma(r, th), Av = A(r, th);
    if (kind == 0)
      return (S * (S + K<R>::TWO * r) * dr_A -
              K<R>::TWO * Av * (r * dr_S + S * (dr_S + dr))) /
             (sqr(S * (S + K<R>::TWO * r))) * sqr(dr_inv);
    return (std::exp(-K<R>::TWO * (x[0] * dchi + chi_min)) / sqr(dchi) *
            (S * (S + K<R>::TWO * r) * dr_A -
             K<R>::TWO * Av * (r * dr_S + S * (dr_S + dx_r))) /
            (sqr(S * (S + K<R>::TWO * r)))) -
           K<R>::TWO * dchi * std::exp(-K<R>::TWO * (x[0] * dchi + chi_min)) / sqr(dchi) * Av /
             (S * (S + K<R>::TWO * r));
  }
  R dr_h22(const R* x) const {
    if (kind == 2) {
      const R r = x[0] * dr + x1_min;
      return -K<R>::TWO / cube(r) * sqr(dtheta_inv) * dr;
    }
    const R r = r_of(x), th = th_of(x);
    const R dx_r = (kind == 0) ? dr : dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S = K<R>::TWO * r * dx_r;
    if (kind == 0) return -dr_S / sqr(Sigma(r, th)) * sqr(dtheta_inv);
    return -dr_S / sqr(Sigma(r, th)) / sqr(deta);
  }
  R dr_h33(const R* x) const {
    if (kind == 2) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -K<R>::TWO / cube(r) / sqr(std::sin(th)) * dr;
    }
    const R r = r_of(x), th = th_of(x);
    const R dx_r = (kind == 0) ? dr : dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S = K<R>::TWO * r * dx_r;
    return -dr_S / sqr(Sigma(r, th)) / sqr(std::sin(th));
  }
  R dr_h13(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    const R r = r_of(x), th = th_of(x);
    if (kind == 0) {
      const R dr_S = K<R>::TWO * r * dr;
      return -a * dr_S / sqr(Sigma(r, th)) * dr_inv;
    }
    const R dx_r = dchi * std::exp(x[0] * dchi + chi_min);
    const R dr_S = K<R>::TWO * r * dx_r;
    return -a * dr_S / sqr(Sigma(r, th)) * (std::exp(-(x[0] * dchi + chi_min)) * dchi_inv) -
           dchi * (std::exp(-(x[0] * dchi + chi_min)) * dchi_inv) * a / Sigma(r, th);
  }
  R dt_h11(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      const R S = Sigma(r, th);
      return (S * (S + K<R>::TWO * r) * dt_A_ks(r, th) -
              K<R>::TWO * A(r, th) * dt_Sigma_ks(th) * (r + S)) /
             (sqr(S * (S + K<R>::TWO * r))) * sqr(dr_inv);
    }
    const R r   = r0 + std::exp(x[0] * dchi + chi_min);
    const R eta = x[1] * deta + eta_min;
    const R th  = eta2theta(eta);
    const R S   = Sigma(r, th);
    return std::exp(-K<R>::TWO * (x[0] * dchi + chi_min)) / sqr(dchi) *
           (S * (S + K<R>::TWO * r) * dt_A_q(r, eta) -
            K<R>::TWO * A(r, th) * dt_Sigma_q(eta) * (r + S)) /
           (sqr(S * (S + K<R>::TWO * r)));
  }
  R dt_h22(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -dt_Sigma_ks(th) / sqr(Sigma(r, th)) * sqr(dtheta_inv);
    }
    const R r   = r0 + std::exp(x[0] * dchi + chi_min);
    const R eta = x[1] * deta + eta_min;
    return -dt_Sigma_q(eta) / sqr(Sigma(r, eta2theta(eta))) / sqr(deta);
  }
  R dt_h33(const R* x) const {
    if (kind == 2) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -K<R>::TWO * std::cos(th) / sqr(r) / cube(std::sin(th)) * dtheta;
    }
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -K<R>::TWO * dtheta * std::cos(th) * (Sigma(r, th) - sqr(a) * sqr(std::sin(th))) /
             cube(std::sin(th)) / sqr(Sigma(r, th));
    }
    const R r   = r0 + std::exp(x[0] * dchi + chi_min);
    const R eta = x[1] * deta + eta_min;
    const R th  = eta2theta(eta);
    return -(dt_Sigma_q(eta) + K<R>::TWO * std::cos(th) / std::sin(th) * Sigma(r, th) * dx_dt(eta)) /
           sqr(Sigma(r, th) * std::sin(th));
  }
  R dt_h13(const R* x) const {
    if (kind == 2) return K<R>::ZERO;
    if (kind == 0) {
      const R r = x[0] * dr + x1_min, th = x[1] * dtheta + x2_min;
      return -a * dt_Sigma_ks(th) / sqr(Sigma(r, th)) * dr_inv;
    }
    const R r   = r0 + std::exp(x[0] * dchi + chi_min);
    const R eta = x[1] * deta + eta_min;
    return -a * dt_Sigma_q(eta) / sqr(Sigma(r, eta2theta(eta))) *
           (std::exp(-(x[0] * dchi + chi_min)) * dchi_inv);
  }

  // ---- vector transformations (kerr_schild.h:599-666, kerr_schild_0.h) ----
  void cov_to_cntrv(const R* x, const R* vi, R* vo) const {
    vo[0] = vi[0] * h11(x) + vi[2] * h13(x);
    vo[1] = vi[1] * h22(x);
    vo[2] = vi[0] * h13(x) + vi[2] * h33(x);
  }
  void cntrv_to_cov(const R* x, const R* vi, R* vo) const {
    vo[0] = vi[0] * h_11(x) + vi[2] * h_13(x);
    vo[1] = vi[1] * h_22(x);
    vo[2] = vi[0] * h_13(x) + vi[2] * h_33(x);
  }
  void cntrv_to_hat(const R* x, const R* vi, R* vo) const {
    if (kind == 2) {
      vo[0] = vi[0] / std::sqrt(h11(x));
      vo[1] = vi[1] * std::sqrt(h_22(x));
      vo[2] = vi[2] * std::sqrt(h_33(x));
      return;
    }
    vo[0] = vi[0] / std::sqrt(h11(x));
    vo[1] = vi[1] * std::sqrt(h_22(x));
    vo[2] = vi[2] * std::sqrt(h_33(x)) + vi[0] * h_13(x) / std::sqrt(h_33(x));
  }
  void cov_to_hat(const R* x, const R* vi, R* vo) const {
    if (kind == 2) {
      vo[0] = vi[0] * std::sqrt(h11(x));
      vo[1] = vi[1] / std::sqrt(h_22(x));
      vo[2] = vi[2] / std::sqrt(h_33(x));
      return;
    }
    const R A0 = std::sqrt(h11(x));
    vo[0] = vi[0] * A0 - vi[2] * A0 * h_13(x) / h_33(x);
    vo[1] = vi[1] / std::sqrt(h_22(x));
    vo[2] = vi[2] / std::sqrt(h_33(x));
  }
  void hat_to_cov(const R* x, const R* vi, R* vo) const {
    if (kind == 2) {
      vo[0] = vi[0] / std::sqrt(h11(x));
      vo[1] = vi[1] * std::sqrt(h_22(x));
      vo[2] = vi[2] * std::sqrt(h_33(x));
      return;
    }
    vo[0] = vi[0] / std::sqrt(h11(x)) + vi[2] * h_13(x) / std::sqrt(h_33(x));
    vo[1] = vi[1] * std::sqrt(h_22(x));
    vo[2] = vi[2] * std::sqrt(h_33(x));
  }
  void hat_to_cntrv(const R* x, const R* vi, R* vo) const {
    if (kind == 2) {
      vo[0] = vi[0] * std::sqrt(h11(x));
      vo[1] = vi[1] / std::sqrt(h_22(x));
      vo[2] = vi[2] / std::sqrt(h_33(x));
      return;
    }
    const R A0 = std::sqrt(h11(x));
    vo[0] = vi[0] * A0;
    vo[1] = vi[1] / std::sqrt(h_22(x));
    vo[2] = vi[2] / std::sqrt(h_33(x)) - vi[0] * A0 * h_13(x) / h_33(x);
  }
};

/* ---- GR pusher: src/kernels/pushers/gr.hpp ------------------------------ */
template <typename R>
struct GRPusher {
  const ort_gr_params_t* prm;
  GRMetric<R>            metric;
  Fld<R>                 DB, DB0;
  Prt<R>                 P;
  R                      dt, ndth;
  int                    order;

  GRPusher(const ort_gr_params_t* p, ort_particles_t* prtls, const ort_field_t* em,
           const ort_field_t* em0, int order_)
    : prm(p), metric(p), DB(em), DB0(em0), P(prtls), dt(static_cast<R>(p->dt)), order(order_) {
    // gr.hpp:108-109
    ndth = K<R>::HALF * (p->charge / p->mass) * static_cast<R>(p->omegaB0) * dt;
  }

  // computeGamma (gr.hpp:236-254)
  R gamma(bool massless, const R* ucov, const R* ucnt) const {
    const R s = ucov[0] * ucnt[0] + ucov[1] * ucnt[1] + ucov[2] * ucnt[2];
    return massless ? std::sqrt(s) : std::sqrt(K<R>::ONE + s);
  }

  // theta clamp of the midpoint (gr.hpp:355-368 and 742-756)
  R clamp_theta(R theta_cd) const {
    const R      small = static_cast<R>(SMALL_ANGLE_GR_D);
    const double large = PI_D - static_cast<double>(small);
    const R      th_ph = metric.theta_cd2ph(theta_cd);
    if (th_ph < small) return metric.theta_ph2cd(small);
    if (static_cast<double>(th_ph) >= large) return metric.theta_ph2cd(static_cast<R>(large));
    return theta_cd;
  }

  // EMHalfPush (gr.hpp:177-217)
  void em_half_push(const R* xp, const R* vp, const R* Dh, const R* Bh, R* vp_upd) const {
    R D0[3] = { Dh[0], Dh[1], Dh[2] }, B0[3] = { Bh[0], Bh[1], Bh[2] };
    R vh[3] = { K<R>::ZERO, K<R>::ZERO, K<R>::ZERO }, vuh[3];
    metric.cov_to_hat(xp, vp, vuh);
    R COEFF = ndth * K<R>::HALF * metric.alpha(xp);
    D0[0] *= COEFF; D0[1] *= COEFF; D0[2] *= COEFF;
    vuh[0] += D0[0]; vuh[1] += D0[1]; vuh[2] += D0[2];
    COEFF *= K<R>::ONE / std::sqrt(K<R>::ONE + sqr(vuh[0]) + sqr(vuh[1]) + sqr(vuh[2]));
    B0[0] *= COEFF; B0[1] *= COEFF; B0[2] *= COEFF;
    COEFF = K<R>::TWO / (K<R>::ONE + sqr(B0[0]) + sqr(B0[1]) + sqr(B0[2]));
    vh[0] = (vuh[0] + vuh[1] * B0[2] - vuh[2] * B0[1]) * COEFF;
    vh[1] = (vuh[1] + vuh[2] * B0[0] - vuh[0] * B0[2]) * COEFF;
    vh[2] = (vuh[2] + vuh[0] * B0[1] - vuh[1] * B0[0]) * COEFF;
    vuh[0] += vh[1] * B0[2] - vh[2] * B0[1] + D0[0];
    vuh[1] += vh[2] * B0[0] - vh[0] * B0[2] + D0[1];
    vuh[2] += vh[0] * B0[1] - vh[1] * B0[0] + D0[2];
    metric.hat_to_cov(xp, vuh, vp_upd);
  }

  // GeodesicMomentumPush, 2D (gr.hpp:260-331)
  void momentum_push(bool massless, const R* xp, const R* vp, R* vp_upd) const {
    R vm[3] = { K<R>::ZERO, K<R>::ZERO, K<R>::ZERO }, vmc[3] = { K<R>::ZERO, K<R>::ZERO, K<R>::ZERO };
    vp_upd[0] = vp[0];
    vp_upd[1] = vp[1];
    vp_upd[2] = vp[2];
    for (int it = 0; it < prm->niter; ++it) {
      vm[0] = K<R>::HALF * (vp[0] + vp_upd[0]);
      vm[1] = K<R>::HALF * (vp[1] + vp_upd[1]);
      vm[2] = vp[2];
      metric.cov_to_cntrv(xp, vm, vmc);
      const R u0 = gamma(massless, vm, vmc) / metric.alpha(xp);
      vp_upd[0] = vp[0] + dt * (-metric.alpha(xp) * u0 * metric.dr_alpha(xp) +
                                vm[0] * metric.dr_beta1(xp) -
                                (K<R>::HALF / u0) *
                                  (metric.dr_h11(xp) * sqr(vm[0]) + metric.dr_h22(xp) * sqr(vm[1]) +
                                   metric.dr_h33(xp) * sqr(vm[2]) +
                                   K<R>::TWO * metric.dr_h13(xp) * vm[0] * vm[2]));
      vp_upd[1] = vp[1] + dt * (-metric.alpha(xp) * u0 * metric.dt_alpha(xp) +
                                vm[0] * metric.dt_beta1(xp) -
                                (K<R>::HALF / u0) *
                                  (metric.dt_h11(xp) * sqr(vm[0]) + metric.dt_h22(xp) * sqr(vm[1]) +
                                   metric.dt_h33(xp) * sqr(vm[2]) +
                                   K<R>::TWO * metric.dt_h13(xp) * vm[0] * vm[2]));
    }
  }

  // GeodesicCoordinatePush, 2D (gr.hpp:333-383)
  void coordinate_push(bool massless, const R* xp, const R* vp, R* xp_upd) const {
    R vc[3] = { K<R>::ZERO, K<R>::ZERO, K<R>::ZERO }, xm[2] = { K<R>::ZERO, K<R>::ZERO };
    xp_upd[0] = xp[0];
    xp_upd[1] = xp[1];
    for (int it = 0; it < prm->niter; ++it) {
      xm[0] = K<R>::HALF * (xp[0] + xp_upd[0]);
      xm[1] = K<R>::HALF * (xp[1] + xp_upd[1]);
      const R xm_[2] = { xm[0], clamp_theta(xm[1]) };
      metric.cov_to_cntrv(xm_, vp, vc);
      const R g  = gamma(massless, vp, vc);
      const R u0 = g / metric.alpha(xm_);
      xp_upd[0] = xp[0] + dt * (vc[0] / u0 - metric.beta1(xm_));
      xp_upd[1] = xp[1] + dt * (vc[1] / u0);
    }
  }

  // interpolateFields (gr.hpp:482-663): D from em, B from em0
  void interpolate(int64_t p, R* e0, R* b0) const {
    const int i = P.i1[p] + DB.ngh, j = P.i2[p] + DB.ngh;
    const R   dx1_ = static_cast<R>(P.dx1[p]), dx2_ = static_cast<R>(P.dx2[p]);
    if (order == 0) {
      const int indx = static_cast<int>(dx1_ + K<R>::HALF);
      const int indy = static_cast<int>(dx2_ + K<R>::HALF);
      const R ponpmx = K<R>::ONE - dx1_, ponppx = dx1_;
      const R ponpmy = K<R>::ONE - dx2_, ponppy = dx2_;
      const R pondmx = static_cast<R>(indx + 1) - (dx1_ + K<R>::HALF);
      const R pondpx = K<R>::ONE - pondmx;
      const R pondmy = static_cast<R>(indy + 1) - (dx2_ + K<R>::HALF);
      const R pondpy = K<R>::ONE - pondmy;
      R c00, c10;
      // (dual, primal)
      auto dp = [&](const Fld<R>& F, int c) {
        c00 = F(i - 1 + indx, j, c) * pondmx + F(i + indx, j, c) * pondpx;
        c10 = F(i - 1 + indx, j + 1, c) * pondmx + F(i + indx, j + 1, c) * pondpx;
        return c00 * ponpmy + c10 * ponppy;
      };
      // (primal, dual)
      auto pd = [&](const Fld<R>& F, int c) {
        c00 = F(i, j - 1 + indy, c) * ponpmx + F(i + 1, j - 1 + indy, c) * ponppx;
        c10 = F(i, j + indy, c) * ponpmx + F(i + 1, j + indy, c) * ponppx;
        return c00 * pondmy + c10 * pondpy;
      };
      // (primal, primal)
      auto pp = [&](const Fld<R>& F, int c) {
        c00 = F(i, j, c) * ponpmx + F(i + 1, j, c) * ponppx;
        c10 = F(i, j + 1, c) * ponpmx + F(i + 1, j + 1, c) * ponppx;
        return c00 * ponpmy + c10 * ponppy;
      };
      // (dual, dual)
      auto dd = [&](const Fld<R>& F, int c) {
        c00 = F(i - 1 + indx, j - 1 + indy, c) * pondmx + F(i + indx, j - 1 + indy, c) * pondpx;
        c10 = F(i - 1 + indx, j + indy, c) * pondmx + F(i + indx, j + indy, c) * pondpx;
        return c00 * pondmy + c10 * pondpy;
      };
      e0[0] = dp(DB, 0);
      e0[1] = pd(DB, 1);
      e0[2] = pp(DB, 2);
      b0[0] = pd(DB0, 3);
      b0[1] = dp(DB0, 4);
      b0[2] = dd(DB0, 5);
      return;
    }
    if (order == 1) interp_shape<1>(i, j, dx1_, dx2_, e0, b0);
    else interp_shape<2>(i, j, dx1_, dx2_, e0, b0);
  }

  template <int O>
  static R gather(const Fld<R>& F, int c, int imin, int jmin, const R* S1, const R* S2) {
    R e = K<R>::ZERO;
    for (int idx2 = 0; idx2 < O + 1; idx2++) {
      R c0 = K<R>::ZERO;
      for (int idx1 = 0; idx1 < O + 1; idx1++) c0 += S1[idx1] * F(imin + idx1, jmin + idx2, c);
      e += c0 * S2[idx2];
    }
    return e;
  }

  template <int O>
  void interp_shape(int i, int j, R dx1_, R dx2_, R* e0, R* b0) const {
    R   S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1];
    int ip, id, jp, jd;
    shape_order<R, false, O>(i, dx1_, ip, S1p);
    shape_order<R, false, O>(j, dx2_, jp, S2p);
    shape_order<R, true, O>(i, dx1_, id, S1d);
    shape_order<R, true, O>(j, dx2_, jd, S2d);
    e0[0] = gather<O>(DB, 0, id, jp, S1d, S2p);
    e0[1] = gather<O>(DB, 1, ip, jd, S1p, S2d);
    e0[2] = gather<O>(DB, 2, ip, jp, S1p, S2p);
    b0[0] = gather<O>(DB0, 3, ip, jd, S1p, S2d);
    b0[1] = gather<O>(DB0, 4, id, jp, S1d, S2p);
    b0[2] = gather<O>(DB0, 5, id, jd, S1d, S2d);
  }

  // from_Xi_to_i_di (gr.hpp:36-45)
  static void to_i_di(R X, int32_t& i, float& di) {
    i  = static_cast<int>(X + 1) - 1;
    di = static_cast<float>(X) - static_cast<float>(i);
  }

  // boundaryConditions (gr.hpp:818-865), single domain
  void bcs(int64_t p) const {
    if ((P.i1[p] < 0 && prm->absorb_i1min) || (P.i1[p] >= prm->ni1 && prm->absorb_i1max))
      P.tag[p] = 0;
    if (P.i2[p] < 0) {
      if (prm->axis_i2min) {
        P.i2[p]  = 0;
        P.dx2[p] = 1.0f - P.dx2[p];
        P.ux2[p] = -P.ux2[p];
      }
    } else if (P.i2[p] >= prm->ni2) {
      if (prm->axis_i2max) {
        P.i2[p]  = prm->ni2 - 1;
        P.dx2[p] = 1.0f - P.dx2[p];
        P.ux2[p] = -P.ux2[p];
      }
    }
  }

  // operator()(Massless_t) gr.hpp:667-723 and operator()(Massive_t) gr.hpp:727-814
  int push(int64_t p) const {
    if (P.tag[p] != 1) return (P.tag[p] != 0) ? 1 : 0;
    P.i1p[p]  = P.i1[p];
    P.i2p[p]  = P.i2[p];
    P.dx1p[p] = P.dx1[p];
    P.dx2p[p] = P.dx2[p];
    const R xp[2] = { static_cast<R>(P.i1[p]) + static_cast<R>(P.dx1[p]),
                      static_cast<R>(P.i2[p]) + static_cast<R>(P.dx2[p]) };
    R vp[3] = { P.ux1[p], P.ux2[p], P.ux3[p] };
    R vp_upd[3], xp_upd[2];
    if (prm->massless) {
      momentum_push(true, xp, vp, vp_upd);
      coordinate_push(true, xp, vp_upd, xp_upd);
    } else {
      const R xp_[2] = { xp[0], clamp_theta(xp[1]) };
      R Dc[3], Bc[3], Dh[3], Bh[3];
      interpolate(p, Dc, Bc);
      metric.cntrv_to_hat(xp, Dc, Dh);
      metric.cntrv_to_hat(xp, Bc, Bh);
      em_half_push(xp, vp, Dh, Bh, vp_upd);
      vp[0] = vp_upd[0]; vp[1] = vp_upd[1]; vp[2] = vp_upd[2];
      momentum_push(false, xp_, vp, vp_upd);
      vp[0] = vp_upd[0]; vp[1] = vp_upd[1]; vp[2] = vp_upd[2];
      em_half_push(xp, vp, Dh, Bh, vp_upd);
      coordinate_push(false, xp, vp_upd, xp_upd);
    }
    // UpdatePhi (gr.hpp:459-480) takes phi by value: the stored phi is unchanged
    int32_t i1_, i2_;
    float   dx1_, dx2_;
    to_i_di(xp_upd[0], i1_, dx1_);
    to_i_di(xp_upd[1], i2_, dx2_);
    P.i1[p]  = i1_;
    P.dx1[p] = dx1_;
    P.i2[p]  = i2_;
    P.dx2[p] = dx2_;
    P.ux1[p] = vp_upd[0];
    P.ux2[p] = vp_upd[1];
    P.ux3[p] = vp_upd[2];
    bcs(p);
    return 0;
  }
};

template <typename R>
int gr_push(const ort_gr_params_t* prm, ort_particles_t* prtls, const ort_field_t* em,
            const ort_field_t* em0, int32_t order) {
  if (prm == nullptr || prtls == nullptr || order < 0 || order > 2) return 1;
  if (!prm->massless && (em == nullptr || em0 == nullptr)) return 1;
  const ort_field_t* e  = em ? em : em0;
  const ort_field_t* e0 = em0 ? em0 : em;
  ort_field_t dummy{};
  dummy.dim = 2;
  dummy.ngh = 2;
  dummy.ncomp = 6;
  GRPusher<R> g(prm, prtls, e ? e : &dummy, e0 ? e0 : &dummy, order);
  int err = 0;
  for (int64_t p = 0; p < g.P.n; ++p) err |= g.push(p);
  return err;
}

/* ---- GR field solver: src/kernels/faraday_gr.hpp:32-106,
 * ampere_gr.hpp:33-181, aux_fields_gr.hpp:31-306, with the ranges of
 * src/engines/grpic/fieldsolvers.h:53-266.  `mutate` plants a mistake (for
 * the tests that show a check sees it): 1 drops beta x B from aux E, 2 flips
 * the sign of the beta x D term of aux H, 3 drops the 1/sqrt(det h) of B^r in
 * Faraday. */
enum { GRF_AUX_E = 0, GRF_AUX_H = 1, GRF_FARADAY = 2, GRF_AMPERE = 3, GRF_CUR_AMPERE = 4,
       GRF_TAVG_DB = 5, GRF_TAVG_J = 6 };

template <typename R>
int gr_field(int op, const ort_gr_params_t* prm, ort_field_t* fa, ort_field_t* fb,
             ort_field_t* fc, double coeff_d, int axis_min, int axis_max, int mutate) {
  GRMetric<R> m(prm);
  const R     coeff = static_cast<R>(coeff_d);
  const R     H = K<R>::HALF, ONE = K<R>::ONE, TWO = K<R>::TWO;
  Fld<R>      A(fa);
  const int   G = A.ngh, n1 = A.n[0], n2 = A.n[1];
  auto pt = [](R a_, R b_) { return std::array<R, 2>{ a_, b_ }; };
  auto sdh  = [&](R a_, R b_) { auto x = pt(a_, b_); return m.sqrt_det_h(x.data()); };
  auto sdht = [&](R a_, R b_) { auto x = pt(a_, b_); return m.sqrt_det_h_tilde(x.data()); };
  auto h_11 = [&](R a_, R b_) { auto x = pt(a_, b_); return m.h_11(x.data()); };
  auto h_22 = [&](R a_, R b_) { auto x = pt(a_, b_); return m.h_22(x.data()); };
  auto h_33 = [&](R a_, R b_) { auto x = pt(a_, b_); return m.h_33(x.data()); };
  auto h_13 = [&](R a_, R b_) { auto x = pt(a_, b_); return m.h_13(x.data()); };
  auto alp  = [&](R a_, R b_) { auto x = pt(a_, b_); return m.alpha(x.data()); };
  auto bet  = [&](R a_, R b_) { auto x = pt(a_, b_); return m.beta1(x.data()); };
  const int i2top = G + n2 + (axis_max ? 1 : 0);
  if (op == GRF_AUX_E) {
    Fld<R> Df(fa), Bf(fb), Ef(fc);
    for (int i2 = G; i2 < i2top; ++i2)
      for (int i1 = G - 1; i1 < G + n1; ++i1) {
        const R i1_ = static_cast<R>(i1 - G), i2_ = static_cast<R>(i2 - G);
        R w1 = sdht(i1_ - H, i2_), w2 = sdht(i1_ + H, i2_), ww = TWO * sdht(i1_, i2_);
        const R D1_half = (w1 * h_13(i1_ - H, i2_) * Df(i1 - 1, i2, 0) +
                           w2 * h_13(i1_ + H, i2_) * Df(i1, i2, 0)) / ww;
        const R B2_half = (w1 * sdh(i1_ - H, i2_) * bet(i1_ - H, i2_) * Bf(i1 - 1, i2, 4) +
                           w2 * sdh(i1_ + H, i2_) * bet(i1_ + H, i2_) * Bf(i1, i2, 4)) / ww;
        w1 = sdht(i1_ - H, i2_ + H);
        w2 = sdht(i1_ + H, i2_ + H);
        ww = TWO * sdht(i1_, i2_ + H);
        const R B3_half =
          (w1 * sdh(i1_ - H, i2_ + H) * bet(i1_ - H, i2_ + H) * Bf(i1 - 1, i2, 5) +
           w2 * sdh(i1_ + H, i2_ + H) * bet(i1_ + H, i2_ + H) * Bf(i1, i2, 5)) / ww;
        w1 = sdht(i1_, i2_);
        w2 = sdht(i1_ + ONE, i2_);
        ww = TWO * sdht(i1_ + H, i2_);
        const R D3_half = (w1 * h_13(i1_, i2_) * Df(i1, i2, 2) +
                           w2 * h_13(i1_ + ONE, i2_) * Df(i1 + 1, i2, 2)) / ww;
        const R D1_cov = h_11(i1_ + H, i2_) * Df(i1, i2, 0) + D3_half;
        const R D2_cov = h_22(i1_, i2_ + H) * Df(i1, i2, 1);
        const R D3_cov = h_33(i1_, i2_) * Df(i1, i2, 2) + D1_half;
        const R keep = mutate == 1 ? K<R>::ZERO : ONE;
        Ef(i1, i2, 0) = alp(i1_ + H, i2_) * D1_cov;
        Ef(i1, i2, 1) = alp(i1_, i2_ + H) * D2_cov - keep * B3_half;
        Ef(i1, i2, 2) = alp(i1_, i2_) * D3_cov + keep * B2_half;
      }
    return 0;
  }
  if (op == GRF_AUX_H) {
    Fld<R> Df(fa), Bf(fb), Hf(fc);
    for (int i2 = G; i2 < i2top; ++i2)
      for (int i1 = G - 1; i1 < G + n1; ++i1) {
        const R i1_ = static_cast<R>(i1 - G), i2_ = static_cast<R>(i2 - G);
        R w1 = sdht(i1_, i2_ + H), w2 = sdht(i1_ + ONE, i2_ + H), ww = TWO * sdht(i1_ + H, i2_ + H);
        const R B1_half = (w1 * h_13(i1_, i2_ + H) * Bf(i1, i2, 3) +
                           w2 * h_13(i1_ + ONE, i2_ + H) * Bf(i1 + 1, i2, 3)) / ww;
        const R D2_half =
          (w1 * sdh(i1_, i2_ + H) * bet(i1_, i2_ + H) * Df(i1, i2, 1) +
           w2 * sdh(i1_ + ONE, i2_ + H) * bet(i1_ + ONE, i2_ + H) * Df(i1 + 1, i2, 1)) / ww;
        w1 = sdht(i1_, i2_);
        w2 = sdht(i1_ + ONE, i2_);
        ww = TWO * sdht(i1_ + H, i2_);
        const R D3_half = (w1 * sdh(i1_, i2_) * bet(i1_, i2_) * Df(i1, i2, 2) +
                           w2 * sdh(i1_ + ONE, i2_) * bet(i1_ + ONE, i2_) * Df(i1 + 1, i2, 2)) /
                          ww;
        w1 = sdht(i1_ - H, i2_ + H);
        w2 = sdht(i1_ + H, i2_ + H);
        ww = TWO * sdht(i1_, i2_ + H);
        const R B3_half = (w1 * h_13(i1_ - H, i2_ + H) * Bf(i1 - 1, i2, 5) +
                           w2 * h_13(i1_ + H, i2_ + H) * Bf(i1, i2, 5)) / ww;
        const R B1_cov = h_11(i1_, i2_ + H) * Bf(i1, i2, 3) + B3_half;
        const R B2_cov = h_22(i1_ + H, i2_) * Bf(i1, i2, 4);
        const R B3_cov = h_33(i1_ + H, i2_ + H) * Bf(i1, i2, 5) + B1_half;
        const R sgn = mutate == 2 ? -ONE : ONE;
        Hf(i1, i2, 3) = alp(i1_, i2_ + H) * B1_cov;
        Hf(i1, i2, 4) = alp(i1_ + H, i2_) * B2_cov + sgn * D3_half;
        Hf(i1, i2, 5) = alp(i1_ + H, i2_ + H) * B3_cov - sgn * D2_half;
      }
    return 0;
  }
  if (op == GRF_FARADAY) {
    Fld<R> Bin(fa), Bout(fb), E(fc);
    for (int i2 = G; i2 < G + n2; ++i2)
      for (int i1 = G; i1 < G + n1; ++i1) {
        const R i1_ = static_cast<R>(i1 - G), i2_ = static_cast<R>(i2 - G);
        const R inv_0p = mutate == 3 ? ONE : ONE / sdh(i1_, i2_ + H);
        const R inv_pp = ONE / sdh(i1_ + H, i2_ + H);
        const R b1 = Bin(i1, i2, 3), b2 = Bin(i1, i2, 4), b3 = Bin(i1, i2, 5);
        Bout(i1, i2, 3) = b1 + coeff * inv_0p * (E(i1, i2, 2) - E(i1, i2 + 1, 2));
        if (i2 == G && axis_min) Bout(i1, i2, 4) = K<R>::ZERO;
        else Bout(i1, i2, 4) = b2 + coeff * (ONE / sdh(i1_ + H, i2_)) * (E(i1 + 1, i2, 2) - E(i1, i2, 2));
        Bout(i1, i2, 5) = b3 + coeff * inv_pp * (E(i1, i2 + 1, 0) - E(i1, i2, 0) + E(i1, i2, 1) -
                                                 E(i1 + 1, i2, 1));
      }
    return 0;
  }
  if (op == GRF_AMPERE || op == GRF_CUR_AMPERE) {
    // rows [G, G + n2]; the axis rows first (they read the row next to them
    // as it was before the update)
    std::vector<int> rows;
    if (axis_min) rows.push_back(G);
    if (axis_max) rows.push_back(G + n2);
    for (int i2 = G; i2 <= G + n2; ++i2)
      if (!((i2 == G && axis_min) || (i2 == G + n2 && axis_max))) rows.push_back(i2);
    for (int i2 : rows)
      for (int i1 = G; i1 < G + n1; ++i1) {
        const R i1_ = static_cast<R>(i1 - G), i2_ = static_cast<R>(i2 - G);
        const bool at_min = (i2 == G) && axis_min, at_max = (i2 == G + n2) && axis_max;
        if (op == GRF_AMPERE) {
          Fld<R> Din(fa), Dout(fb), Hh(fc);
          if (at_min) {
            const R d1 = Din(i1, i2 + 1, 0), d2 = Din(i1, i2, 1);
            Dout(i1, i2, 0) = d1;
            Dout(i1, i2, 1) = d2 + coeff * (ONE / sdh(i1_, H)) * (Hh(i1 - 1, i2, 5) - Hh(i1, i2, 5));
          } else if (at_max) {
            Dout(i1, i2, 0) = Din(i1, i2 - 1, 0);
          } else {
            const R d1 = Din(i1, i2, 0), d2 = Din(i1, i2, 1), d3 = Din(i1, i2, 2);
            Dout(i1, i2, 0) = d1 + coeff * (ONE / sdh(i1_ + H, i2_)) * (Hh(i1, i2, 5) - Hh(i1, i2 - 1, 5));
            Dout(i1, i2, 1) = d2 + coeff * (ONE / sdh(i1_, i2_ + H)) * (Hh(i1 - 1, i2, 5) - Hh(i1, i2, 5));
            Dout(i1, i2, 2) = d3 + coeff * (ONE / sdh(i1_, i2_)) *
                                     ((Hh(i1, i2 - 1, 3) - Hh(i1, i2, 3)) +
                                      (Hh(i1, i2, 4) - Hh(i1 - 1, i2, 4)));
          }
        } else {
          Fld<R> Df(fa), J(fb);
          if (at_min) {
            Df(i1, i2, 0) += J(i1, i2, 0) * H * coeff / m.polar_area(i1_ + H);
            Df(i1, i2, 1) += J(i1, i2, 1) * coeff * (ONE / sdh(i1_, i2_ + H));
          } else if (at_max) {
            Df(i1, i2, 0) += J(i1, i2, 0) * H * coeff / m.polar_area(i1_ + H);
          } else {
            Df(i1, i2, 0) += J(i1, i2, 0) * coeff * (ONE / sdh(i1_ + H, i2_));
            Df(i1, i2, 1) += J(i1, i2, 1) * coeff * (ONE / sdh(i1_, i2_ + H));
            Df(i1, i2, 2) += J(i1, i2, 2) * coeff * (ONE / sdh(i1_, i2_));
          }
        }
      }
    return 0;
  }
  if (op == GRF_TAVG_DB || op == GRF_TAVG_J) {
    Fld<R> src(fa), dst(fb);
    const int nc = op == GRF_TAVG_DB ? 6 : 3;
    for (int c = 0; c < nc; ++c)
      for (int i2 = G; i2 < G + n2; ++i2)
        for (int i1 = G; i1 < G + n1; ++i1) dst(i1, i2, c) = H * (dst(i1, i2, c) + src(i1, i2, c));
    return 0;
  }
  return 1;
}

/* ---- GR field boundaries and the spherical current filter ----------------
 * op 0: MatchBoundaries_kernel<GRPIC, M, FS, in::x1>, 2D
 *       (src/kernels/fields_bcs.hpp:176-345): over x1 cells [i1_lo, i1_hi)
 *       and every x2 row (ghosts included), each component the target
 *       defines and the tags select is blended towards the target,
 *       F = s F + (1 - s) T with s = tanh(4 |r - r_edge| / ds), r taken at
 *       the component's own x1 position; D^3 on and beyond an axis row is
 *       only damped (s F).
 * op 1: gr::HorizonBoundaries_kernel (fields_bcs.hpp:1188-1240): rows
 *       x1 = 0 .. 2 + nfilter copy row G + 1 + nfilter, x2 rows [G, G + n2].
 * op 2: AxisBoundaries_kernel (fields_bcs.hpp:825-867), every x1 column.
 * op 3: gr::AbsorbCurrents_kernel<M, 1> (fields_bcs.hpp:1243-1280): J *=
 *       tanh(|r(x1) - r_edge| / (ds / 4)) over [i1_lo, i1_hi) x every row.
 * op 4: DigitalFilter_kernel<2D, non-Cartesian> (digital_filter.hpp:196-
 *       266) over x1 [G, G + n1), x2 [G, G + n2] (currents.h:80-110):
 *       fa = array, fb = buffer.
 * Tags: the reference's BC bits, D and E share bits 0-2, B and H bits 3-5
 * (src/global/global.h:306-329). */
template <typename R>
int gr_bc(int op, const ort_gr_params_t* prm, ort_field_t* fa, ort_field_t* fb,
          const ort_gr_bc_args_t* g, int mutate) {
  GRMetric<R> m(prm);
  Fld<R>      F(fa);
  const int   G = F.ngh, n1 = F.n[0], n2 = F.n[1];
  const int   N1 = n1 + 2 * G, N2 = n2 + 2 * G;
  const R     ONE = K<R>::ONE, ZERO = K<R>::ZERO, H = K<R>::HALF;
  const bool  setD = (g->tags & 0x7u) != 0, setB = (g->tags & 0x38u) != 0;
  auto r_at = [&](R x1) { const R x[2] = { x1, ZERO }; return m.r_of(x); };
  if (op == 0) {
    Fld<R> T(fb);
    const R four = static_cast<R>(4.0);
    auto shape = [&](R x1) {
      return std::tanh(std::fabs(r_at(x1) - static_cast<R>(g->xg_edge)) * four /
                       static_cast<R>(g->ds));
    };
    auto has = [&](int c) { return (g->comps >> c) & 1u; };
    for (int i2 = 0; i2 < N2; ++i2)
      for (int i1 = g->i1_lo; i1 < g->i1_hi; ++i1) {
        const R i1_ = static_cast<R>(i1 - G);
        auto blend = [&](int c, R s_) { F(i1, i2, c) = s_ * F(i1, i2, c) + (ONE - s_) * T(i1, i2, c); };
        if (has(0) || has(4)) {   // (i1 + 1/2, i2): D^1, B^2
          const R s_ = shape(i1_ + H);
          if (has(0) && setD) blend(0, s_);
          if (has(4) && setB) blend(4, s_);
        }
        if (has(1) || has(3)) {   // (i1, i2 + 1/2): D^2, B^1
          const R s_ = shape(i1_);
          if (has(1) && setD) blend(1, s_);
          if (has(3) && setB) blend(3, s_);
        }
        if (has(2) && setD) {     // (i1, i2): D^3
          const R s_ = shape(i1_);
          F(i1, i2, 2) = s_ * F(i1, i2, 2);
          if ((!g->axis_min || i2 > G) && (!g->axis_max || i2 < N2 - G))
            F(i1, i2, 2) += (ONE - s_) * T(i1, i2, 2);
        }
        if (has(5) && setB) {     // (i1 + 1/2, i2 + 1/2): B^3
          const R s_ = shape(mutate == 1 ? i1_ : i1_ + H);
          blend(5, s_);
        }
      }
    return 0;
  }
  if (op == 1) {
    const int src = G + 1 + g->nfilter;
    for (int i2 = G; i2 <= G + n2; ++i2)
      for (int i = 0; i <= 2 + g->nfilter; ++i) {
        if (setD)
          for (int c = 0; c < 3; ++c) F(i, i2, c) = F(src, i2, c);
        if (setB)
          for (int c = 3; c < 6; ++c) F(i, i2, c) = F(src, i2, c);
      }
    return 0;
  }
  if (op == 2) {
    const R sgn = mutate == 2 ? ONE : -ONE;
    for (int i1 = 0; i1 < N1; ++i1) {
      if (!g->upper) {
        const int e = G;
        if (setD) {
          F(i1, e - 1, 1) = sgn * F(i1, e, 1);
          F(i1, e, 2)     = ZERO;
          F(i1, e - 1, 2) = F(i1, e + 1, 2);
        }
        if (setB) {
          F(i1, e - 1, 3) = F(i1, e, 3);
          F(i1, e, 4)     = ZERO;
          F(i1, e - 1, 4) = -F(i1, e + 1, 4);
          F(i1, e - 1, 5) = F(i1, e, 5);
        }
      } else {
        const int e = G + n2;
        if (setD) {
          F(i1, e, 1)     = sgn * F(i1, e - 1, 1);
          F(i1, e, 2)     = ZERO;
          F(i1, e + 1, 2) = F(i1, e - 1, 2);
        }
        if (setB) {
          F(i1, e, 3)     = F(i1, e - 1, 3);
          F(i1, e, 4)     = ZERO;
          F(i1, e + 1, 4) = -F(i1, e - 1, 4);
          F(i1, e, 5)     = F(i1, e - 1, 5);
        }
      }
    }
    return 0;
  }
  if (op == 3) {
    const R q = static_cast<R>(0.25) * static_cast<R>(g->ds);
    for (int i2 = 0; i2 < N2; ++i2)
      for (int i1 = g->i1_lo; i1 < g->i1_hi; ++i1) {
        const R f = std::tanh(std::fabs(r_at(static_cast<R>(i1 - G)) - static_cast<R>(g->xg_edge)) / q);
        for (int c = 0; c < 3; ++c) F(i1, i2, c) *= f;
      }
    return 0;
  }
  if (op == 4) {
    Fld<R> B(fb);
    const R Q = K<R>::HALF * K<R>::HALF, E8 = Q * H, S16 = E8 * H;
    // binomial filter along x1 of buffer component c at (i, j)
    auto f1 = [&](int c, int i, int j) { return H * B(i, j, c) + Q * (B(i - 1, j, c) + B(i + 1, j, c)); };
    const int i2_min = G, i2_max = G + n2;
    for (int i2 = G; i2 <= G + n2; ++i2)
      for (int i1 = G; i1 < G + n1; ++i1) {
        if (g->axis_min && i2 == i2_min) {
          F(i1, i2, 0) = H * f1(0, i1, i2) + H * f1(0, i1, i2 + 1);
          F(i1, i2, 2) = ZERO;
          F(i1, i2, 1) = Q * (f1(1, i1, i2) + f1(1, i1, i2 + 1));
        } else if (g->axis_min && i2 == i2_min + 1) {
          F(i1, i2, 0) = H * f1(0, i1, i2) + Q * (f1(0, i1, i2 + 1) + f1(0, i1, i2 - 1));
          F(i1, i2, 2) = H * f1(2, i1, i2) + Q * f1(2, i1, i2 + 1);
          F(i1, i2, 1) = H * f1(1, i1, i2) + Q * (f1(1, i1, i2 - 1) + f1(1, i1, i2 + 1));
        } else if (g->axis_max && i2 == i2_max - 1) {
          F(i1, i2, 0) = H * f1(0, i1, i2) + Q * (f1(0, i1, i2 - 1) + f1(0, i1, i2 + 1));
          F(i1, i2, 2) = H * f1(2, i1, i2) + Q * f1(2, i1, i2 - 1);
          F(i1, i2, 1) = Q * (f1(1, i1, i2) + f1(1, i1, i2 - 1));
        } else if (g->axis_max && i2 == i2_max) {
          F(i1, i2, 0) = H * f1(0, i1, i2) + H * f1(0, i1, i2 - 1);
          F(i1, i2, 2) = ZERO;
        } else {
          for (int c = 0; c < 3; ++c)
            F(i1, i2, c) = Q * B(i1, i2, c) +
                           E8 * (B(i1 - 1, i2, c) + B(i1 + 1, i2, c) + B(i1, i2 - 1, c) +
                                 B(i1, i2 + 1, c)) +
                           S16 * (B(i1 - 1, i2 - 1, c) + B(i1 + 1, i2 + 1, c) +
                                  B(i1 - 1, i2 + 1, c) + B(i1 + 1, i2 - 1, c));
        }
      }
    return 0;
  }
  return 1;
}

/* ---- velocity of the GR deposit (currents_deposit.hpp:112-166, GRPIC) ------
 * Per particle: theta clamped to [SMALL_ANGLE_GR, pi - SMALL_ANGLE_GR] as
 * the pusher does, v^i = h^ij u_j (kerr_schild.h:638-642), and
 * v^i * alpha / sqrt(1 + u_i v^i); the out-of-plane component (set to zero
 * if not finite) is what the 2D deposit uses.  Dead particles get 0. */
template <typename R>
int gr_deposit_velocity(const ort_gr_params_t* prm, const ort_particles_t* prtls, double* vz) {
  GRMetric<R> m(prm);
  Prt<R>      P(prtls);
  const R small = static_cast<R>(SMALL_ANGLE_GR_D);
  const R large = static_cast<R>(PI_D - SMALL_ANGLE_GR_D);
  for (int64_t p = 0; p < P.n; ++p) {
    vz[p] = 0.0;
    if (P.tag[p] != TAG_DEAD) {
      R x[2] = { static_cast<R>(P.i1[p]) + static_cast<R>(P.dx1[p]),
                 static_cast<R>(P.i2[p]) + static_cast<R>(P.dx2[p]) };
      const R th = m.theta_cd2ph(x[1]);
      if (th < small) x[1] = m.theta_ph2cd(small);
      else if (th >= large) x[1] = m.theta_ph2cd(large);
      const R u[3] = { P.ux1[p], P.ux2[p], P.ux3[p] };
      const R v[3] = { u[0] * m.h11(x) + u[2] * m.h13(x), u[1] * m.h22(x),
                       u[0] * m.h13(x) + u[2] * m.h33(x) };
      const R inv_energy = m.alpha(x) / std::sqrt(K<R>::ONE + u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
      R v3 = v[2];
      if (std::isnan(v3) || std::isinf(v3)) v3 = K<R>::ZERO;
      vz[p] = static_cast<double>(v3 * inv_energy);
    }
  }
  return 0;
}

// Metric quantities at code coordinates, for the tests: per point, NQ values
constexpr int NQ = 27;
template <typename R>
int gr_metric(const ort_gr_params_t* prm, int64_t n, const double* x1, const double* x2,
              double* out) {
  GRMetric<R> m(prm);
  for (int64_t k = 0; k < n; ++k) {
    const R x[2] = { static_cast<R>(x1[k]), static_cast<R>(x2[k]) };
    double* o = out + NQ * k;
    o[0]  = m.h_11(x);     o[1]  = m.h_22(x);     o[2]  = m.h_33(x);     o[3]  = m.h_13(x);
    o[4]  = m.h11(x);      o[5]  = m.h22(x);      o[6]  = m.h33(x);      o[7]  = m.h13(x);
    o[8]  = m.alpha(x);    o[9]  = m.beta1(x);
    o[10] = m.dr_alpha(x); o[11] = m.dt_alpha(x); o[12] = m.dr_beta1(x); o[13] = m.dt_beta1(x);
    o[14] = m.dr_h11(x);   o[15] = m.dr_h22(x);   o[16] = m.dr_h33(x);   o[17] = m.dr_h13(x);
    o[18] = m.dt_h11(x);   o[19] = m.dt_h22(x);   o[20] = m.dt_h33(x);   o[21] = m.dt_h13(x);
    o[22] = m.r_of(x);     o[23] = m.th_of(x);
    o[24] = m.sqrt_det_h(x); o[25] = m.sqrt_det_h_tilde(x); o[26] = m.polar_area(x[0]);
  }
  return 0;
}

// vector transformations, for the tests: which = 0 D->U, 1 U->D, 2 D->T,
// 3 T->D, 4 U->T, 5 T->U
template <typename R>
int gr_transform(const ort_gr_params_t* prm, int32_t which, int64_t n, const double* x1,
                 const double* x2, const double* vin, double* vout) {
  GRMetric<R> m(prm);
  for (int64_t k = 0; k < n; ++k) {
    const R x[2] = { static_cast<R>(x1[k]), static_cast<R>(x2[k]) };
    const R vi[3] = { static_cast<R>(vin[3 * k]), static_cast<R>(vin[3 * k + 1]),
                      static_cast<R>(vin[3 * k + 2]) };
    R vo[3];
    switch (which) {
      case 0: m.cov_to_cntrv(x, vi, vo); break;
      case 1: m.cntrv_to_cov(x, vi, vo); break;
      case 2: m.cov_to_hat(x, vi, vo); break;
      case 3: m.hat_to_cov(x, vi, vo); break;
      case 4: m.cntrv_to_hat(x, vi, vo); break;
      case 5: m.hat_to_cntrv(x, vi, vo); break;
      default: return 1;
    }
    for (int c = 0; c < 3; ++c) vout[3 * k + c] = static_cast<double>(vo[c]);
  }
  return 0;
}

}  // namespace

#define ORT_GR_DEFINE(SUF, R)                                                               \
  extern "C" int ntt_oracle_gr_push_##SUF(const ort_gr_params_t* prm,                       \
                                          ort_particles_t* prtls, const ort_field_t* em,    \
                                          const ort_field_t* em0, int32_t order) {          \
    return gr_push<R>(prm, prtls, em, em0, order);                                          \
  }                                                                                         \
  extern "C" int ntt_oracle_gr_metric_##SUF(const ort_gr_params_t* prm, int64_t n,          \
                                            const double* x1, const double* x2,             \
                                            double* out) {                                  \
    return gr_metric<R>(prm, n, x1, x2, out);                                               \
  }                                                                                         \
  extern "C" int ntt_oracle_gr_field_##SUF(int32_t op, const ort_gr_params_t* prm,          \
                                           ort_field_t* fa, ort_field_t* fb,                \
                                           ort_field_t* fc, double coeff, int32_t axis_min, \
                                           int32_t axis_max, int32_t mutate) {               \
    return gr_field<R>(op, prm, fa, fb, fc, coeff, axis_min, axis_max, mutate);             \
  }                                                                                         \
  extern "C" int ntt_oracle_gr_bc_##SUF(int32_t op, const ort_gr_params_t* prm,             \
                                        ort_field_t* fa, ort_field_t* fb,                   \
                                        const ort_gr_bc_args_t* args, int32_t mutate) {    \
    return gr_bc<R>(op, prm, fa, fb, args, mutate);                                         \
  }                                                                                         \
  extern "C" int ntt_oracle_gr_deposit_velocity_##SUF(const ort_gr_params_t* prm,           \
                                                      const ort_particles_t* prtls,         \
                                                      double* vz) {                         \
    return gr_deposit_velocity<R>(prm, prtls, vz);                                          \
  }                                                                                         \
  extern "C" int ntt_oracle_gr_transform_##SUF(const ort_gr_params_t* prm, int32_t which,   \
                                               int64_t n, const double* x1,                 \
                                               const double* x2, const double* vin,         \
                                               double* vout) {                              \
    return gr_transform<R>(prm, which, n, x1, x2, vin, vout);                               \
  }

ORT_GR_DEFINE(f32, float)
ORT_GR_DEFINE(f64, double)
