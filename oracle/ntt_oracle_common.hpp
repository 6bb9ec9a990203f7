/*
 * ntt_oracle_common.hpp -- pieces shared by the oracle's translation units:
 * constants, host field/particle accessors and the B-spline shapes.
 *
 * TEST INFRASTRUCTURE ONLY (see ntt_oracle.h).
 */
#ifndef NTT_ORACLE_COMMON_HPP
#define NTT_ORACLE_COMMON_HPP

#include "ntt_oracle.h"

#include <cmath>
#include <cstdint>

namespace ort {

/* ---- constants: src/global/utils/numeric.h:30-68 ---------------------- */
template <typename R> struct K;
template <> struct K<float> {
  static constexpr float ONE = 1.0f, TWO = 2.0f, THREE = 3.0f, FOUR = 4.0f;
  static constexpr float ZERO = 0.0f, HALF = 0.5f, THIRD = 0.333333f;
  static constexpr float THREE_FOURTHS = 0.75f, THREE_HALFS = 1.5f;
  static constexpr float INV_2 = 0.5f, INV_4 = 0.25f, INV_8 = 0.125f;
  static constexpr float INV_16 = 0.0625f, INV_32 = 0.03125f, INV_64 = 0.015625f;
};
template <> struct K<double> {
  static constexpr double ONE = 1.0, TWO = 2.0, THREE = 3.0, FOUR = 4.0;
  static constexpr double ZERO = 0.0, HALF = 0.5, THIRD = 0.3333333333333333;
  static constexpr double THREE_FOURTHS = 0.75, THREE_HALFS = 1.5;
  static constexpr double INV_2 = 0.5, INV_4 = 0.25, INV_8 = 0.125;
  static constexpr double INV_16 = 0.0625, INV_32 = 0.03125, INV_64 = 0.015625;
};

enum { TAG_DEAD = 0, TAG_ALIVE = 1 };
enum { P_PHOTON = 1, P_BORIS = 2, P_VAY = 4, P_GCA = 8 };
enum { BC_PERIODIC = 1, BC_ABSORB = 2, BC_ATMOSPHERE = 3, BC_REFLECT = 5,
       BC_HORIZON = 6, BC_AXIS = 7 };

/* ---- field accessor (Kokkos LayoutLeft, component slowest) ------------- */
template <typename R> struct Fld {
  R*      d;
  int     dim, ngh, nc;
  int     n[3];
  int64_t N1, N2, N3;
  explicit Fld(const ort_field_t* f)
    : d(static_cast<R*>(f->data)), dim(f->dim), ngh(f->ngh), nc(f->ncomp) {
    for (int k = 0; k < 3; ++k) n[k] = f->n[k];
    N1 = n[0] + 2 * ngh;
    N2 = dim >= 2 ? n[1] + 2 * ngh : 1;
    N3 = dim >= 3 ? n[2] + 2 * ngh : 1;
  }
  int64_t size() const { return N1 * N2 * N3 * nc; }
  R& operator()(int64_t i, int64_t j, int c) const { return d[i + N1 * (j + N2 * (int64_t)c)]; }
  R& operator()(int64_t i, int64_t j, int64_t k, int c) const {
    return d[i + N1 * (j + N2 * (k + N3 * (int64_t)c))];
  }
  R& at(int64_t i, int64_t j, int64_t k, int c) const { return d[i + N1 * (j + N2 * (k + N3 * (int64_t)c))]; }
};

template <typename R> struct Prt {
  int32_t *i1, *i2, *i3, *i1p, *i2p, *i3p;
  float   *dx1, *dx2, *dx3, *dx1p, *dx2p, *dx3p;
  R       *ux1, *ux2, *ux3, *w, *phi;
  int16_t *tag;
  int64_t  n;
  explicit Prt(const ort_particles_t* p)
    : i1(p->i1), i2(p->i2), i3(p->i3), i1p(p->i1_prev), i2p(p->i2_prev),
      i3p(p->i3_prev), dx1(p->dx1), dx2(p->dx2), dx3(p->dx3),
      dx1p(p->dx1_prev), dx2p(p->dx2_prev), dx3p(p->dx3_prev),
      ux1(static_cast<R*>(p->ux1)), ux2(static_cast<R*>(p->ux2)),
      ux3(static_cast<R*>(p->ux3)), w(static_cast<R*>(p->weight)),
      phi(static_cast<R*>(p->phi)), tag(p->tag), n(p->npart) {}
};

/* ---- B-spline shapes: src/kernels/particle_shapes.hpp:544-588 ---------- */
template <typename R, bool STAG, int O>
inline void shape_order(int i, R di, int& i_min, R* S) {
  using C = K<R>;
  if constexpr (O == 1) {
    if constexpr (!STAG) {
      i_min = i;
      S[0]  = C::ONE - di;
      S[1]  = di;
    } else {
      if (di < C::HALF) {
        i_min = i - 1;
        S[0]  = C::HALF - di;
        S[1]  = C::ONE - S[0];
      } else {
        i_min = i;
        S[0]  = C::THREE_HALFS - di;
        S[1]  = C::ONE - S[0];
      }
    }
  } else if constexpr (O == 2) {
    if constexpr (!STAG) {
      if (di < C::HALF) {
        i_min = i - 1;
        S[0]  = C::HALF * ((C::HALF - di) * (C::HALF - di));
        S[1]  = C::THREE_FOURTHS - di * di;
        S[2]  = C::ONE - S[0] - S[1];
      } else {
        i_min = i;
        S[0]  = C::HALF * ((C::THREE_HALFS - di) * (C::THREE_HALFS - di));
        S[1]  = C::THREE_FOURTHS - (C::ONE - di) * (C::ONE - di);
        S[2]  = C::ONE - S[0] - S[1];
      }
    } else {
      i_min = i - 1;
      S[0]  = C::HALF * ((C::ONE - di) * (C::ONE - di));
      S[2]  = C::HALF * (di * di);
      S[1]  = C::ONE - S[0] - S[2];
    }
  }
}

/* ---- merged initial/final deposit stencils: particle_shapes.hpp:934-1024 */
template <typename R, int O>
inline void for_deposit(int i_init, R di_init, int i_fin, R di_fin, int& i_min,
                        int& i_max, R* iS, R* fS) {
  int i_init_min, i_fin_min;
  R   iS_[O + 1], fS_[O + 1];
  shape_order<R, false, O>(i_init, di_init, i_init_min, iS_);
  shape_order<R, false, O>(i_fin, di_fin, i_fin_min, fS_);
  if (i_init_min < i_fin_min) {
    i_min = i_init_min;
    i_max = i_min + O + 1;
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = K<R>::ZERO;
    fS[0]     = K<R>::ZERO;
    for (int j = 0; j < O + 1; j++) fS[j + 1] = fS_[j];
  } else if (i_init_min > i_fin_min) {
    i_min = i_fin_min;
    i_max = i_min + O + 1;
    iS[0] = K<R>::ZERO;
    for (int j = 0; j < O + 1; j++) iS[j + 1] = iS_[j];
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = K<R>::ZERO;
  } else {
    i_min = i_init_min;
    i_max = i_min + O;
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = K<R>::ZERO;
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = K<R>::ZERO;
  }
}

}  // namespace ort

#endif
