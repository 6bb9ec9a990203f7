// sr_particle.hpp -- per-particle device routines of the SR hot path:
// B-spline shapes, field interpolation, Boris/Vay momentum update, Cartesian
// position update with boundary conditions, and the charge-conserving current
// deposit.  Every routine is templated on a field accessor so that the same
// code runs against HBM (FieldView) and against LDS-staged windows.
//
// Reference: src/kernels/pushers/sr.hpp, src/kernels/particle_shapes.hpp,
// src/kernels/currents_deposit.hpp (line numbers at each routine).
#pragma once

#include "ntt_common.hpp"

namespace ntt {

enum PusherAlg : int { ALG_BORIS = 0, ALG_VAY = 1, ALG_PHOTON = 2 };

// ---- shapes: particle_shapes.hpp:544-588 --------------------------------
// SEL: the two branches on di < 1/2 taken as selects of their operands (the
// same operations on the same values, so the same results) rather than as
// divergent branches.  Measured at 512 x 256 x 256, order 2: the deposit's
// stencils run faster with selects (15.4 against 15.6 ms per launch), the
// push's field gather with branches (15.9 against 16.7 ms).
template <bool STAG, int O, bool SEL = false>
__device__ __forceinline__ void shape_order(int i, float di, int& i_min, float* S) {
  if constexpr (O == 1) {
    if constexpr (!STAG) {
      i_min = i;
      S[0]  = ONE - di;
      S[1]  = di;
    } else if constexpr (SEL) {
      const bool lo = di < HALF;
      i_min = lo ? i - 1 : i;
      S[0]  = (lo ? HALF : THREE_HALFS) - di;
      S[1]  = ONE - S[0];
    } else {
      if (di < HALF) {
        i_min = i - 1;
        S[0]  = HALF - di;
      } else {
        i_min = i;
        S[0]  = THREE_HALFS - di;
      }
      S[1] = ONE - S[0];
    }
  } else if constexpr (O == 2) {
    if constexpr (!STAG) {
      if constexpr (SEL) {
        const bool  lo = di < HALF;
        const float t  = (lo ? HALF : THREE_HALFS) - di;  // 1/2 - di or 3/2 - di
        const float v  = lo ? di : ONE - di;               // di or 1 - di
        i_min = lo ? i - 1 : i;
        S[0]  = HALF * (t * t);
        S[1]  = THREE_FOURTHS - v * v;
      } else if (di < HALF) {
        i_min = i - 1;
        S[0]  = HALF * ((HALF - di) * (HALF - di));
        S[1]  = THREE_FOURTHS - di * di;
      } else {
        i_min = i;
        S[0]  = HALF * ((THREE_HALFS - di) * (THREE_HALFS - di));
        S[1]  = THREE_FOURTHS - (ONE - di) * (ONE - di);
      }
      S[2] = ONE - S[0] - S[1];
    } else {
      i_min = i - 1;
      S[0]  = HALF * ((ONE - di) * (ONE - di));
      S[2]  = HALF * (di * di);
      S[1]  = ONE - S[0] - S[2];
    }
  }
}

// ---- merged deposit stencils: particle_shapes.hpp:934-1024 --------------
// Returns i_min and the number of extra slots (i_max - i_min).
template <int O>
__device__ __forceinline__ void for_deposit(int i_init, float di_init, int i_fin, float di_fin,
                                            int& i_min, int& di, float* iS, float* fS) {
  int   a, b;
  float iS_[O + 1], fS_[O + 1];
  shape_order<false, O, true>(i_init, di_init, a, iS_);
  shape_order<false, O, true>(i_fin, di_fin, b, fS_);
  if (a < b) {
    i_min = a;
    di    = O + 1;
#pragma unroll
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = ZERO;
    fS[0]     = ZERO;
#pragma unroll
    for (int j = 0; j < O + 1; j++) fS[j + 1] = fS_[j];
  } else if (a > b) {
    i_min = b;
    di    = O + 1;
    iS[0] = ZERO;
#pragma unroll
    for (int j = 0; j < O + 1; j++) iS[j + 1] = iS_[j];
#pragma unroll
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = ZERO;
  } else {
    i_min = a;
    di    = O;
#pragma unroll
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = ZERO;
#pragma unroll
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = ZERO;
  }
}

// ---- field accessors --------------------------------------------------------
// Global (HBM) accessor over the ghost-extended em array.
struct GlobalEM {
  const float* __restrict__ d;
  int32_t N1, N2;
  int64_t cstride;
  __device__ __forceinline__ float operator()(int c, int i, int j, int k) const {
    return d[static_cast<int64_t>(i) +
             static_cast<int64_t>(N1) * (static_cast<int64_t>(j) + static_cast<int64_t>(N2) * k) +
             cstride * c];
  }
};

// ---- interpolation, order 0 "direct": sr.hpp:905-1117 ------------------------
template <class A>
__device__ __forceinline__ float lin2(const A& F, int c, int ia, int ja, int k, float wx0,
                                      float wx1, float wy0, float wy1) {
  const float c000 = F(c, ia, ja, k), c100 = F(c, ia + 1, ja, k);
  const float c010 = F(c, ia, ja + 1, k), c110 = F(c, ia + 1, ja + 1, k);
  const float c00 = c000 * wx0 + c100 * wx1;
  const float c10 = c010 * wx0 + c110 * wx1;
  return c00 * wy0 + c10 * wy1;
}

template <class A>
__device__ __forceinline__ float lin3(const A& F, int c, int ia, int ja, int ka, float wx0,
                                      float wx1, float wy0, float wy1, float wz0, float wz1) {
  const float c000 = F(c, ia, ja, ka), c100 = F(c, ia + 1, ja, ka);
  const float c010 = F(c, ia, ja + 1, ka), c110 = F(c, ia + 1, ja + 1, ka);
  const float c001 = F(c, ia, ja, ka + 1), c101 = F(c, ia + 1, ja, ka + 1);
  const float c011 = F(c, ia, ja + 1, ka + 1), c111 = F(c, ia + 1, ja + 1, ka + 1);
  const float c00 = c000 * wx0 + c100 * wx1;
  const float c10 = c010 * wx0 + c110 * wx1;
  const float c0  = c00 * wy0 + c10 * wy1;
  const float c01 = c001 * wx0 + c101 * wx1;
  const float c11 = c011 * wx0 + c111 * wx1;
  const float c1  = c01 * wy0 + c11 * wy1;
  return c0 * wz0 + c1 * wz1;
}

// ---- interpolation, B-spline orders: sr.hpp:1170-1366 ------------------------
template <int O, class A>
__device__ __forceinline__ float spl2(const A& F, int c, int imin, int jmin, const float* S1,
                                      const float* S2) {
  float e = ZERO;
#pragma unroll
  for (int b = 0; b < O + 1; b++) {
    float c0 = ZERO;
#pragma unroll
    for (int a = 0; a < O + 1; a++) c0 += S1[a] * F(c, imin + a, jmin + b, 0);
    e += c0 * S2[b];
  }
  return e;
}

template <int O, class A>
__device__ __forceinline__ float spl3(const A& F, int c, int imin, int jmin, int kmin,
                                      const float* S1, const float* S2, const float* S3) {
  float e = ZERO;
#pragma unroll
  for (int g = 0; g < O + 1; g++) {
    float c0 = ZERO;
#pragma unroll
    for (int b = 0; b < O + 1; b++) {
      float c00 = ZERO;
#pragma unroll
      for (int a = 0; a < O + 1; a++) c00 += S1[a] * F(c, imin + a, jmin + b, kmin + g);
      c0 += c00 * S2[b];
    }
    e += c0 * S3[g];
  }
  return e;
}

// Interpolates E and B at the particle.  (i, j, k) are ghost-extended cell
// indices (particle index + ngh); k is ignored in 2D.
template <int D, int O, class A>
__device__ __forceinline__ void interpolate(const A& F, int i, int j, int k, float dx1,
                                            float dx2, float dx3, float* e0, float* b0) {
  if constexpr (O == 0) {
    const int indx = static_cast<int>(dx1 + HALF);
    const int indy = static_cast<int>(dx2 + HALF);
    const float ponpmx = ONE - dx1, ponppx = dx1;
    const float ponpmy = ONE - dx2, ponppy = dx2;
    const float pondmx = static_cast<float>(indx + 1) - (dx1 + HALF), pondpx = ONE - pondmx;
    const float pondmy = static_cast<float>(indy + 1) - (dx2 + HALF), pondpy = ONE - pondmy;
    if constexpr (D == 2) {
      e0[0] = lin2(F, 0, i - 1 + indx, j, 0, pondmx, pondpx, ponpmy, ponppy);
      e0[1] = lin2(F, 1, i, j - 1 + indy, 0, ponpmx, ponppx, pondmy, pondpy);
      e0[2] = lin2(F, 2, i, j, 0, ponpmx, ponppx, ponpmy, ponppy);
      b0[0] = lin2(F, 3, i, j - 1 + indy, 0, ponpmx, ponppx, pondmy, pondpy);
      b0[1] = lin2(F, 4, i - 1 + indx, j, 0, pondmx, pondpx, ponpmy, ponppy);
      b0[2] = lin2(F, 5, i - 1 + indx, j - 1 + indy, 0, pondmx, pondpx, pondmy, pondpy);
    } else {
      const int   indz   = static_cast<int>(dx3 + HALF);
      const float ponpmz = ONE - dx3, ponppz = dx3;
      const float pondmz = static_cast<float>(indz + 1) - (dx3 + HALF), pondpz = ONE - pondmz;
      e0[0] = lin3(F, 0, i - 1 + indx, j, k, pondmx, pondpx, ponpmy, ponppy, ponpmz, ponppz);
      e0[1] = lin3(F, 1, i, j - 1 + indy, k, ponpmx, ponppx, pondmy, pondpy, ponpmz, ponppz);
      e0[2] = lin3(F, 2, i, j, k - 1 + indz, ponpmx, ponppx, ponpmy, ponppy, pondmz, pondpz);
      b0[0] = lin3(F, 3, i, j - 1 + indy, k - 1 + indz, ponpmx, ponppx, pondmy, pondpy, pondmz, pondpz);
      b0[1] = lin3(F, 4, i - 1 + indx, j, k - 1 + indz, pondmx, pondpx, ponpmy, ponppy, pondmz, pondpz);
      b0[2] = lin3(F, 5, i - 1 + indx, j - 1 + indy, k, pondmx, pondpx, pondmy, pondpy, ponpmz, ponppz);
    }
  } else {
    float S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1];
    int   ip, id, jp, jd;
    shape_order<false, O>(i, dx1, ip, S1p);
    shape_order<false, O>(j, dx2, jp, S2p);
    shape_order<true, O>(i, dx1, id, S1d);
    shape_order<true, O>(j, dx2, jd, S2d);
    if constexpr (D == 2) {
      e0[0] = spl2<O>(F, 0, id, jp, S1d, S2p);
      e0[1] = spl2<O>(F, 1, ip, jd, S1p, S2d);
      e0[2] = spl2<O>(F, 2, ip, jp, S1p, S2p);
      b0[0] = spl2<O>(F, 3, ip, jd, S1p, S2d);
      b0[1] = spl2<O>(F, 4, id, jp, S1d, S2p);
      b0[2] = spl2<O>(F, 5, id, jd, S1d, S2d);
    } else {
      float S3p[O + 1], S3d[O + 1];
      int   kp, kd;
      shape_order<false, O>(k, dx3, kp, S3p);
      shape_order<true, O>(k, dx3, kd, S3d);
      e0[0] = spl3<O>(F, 0, id, jp, kp, S1d, S2p, S3p);
      e0[1] = spl3<O>(F, 1, ip, jd, kp, S1p, S2d, S3p);
      e0[2] = spl3<O>(F, 2, ip, jp, kd, S1p, S2p, S3d);
      b0[0] = spl3<O>(F, 3, ip, jd, kd, S1p, S2d, S3d);
      b0[1] = spl3<O>(F, 4, id, jp, kd, S1d, S2p, S3d);
      b0[2] = spl3<O>(F, 5, id, jd, kp, S1d, S2d, S3p);
    }
  }
}

// Ghost-extended index range [lo, hi] of the interpolation stencil around
// cell i (ghost-extended), used for bounds checks: orders 0-2 read at most
// [i - 1, i + 2].
template <int O>
__device__ __forceinline__ void interp_reach(int& lo, int& hi) {
  if constexpr (O == 0) { lo = -1; hi = 1; }
  else if constexpr (O == 1) { lo = -1; hi = 1; }
  else { lo = -1; hi = 2; }
}

// ---- momentum update: sr.hpp:337-437 -----------------------------------------
__device__ __forceinline__ void boris(float coeff, float* u, float* e0, float* b0) {
  float COEFF = coeff;
  e0[0] *= COEFF; e0[1] *= COEFF; e0[2] *= COEFF;
  float u0[3] = { u[0] + e0[0], u[1] + e0[1], u[2] + e0[2] };
  COEFF *= ONE / sqrtf(ONE + (u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]));
  b0[0] *= COEFF; b0[1] *= COEFF; b0[2] *= COEFF;
  COEFF = TWO / (ONE + (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]));
  const float u1[3] = { (u0[0] + (u0[1] * b0[2] - u0[2] * b0[1])) * COEFF,
                        (u0[1] + (u0[2] * b0[0] - u0[0] * b0[2])) * COEFF,
                        (u0[2] + (u0[0] * b0[1] - u0[1] * b0[0])) * COEFF };
  u0[0] += (u1[1] * b0[2] - u1[2] * b0[1]) + e0[0];
  u0[1] += (u1[2] * b0[0] - u1[0] * b0[2]) + e0[1];
  u0[2] += (u1[0] * b0[1] - u1[1] * b0[0]) + e0[2];
  u[0] = u0[0]; u[1] = u0[1]; u[2] = u0[2];
}

__device__ __forceinline__ void vay(float coeff, float* u, float* e0, float* b0) {
  float COEFF = coeff;
  e0[0] *= COEFF; e0[1] *= COEFF; e0[2] *= COEFF;
  b0[0] *= COEFF; b0[1] *= COEFF; b0[2] *= COEFF;
  COEFF = ONE / sqrtf(ONE + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  const float u1[3] = { (u[0] + TWO * e0[0] + (u[1] * b0[2] - u[2] * b0[1]) * COEFF),
                        (u[1] + TWO * e0[1] + (u[2] * b0[0] - u[0] * b0[2]) * COEFF),
                        (u[2] + TWO * e0[2] + (u[0] * b0[1] - u[1] * b0[0]) * COEFF) };
  COEFF = u1[0] * b0[0] + u1[1] * b0[1] + u1[2] * b0[2];
  float COEFF2 = ONE + (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]) -
                 (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]);
  COEFF = ONE / sqrtf(INV_2 * (COEFF2 + sqrtf(COEFF2 * COEFF2 +
                                              FOUR * (b0[0] * b0[0] + b0[1] * b0[1] +
                                                      b0[2] * b0[2] + COEFF * COEFF))));
  COEFF2 = ONE / (ONE + (b0[0] * COEFF) * (b0[0] * COEFF) + (b0[1] * COEFF) * (b0[1] * COEFF) +
                  (b0[2] * COEFF) * (b0[2] * COEFF));
  const float ub = u1[0] * b0[0] + u1[1] * b0[1] + u1[2] * b0[2];
  u[0] = COEFF2 * (u1[0] + COEFF * ub * (b0[0] * COEFF) + u1[1] * b0[2] * COEFF -
                   u1[2] * b0[1] * COEFF);
  u[1] = COEFF2 * (u1[1] + COEFF * ub * (b0[1] * COEFF) + u1[2] * b0[0] * COEFF -
                   u1[0] * b0[2] * COEFF);
  u[2] = COEFF2 * (u1[2] + COEFF * ub * (b0[2] * COEFF) + u1[0] * b0[1] * COEFF -
                   u1[1] * b0[0] * COEFF);
}

// ---- position update, Cartesian: sr.hpp:528-572 ------------------------------
__device__ __forceinline__ void push_coord(int& i, float& dx, float v_dt) {
  dx += v_dt;
  i += static_cast<int>(dx >= ONE) - static_cast<int>(dx < ZERO);
  dx -= static_cast<float>(dx >= ONE);
  dx += static_cast<float>(dx < ZERO);
}

// ---- boundary conditions, Cartesian: sr.hpp:659-790 --------------------------
struct PrtlBCs {
  uint8_t lo[3], hi[3];
  int32_t ni[3];
};

__device__ __forceinline__ void bc_dir(int& i, int& ip, float& dx, float& u, int16_t& tag, int ni,
                                       uint8_t bmin, uint8_t bmax, bool allow_axis) {
  bool invert = false;
  if (i < 0) {
    if (bmin == NTT_PRTL_BC_PERIODIC) { i += ni; ip += ni; }
    else if (bmin == NTT_PRTL_BC_ABSORB || bmin == NTT_PRTL_BC_ATMOSPHERE) tag = NTT_TAG_DEAD;
    else if (bmin == NTT_PRTL_BC_REFLECT) { i = 0; dx = ONE - dx; invert = true; }
    else if (allow_axis && bmin == NTT_PRTL_BC_AXIS) { i = 0; dx = ONE - dx; }
  } else if (i >= ni) {
    if (bmax == NTT_PRTL_BC_PERIODIC) { i -= ni; ip -= ni; }
    else if (bmax == NTT_PRTL_BC_ABSORB || bmax == NTT_PRTL_BC_ATMOSPHERE) tag = NTT_TAG_DEAD;
    else if (bmax == NTT_PRTL_BC_REFLECT) { i = ni - 1; dx = ONE - dx; invert = true; }
    else if (allow_axis && bmax == NTT_PRTL_BC_AXIS) { i = ni - 1; dx = ONE - dx; }
  }
  if (invert) u = -u;
}

// Per-particle state of one SR update, held in registers.
struct PState {
  int   i[3], ip[3];
  float dx[3], dxp[3];
  float u[3];
  int16_t tag;
};

// Momentum and position update of a particle (alive check done by the
// caller): interpolation, momentum update, position update -- without the
// boundary conditions.  sr.hpp:117-332 (no-policy path) + 526-572.
template <int D, int O, int ALG, class A>
__device__ __forceinline__ void sr_move(const A& F, int ngh, float ndth, float dt, float dxm,
                                        PState& s) {
  if constexpr (ALG != ALG_PHOTON) {
    float e0[3], b0[3];
    interpolate<D, O>(F, s.i[0] + ngh, s.i[1] + ngh, D == 3 ? s.i[2] + ngh : 0, s.dx[0], s.dx[1],
                      D == 3 ? s.dx[2] : 0.0f, e0, b0);
    // transform_xyz<U, XYZ> (minkowski.h:282-289): in-plane components times
    // dx (the identity for dx = 1, skipped by a branch uniform over the launch)
    if (dxm != ONE) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        e0[d] *= dxm;
        b0[d] *= dxm;
      }
    }
    if constexpr (ALG == ALG_BORIS) boris(ndth, s.u, e0, b0);
    else vay(ndth, s.u, e0, b0);
  }
  const float ux1 = s.u[0], ux2 = s.u[1], ux3 = s.u[2];
  const float dt_inv_energy = (ALG != ALG_PHOTON)
                                ? (dt / sqrtf(ONE + ux1 * ux1 + ux2 * ux2 + ux3 * ux3))
                                : (dt / sqrtf(ux1 * ux1 + ux2 * ux2 + ux3 * ux3));
  // transform<d, XYZ, U>: u / dx (minkowski.h:277-281), exact without the
  // division for dx = 1
  float ud[3] = { s.u[0], s.u[1], s.u[2] };
  if (dxm != ONE) {
#pragma unroll
    for (int d = 0; d < D; ++d) ud[d] = s.u[d] / dxm;
  }
#pragma unroll
  for (int d = 0; d < D; ++d) {
    s.ip[d]  = s.i[d];
    s.dxp[d] = s.dx[d];
    push_coord(s.i[d], s.dx[d], ud[d] * dt_inv_energy);
  }
}

// boundaryConditions, Cartesian (sr.hpp:659-790)
template <int D>
__device__ __forceinline__ void sr_apply_bcs(const PrtlBCs& bc, PState& s) {
#pragma unroll
  for (int d = 0; d < D; ++d)
    bc_dir(s.i[d], s.ip[d], s.dx[d], s.u[d], s.tag, bc.ni[d], bc.lo[d], bc.hi[d], d == 1);
}

// periodic wrap of i and i_prev in every direction (sr.hpp:663-668 and the
// analogous branches), the only boundary the fused step path accepts
template <int D>
__device__ __forceinline__ void sr_wrap_periodic(const int32_t* ni, PState& s) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (s.i[d] < 0) {
      s.i[d] += ni[d];
      s.ip[d] += ni[d];
    } else if (s.i[d] >= ni[d]) {
      s.i[d] -= ni[d];
      s.ip[d] -= ni[d];
    }
  }
}

// One full SR update of a particle: sr_move followed by the boundaries.
template <int D, int O, int ALG, class A>
__device__ __forceinline__ void sr_update(const A& F, int ngh, float ndth, float dt, float dxm,
                                          const PrtlBCs& bc, PState& s) {
  sr_move<D, O, ALG>(F, ngh, ndth, dt, dxm, s);
  sr_apply_bcs<D>(bc, s);
}

// ---- current deposit: currents_deposit.hpp:108-761 ---------------------------
// A "J sink" receives additions at ghost-extended indices: add(i, j, k, comp, v)
// at absolute indices, or set_origin(i0, j0, k0) once per particle followed by
// add_rel(a, b, g, comp, v) at (i0 + a, j0 + b, k0 + g).

// velocity for the out-of-plane 2D component: currents_deposit.hpp:112-166
__device__ __forceinline__ float vp_out_of_plane(const float* u) {
  const float inv_energy = ONE / sqrtf(ONE + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  float v = u[2];
  if (isnan(v) || isinf(v)) v = ZERO;
  return v * inv_energy;
}

// zig-zag deposit, O = 0: currents_deposit.hpp:171-405
template <int D, class Sink>
__device__ __forceinline__ void deposit_zigzag(Sink& J, int G, const PState& s, float vz,
                                               float coeff, float inv_dt) {
  float W1[3], W2[3], F1[3], F2[3];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int   i = s.i[d], ip = s.ip[d];
    const float dxp_r = static_cast<float>(i == ip) * (s.dx[d] + s.dxp[d]) * INV_2;
    const int   gt    = static_cast<int>(i > ip);
    W1[d] = INV_2 * (dxp_r + s.dxp[d] + static_cast<float>(gt));
    W2[d] = INV_2 * (s.dx[d] + dxp_r + static_cast<float>(gt + ip - i));
    F1[d] = (static_cast<float>(gt) + dxp_r - s.dxp[d]) * coeff * inv_dt;
    F2[d] = (static_cast<float>(i - ip - gt) + s.dx[d] - dxp_r) * coeff * inv_dt;
  }
  const int a = s.ip[0] + G, b = s.ip[1] + G, A = s.i[0] + G, B = s.i[1] + G;
  if constexpr (D == 2) {
    const float Fx3_1 = HALF * vz * coeff;
    const float Fx3_2 = HALF * vz * coeff;
    J.add(a, b, 0, 0, F1[0] * (ONE - W1[1]));
    J.add(a, b + 1, 0, 0, F1[0] * W1[1]);
    J.add(A, B, 0, 0, F2[0] * (ONE - W2[1]));
    J.add(A, B + 1, 0, 0, F2[0] * W2[1]);
    J.add(a, b, 0, 1, F1[1] * (ONE - W1[0]));
    J.add(a + 1, b, 0, 1, F1[1] * W1[0]);
    J.add(A, B, 0, 1, F2[1] * (ONE - W2[0]));
    J.add(A + 1, B, 0, 1, F2[1] * W2[0]);
    J.add(a, b, 0, 2, Fx3_1 * (ONE - W1[0]) * (ONE - W1[1]));
    J.add(a + 1, b, 0, 2, Fx3_1 * W1[0] * (ONE - W1[1]));
    J.add(a, b + 1, 0, 2, Fx3_1 * (ONE - W1[0]) * W1[1]);
    J.add(a + 1, b + 1, 0, 2, Fx3_1 * W1[0] * W1[1]);
    J.add(A, B, 0, 2, Fx3_2 * (ONE - W2[0]) * (ONE - W2[1]));
    J.add(A + 1, B, 0, 2, Fx3_2 * W2[0] * (ONE - W2[1]));
    J.add(A, B + 1, 0, 2, Fx3_2 * (ONE - W2[0]) * W2[1]);
    J.add(A + 1, B + 1, 0, 2, Fx3_2 * W2[0] * W2[1]);
  } else {
    const int c = s.ip[2] + G, C = s.i[2] + G;
    J.add(a, b, c, 0, F1[0] * (ONE - W1[1]) * (ONE - W1[2]));
    J.add(a, b + 1, c, 0, F1[0] * W1[1] * (ONE - W1[2]));
    J.add(a, b, c + 1, 0, F1[0] * (ONE - W1[1]) * W1[2]);
    J.add(a, b + 1, c + 1, 0, F1[0] * W1[1] * W1[2]);
    J.add(A, B, C, 0, F2[0] * (ONE - W2[1]) * (ONE - W2[2]));
    J.add(A, B + 1, C, 0, F2[0] * W2[1] * (ONE - W2[2]));
    J.add(A, B, C + 1, 0, F2[0] * (ONE - W2[1]) * W2[2]);
    J.add(A, B + 1, C + 1, 0, F2[0] * W2[1] * W2[2]);
    J.add(a, b, c, 1, F1[1] * (ONE - W1[0]) * (ONE - W1[2]));
    J.add(a + 1, b, c, 1, F1[1] * W1[0] * (ONE - W1[2]));
    J.add(a, b, c + 1, 1, F1[1] * (ONE - W1[0]) * W1[2]);
    J.add(a + 1, b, c + 1, 1, F1[1] * W1[0] * W1[2]);
    J.add(A, B, C, 1, F2[1] * (ONE - W2[0]) * (ONE - W2[2]));
    J.add(A + 1, B, C, 1, F2[1] * W2[0] * (ONE - W2[2]));
    J.add(A, B, C + 1, 1, F2[1] * (ONE - W2[0]) * W2[2]);
    J.add(A + 1, B, C + 1, 1, F2[1] * W2[0] * W2[2]);
    J.add(a, b, c, 2, F1[2] * (ONE - W1[0]) * (ONE - W1[1]));
    J.add(a + 1, b, c, 2, F1[2] * W1[0] * (ONE - W1[1]));
    J.add(a, b + 1, c, 2, F1[2] * (ONE - W1[0]) * W1[1]);
    J.add(a + 1, b + 1, c, 2, F1[2] * W1[0] * W1[1]);
    J.add(A, B, C, 2, F2[2] * (ONE - W2[0]) * (ONE - W2[1]));
    J.add(A + 1, B, C, 2, F2[2] * W2[0] * (ONE - W2[1]));
    J.add(A, B + 1, C, 2, F2[2] * (ONE - W2[0]) * W2[1]);
    J.add(A + 1, B + 1, C, 2, F2[2] * W2[0] * W2[1]);
  }
}

// Esirkepov deposit, O = 1, 2 in 2D: currents_deposit.hpp:476-555
template <int O, class Sink>
__device__ __forceinline__ void deposit_esirkepov_2d(Sink& J, int G, const PState& s, float vz,
                                                     float coeff, float inv_dt) {
  constexpr int M = O + 2;
  float iS1[M], fS1[M], iS2[M], fS2[M];
  int   i1_min, d1, i2_min, d2;
  for_deposit<O>(s.ip[0], s.dxp[0], s.i[0], s.dx[0], i1_min, d1, iS1, fS1);
  for_deposit<O>(s.ip[1], s.dxp[1], s.i[1], s.dx[1], i2_min, d2, iS2, fS2);
  const float Qdt = coeff * inv_dt;
  const float QVx3 = coeff * vz;
  i1_min += G;
  i2_min += G;
  J.set_origin(i1_min, i2_min, 0);
  // jx1: prefix sum over i of -Q*Wx1.  The per-slot coefficients
  // -Qdt/2 (fS - iS) and the transverse sums are formed once, so every slot
  // is one FMA (a reassociation of the reference's product: round-off only)
  float c1[M], c2[M], s1[M], s2[M];
#pragma unroll
  for (int a = 0; a < M; ++a) {
    c1[a] = -Qdt * (HALF * (fS1[a] - iS1[a]));
    s1[a] = fS1[a] + iS1[a];
  }
#pragma unroll
  for (int b = 0; b < M; ++b) {
    c2[b] = -Qdt * (HALF * (fS2[b] - iS2[b]));
    s2[b] = fS2[b] + iS2[b];
  }
  float jx1[M][M], jx2[M][M];
#pragma unroll
  for (int b = 0; b < M; ++b) jx1[0][b] = c1[0] * s2[b];
#pragma unroll
  for (int a = 1; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b) jx1[a][b] = fmaf(c1[a], s2[b], jx1[a - 1][b]);
#pragma unroll
  for (int a = 0; a < M; ++a) jx2[a][0] = c2[0] * s1[a];
#pragma unroll
  for (int b = 1; b < M; ++b)
#pragma unroll
    for (int a = 0; a < M; ++a) jx2[a][b] = fmaf(c2[b], s1[a], jx2[a][b - 1]);
  // jx3 = QVx3 * THIRD * (fS2 (iS1/2 + fS1) + iS2 (fS1/2 + iS1))
  const float q3 = QVx3 * THIRD;
  float p1[M], r1[M];
#pragma unroll
  for (int a = 0; a < M; ++a) {
    p1[a] = HALF * iS1[a] + fS1[a];
    r1[a] = HALF * fS1[a] + iS1[a];
  }
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b) {
      if (a < d1 && b <= d2) J.add_rel(a, b, 0, 0, jx1[a][b]);
      if (a <= d1 && b < d2) J.add_rel(a, b, 0, 1, jx2[a][b]);
      if (a <= d1 && b <= d2) J.add_rel(a, b, 0, 2, q3 * fmaf(fS2[b], p1[a], iS2[b] * r1[a]));
    }
}

// Esirkepov deposit, O = 1, 2 in 3D: currents_deposit.hpp:557-754
template <int O, class Sink>
__device__ __forceinline__ void deposit_esirkepov_3d(Sink& J, int G, const PState& s, float coeff,
                                                     float inv_dt) {
  constexpr int M = O + 2;
  float iS1[M], fS1[M], iS2[M], fS2[M], iS3[M], fS3[M];
  int   i1_min, d1, i2_min, d2, i3_min, d3;
  for_deposit<O>(s.ip[0], s.dxp[0], s.i[0], s.dx[0], i1_min, d1, iS1, fS1);
  for_deposit<O>(s.ip[1], s.dxp[1], s.i[1], s.dx[1], i2_min, d2, iS2, fS2);
  for_deposit<O>(s.ip[2], s.dxp[2], s.i[2], s.dx[2], i3_min, d3, iS3, fS3);
  const float Qdt = coeff * inv_dt;
  i1_min += G;
  i2_min += G;
  i3_min += G;
  J.set_origin(i1_min, i2_min, i3_min);
  // Separable form of the reference's weights: the current along x at slot
  // (a, b, g) is C1[a] * T23[b][g], with C1[a] the prefix over a' <= a of
  // -Qdt/3 (fS1 - iS1)[a'] and the transverse weight
  //   T23 = iS2 iS3 + fS2 fS3 + (iS2 fS3 + fS2 iS3) / 2
  //       = (iS2 + fS2 / 2) iS3 + (fS2 + iS2 / 2) fS3,
  // two VALU per transverse pair and one per slot (a reassociation of the
  // reference's products and prefix sums: round-off only).  The prefix over
  // all M slots vanishes (charge conservation), so the longitudinal slots
  // run over a <= O only, the last of them when the particle changed its
  // stencil along that direction (a < d).
  const float q3 = -Qdt * THIRD;
  float C1[O + 1], C2[O + 1], C3[O + 1];
  C1[0] = q3 * (fS1[0] - iS1[0]);
  C2[0] = q3 * (fS2[0] - iS2[0]);
  C3[0] = q3 * (fS3[0] - iS3[0]);
#pragma unroll
  for (int a = 1; a <= O; ++a) {
    C1[a] = fmaf(q3, fS1[a] - iS1[a], C1[a - 1]);
    C2[a] = fmaf(q3, fS2[a] - iS2[a], C2[a - 1]);
    C3[a] = fmaf(q3, fS3[a] - iS3[a], C3[a - 1]);
  }
  float P1[M], Q1[M], P2[M], Q2[M];
#pragma unroll
  for (int a = 0; a < M; ++a) {
    P1[a] = fmaf(HALF, fS1[a], iS1[a]);
    Q1[a] = fmaf(HALF, iS1[a], fS1[a]);
    P2[a] = fmaf(HALF, fS2[a], iS2[a]);
    Q2[a] = fmaf(HALF, iS2[a], fS2[a]);
  }
  // jx1
#pragma unroll
  for (int b = 0; b < M; ++b) {
#pragma unroll
    for (int g = 0; g < M; ++g) {
      const float t = fmaf(P2[b], iS3[g], Q2[b] * fS3[g]);
#pragma unroll
      for (int a = 0; a <= O; ++a)
        if (a < d1 && b <= d2 && g <= d3) J.add_rel(a, b, g, 0, C1[a] * t);
    }
  }
  // jx2
#pragma unroll
  for (int a = 0; a < M; ++a) {
#pragma unroll
    for (int g = 0; g < M; ++g) {
      const float t = fmaf(P1[a], iS3[g], Q1[a] * fS3[g]);
#pragma unroll
      for (int b = 0; b <= O; ++b)
        if (a <= d1 && b < d2 && g <= d3) J.add_rel(a, b, g, 1, C2[b] * t);
    }
  }
  // jx3
#pragma unroll
  for (int a = 0; a < M; ++a) {
#pragma unroll
    for (int b = 0; b < M; ++b) {
      const float t = fmaf(P1[a], iS2[b], Q1[a] * fS2[b]);
#pragma unroll
      for (int g = 0; g <= O; ++g)
        if (a <= d1 && b <= d2 && g < d3) J.add_rel(a, b, g, 2, C3[g] * t);
    }
  }
}

// Whether a particle's merged deposit stencil keeps its O + 1 cells along
// direction d (for_deposit's i_min of the initial and final shapes agree),
// i.e. the particle does not need the extra slot of a stencil that shifted.
template <int O>
__device__ __forceinline__ bool deposit_stays(int i, float dx, int ip, float dxp) {
  if constexpr (O == 2) {
    const int a = (dxp < HALF) ? ip - 1 : ip;
    const int b = (dx < HALF) ? i - 1 : i;
    return a == b;
  } else {
    return i == ip;
  }
}

// deposit_esirkepov_3d for a particle whose stencil stays in every direction
// (deposit_stays true for d = 1, 2, 3): the same values, slot for slot, with
// the loop bounds of that case known at compile time (3 (O + 1)^2 O slots, no
// per-slot conditions).
template <int O, class Sink>
__device__ __forceinline__ void deposit_esirkepov_3d_stay(Sink& J, int G, const PState& s,
                                                          float coeff, float inv_dt) {
  constexpr int M = O + 1;
  float iS1[M], fS1[M], iS2[M], fS2[M], iS3[M], fS3[M];
  int   i1_min, i2_min, i3_min, unused;
  shape_order<false, O, true>(s.ip[0], s.dxp[0], i1_min, iS1);
  shape_order<false, O, true>(s.i[0], s.dx[0], unused, fS1);
  shape_order<false, O, true>(s.ip[1], s.dxp[1], i2_min, iS2);
  shape_order<false, O, true>(s.i[1], s.dx[1], unused, fS2);
  shape_order<false, O, true>(s.ip[2], s.dxp[2], i3_min, iS3);
  shape_order<false, O, true>(s.i[2], s.dx[2], unused, fS3);
  const float Qdt = coeff * inv_dt;
  J.set_origin(i1_min + G, i2_min + G, i3_min + G);
  const float q3 = -Qdt * THIRD;
  float C1[O], C2[O], C3[O];
  C1[0] = q3 * (fS1[0] - iS1[0]);
  C2[0] = q3 * (fS2[0] - iS2[0]);
  C3[0] = q3 * (fS3[0] - iS3[0]);
#pragma unroll
  for (int a = 1; a < O; ++a) {
    C1[a] = fmaf(q3, fS1[a] - iS1[a], C1[a - 1]);
    C2[a] = fmaf(q3, fS2[a] - iS2[a], C2[a - 1]);
    C3[a] = fmaf(q3, fS3[a] - iS3[a], C3[a - 1]);
  }
  float P1[M], Q1[M], P2[M], Q2[M];
#pragma unroll
  for (int a = 0; a < M; ++a) {
    P1[a] = fmaf(HALF, fS1[a], iS1[a]);
    Q1[a] = fmaf(HALF, iS1[a], fS1[a]);
    P2[a] = fmaf(HALF, fS2[a], iS2[a]);
    Q2[a] = fmaf(HALF, iS2[a], fS2[a]);
  }
#pragma unroll
  for (int b = 0; b < M; ++b)
#pragma unroll
    for (int g = 0; g < M; ++g) {
      const float t = fmaf(P2[b], iS3[g], Q2[b] * fS3[g]);
#pragma unroll
      for (int a = 0; a < O; ++a) J.add_rel(a, b, g, 0, C1[a] * t);
    }
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int g = 0; g < M; ++g) {
      const float t = fmaf(P1[a], iS3[g], Q1[a] * fS3[g]);
#pragma unroll
      for (int b = 0; b < O; ++b) J.add_rel(a, b, g, 1, C2[b] * t);
    }
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b) {
      const float t = fmaf(P1[a], iS2[b], Q1[a] * fS2[b]);
#pragma unroll
      for (int g = 0; g < O; ++g) J.add_rel(a, b, g, 2, C3[g] * t);
    }
}

template <int D, int O, class Sink>
__device__ __forceinline__ void deposit_particle(Sink& J, int G, const PState& s, float coeff,
                                                 float inv_dt) {
  float vz = ZERO;
  if constexpr (D == 2) vz = vp_out_of_plane(s.u);
  if constexpr (O == 0) deposit_zigzag<D>(J, G, s, vz, coeff, inv_dt);
  else if constexpr (D == 2) deposit_esirkepov_2d<O>(J, G, s, vz, coeff, inv_dt);
  else deposit_esirkepov_3d<O>(J, G, s, coeff, inv_dt);
}

// Exact index range [lo, hi] (particle indices, without ghosts) written by
// the deposit along one direction: the zig-zag touches i_prev, i and the
// node above each; Esirkepov spans [i_min, i_max] of for_deposit.
template <int O>
__device__ __forceinline__ void deposit_extent(int i, float dx, int ip, float dxp, int& lo,
                                               int& hi) {
  if constexpr (O == 0) {
    lo = min(i, ip);
    hi = max(i, ip) + 1;
  } else {
    const int a = (O == 2 && dxp < HALF) ? ip - 1 : ip;
    const int b = (O == 2 && dx < HALF) ? i - 1 : i;
    lo = min(a, b);
    hi = (a == b) ? a + O : lo + O + 1;
  }
}

}  // namespace ntt
