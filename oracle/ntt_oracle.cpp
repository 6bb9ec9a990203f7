/*
 * ntt_oracle.cpp -- CPU restatement of Entity's SR PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline timed beside it.  Product code never links or calls this file.
 *
 * Each function follows the reference file:line cited above it, in the
 * reference's operation order, so that with -ffp-contract=off (and IEEE
 * single/double arithmetic) it reproduces the reference's arithmetic for one
 * particle or cell.  Citations are relative to the Entity source tree.
 *
 * Sequential by design: particles are visited in index order and currents are
 * accumulated in that order, so every result is deterministic.  (The
 * reference accumulates currents through a Kokkos ScatterView whose order is
 * backend-dependent; J is therefore compared with a tolerance.)
 */
#include "ntt_oracle.h"

#include <array>
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ntt_oracle_common.hpp"

// When set (by the *_omp entry points only), the step's loops run with OpenMP
// threads and the deposit goes to per-thread copies of J summed afterwards,
// as Kokkos' OpenMP ScatterView does.  The checker entry points leave it off
// and keep the sequential order.
static thread_local bool t_par = false;

namespace {

using namespace ort;


/* ---- field interpolation: src/kernels/pushers/sr.hpp:851-1370 ---------- */
/* O = 0: "direct" linear interpolation with dual/primal weights, 2D
 * (sr.hpp:905-985) and 3D (sr.hpp:986-1117). */
template <typename R>
void interp_o0_2d(const Fld<R>& EB, int i1, int i2, float fdx1, float fdx2, R* e0, R* b0) {
  using C = K<R>;
  const int i = i1 + EB.ngh, j = i2 + EB.ngh;
  const R   dx1_ = static_cast<R>(fdx1), dx2_ = static_cast<R>(fdx2);
  int indx = static_cast<int>(dx1_ + C::HALF);
  int indy = static_cast<int>(dx2_ + C::HALF);
  R c000, c100, c010, c110, c00, c10;
  R ponpmx = C::ONE - dx1_, ponppx = dx1_;
  R ponpmy = C::ONE - dx2_, ponppy = dx2_;
  R pondmx = static_cast<R>(indx + 1) - (dx1_ + C::HALF);
  R pondpx = C::ONE - pondmx;
  R pondmy = static_cast<R>(indy + 1) - (dx2_ + C::HALF);
  R pondpy = C::ONE - pondmy;
  // Ex1 (dual, primal)
  c000 = EB(i - 1 + indx, j, 0); c100 = EB(i + indx, j, 0);
  c010 = EB(i - 1 + indx, j + 1, 0); c110 = EB(i + indx, j + 1, 0);
  c00 = c000 * pondmx + c100 * pondpx; c10 = c010 * pondmx + c110 * pondpx;
  e0[0] = c00 * ponpmy + c10 * ponppy;
  // Ex2 (primal, dual)
  c000 = EB(i, j - 1 + indy, 1); c100 = EB(i + 1, j - 1 + indy, 1);
  c010 = EB(i, j + indy, 1); c110 = EB(i + 1, j + indy, 1);
  c00 = c000 * ponpmx + c100 * ponppx; c10 = c010 * ponpmx + c110 * ponppx;
  e0[1] = c00 * pondmy + c10 * pondpy;
  // Ex3 (primal, primal)
  c000 = EB(i, j, 2); c100 = EB(i + 1, j, 2);
  c010 = EB(i, j + 1, 2); c110 = EB(i + 1, j + 1, 2);
  c00 = c000 * ponpmx + c100 * ponppx; c10 = c010 * ponpmx + c110 * ponppx;
  e0[2] = c00 * ponpmy + c10 * ponppy;
  // Bx1 (primal, dual)
  c000 = EB(i, j - 1 + indy, 3); c100 = EB(i + 1, j - 1 + indy, 3);
  c010 = EB(i, j + indy, 3); c110 = EB(i + 1, j + indy, 3);
  c00 = c000 * ponpmx + c100 * ponppx; c10 = c010 * ponpmx + c110 * ponppx;
  b0[0] = c00 * pondmy + c10 * pondpy;
  // Bx2 (dual, primal)
  c000 = EB(i - 1 + indx, j, 4); c100 = EB(i + indx, j, 4);
  c010 = EB(i - 1 + indx, j + 1, 4); c110 = EB(i + indx, j + 1, 4);
  c00 = c000 * pondmx + c100 * pondpx; c10 = c010 * pondmx + c110 * pondpx;
  b0[1] = c00 * ponpmy + c10 * ponppy;
  // Bx3 (dual, dual)
  c000 = EB(i - 1 + indx, j - 1 + indy, 5); c100 = EB(i + indx, j - 1 + indy, 5);
  c010 = EB(i - 1 + indx, j + indy, 5); c110 = EB(i + indx, j + indy, 5);
  c00 = c000 * pondmx + c100 * pondpx; c10 = c010 * pondmx + c110 * pondpx;
  b0[2] = c00 * pondmy + c10 * pondpy;
}

template <typename R>
inline R trilin(const Fld<R>& EB, int c, int ia, int ja, int ka, R wx0, R wx1,
                R wy0, R wy1, R wz0, R wz1) {
  // sr.hpp:1013-1027 pattern: x first, then y, then z.
  const R c000 = EB.at(ia, ja, ka, c), c100 = EB.at(ia + 1, ja, ka, c);
  const R c010 = EB.at(ia, ja + 1, ka, c), c110 = EB.at(ia + 1, ja + 1, ka, c);
  const R c001 = EB.at(ia, ja, ka + 1, c), c101 = EB.at(ia + 1, ja, ka + 1, c);
  const R c011 = EB.at(ia, ja + 1, ka + 1, c), c111 = EB.at(ia + 1, ja + 1, ka + 1, c);
  const R c00 = c000 * wx0 + c100 * wx1;
  const R c10 = c010 * wx0 + c110 * wx1;
  const R c0  = c00 * wy0 + c10 * wy1;
  const R c01 = c001 * wx0 + c101 * wx1;
  const R c11 = c011 * wx0 + c111 * wx1;
  const R c1  = c01 * wy0 + c11 * wy1;
  return c0 * wz0 + c1 * wz1;
}

template <typename R>
void interp_o0_3d(const Fld<R>& EB, int i1, int i2, int i3, float fdx1, float fdx2,
                  float fdx3, R* e0, R* b0) {
  using C = K<R>;
  const int i = i1 + EB.ngh, j = i2 + EB.ngh, k = i3 + EB.ngh;
  const R   dx1_ = fdx1, dx2_ = fdx2, dx3_ = fdx3;
  int indx = static_cast<int>(dx1_ + C::HALF);
  int indy = static_cast<int>(dx2_ + C::HALF);
  int indz = static_cast<int>(dx3_ + C::HALF);
  R ponpmx = C::ONE - dx1_, ponppx = dx1_;
  R ponpmy = C::ONE - dx2_, ponppy = dx2_;
  R ponpmz = C::ONE - dx3_, ponppz = dx3_;
  R pondmx = static_cast<R>(indx + 1) - (dx1_ + C::HALF), pondpx = C::ONE - pondmx;
  R pondmy = static_cast<R>(indy + 1) - (dx2_ + C::HALF), pondpy = C::ONE - pondmy;
  R pondmz = static_cast<R>(indz + 1) - (dx3_ + C::HALF), pondpz = C::ONE - pondmz;
  // Ex1 (dual, primal, primal)
  e0[0] = trilin(EB, 0, i - 1 + indx, j, k, pondmx, pondpx, ponpmy, ponppy, ponpmz, ponppz);
  // Ex2 (primal, dual, primal)
  e0[1] = trilin(EB, 1, i, j - 1 + indy, k, ponpmx, ponppx, pondmy, pondpy, ponpmz, ponppz);
  // Ex3 (primal, primal, dual)
  e0[2] = trilin(EB, 2, i, j, k - 1 + indz, ponpmx, ponppx, ponpmy, ponppy, pondmz, pondpz);
  // Bx1 (primal, dual, dual)
  b0[0] = trilin(EB, 3, i, j - 1 + indy, k - 1 + indz, ponpmx, ponppx, pondmy, pondpy, pondmz, pondpz);
  // Bx2 (dual, primal, dual)
  b0[1] = trilin(EB, 4, i - 1 + indx, j, k - 1 + indz, pondmx, pondpx, ponpmy, ponppy, pondmz, pondpz);
  // Bx3 (dual, dual, primal)
  b0[2] = trilin(EB, 5, i - 1 + indx, j - 1 + indy, k, pondmx, pondpx, pondmy, pondpy, ponpmz, ponppz);
}

/* O >= 1: tensor-product B-spline interpolation, sr.hpp:1170-1366.  The
 * accumulation order is x innermost, then y, then z. */
template <typename R, int O>
R gather2d(const Fld<R>& EB, int c, int imin, int jmin, const R* S1, const R* S2) {
  R e = K<R>::ZERO;
  for (int idx2 = 0; idx2 < O + 1; idx2++) {
    R c0 = K<R>::ZERO;
    for (int idx1 = 0; idx1 < O + 1; idx1++) c0 += S1[idx1] * EB(imin + idx1, jmin + idx2, c);
    e += c0 * S2[idx2];
  }
  return e;
}

template <typename R, int O>
R gather3d(const Fld<R>& EB, int c, int imin, int jmin, int kmin, const R* S1,
           const R* S2, const R* S3) {
  R e = K<R>::ZERO;
  for (int idx3 = 0; idx3 < O + 1; idx3++) {
    R c0 = K<R>::ZERO;
    for (int idx2 = 0; idx2 < O + 1; idx2++) {
      R c00 = K<R>::ZERO;
      for (int idx1 = 0; idx1 < O + 1; idx1++)
        c00 += S1[idx1] * EB.at(imin + idx1, jmin + idx2, kmin + idx3, c);
      c0 += c00 * S2[idx2];
    }
    e += c0 * S3[idx3];
  }
  return e;
}

template <typename R, int O>
void interp_shape_2d(const Fld<R>& EB, int i1, int i2, float fdx1, float fdx2, R* e0, R* b0) {
  const int i = i1 + EB.ngh, j = i2 + EB.ngh;
  const R   dx1_ = fdx1, dx2_ = fdx2;
  R   S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1];
  int ip, id, jp, jd;
  shape_order<R, false, O>(i, dx1_, ip, S1p);
  shape_order<R, false, O>(j, dx2_, jp, S2p);
  shape_order<R, true, O>(i, dx1_, id, S1d);
  shape_order<R, true, O>(j, dx2_, jd, S2d);
  e0[0] = gather2d<R, O>(EB, 0, id, jp, S1d, S2p);
  e0[1] = gather2d<R, O>(EB, 1, ip, jd, S1p, S2d);
  e0[2] = gather2d<R, O>(EB, 2, ip, jp, S1p, S2p);
  b0[0] = gather2d<R, O>(EB, 3, ip, jd, S1p, S2d);
  b0[1] = gather2d<R, O>(EB, 4, id, jp, S1d, S2p);
  b0[2] = gather2d<R, O>(EB, 5, id, jd, S1d, S2d);
}

template <typename R, int O>
void interp_shape_3d(const Fld<R>& EB, int i1, int i2, int i3, float fdx1, float fdx2,
                     float fdx3, R* e0, R* b0) {
  const int i = i1 + EB.ngh, j = i2 + EB.ngh, k = i3 + EB.ngh;
  const R   dx1_ = fdx1, dx2_ = fdx2, dx3_ = fdx3;
  R   S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1], S3p[O + 1], S3d[O + 1];
  int ip, id, jp, jd, kp, kd;
  shape_order<R, false, O>(i, dx1_, ip, S1p);
  shape_order<R, false, O>(j, dx2_, jp, S2p);
  shape_order<R, false, O>(k, dx3_, kp, S3p);
  shape_order<R, true, O>(i, dx1_, id, S1d);
  shape_order<R, true, O>(j, dx2_, jd, S2d);
  shape_order<R, true, O>(k, dx3_, kd, S3d);
  e0[0] = gather3d<R, O>(EB, 0, id, jp, kp, S1d, S2p, S3p);
  e0[1] = gather3d<R, O>(EB, 1, ip, jd, kp, S1p, S2d, S3p);
  e0[2] = gather3d<R, O>(EB, 2, ip, jp, kd, S1p, S2p, S3d);
  b0[0] = gather3d<R, O>(EB, 3, ip, jd, kd, S1p, S2d, S3d);
  b0[1] = gather3d<R, O>(EB, 4, id, jp, kd, S1d, S2p, S3d);
  b0[2] = gather3d<R, O>(EB, 5, id, jd, kp, S1d, S2d, S3p);
}

template <typename R>
void interpolate(const Fld<R>& EB, int order, const Prt<R>& P, int64_t p, R* e0, R* b0) {
  if (EB.dim == 2) {
    if (order == 0) interp_o0_2d<R>(EB, P.i1[p], P.i2[p], P.dx1[p], P.dx2[p], e0, b0);
    else if (order == 1) interp_shape_2d<R, 1>(EB, P.i1[p], P.i2[p], P.dx1[p], P.dx2[p], e0, b0);
    else interp_shape_2d<R, 2>(EB, P.i1[p], P.i2[p], P.dx1[p], P.dx2[p], e0, b0);
  } else {
    if (order == 0)
      interp_o0_3d<R>(EB, P.i1[p], P.i2[p], P.i3[p], P.dx1[p], P.dx2[p], P.dx3[p], e0, b0);
    else if (order == 1)
      interp_shape_3d<R, 1>(EB, P.i1[p], P.i2[p], P.i3[p], P.dx1[p], P.dx2[p], P.dx3[p], e0, b0);
    else
      interp_shape_3d<R, 2>(EB, P.i1[p], P.i2[p], P.i3[p], P.dx1[p], P.dx2[p], P.dx3[p], e0, b0);
  }
}

/* ---- velocity pushers: src/kernels/pushers/sr.hpp:337-437 -------------- */
template <typename R>
void boris(R coeff_in, R* u, R* e0, R* b0) {
  using C = K<R>;
  R COEFF = coeff_in;
  e0[0] *= COEFF; e0[1] *= COEFF; e0[2] *= COEFF;
  R u0[3] = { u[0] + e0[0], u[1] + e0[1], u[2] + e0[2] };
  COEFF *= C::ONE / std::sqrt(C::ONE + (u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]));
  b0[0] *= COEFF; b0[1] *= COEFF; b0[2] *= COEFF;
  COEFF = C::TWO / (C::ONE + (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]));
  const R u1[3] = { (u0[0] + (u0[1] * b0[2] - u0[2] * b0[1])) * COEFF,
                    (u0[1] + (u0[2] * b0[0] - u0[0] * b0[2])) * COEFF,
                    (u0[2] + (u0[0] * b0[1] - u0[1] * b0[0])) * COEFF };
  u0[0] += (u1[1] * b0[2] - u1[2] * b0[1]) + e0[0];
  u0[1] += (u1[2] * b0[0] - u1[0] * b0[2]) + e0[1];
  u0[2] += (u1[0] * b0[1] - u1[1] * b0[0]) + e0[2];
  u[0] = u0[0]; u[1] = u0[1]; u[2] = u0[2];
}

template <typename R>
void vay(R coeff_in, R* u, R* e0, R* b0) {
  using C = K<R>;
  R COEFF = coeff_in;
  e0[0] *= COEFF; e0[1] *= COEFF; e0[2] *= COEFF;
  b0[0] *= COEFF; b0[1] *= COEFF; b0[2] *= COEFF;
  COEFF = C::ONE / std::sqrt(C::ONE + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  const R u1[3] = {
    (u[0] + C::TWO * e0[0] + (u[1] * b0[2] - u[2] * b0[1]) * COEFF),
    (u[1] + C::TWO * e0[1] + (u[2] * b0[0] - u[0] * b0[2]) * COEFF),
    (u[2] + C::TWO * e0[2] + (u[0] * b0[1] - u[1] * b0[0]) * COEFF)
  };
  COEFF    = u1[0] * b0[0] + u1[1] * b0[1] + u1[2] * b0[2];
  R COEFF2 = C::ONE + (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]) -
             (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]);
  COEFF = C::ONE /
          std::sqrt(C::INV_2 *
                    (COEFF2 + std::sqrt(COEFF2 * COEFF2 +
                                        C::FOUR * (b0[0] * b0[0] + b0[1] * b0[1] +
                                                   b0[2] * b0[2] + COEFF * COEFF))));
  COEFF2 = C::ONE / (C::ONE + (b0[0] * COEFF) * (b0[0] * COEFF) +
                     (b0[1] * COEFF) * (b0[1] * COEFF) + (b0[2] * COEFF) * (b0[2] * COEFF));
  const R ub = u1[0] * b0[0] + u1[1] * b0[1] + u1[2] * b0[2];
  u[0] = COEFF2 * (u1[0] + COEFF * ub * (b0[0] * COEFF) + u1[1] * b0[2] * COEFF -
                   u1[2] * b0[1] * COEFF);
  u[1] = COEFF2 * (u1[1] + COEFF * ub * (b0[1] * COEFF) + u1[2] * b0[0] * COEFF -
                   u1[0] * b0[2] * COEFF);
  u[2] = COEFF2 * (u1[2] + COEFF * ub * (b0[2] * COEFF) + u1[0] * b0[1] * COEFF -
                   u1[1] * b0[0] * COEFF);
}

/* ---- Cartesian position push: sr.hpp:526-572 ---------------------------- */
template <typename R>
inline void push_coord(int32_t& i, int32_t& ip, float& dx, float& dxp, R u_over_dx, R dt_inv_energy) {
  using C = K<R>;
  ip  = i;
  dxp = dx;
  dx += u_over_dx * dt_inv_energy;  // float += real (computed in real, stored as float)
  i += static_cast<int>(dx >= C::ONE) - static_cast<int>(dx < C::ZERO);
  dx -= (dx >= C::ONE);
  dx += (dx < C::ZERO);
}

/* ---- particle boundary conditions, Cartesian: sr.hpp:659-814 ----------- */
template <typename R>
inline void bc_dir(int32_t& i, int32_t& ip, float& dx, R& u, int16_t& tag, int ni,
                   uint8_t bmin, uint8_t bmax, bool allow_axis) {
  using C = K<R>;
  bool invert = false;
  if (i < 0) {
    if (bmin == BC_PERIODIC) { i += ni; ip += ni; }
    else if (bmin == BC_ABSORB || bmin == BC_ATMOSPHERE) tag = TAG_DEAD;
    else if (bmin == BC_REFLECT) { i = 0; dx = C::ONE - dx; invert = true; }
    else if (allow_axis && bmin == BC_AXIS) { i = 0; dx = C::ONE - dx; }
  } else if (i >= ni) {
    if (bmax == BC_PERIODIC) { i -= ni; ip -= ni; }
    else if (bmax == BC_ABSORB || bmax == BC_ATMOSPHERE) tag = TAG_DEAD;
    else if (bmax == BC_REFLECT) { i = ni - 1; dx = C::ONE - dx; invert = true; }
    else if (allow_axis && bmax == BC_AXIS) { i = ni - 1; dx = C::ONE - dx; }
  }
  if (invert) u = -u;
}

/* ---- SR pusher, one particle: sr.hpp:117-332 (no-policy path) ---------- */
template <typename R>
int sr_push_one(const ort_sr_params_t* prm, const Fld<R>& EB, Prt<R>& P, int64_t p) {
  using C = K<R>;
  if (P.tag[p] != TAG_ALIVE) {
    return (P.tag[p] != TAG_DEAD) ? 1 : 0;  // KernelError "Invalid particle tag"
  }
  const int D = EB.dim;
  const R   dxm = static_cast<R>(prm->dx);
  const bool massive = (prm->pusher_flags != P_PHOTON);
  if (prm->pusher_flags & P_GCA) return 4;  // hybrid GCA mode is not restated
  if (massive) {
    R e0[3], b0[3];
    interpolate<R>(EB, prm->shape_order, P, p, e0, b0);
    // transform_xyz<U, XYZ>: v * sqrt(h_ii) = v * dx for i <= D (minkowski.h:282-289)
    R ec[3], bc[3];
    for (int d = 0; d < 3; ++d) {
      ec[d] = (d < D) ? e0[d] * dxm : e0[d];
      bc[d] = (d < D) ? b0[d] * dxm : b0[d];
    }
    // normalized_dt_half (sr.hpp:111)
    const R ndth = C::HALF * static_cast<R>(prm->charge / prm->mass) *
                   static_cast<R>(prm->omegaB0) * static_cast<R>(prm->dt);
    R u[3] = { P.ux1[p], P.ux2[p], P.ux3[p] };
    if (prm->pusher_flags & P_BORIS) boris<R>(ndth, u, ec, bc);
    else if (prm->pusher_flags & P_VAY) vay<R>(ndth, u, ec, bc);
    else return 4;  // KernelError "Invalid pusher algorithm"
    P.ux1[p] = u[0]; P.ux2[p] = u[1]; P.ux3[p] = u[2];
  }
  // positionPush, Cartesian branch (sr.hpp:528-572)
  const R dt = static_cast<R>(prm->dt);
  const R ux1 = P.ux1[p], ux2 = P.ux2[p], ux3 = P.ux3[p];
  const R dt_inv_energy = massive
    ? (dt / std::sqrt(C::ONE + ux1 * ux1 + ux2 * ux2 + ux3 * ux3))
    : (dt / std::sqrt(ux1 * ux1 + ux2 * ux2 + ux3 * ux3));
  // transform<i, XYZ, U>: v / sqrt(h_ii) = v / dx (minkowski.h:277-281)
  push_coord<R>(P.i1[p], P.i1p[p], P.dx1[p], P.dx1p[p], ux1 / dxm, dt_inv_energy);
  if (D >= 2) push_coord<R>(P.i2[p], P.i2p[p], P.dx2[p], P.dx2p[p], ux2 / dxm, dt_inv_energy);
  if (D >= 3) push_coord<R>(P.i3[p], P.i3p[p], P.dx3[p], P.dx3p[p], ux3 / dxm, dt_inv_energy);
  // boundaryConditions (sr.hpp:659-790)
  bc_dir<R>(P.i1[p], P.i1p[p], P.dx1[p], P.ux1[p], P.tag[p], prm->ni1, prm->prtl_bc[0][0],
            prm->prtl_bc[0][1], false);
  if (D >= 2)
    bc_dir<R>(P.i2[p], P.i2p[p], P.dx2[p], P.ux2[p], P.tag[p], prm->ni2, prm->prtl_bc[1][0],
              prm->prtl_bc[1][1], true);
  if (D >= 3)
    bc_dir<R>(P.i3[p], P.i3p[p], P.dx3[p], P.ux3[p], P.tag[p], prm->ni3, prm->prtl_bc[2][0],
              prm->prtl_bc[2][1], false);
  return 0;
}

template <typename R>
int sr_push(const ort_sr_params_t* prm, ort_particles_t* prtls, const ort_field_t* em) {
  Fld<R> EB(em);
  Prt<R> P(prtls);
  if (prm->shape_order < 0 || prm->shape_order > 2) return 8;
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err) if (t_par)
  for (int64_t p = 0; p < P.n; ++p) err |= sr_push_one<R>(prm, EB, P, p);
  return err;
}

/* ---- current deposit: src/kernels/currents_deposit.hpp:108-761 --------- */
template <typename R>
struct JAcc {
  Fld<R>* J;
  int     err = 0;
  inline void add(int i, int j, int k, int c, R v) {
    if (i < 0 || i >= J->N1 || j < 0 || j >= J->N2 || k < 0 || k >= J->N3) {
      err |= 2;
      return;
    }
    J->at(i, j, k, c) += v;
  }
};

template <typename R>
void deposit_zigzag(JAcc<R>& J, int D, int G, const Prt<R>& P, int64_t p, const R* vp,
                    R coeff, R inv_dt) {
  using C = K<R>;
  // dimension-1 quantities (currents_deposit.hpp:171-195)
  const int i1 = P.i1[p], i1p = P.i1p[p];
  const float dxp_r_1 = static_cast<float>(i1 == i1p) * (P.dx1[p] + P.dx1p[p]) *
                        static_cast<float>(C::INV_2);
  const R Wx1_1 = C::INV_2 * (dxp_r_1 + P.dx1p[p] + static_cast<R>(i1 > i1p));
  const R Wx1_2 = C::INV_2 * (P.dx1[p] + dxp_r_1 +
                              static_cast<R>(static_cast<int>(i1 > i1p) + i1p - i1));
  const R Fx1_1 = (static_cast<R>(i1 > i1p) + dxp_r_1 - P.dx1p[p]) * coeff * inv_dt;
  const R Fx1_2 = (static_cast<R>(i1 - i1p - static_cast<int>(i1 > i1p)) + P.dx1[p] -
                   dxp_r_1) * coeff * inv_dt;
  const int i2 = P.i2[p], i2p = P.i2p[p];
  const float dxp_r_2 = static_cast<float>(i2 == i2p) * (P.dx2[p] + P.dx2p[p]) *
                        static_cast<float>(C::INV_2);
  const R Wx2_1 = C::INV_2 * (dxp_r_2 + P.dx2p[p] + static_cast<R>(i2 > i2p));
  const R Wx2_2 = C::INV_2 * (P.dx2[p] + dxp_r_2 +
                              static_cast<R>(static_cast<int>(i2 > i2p) + i2p - i2));
  const R Fx2_1 = (static_cast<R>(i2 > i2p) + dxp_r_2 - P.dx2p[p]) * coeff * inv_dt;
  const R Fx2_2 = (static_cast<R>(i2 - i2p - static_cast<int>(i2 > i2p)) + P.dx2[p] -
                   dxp_r_2) * coeff * inv_dt;
  const int a = i1p + G, b = i2p + G, A = i1 + G, B = i2 + G;
  if (D == 2) {
    // currents_deposit.hpp:248-305
    const R Fx3_1 = C::HALF * vp[2] * coeff;
    const R Fx3_2 = C::HALF * vp[2] * coeff;
    J.add(a, b, 0, 0, Fx1_1 * (C::ONE - Wx2_1));
    J.add(a, b + 1, 0, 0, Fx1_1 * Wx2_1);
    J.add(A, B, 0, 0, Fx1_2 * (C::ONE - Wx2_2));
    J.add(A, B + 1, 0, 0, Fx1_2 * Wx2_2);
    J.add(a, b, 0, 1, Fx2_1 * (C::ONE - Wx1_1));
    J.add(a + 1, b, 0, 1, Fx2_1 * Wx1_1);
    J.add(A, B, 0, 1, Fx2_2 * (C::ONE - Wx1_2));
    J.add(A + 1, B, 0, 1, Fx2_2 * Wx1_2);
    J.add(a, b, 0, 2, Fx3_1 * (C::ONE - Wx1_1) * (C::ONE - Wx2_1));
    J.add(a + 1, b, 0, 2, Fx3_1 * Wx1_1 * (C::ONE - Wx2_1));
    J.add(a, b + 1, 0, 2, Fx3_1 * (C::ONE - Wx1_1) * Wx2_1);
    J.add(a + 1, b + 1, 0, 2, Fx3_1 * Wx1_1 * Wx2_1);
    J.add(A, B, 0, 2, Fx3_2 * (C::ONE - Wx1_2) * (C::ONE - Wx2_2));
    J.add(A + 1, B, 0, 2, Fx3_2 * Wx1_2 * (C::ONE - Wx2_2));
    J.add(A, B + 1, 0, 2, Fx3_2 * (C::ONE - Wx1_2) * Wx2_2);
    J.add(A + 1, B + 1, 0, 2, Fx3_2 * Wx1_2 * Wx2_2);
  } else {
    // currents_deposit.hpp:306-402
    const int i3 = P.i3[p], i3p = P.i3p[p];
    const float dxp_r_3 = static_cast<float>(i3 == i3p) * (P.dx3[p] + P.dx3p[p]) *
                          static_cast<float>(C::INV_2);
    const R Wx3_1 = C::INV_2 * (dxp_r_3 + P.dx3p[p] + static_cast<R>(i3 > i3p));
    const R Wx3_2 = C::INV_2 * (P.dx3[p] + dxp_r_3 +
                                static_cast<R>(static_cast<int>(i3 > i3p) + i3p - i3));
    const R Fx3_1 = (static_cast<R>(i3 > i3p) + dxp_r_3 - P.dx3p[p]) * coeff * inv_dt;
    const R Fx3_2 = (static_cast<R>(i3 - i3p - static_cast<int>(i3 > i3p)) + P.dx3[p] -
                     dxp_r_3) * coeff * inv_dt;
    const int c = i3p + G, Cc = i3 + G;
    J.add(a, b, c, 0, Fx1_1 * (C::ONE - Wx2_1) * (C::ONE - Wx3_1));
    J.add(a, b + 1, c, 0, Fx1_1 * Wx2_1 * (C::ONE - Wx3_1));
    J.add(a, b, c + 1, 0, Fx1_1 * (C::ONE - Wx2_1) * Wx3_1);
    J.add(a, b + 1, c + 1, 0, Fx1_1 * Wx2_1 * Wx3_1);
    J.add(A, B, Cc, 0, Fx1_2 * (C::ONE - Wx2_2) * (C::ONE - Wx3_2));
    J.add(A, B + 1, Cc, 0, Fx1_2 * Wx2_2 * (C::ONE - Wx3_2));
    J.add(A, B, Cc + 1, 0, Fx1_2 * (C::ONE - Wx2_2) * Wx3_2);
    J.add(A, B + 1, Cc + 1, 0, Fx1_2 * Wx2_2 * Wx3_2);

    J.add(a, b, c, 1, Fx2_1 * (C::ONE - Wx1_1) * (C::ONE - Wx3_1));
    J.add(a + 1, b, c, 1, Fx2_1 * Wx1_1 * (C::ONE - Wx3_1));
    J.add(a, b, c + 1, 1, Fx2_1 * (C::ONE - Wx1_1) * Wx3_1);
    J.add(a + 1, b, c + 1, 1, Fx2_1 * Wx1_1 * Wx3_1);
    J.add(A, B, Cc, 1, Fx2_2 * (C::ONE - Wx1_2) * (C::ONE - Wx3_2));
    J.add(A + 1, B, Cc, 1, Fx2_2 * Wx1_2 * (C::ONE - Wx3_2));
    J.add(A, B, Cc + 1, 1, Fx2_2 * (C::ONE - Wx1_2) * Wx3_2);
    J.add(A + 1, B, Cc + 1, 1, Fx2_2 * Wx1_2 * Wx3_2);

    J.add(a, b, c, 2, Fx3_1 * (C::ONE - Wx1_1) * (C::ONE - Wx2_1));
    J.add(a + 1, b, c, 2, Fx3_1 * Wx1_1 * (C::ONE - Wx2_1));
    J.add(a, b + 1, c, 2, Fx3_1 * (C::ONE - Wx1_1) * Wx2_1);
    J.add(a + 1, b + 1, c, 2, Fx3_1 * Wx1_1 * Wx2_1);
    J.add(A, B, Cc, 2, Fx3_2 * (C::ONE - Wx1_2) * (C::ONE - Wx2_2));
    J.add(A + 1, B, Cc, 2, Fx3_2 * Wx1_2 * (C::ONE - Wx2_2));
    J.add(A, B + 1, Cc, 2, Fx3_2 * (C::ONE - Wx1_2) * Wx2_2);
    J.add(A + 1, B + 1, Cc, 2, Fx3_2 * Wx1_2 * Wx2_2);
  }
}

template <typename R, int O>
void deposit_esirkepov_2d(JAcc<R>& J, int G, const Prt<R>& P, int64_t p, const R* vp,
                          R coeff, R inv_dt) {
  using C = K<R>;
  // currents_deposit.hpp:406-555
  R iS_x1[O + 2], fS_x1[O + 2], iS_x2[O + 2], fS_x2[O + 2];
  int i1_min, i1_max, i2_min, i2_max;
  for_deposit<R, O>(P.i1p[p], static_cast<R>(P.dx1p[p]), P.i1[p], static_cast<R>(P.dx1[p]),
                    i1_min, i1_max, iS_x1, fS_x1);
  for_deposit<R, O>(P.i2p[p], static_cast<R>(P.dx2p[p]), P.i2[p], static_cast<R>(P.dx2[p]),
                    i2_min, i2_max, iS_x2, fS_x2);
  R Wx1[O + 2][O + 2], Wx2[O + 2][O + 2], Wx3[O + 2][O + 2];
  for (int i = 0; i < O + 2; ++i) {
    for (int j = 0; j < O + 2; ++j) {
      Wx1[i][j] = C::HALF * (fS_x1[i] - iS_x1[i]) * (fS_x2[j] + iS_x2[j]);
      Wx2[i][j] = C::HALF * (fS_x1[i] + iS_x1[i]) * (fS_x2[j] - iS_x2[j]);
      Wx3[i][j] = C::THIRD * (fS_x2[j] * (C::HALF * iS_x1[i] + fS_x1[i]) +
                              iS_x2[j] * (C::HALF * fS_x1[i] + iS_x1[i]));
    }
  }
  R jx1[O + 2][O + 2], jx2[O + 2][O + 2];
  const R Qdx1dt = coeff * inv_dt;
  const R Qdx2dt = coeff * inv_dt;
  const R QVx3   = coeff * vp[2];
  for (int j = 0; j < O + 2; ++j) jx1[0][j] = -Qdx1dt * Wx1[0][j];
  for (int i = 1; i < O + 2; ++i)
    for (int j = 0; j < O + 2; ++j) jx1[i][j] = jx1[i - 1][j] - Qdx1dt * Wx1[i][j];
  for (int i = 0; i < O + 2; ++i) jx2[i][0] = -Qdx2dt * Wx2[i][0];
  for (int j = 1; j < O + 2; ++j)
    for (int i = 0; i < O + 2; ++i) jx2[i][j] = jx2[i][j - 1] - Qdx2dt * Wx2[i][j];
  i1_min += G; i2_min += G; i1_max += G; i2_max += G;
  const int di_x1 = i1_max - i1_min, di_x2 = i2_max - i2_min;
  for (int i = 0; i < di_x1; ++i)
    for (int j = 0; j <= di_x2; ++j) J.add(i1_min + i, i2_min + j, 0, 0, jx1[i][j]);
  for (int i = 0; i <= di_x1; ++i)
    for (int j = 0; j < di_x2; ++j) J.add(i1_min + i, i2_min + j, 0, 1, jx2[i][j]);
  for (int i = 0; i <= di_x1; ++i)
    for (int j = 0; j <= di_x2; ++j) J.add(i1_min + i, i2_min + j, 0, 2, QVx3 * Wx3[i][j]);
}

template <typename R, int O>
void deposit_esirkepov_3d(JAcc<R>& J, int G, const Prt<R>& P, int64_t p, R coeff, R inv_dt) {
  using C = K<R>;
  // currents_deposit.hpp:557-754
  R iS_x1[O + 2], fS_x1[O + 2], iS_x2[O + 2], fS_x2[O + 2], iS_x3[O + 2], fS_x3[O + 2];
  int i1_min, i1_max, i2_min, i2_max, i3_min, i3_max;
  for_deposit<R, O>(P.i1p[p], static_cast<R>(P.dx1p[p]), P.i1[p], static_cast<R>(P.dx1[p]),
                    i1_min, i1_max, iS_x1, fS_x1);
  for_deposit<R, O>(P.i2p[p], static_cast<R>(P.dx2p[p]), P.i2[p], static_cast<R>(P.dx2[p]),
                    i2_min, i2_max, iS_x2, fS_x2);
  for_deposit<R, O>(P.i3p[p], static_cast<R>(P.dx3p[p]), P.i3[p], static_cast<R>(P.dx3[p]),
                    i3_min, i3_max, iS_x3, fS_x3);
  constexpr int M = O + 2;
  R Wx1[M][M][M], Wx2[M][M][M], Wx3[M][M][M];
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k) {
        Wx1[i][j][k] = C::THIRD * (fS_x1[i] - iS_x1[i]) *
                       ((iS_x2[j] * iS_x3[k] + fS_x2[j] * fS_x3[k]) +
                        C::HALF * (iS_x3[k] * fS_x2[j] + iS_x2[j] * fS_x3[k]));
        Wx2[i][j][k] = C::THIRD * (fS_x2[j] - iS_x2[j]) *
                       (iS_x1[i] * iS_x3[k] + fS_x1[i] * fS_x3[k] +
                        C::HALF * (iS_x3[k] * fS_x1[i] + iS_x1[i] * fS_x3[k]));
        Wx3[i][j][k] = C::THIRD * (fS_x3[k] - iS_x3[k]) *
                       (iS_x1[i] * iS_x2[j] + fS_x1[i] * fS_x2[j] +
                        C::HALF * (iS_x1[i] * fS_x2[j] + iS_x2[j] * fS_x1[i]));
      }
  R jx1[M][M][M], jx2[M][M][M], jx3[M][M][M];
  const R Qdxdt = coeff * inv_dt, Qdydt = coeff * inv_dt, Qdzdt = coeff * inv_dt;
  for (int j = 0; j < M; ++j)
    for (int k = 0; k < M; ++k) jx1[0][j][k] = -Qdxdt * Wx1[0][j][k];
  for (int i = 1; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k) jx1[i][j][k] = jx1[i - 1][j][k] - Qdxdt * Wx1[i][j][k];
  for (int i = 0; i < M; ++i)
    for (int k = 0; k < M; ++k) jx2[i][0][k] = -Qdydt * Wx2[i][0][k];
  for (int i = 0; i < M; ++i)
    for (int j = 1; j < M; ++j)
      for (int k = 0; k < M; ++k) jx2[i][j][k] = jx2[i][j - 1][k] - Qdydt * Wx2[i][j][k];
  // the reference uses Qdydt for the first jx3 slice (currents_deposit.hpp:700)
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j) jx3[i][j][0] = -Qdydt * Wx3[i][j][0];
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 1; k < M; ++k) jx3[i][j][k] = jx3[i][j][k - 1] - Qdzdt * Wx3[i][j][k];
  i1_min += G; i2_min += G; i3_min += G; i1_max += G; i2_max += G; i3_max += G;
  const int d1 = i1_max - i1_min, d2 = i2_max - i2_min, d3 = i3_max - i3_min;
  for (int i = 0; i < d1; ++i)
    for (int j = 0; j <= d2; ++j)
      for (int k = 0; k <= d3; ++k) J.add(i1_min + i, i2_min + j, i3_min + k, 0, jx1[i][j][k]);
  for (int i = 0; i <= d1; ++i)
    for (int j = 0; j < d2; ++j)
      for (int k = 0; k <= d3; ++k) J.add(i1_min + i, i2_min + j, i3_min + k, 1, jx2[i][j][k]);
  for (int i = 0; i <= d1; ++i)
    for (int j = 0; j <= d2; ++j)
      for (int k = 0; k < d3; ++k) J.add(i1_min + i, i2_min + j, i3_min + k, 2, jx3[i][j][k]);
}

// One particle's deposit into the accumulator (the loop body of deposit()).
template <typename R>
void deposit_one(JAcc<R>& acc, const Prt<R>& P, int64_t p, int D, int G, R charge, R inv_dt,
                 R dxm, int order) {
  using C = K<R>;
  R vp[3];
  const R u[3] = { P.ux1[p], P.ux2[p], P.ux3[p] };
  for (int d = 0; d < 3; ++d) vp[d] = (d < D) ? u[d] / dxm : u[d];
  const R inv_energy = C::ONE / std::sqrt(C::ONE + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  if (std::isnan(vp[2]) || std::isinf(vp[2])) vp[2] = C::ZERO;
  vp[0] *= inv_energy; vp[1] *= inv_energy; vp[2] *= inv_energy;
  const R coeff = P.w[p] * charge;
  if (order == 0) deposit_zigzag<R>(acc, D, G, P, p, vp, coeff, inv_dt);
  else if (D == 2 && order == 1) deposit_esirkepov_2d<R, 1>(acc, G, P, p, vp, coeff, inv_dt);
  else if (D == 2 && order == 2) deposit_esirkepov_2d<R, 2>(acc, G, P, p, vp, coeff, inv_dt);
  else if (D == 3 && order == 1) deposit_esirkepov_3d<R, 1>(acc, G, P, p, coeff, inv_dt);
  else deposit_esirkepov_3d<R, 2>(acc, G, P, p, coeff, inv_dt);
}

// The deposit over OpenMP threads: each thread adds its share of the
// particles into a private copy of J, and the copies are summed into J.
template <typename R>
int deposit_parallel(const Prt<R>& P, Fld<R>& J, R charge, R inv_dt, R dxm, int order) {
  const int     D = J.dim, G = J.ngh;
  const int64_t size = J.size();
  int           err  = 0;
  std::vector<std::vector<R>> copies;
#pragma omp parallel reduction(| : err)
  {
#pragma omp single
    copies.resize(static_cast<size_t>(omp_get_num_threads()));
    const int t = omp_get_thread_num();
    copies[t].assign(static_cast<size_t>(size), R(0));
    Fld<R> Jt = J;
    Jt.d = copies[t].data();
    JAcc<R> acc{ &Jt };
#pragma omp for schedule(static)
    for (int64_t p = 0; p < P.n; ++p)
      if (P.tag[p] != TAG_DEAD) deposit_one<R>(acc, P, p, D, G, charge, inv_dt, dxm, order);
    err |= acc.err;
#pragma omp for schedule(static)
    for (int64_t k = 0; k < size; ++k) {
      R v = J.d[k];
      for (const auto& c : copies) v += c[k];
      J.d[k] = v;
    }
  }
  return err;
}

template <typename R>
int deposit(const ort_particles_t* prtls, ort_field_t* cur, double charge_d, double dt_d,
            double dx_d, int order) {
  using C = K<R>;
  Fld<R> J(cur);
  Prt<R> P(prtls);
  JAcc<R> acc{ &J };
  if (order < 0 || order > 2) return 8;
  const R charge = static_cast<R>(charge_d);
  const R inv_dt = C::ONE / static_cast<R>(dt_d);
  const R dxm    = static_cast<R>(dx_d);
  const int D = J.dim, G = J.ngh;
  if (t_par) return deposit_parallel<R>(P, J, charge, inv_dt, dxm, order);
  for (int64_t p = 0; p < P.n; ++p) {
    if (P.tag[p] == TAG_DEAD) continue;
    // velocity for the out-of-plane component (currents_deposit.hpp:112-166)
    R vp[3];
    const R u[3] = { P.ux1[p], P.ux2[p], P.ux3[p] };
    for (int d = 0; d < 3; ++d) vp[d] = (d < D) ? u[d] / dxm : u[d];
    const R inv_energy = C::ONE / std::sqrt(C::ONE + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    if (std::isnan(vp[2]) || std::isinf(vp[2])) vp[2] = C::ZERO;
    vp[0] *= inv_energy; vp[1] *= inv_energy; vp[2] *= inv_energy;
    const R coeff = P.w[p] * charge;
    if (order == 0) deposit_zigzag<R>(acc, D, G, P, p, vp, coeff, inv_dt);
    else if (D == 2 && order == 1) deposit_esirkepov_2d<R, 1>(acc, G, P, p, vp, coeff, inv_dt);
    else if (D == 2 && order == 2) deposit_esirkepov_2d<R, 2>(acc, G, P, p, vp, coeff, inv_dt);
    else if (D == 3 && order == 1) deposit_esirkepov_3d<R, 1>(acc, G, P, p, coeff, inv_dt);
    else deposit_esirkepov_3d<R, 2>(acc, G, P, p, coeff, inv_dt);
  }
  return acc.err;
}

// The 2D deposit with the out-of-plane velocity given per particle: the
// GRPIC instantiation of DepositCurrents_kernel (currents_deposit.hpp:
// 131-166 compute it from the metric; the rest is the same scheme).
template <typename R>
int deposit_vz(const ort_particles_t* prtls, ort_field_t* cur, double charge_d, double dt_d,
               int order, const double* vz) {
  using C = K<R>;
  Fld<R> J(cur);
  Prt<R> P(prtls);
  JAcc<R> acc{ &J };
  if (order < 0 || order > 2 || J.dim != 2) return 8;
  const R charge = static_cast<R>(charge_d);
  const R inv_dt = C::ONE / static_cast<R>(dt_d);
  const int G = J.ngh;
  for (int64_t p = 0; p < P.n; ++p) {
    if (P.tag[p] == TAG_DEAD) continue;
    const R vp[3] = { C::ZERO, C::ZERO, static_cast<R>(vz[p]) };
    const R coeff = P.w[p] * charge;
    if (order == 0) deposit_zigzag<R>(acc, 2, G, P, p, vp, coeff, inv_dt);
    else if (order == 1) deposit_esirkepov_2d<R, 1>(acc, G, P, p, vp, coeff, inv_dt);
    else deposit_esirkepov_2d<R, 2>(acc, G, P, p, vp, coeff, inv_dt);
  }
  return acc.err;
}

/* ---- Faraday (Minkowski): src/kernels/faraday_mink.hpp:71-166 ---------- */
template <typename R>
int faraday(ort_field_t* em, double c1d, double c2d, const ort_fd_coeffs_t* fd) {
  using C = K<R>;
  Fld<R> EB(em);
  const R coeff1 = c1d, coeff2 = c2d;
  const R deltax = fd ? fd->delta_x : 0, deltay = fd ? fd->delta_y : 0, deltaz = fd ? fd->delta_z : 0;
  const R betaxy = fd ? fd->beta_xy : 0, betayx = fd ? fd->beta_yx : 0;
  const R betaxz = fd ? fd->beta_xz : 0, betazx = fd ? fd->beta_zx : 0;
  const R betayz = fd ? fd->beta_yz : 0, betazy = fd ? fd->beta_zy : 0;
  const int G = EB.ngh;
  if (EB.dim == 2) {
    const R alphax = C::ONE - C::TWO * betaxy - C::THREE * deltax;
    const R alphay = C::ONE - C::TWO * betayx - C::THREE * deltay;
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t i2 = G; i2 < G + EB.n[1]; ++i2)
      for (int64_t i1 = G; i1 < G + EB.n[0]; ++i1) {
        auto E = [&](int64_t a, int64_t b, int c) -> R { return EB(a, b, c); };
        EB(i1, i2, 3) += coeff1 * (-alphay * (E(i1, i2 + 1, 2) - E(i1, i2, 2)) -
                                   deltay * (E(i1, i2 + 2, 2) - E(i1, i2 - 1, 2)) -
                                   betayx * (E(i1 + 1, i2 + 1, 2) - E(i1 + 1, i2, 2)) -
                                   betayx * (E(i1 - 1, i2 + 1, 2) - E(i1 - 1, i2, 2)));
        EB(i1, i2, 4) += coeff1 * (+alphax * (E(i1 + 1, i2, 2) - E(i1, i2, 2)) +
                                   deltax * (E(i1 + 2, i2, 2) - E(i1 - 1, i2, 2)) +
                                   betaxy * (E(i1 + 1, i2 + 1, 2) - E(i1, i2 + 1, 2)) +
                                   betaxy * (E(i1 + 1, i2 - 1, 2) - E(i1, i2 - 1, 2)));
        EB(i1, i2, 5) += coeff2 * (+alphay * (E(i1, i2 + 1, 0) - E(i1, i2, 0)) +
                                   deltay * (E(i1, i2 + 2, 0) - E(i1, i2 - 1, 0)) +
                                   betayx * (E(i1 + 1, i2 + 1, 0) - E(i1 + 1, i2, 0)) +
                                   betayx * (E(i1 - 1, i2 + 1, 0) - E(i1 - 1, i2, 0)) -
                                   alphax * (E(i1 + 1, i2, 1) - E(i1, i2, 1)) -
                                   deltax * (E(i1 + 2, i2, 1) - E(i1 - 1, i2, 1)) -
                                   betaxy * (E(i1 + 1, i2 + 1, 1) - E(i1, i2 + 1, 1)) -
                                   betaxy * (E(i1 + 1, i2 - 1, 1) - E(i1, i2 - 1, 1)));
      }
  } else if (EB.dim == 3) {
    const R alphax = C::ONE - C::TWO * betaxy - C::TWO * betaxz - C::THREE * deltax;
    const R alphay = C::ONE - C::TWO * betayx - C::TWO * betayz - C::THREE * deltay;
    const R alphaz = C::ONE - C::TWO * betazx - C::TWO * betazy - C::THREE * deltaz;
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t i3 = G; i3 < G + EB.n[2]; ++i3)
      for (int64_t i2 = G; i2 < G + EB.n[1]; ++i2)
        for (int64_t i1 = G; i1 < G + EB.n[0]; ++i1) {
          auto E = [&](int64_t a, int64_t b, int64_t c, int comp) -> R { return EB.at(a, b, c, comp); };
          EB.at(i1, i2, i3, 3) += coeff1 * (
            +alphaz * (E(i1, i2, i3 + 1, 1) - E(i1, i2, i3, 1))
            + deltaz * (E(i1, i2, i3 + 2, 1) - E(i1, i2, i3 - 1, 1))
            + betazx * (E(i1 + 1, i2, i3 + 1, 1) - E(i1 + 1, i2, i3, 1))
            + betazx * (E(i1 - 1, i2, i3 + 1, 1) - E(i1 - 1, i2, i3, 1))
            + betazy * (E(i1, i2 + 1, i3 + 1, 1) - E(i1, i2 + 1, i3, 1))
            + betazy * (E(i1, i2 - 1, i3 + 1, 1) - E(i1, i2 - 1, i3, 1))
            - alphay * (E(i1, i2 + 1, i3, 2) - E(i1, i2, i3, 2))
            - deltay * (E(i1, i2 + 2, i3, 2) - E(i1, i2 - 1, i3, 2))
            - betayx * (E(i1 + 1, i2 + 1, i3, 2) - E(i1 + 1, i2, i3, 2))
            - betayx * (E(i1 - 1, i2 + 1, i3, 2) - E(i1 - 1, i2, i3, 2))
            - betayz * (E(i1, i2 + 1, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
            - betayz * (E(i1, i2 + 1, i3 - 1, 2) - E(i1, i2, i3 - 1, 2)));
          EB.at(i1, i2, i3, 4) += coeff1 * (
            +alphax * (E(i1 + 1, i2, i3, 2) - E(i1, i2, i3, 2))
            + deltax * (E(i1 + 2, i2, i3, 2) - E(i1 - 1, i2, i3, 2))
            + betaxy * (E(i1 + 1, i2 + 1, i3, 2) - E(i1, i2 + 1, i3, 2))
            + betaxy * (E(i1 + 1, i2 - 1, i3, 2) - E(i1, i2 - 1, i3, 2))
            + betaxz * (E(i1 + 1, i2, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
            + betaxz * (E(i1 + 1, i2, i3 - 1, 2) - E(i1, i2, i3 - 1, 2))
            - alphaz * (E(i1, i2, i3 + 1, 0) - E(i1, i2, i3, 0))
            - deltaz * (E(i1, i2, i3 + 2, 0) - E(i1, i2, i3 - 1, 0))
            - betazx * (E(i1 + 1, i2, i3 + 1, 0) - E(i1 + 1, i2, i3, 0))
            - betazx * (E(i1 - 1, i2, i3 + 1, 0) - E(i1 - 1, i2, i3, 0))
            - betazy * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2 + 1, i3, 0))
            - betazy * (E(i1, i2 - 1, i3 + 1, 0) - E(i1, i2 - 1, i3, 0)));
          EB.at(i1, i2, i3, 5) += coeff1 * (
            +alphay * (E(i1, i2 + 1, i3, 0) - E(i1, i2, i3, 0))
            + deltay * (E(i1, i2 + 2, i3, 0) - E(i1, i2 - 1, i3, 0))
            + betayx * (E(i1 + 1, i2 + 1, i3, 0) - E(i1 + 1, i2, i3, 0))
            + betayx * (E(i1 - 1, i2 + 1, i3, 0) - E(i1 - 1, i2, i3, 0))
            + betayz * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2, i3 + 1, 0))
            + betayz * (E(i1, i2 + 1, i3 - 1, 0) - E(i1, i2, i3 - 1, 0))
            - alphax * (E(i1 + 1, i2, i3, 1) - E(i1, i2, i3, 1))
            - deltax * (E(i1 + 2, i2, i3, 1) - E(i1 - 1, i2, i3, 1))
            - betaxy * (E(i1 + 1, i2 + 1, i3, 1) - E(i1, i2 + 1, i3, 1))
            - betaxy * (E(i1 + 1, i2 - 1, i3, 1) - E(i1, i2 - 1, i3, 1))
            - betaxz * (E(i1 + 1, i2, i3 + 1, 1) - E(i1, i2, i3 + 1, 1))
            - betaxz * (E(i1 + 1, i2, i3 - 1, 1) - E(i1, i2, i3 - 1, 1)));
        }
  } else {
    return 8;
  }
  return 0;
}

/* ---- Ampere (Minkowski): src/kernels/ampere_mink.hpp:57-86 ------------- */
template <typename R>
int ampere(ort_field_t* em, double c1d, double c2d) {
  Fld<R> EB(em);
  const R coeff1 = c1d, coeff2 = c2d;
  const int G = EB.ngh;
  if (EB.dim == 2) {
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t i2 = G; i2 < G + EB.n[1]; ++i2)
      for (int64_t i1 = G; i1 < G + EB.n[0]; ++i1) {
        EB(i1, i2, 0) += coeff1 * (EB(i1, i2, 5) - EB(i1, i2 - 1, 5));
        EB(i1, i2, 1) += coeff1 * (EB(i1 - 1, i2, 5) - EB(i1, i2, 5));
        EB(i1, i2, 2) += coeff2 * (EB(i1, i2 - 1, 3) - EB(i1, i2, 3) + EB(i1, i2, 4) -
                                   EB(i1 - 1, i2, 4));
      }
  } else if (EB.dim == 3) {
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t i3 = G; i3 < G + EB.n[2]; ++i3)
      for (int64_t i2 = G; i2 < G + EB.n[1]; ++i2)
        for (int64_t i1 = G; i1 < G + EB.n[0]; ++i1) {
          EB.at(i1, i2, i3, 0) += coeff1 * (EB.at(i1, i2, i3 - 1, 4) - EB.at(i1, i2, i3, 4) +
                                            EB.at(i1, i2, i3, 5) - EB.at(i1, i2 - 1, i3, 5));
          EB.at(i1, i2, i3, 1) += coeff1 * (EB.at(i1 - 1, i2, i3, 5) - EB.at(i1, i2, i3, 5) +
                                            EB.at(i1, i2, i3, 3) - EB.at(i1, i2, i3 - 1, 3));
          EB.at(i1, i2, i3, 2) += coeff1 * (EB.at(i1, i2 - 1, i3, 3) - EB.at(i1, i2, i3, 3) +
                                            EB.at(i1, i2, i3, 4) - EB.at(i1 - 1, i2, i3, 4));
        }
  } else {
    return 8;
  }
  return 0;
}

/* ---- E += coeff*J; J /= ppc0: src/kernels/ampere_mink.hpp:156-210 ------ */
template <typename R>
int currents_ampere(ort_field_t* em, ort_field_t* cur, double coeff_d, double ppc0_d) {
  Fld<R> E(em), J(cur);
  const R coeff = coeff_d, ppc0 = ppc0_d;
  const int G = E.ngh;
  const int64_t n3 = E.dim >= 3 ? E.n[2] : 1;
  const int64_t k0 = E.dim >= 3 ? G : 0;
#pragma omp parallel for schedule(static) if (t_par)
  for (int64_t i3 = k0; i3 < k0 + n3; ++i3)
    for (int64_t i2 = G; i2 < G + E.n[1]; ++i2)
      for (int64_t i1 = G; i1 < G + E.n[0]; ++i1)
        for (int c = 0; c < 3; ++c) E.at(i1, i2, i3, c) += J.at(i1, i2, i3, c) * coeff;
#pragma omp parallel for schedule(static) if (t_par)
  for (int64_t i3 = k0; i3 < k0 + n3; ++i3)
    for (int64_t i2 = G; i2 < G + E.n[1]; ++i2)
      for (int64_t i1 = G; i1 < G + E.n[0]; ++i1)
        for (int c = 0; c < 3; ++c) J.at(i1, i2, i3, c) /= ppc0;
  return 0;
}

/* ---- digital filter, periodic interior: digital_filter.hpp:182-194 (2D)
 *      and :349-378 (3D, with the reference's 1/32 group as written) ----- */
template <typename R>
int filter(ort_field_t* arr, const ort_field_t* buf) {
  using C = K<R>;
  Fld<R> A(arr), B(const_cast<ort_field_t*>(buf));
  const int G = A.ngh;
  if (A.dim == 2) {
    for (int c = 0; c < 3; ++c)
#pragma omp parallel for schedule(static) if (t_par)
      for (int64_t i2 = G; i2 < G + A.n[1]; ++i2)
        for (int64_t i1 = G; i1 < G + A.n[0]; ++i1)
          A(i1, i2, c) = C::INV_4 * B(i1, i2, c) +
                         C::INV_8 * (B(i1 - 1, i2, c) + B(i1 + 1, i2, c) + B(i1, i2 - 1, c) +
                                     B(i1, i2 + 1, c)) +
                         C::INV_16 * (B(i1 - 1, i2 - 1, c) + B(i1 + 1, i2 + 1, c) +
                                      B(i1 - 1, i2 + 1, c) + B(i1 + 1, i2 - 1, c));
  } else if (A.dim == 3) {
    for (int c = 0; c < 3; ++c)
#pragma omp parallel for schedule(static) if (t_par)
      for (int64_t k = G; k < G + A.n[2]; ++k)
        for (int64_t j = G; j < G + A.n[1]; ++j)
          for (int64_t i = G; i < G + A.n[0]; ++i) {
            auto b = [&](int64_t x, int64_t y, int64_t z) -> R { return B.at(x, y, z, c); };
            A.at(i, j, k, c) =
              C::INV_8 * b(i, j, k) +
              C::INV_16 * (b(i - 1, j, k) + b(i + 1, j, k) + b(i, j - 1, k) + b(i, j + 1, k) +
                           b(i, j, k - 1) + b(i, j, k + 1)) +
              C::INV_32 * (b(i - 1, j - 1, k) + b(i + 1, j + 1, k) + b(i - 1, j + 1, k) +
                           b(i + 1, j - 1, k) + b(i, j - 1, k - 1) + b(i, j + 1, k + 1) +
                           b(i, j, k - 1) + b(i, j, k + 1) + b(i - 1, j, k - 1) +
                           b(i + 1, j, k + 1) + b(i - 1, j, k + 1) + b(i + 1, j, k - 1)) +
              C::INV_64 * (b(i - 1, j - 1, k - 1) + b(i + 1, j + 1, k + 1) +
                           b(i - 1, j + 1, k + 1) + b(i + 1, j - 1, k - 1) +
                           b(i - 1, j - 1, k + 1) + b(i + 1, j + 1, k - 1) +
                           b(i - 1, j + 1, k - 1) + b(i + 1, j - 1, k + 1));
          }
  } else {
    return 8;
  }
  return 0;
}

/* ---- periodic ghost fill: metadomain_comm.cpp:204-367 + comm_nompi.hpp:47-63.
 * For every direction, the ghost slab receives the opposite active slab; the
 * result is the periodic extension of the active region. */
template <typename R>
int communicate(ort_field_t* fld, int c0, int c1) {
  Fld<R> F(fld);
  const int G = F.ngh, D = F.dim;
  auto wrap = [&](int64_t i, int d) -> int64_t {
    const int64_t n = F.n[d];
    if (i < G) return i + n;
    if (i >= G + n) return i - n;
    return i;
  };
  const int64_t N3 = D >= 3 ? F.N3 : 1;
  for (int c = c0; c < c1; ++c)
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t k = 0; k < N3; ++k)
      for (int64_t j = 0; j < F.N2; ++j)
        for (int64_t i = 0; i < F.N1; ++i) {
          const int64_t si = wrap(i, 0), sj = wrap(j, 1), sk = D >= 3 ? wrap(k, 2) : k;
          if (si != i || sj != j || sk != k) F.at(i, j, k, c) = F.at(si, sj, sk, c);
        }
  return 0;
}

/* ---- periodic J fold: metadomain_comm.cpp:408-562 (comm_nompi.hpp:64-115,
 * AddBufferedFields :369-406).  buff = 0; for each direction d in
 * Directions<D>::all order, buff(recv slab) += cur(send slab); then
 * cur(active) += buff(active). */
template <typename R>
int synchronize(ort_field_t* cur, ort_field_t* buff) {
  Fld<R> J(cur), B(buff);
  const int G = J.ngh, D = J.dim;
  std::memset(B.d, 0, sizeof(R) * B.size());
  // direction lists: src/global/arch/directions.h:180-233 (lexicographic, -1..1)
  std::vector<std::array<int, 3>> dirs;
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int c = -1; c <= 1; ++c) {
        if (D == 2 && c != 0) continue;
        if (a == 0 && b == 0 && c == 0) continue;
        dirs.push_back({ a, b, c });
      }
  for (const auto& dir : dirs) {
    int64_t lo[3], hi[3], off[3];
    for (int d = 0; d < 3; ++d) {
      if (d >= D) { lo[d] = 0; hi[d] = 1; off[d] = 0; continue; }
      const int64_t n = J.n[d], imin = G, imax = G + n;
      int64_t s0, r0, r1;
      // send slab (synchronize=true): dir 0 -> [imin-G, imax+G), +1 -> [imax-G, imax+G), -1 -> [imin-G, imin+G)
      if (dir[d] == 0) s0 = imin - G;
      else if (dir[d] == 1) s0 = imax - G;
      else s0 = imin - G;
      // recv slab: -dir 0 -> full, -dir == 1 -> [imax-G, imax+G), -dir == -1 -> [imin-G, imin+G)
      if (-dir[d] == 0) { r0 = imin - G; r1 = imax + G; }
      else if (-dir[d] == 1) { r0 = imax - G; r1 = imax + G; }
      else { r0 = imin - G; r1 = imin + G; }
      lo[d] = r0; hi[d] = r1; off[d] = r0 - s0;
    }
    for (int c = 0; c < 3; ++c)
#pragma omp parallel for schedule(static) if (t_par)
      for (int64_t k = lo[2]; k < hi[2]; ++k)
        for (int64_t j = lo[1]; j < hi[1]; ++j)
          for (int64_t i = lo[0]; i < hi[0]; ++i)
            B.at(i, j, k, c) += J.at(i - off[0], j - off[1], k - off[2], c);
  }
  const int64_t k0 = D >= 3 ? G : 0, k1 = D >= 3 ? G + J.n[2] : 1;
  for (int c = 0; c < 3; ++c)
#pragma omp parallel for schedule(static) if (t_par)
    for (int64_t k = k0; k < k1; ++k)
      for (int64_t j = G; j < G + J.n[1]; ++j)
        for (int64_t i = G; i < G + J.n[0]; ++i) J.at(i, j, k, c) += B.at(i, j, k, c);
  return 0;
}

/* ---- one SRPIC step: src/engines/srpic/srpic.hpp:65-188 ---------------- */
template <typename R>
int sr_step(ort_field_t* em, ort_field_t* cur, ort_field_t* buff, ort_species_t* sp,
            int nsp, const ort_step_params_t* prm, const uint8_t prtl_bc[3][2]) {
  using C = K<R>;
  int err = 0;
  const int D = em->dim;
  if (prm->step == 0) err |= communicate<R>(em, 0, 6);
  // srpic::Faraday coefficients (fieldsolvers.h:44-76)
  const R dt = static_cast<R>(prm->dt);
  // dx = sqrt(h_11) as the drivers compute it (fieldsolvers.h:50,110)
  const R dxm = std::sqrt(static_cast<R>(prm->dx) * static_cast<R>(prm->dx));
  auto fd_coeffs = [&](R fraction, R& c1, R& c2) {
    const R dT = fraction * static_cast<R>(prm->correction) * dt;
    if (D == 2) { c1 = dT / (dxm * dxm); c2 = dT; }
    else { c1 = dT / dxm; c2 = C::ZERO; }
  };
  R c1, c2;
  if (prm->fieldsolver_enable) {
    fd_coeffs(C::HALF, c1, c2);
    err |= faraday<R>(em, c1, c2, &prm->fd);
    err |= communicate<R>(em, 3, 6);
  }
  for (int s = 0; s < nsp; ++s) {
    if (sp[s].pusher_flags == 0 || sp[s].prtls.npart == 0) continue;
    ort_sr_params_t pp{};
    pp.mass = sp[s].mass; pp.charge = sp[s].charge;
    pp.dt = prm->dt; pp.omegaB0 = prm->omegaB0; pp.dx = prm->dx;
    pp.ni1 = em->n[0]; pp.ni2 = D >= 2 ? em->n[1] : 0; pp.ni3 = D >= 3 ? em->n[2] : 0;
    pp.pusher_flags = sp[s].pusher_flags;
    std::memcpy(pp.prtl_bc, prtl_bc, 6);
    pp.shape_order = prm->shape_order;
    err |= sr_push<R>(&pp, &sp[s].prtls, em);
  }
  if (prm->deposit_enable) {
    Fld<R> J(cur);
    std::memset(J.d, 0, sizeof(R) * J.size());
    for (int s = 0; s < nsp; ++s) {
      if (sp[s].pusher_flags == 0 || sp[s].prtls.npart == 0 || sp[s].charge == 0.0f) continue;
      err |= deposit<R>(&sp[s].prtls, cur, sp[s].charge, prm->dt, prm->dx, prm->shape_order);
    }
    err |= synchronize<R>(cur, buff);
    err |= communicate<R>(cur, 0, 3);
    // srpic::CurrentsFilter (currents.h:89-119)
    Fld<R> Bf(buff);
    for (int f = 0; f < prm->n_filters; ++f) {
      std::memcpy(Bf.d, J.d, sizeof(R) * J.size());
      err |= filter<R>(cur, buff);
      err |= communicate<R>(cur, 0, 3);
    }
  }
  if (prm->fieldsolver_enable) {
    fd_coeffs(C::HALF, c1, c2);
    err |= faraday<R>(em, c1, c2, &prm->fd);
    err |= communicate<R>(em, 3, 6);
    fd_coeffs(C::ONE, c1, c2);
    err |= ampere<R>(em, c1, c2);
    if (prm->deposit_enable) {
      // srpic::CurrentsAmpere (fieldsolvers.h:150-157): coeff = -dt*q0/(B0*V0)
      const R coeff = -dt * static_cast<R>(prm->q0) /
                      (static_cast<R>(prm->B0) * static_cast<R>(prm->V0));
      err |= currents_ampere<R>(em, cur, coeff, static_cast<R>(prm->ppc0));
    }
    err |= communicate<R>(em, 0, 3);
    if (prm->deposit_enable) err |= communicate<R>(cur, 0, 3);
  }
  return err;
}

}  // namespace

#define ORT_DEFINE(SUF, R)                                                                  \
  extern "C" int ntt_oracle_sr_push_##SUF(const ort_sr_params_t* prm,                       \
                                          ort_particles_t* prtls, const ort_field_t* em) {  \
    return sr_push<R>(prm, prtls, em);                                                     \
  }                                                                                         \
  extern "C" int ntt_oracle_interp_##SUF(const ort_sr_params_t* prm,                        \
                                         const ort_particles_t* prtls,                      \
                                         const ort_field_t* em, R* e_out, R* b_out) {       \
    Fld<R> EB(em);                                                                          \
    Prt<R> P(prtls);                                                                        \
    for (int64_t p = 0; p < P.n; ++p) interpolate<R>(EB, prm->shape_order, P, p,            \
                                                     e_out + 3 * p, b_out + 3 * p);         \
    return 0;                                                                               \
  }                                                                                         \
  extern "C" int ntt_oracle_deposit_##SUF(const ort_particles_t* prtls, ort_field_t* cur,   \
                                          double charge, double dt, double dx,              \
                                          int32_t order) {                                  \
    return deposit<R>(prtls, cur, charge, dt, dx, order);                                   \
  }                                                                                         \
  extern "C" int ntt_oracle_deposit_vz_##SUF(const ort_particles_t* prtls, ort_field_t* cur,\
                                             double charge, double dt, int32_t order,       \
                                             const double* vz) {                            \
    return deposit_vz<R>(prtls, cur, charge, dt, order, vz);                                \
  }                                                                                         \
  extern "C" int ntt_oracle_faraday_##SUF(ort_field_t* em, double c1, double c2,            \
                                          const ort_fd_coeffs_t* fd) {                      \
    return faraday<R>(em, c1, c2, fd);                                                      \
  }                                                                                         \
  extern "C" int ntt_oracle_ampere_##SUF(ort_field_t* em, double c1, double c2) {           \
    return ampere<R>(em, c1, c2);                                                           \
  }                                                                                         \
  extern "C" int ntt_oracle_currents_ampere_##SUF(ort_field_t* em, ort_field_t* cur,        \
                                                  double coeff, double ppc0) {              \
    return currents_ampere<R>(em, cur, coeff, ppc0);                                        \
  }                                                                                         \
  extern "C" int ntt_oracle_filter_##SUF(ort_field_t* a, const ort_field_t* b) {            \
    return filter<R>(a, b);                                                                 \
  }                                                                                         \
  extern "C" int ntt_oracle_communicate_##SUF(ort_field_t* f, int32_t c0, int32_t c1) {     \
    return communicate<R>(f, c0, c1);                                                       \
  }                                                                                         \
  extern "C" int ntt_oracle_synchronize_##SUF(ort_field_t* cur, ort_field_t* buff) {        \
    return synchronize<R>(cur, buff);                                                       \
  }                                                                                         \
  extern "C" int ntt_oracle_sr_step_##SUF(ort_field_t* em, ort_field_t* cur,                \
                                          ort_field_t* buff, ort_species_t* species,        \
                                          int32_t nspecies, const ort_step_params_t* prm,   \
                                          const uint8_t prtl_bc[3][2]) {                    \
    return sr_step<R>(em, cur, buff, species, nspecies, prm, prtl_bc);                      \
  }

ORT_DEFINE(f32, float)
ORT_DEFINE(f64, double)

// The step with OpenMP threads (the CPU baseline; results differ from the
// sequential step in the current's summation order only).
extern "C" int ntt_oracle_sr_step_omp_f32(ort_field_t* em, ort_field_t* cur, ort_field_t* buff,
                                          ort_species_t* species, int32_t nspecies,
                                          const ort_step_params_t* prm,
                                          const uint8_t prtl_bc[3][2]) {
  t_par = true;
  const int r = sr_step<float>(em, cur, buff, species, nspecies, prm, prtl_bc);
  t_par = false;
  return r;
}

// The same in double: the finite-difference reference of gradient checks.
extern "C" int ntt_oracle_sr_step_omp_f64(ort_field_t* em, ort_field_t* cur, ort_field_t* buff,
                                          ort_species_t* species, int32_t nspecies,
                                          const ort_step_params_t* prm,
                                          const uint8_t prtl_bc[3][2]) {
  t_par = true;
  const int r = sr_step<double>(em, cur, buff, species, nspecies, prm, prtl_bc);
  t_par = false;
  return r;
}

extern "C" int ntt_oracle_omp_threads(void) { return omp_get_max_threads(); }
