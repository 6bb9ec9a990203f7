// adjoint.hip -- reverse-mode (vector-Jacobian) kernels of the 2D periodic SR
// step: particle push, charge-conserving deposit, Yee curls and the digital
// filter.  No reference counterpart: the reference has no adjoint.
//
// * Particle kernels differentiate a restatement of the forward per-particle
//   arithmetic (shape functions, gather, Boris, position update, Esirkepov
//   deposit -- sr.hpp:337-572 and currents_deposit.hpp:476-555 as in
//   sr_particle.hpp) written once over a scalar type T, evaluated with
//   forward-mode dual numbers: the Jacobian of one particle's update is exact
//   for the branch the particle takes (cell crossings and the B-spline
//   pieces are selected by the primal values, as in the forward).  A
//   particle's vector-Jacobian product is then a dot product with the
//   incoming adjoint.
// * The adjoint of a gather is a scatter (atomic adds into the field
//   adjoint); the adjoint of the deposit is a gather of the current adjoint.
// * The Yee curls and the 2D binomial filter are linear: their adjoints are
//   the transposed stencils (the filter is symmetric).
// * Field adjoints use an active-only layout [comp][n2][n1] with periodic
//   wrap; particle adjoints are per-particle arrays.  The real type is float
//   or double (precision 32 / 64).
#include "internal.hpp"

namespace ntt {
namespace {

template <int N, typename R>
struct Dual {
  R v;
  R d[N];
  __device__ __forceinline__ Dual() {}
  __device__ __forceinline__ Dual(R x) : v(x) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = R(0);
  }
  __device__ __forceinline__ static Dual var(R x, int k) {
    Dual r(x);
    r.d[k] = R(1);
    return r;
  }
};

template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator+(const Dual<N, R>& a, const Dual<N, R>& b) {
  Dual<N, R> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator-(const Dual<N, R>& a, const Dual<N, R>& b) {
  Dual<N, R> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator-(const Dual<N, R>& a) {
  Dual<N, R> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator*(const Dual<N, R>& a, const Dual<N, R>& b) {
  Dual<N, R> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator/(const Dual<N, R>& a, const Dual<N, R>& b) {
  Dual<N, R> r;
  const R inv = R(1) / b.v;
  r.v = a.v * inv;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> dsqrt(const Dual<N, R>& a) {
  Dual<N, R> r;
  r.v = sqrt(a.v);
  const R h = R(0.5) / r.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
  return r;
}
// mixed operands
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator*(R s, const Dual<N, R>& a) {
  Dual<N, R> r;
  r.v = s * a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator*(const Dual<N, R>& a, R s) { return s * a; }
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator+(const Dual<N, R>& a, R s) {
  Dual<N, R> r = a;
  r.v += s;
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator+(R s, const Dual<N, R>& a) { return a + s; }
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator-(R s, const Dual<N, R>& a) {
  Dual<N, R> r = -a;
  r.v += s;
  return r;
}
template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator-(const Dual<N, R>& a, R s) {
  Dual<N, R> r = a;
  r.v -= s;
  return r;
}

template <int N, typename R>
__device__ __forceinline__ Dual<N, R> operator/(R s, const Dual<N, R>& a) {
  return Dual<N, R>(s) / a;
}

template <typename R> __device__ __forceinline__ R val(R x) { return x; }
template <int N, typename R> __device__ __forceinline__ R val(const Dual<N, R>& x) { return x.v; }
template <typename R> __device__ __forceinline__ R tsqrt(R x) { return sqrt(x); }
template <int N, typename R> __device__ __forceinline__ Dual<N, R> tsqrt(const Dual<N, R>& x) {
  return dsqrt(x);
}

// ---- forward arithmetic over a scalar type T --------------------------------
// prtl_shape::order<STAG, O> (particle_shapes.hpp:544-588), O = 1, 2
template <bool STAG, int O, typename T, typename R>
__device__ __forceinline__ void shape_t(int i, const T& di, int& imin, T* S) {
  const R half(0.5), one(1), th(1.5), tf(0.75);
  if constexpr (O == 1) {
    if constexpr (!STAG) {
      imin = i;
      S[0] = one - di;
      S[1] = di;
    } else {
      if (val(di) < half) {
        imin = i - 1;
        S[0] = half - di;
      } else {
        imin = i;
        S[0] = th - di;
      }
      S[1] = one - S[0];
    }
  } else {
    if constexpr (!STAG) {
      if (val(di) < half) {
        imin = i - 1;
        const T a = half - di;
        S[0] = half * (a * a);
        S[1] = tf - di * di;
      } else {
        imin = i;
        const T a = th - di, b = one - di;
        S[0] = half * (a * a);
        S[1] = tf - b * b;
      }
      S[2] = one - S[0] - S[1];
    } else {
      imin = i - 1;
      const T a = one - di;
      S[0] = half * (a * a);
      S[2] = half * (di * di);
      S[1] = one - S[0] - S[2];
    }
  }
}

// for_deposit (particle_shapes.hpp:934-1024)
template <int O, typename T, typename R>
__device__ __forceinline__ void for_deposit_t(int i_init, const T& di_init, int i_fin,
                                              const T& di_fin, int& i_min, int& di, T* iS, T* fS) {
  int a, b;
  T   iS_[O + 1], fS_[O + 1];
  shape_t<false, O, T, R>(i_init, di_init, a, iS_);
  shape_t<false, O, T, R>(i_fin, di_fin, b, fS_);
  const T z(R(0));
  if (a < b) {
    i_min = a;
    di = O + 1;
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = z;
    fS[0] = z;
    for (int j = 0; j < O + 1; j++) fS[j + 1] = fS_[j];
  } else if (a > b) {
    i_min = b;
    di = O + 1;
    iS[0] = z;
    for (int j = 0; j < O + 1; j++) iS[j + 1] = iS_[j];
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = z;
  } else {
    i_min = a;
    di = O;
    for (int j = 0; j < O + 1; j++) iS[j] = iS_[j];
    iS[O + 1] = z;
    for (int j = 0; j < O + 1; j++) fS[j] = fS_[j];
    fS[O + 1] = z;
  }
}

// Boris (sr.hpp:337-368) over T
template <typename T, typename R>
__device__ __forceinline__ void boris_t(R coeff, T* u, T* e0, T* b0) {
  for (int c = 0; c < 3; ++c) e0[c] = coeff * e0[c];
  T u0[3] = { u[0] + e0[0], u[1] + e0[1], u[2] + e0[2] };
  T C = coeff * (R(1) / tsqrt(R(1) + (u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2])));
  for (int c = 0; c < 3; ++c) b0[c] = b0[c] * C;
  C = R(2) / (R(1) + (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]));
  const T u1[3] = { (u0[0] + (u0[1] * b0[2] - u0[2] * b0[1])) * C,
                    (u0[1] + (u0[2] * b0[0] - u0[0] * b0[2])) * C,
                    (u0[2] + (u0[0] * b0[1] - u0[1] * b0[0])) * C };
  u[0] = u0[0] + ((u1[1] * b0[2] - u1[2] * b0[1]) + e0[0]);
  u[1] = u0[1] + ((u1[2] * b0[0] - u1[0] * b0[2]) + e0[1]);
  u[2] = u0[2] + ((u1[0] * b0[1] - u1[1] * b0[0]) + e0[2]);
}

// ---- field access for the adjoint kernels ---------------------------------------
struct EmView {
  const float* d;
  int32_t N1, N2;
  int64_t cs;
  __device__ __forceinline__ float operator()(int c, int i, int j) const {
    return d[static_cast<int64_t>(i) + static_cast<int64_t>(N1) * j + cs * c];
  }
};

// active-only adjoint field [comp][n2][n1] with periodic wrap
template <typename R>
struct AdjField {
  R*      d;
  int32_t n1, n2;
  // indices are a few cells outside [0, n): one conditional add or subtract,
  // and the division only for a lane still outside (n below the reach)
  __device__ __forceinline__ static int wrap(int x, int n) {
    x = (x < 0) ? x + n : x;
    x = (x >= n) ? x - n : x;
    if (static_cast<unsigned>(x) >= static_cast<unsigned>(n)) x = ((x % n) + n) % n;
    return x;
  }
  __device__ __forceinline__ int64_t at(int c, int i, int j) const {
    return static_cast<int64_t>(wrap(i, n1)) + static_cast<int64_t>(n1) * wrap(j, n2) +
           static_cast<int64_t>(n1) * n2 * c;
  }
  __device__ __forceinline__ R get(int c, int i, int j) const { return d[at(c, i, j)]; }
  __device__ __forceinline__ void add(int c, int i, int j, R v) const { atomicAdd(d + at(c, i, j), v); }
};

template <typename R>
struct PAdj {  // per-particle adjoints, indexed by particle ID when ids != nullptr
  R *lx1, *lx2, *lu1, *lu2, *lu3;
  int64_t size;
};

// entry of particle p in the adjoint arrays: its ID, or p; -1 (and a device
// error) for an ID outside [0, size)
__device__ __forceinline__ int64_t adj_index(const int64_t* ids, int64_t p, int64_t size,
                                             uint32_t* err) {
  if (!ids) return p;
  const int64_t q = ids[p];
  if (q < 0 || q >= size) {
    raise_device_error(err, NTT_DEVERR_OUT_OF_BOUNDS);
    return -1;
  }
  return q;
}

// ---- deposit VJP ----------------------------------------------------------------
// Esirkepov 2D (currents_deposit.hpp:476-555) over the 7 inputs dx_prev1,2,
// dx1,2, u1..3 (u enters through the out-of-plane velocity); the adjoint of
// each current slot is read from the current adjoint at its periodic image.
// acc[0..6] = d<lam_j, J_p>/d(dx_prev1, dx_prev2, dx1, dx2, u1, u2, u3).
template <int O, typename R>
__device__ __forceinline__ void deposit_vjp_acc(const ParticleView& P, int64_t p,
                                                const AdjField<R>& lj, R charge, R inv_dt,
                                                R* acc) {
  using D7 = Dual<7, R>;
  constexpr int M = O + 2;
  const D7 xp1 = D7::var(static_cast<R>(P.dx1p[p]), 0), xp2 = D7::var(static_cast<R>(P.dx2p[p]), 1);
  const D7 x1 = D7::var(static_cast<R>(P.dx1[p]), 2), x2 = D7::var(static_cast<R>(P.dx2[p]), 3);
  const D7 u1 = D7::var(static_cast<R>(P.ux1[p]), 4), u2 = D7::var(static_cast<R>(P.ux2[p]), 5),
           u3 = D7::var(static_cast<R>(P.ux3[p]), 6);
  // vp_out_of_plane (currents_deposit.hpp:112-166)
  const D7 vz = u3 / tsqrt(R(1) + (u1 * u1 + u2 * u2 + u3 * u3));
  const R  coeff = static_cast<R>(P.w[p]) * charge;
  D7  iS1[M], fS1[M], iS2[M], fS2[M];
  int i1_min, d1, i2_min, d2;
  for_deposit_t<O, D7, R>(P.i1p[p], xp1, P.i1[p], x1, i1_min, d1, iS1, fS1);
  for_deposit_t<O, D7, R>(P.i2p[p], xp2, P.i2[p], x2, i2_min, d2, iS2, fS2);
  const R Qdt = coeff * inv_dt;
  for (int k = 0; k < 7; ++k) acc[k] = R(0);
  // slots inside the active range are read directly, others through the
  // periodic wrap
  const bool    inner = i1_min >= 0 && i1_min + M <= lj.n1 && i2_min >= 0 && i2_min + M <= lj.n2;
  const int64_t base  = static_cast<int64_t>(i1_min) + static_cast<int64_t>(lj.n1) * i2_min;
  const int64_t cs    = static_cast<int64_t>(lj.n1) * lj.n2;
  auto take = [&](int c, int a, int b, const D7& v) {
    const R l = inner ? lj.d[base + a + static_cast<int64_t>(lj.n1) * b + cs * c]
                      : lj.get(c, i1_min + a, i2_min + b);
    for (int k = 0; k < 7; ++k) acc[k] += l * v.d[k];
  };
  // jx1: prefix over a of -Qdt/2 (fS1 - iS1)(fS2 + iS2)
  for (int b = 0; b < M; ++b) {
    D7 run(R(0));
    for (int a = 0; a < M; ++a) {
      run = run - Qdt * (R(0.5) * ((fS1[a] - iS1[a]) * (fS2[b] + iS2[b])));
      if (a < d1 && b <= d2) take(0, a, b, run);
    }
  }
  for (int a = 0; a < M; ++a) {
    D7 run(R(0));
    for (int b = 0; b < M; ++b) {
      run = run - Qdt * (R(0.5) * ((fS1[a] + iS1[a]) * (fS2[b] - iS2[b])));
      if (a <= d1 && b < d2) take(1, a, b, run);
    }
  }
  const D7 QVx3 = coeff * vz;
  const R  third = static_cast<R>(0.333333f);
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b)
      if (a <= d1 && b <= d2) {
        const D7 w3 = third * (fS2[b] * (R(0.5) * iS1[a] + fS1[a]) +
                               iS2[b] * (R(0.5) * fS1[a] + iS1[a]));
        take(2, a, b, QVx3 * w3);
      }
}

// The same vector-Jacobian product by a hand-written reverse sweep (a few
// times the forward arithmetic instead of seven): the Esirkepov currents are
// prefix sums of separable products, so their adjoint is a suffix sum of the
// current adjoint times the other factor; the shapes' adjoints then go
// through the B-spline derivatives, and the out-of-plane term through
// v_z = u3 / gamma.  Equal to deposit_vjp_acc to round-off.
template <int O, typename R>
__device__ __forceinline__ void shape_d(int i, R di, int& imin, R* S, R* dS) {
  // prtl_shape::order<false, O> (particle_shapes.hpp:544-588) and its derivative
  if constexpr (O == 1) {
    imin  = i;
    S[0]  = R(1) - di;
    S[1]  = di;
    dS[0] = R(-1);
    dS[1] = R(1);
  } else {
    if (di < R(0.5)) {
      imin      = i - 1;
      const R a = R(0.5) - di;
      S[0]      = R(0.5) * a * a;
      S[1]      = R(0.75) - di * di;
      dS[0]     = -a;
      dS[1]     = R(-2) * di;
    } else {
      imin      = i;
      const R a = R(1.5) - di, b = R(1) - di;
      S[0]      = R(0.5) * a * a;
      S[1]      = R(0.75) - b * b;
      dS[0]     = -a;
      dS[1]     = R(2) * b;
    }
    S[2]  = R(1) - S[0] - S[1];
    dS[2] = -dS[0] - dS[1];
  }
}

// prtl_shape::order<true, O> (particle_shapes.hpp:544-588), the staggered
// shape, and its derivative in di
template <int O, typename R>
__device__ __forceinline__ void shape_sd(int i, R di, int& imin, R* S, R* dS) {
  if constexpr (O == 1) {
    if (di < R(0.5)) {
      imin = i - 1;
      S[0] = R(0.5) - di;
    } else {
      imin = i;
      S[0] = R(1.5) - di;
    }
    S[1]  = R(1) - S[0];
    dS[0] = R(-1);
    dS[1] = R(1);
  } else {
    imin      = i - 1;
    const R a = R(1) - di;
    S[0]      = R(0.5) * (a * a);
    S[2]      = R(0.5) * (di * di);
    S[1]      = R(1) - S[0] - S[2];
    dS[0]     = -a;
    dS[2]     = di;
    dS[1]     = -dS[0] - dS[2];
  }
}

template <int O, typename R>
__device__ __forceinline__ void for_deposit_d(int i_init, R di_init, int i_fin, R di_fin,
                                              int& i_min, int& di, R* iS, R* fS, R* diS,
                                              R* dfS) {
  int a, b;
  R   iS_[O + 1], fS_[O + 1], diS_[O + 1], dfS_[O + 1];
  shape_d<O, R>(i_init, di_init, a, iS_, diS_);
  shape_d<O, R>(i_fin, di_fin, b, fS_, dfS_);
#pragma unroll
  for (int j = 0; j < O + 2; ++j) iS[j] = fS[j] = diS[j] = dfS[j] = R(0);
  // particle_shapes.hpp:934-1024: the two stencils on one common range
  const int si = (a > b) ? 1 : 0, sf = (a < b) ? 1 : 0;
  i_min = (a < b) ? a : b;
  di    = (a == b) ? O : O + 1;
  // shifted copies with selects (static indices keep the arrays in registers)
#pragma unroll
  for (int j = 0; j < O + 2; ++j) {
    const R i0 = (j < O + 1) ? iS_[j] : R(0), i1 = (j > 0) ? iS_[j - 1] : R(0);
    const R d0 = (j < O + 1) ? diS_[j] : R(0), d1v = (j > 0) ? diS_[j - 1] : R(0);
    const R f0 = (j < O + 1) ? fS_[j] : R(0), f1 = (j > 0) ? fS_[j - 1] : R(0);
    const R e0 = (j < O + 1) ? dfS_[j] : R(0), e1 = (j > 0) ? dfS_[j - 1] : R(0);
    iS[j]  = si ? i1 : i0;
    diS[j] = si ? d1v : d0;
    fS[j]  = sf ? f1 : f0;
    dfS[j] = sf ? e1 : e0;
  }
}

// Current-adjoint accessors of the reverse deposit: slot (a, b) of component c
// for a particle whose stencil starts at (i1_min, i2_min).
template <typename R, int M>
struct LamGlobal {  // the periodic field in HBM; columns and rows wrapped once
  const R* d;
  int64_t  cs;
  int      col[M], row[M];
  __device__ __forceinline__ LamGlobal(const AdjField<R>& lj, int i1_min, int i2_min)
      : d(lj.d), cs(static_cast<int64_t>(lj.n1) * lj.n2) {
    // offsets within one component stay below n1 * n2 < 2^31
#pragma unroll
    for (int a = 0; a < M; ++a) {
      col[a] = AdjField<R>::wrap(i1_min + a, lj.n1);
      row[a] = AdjField<R>::wrap(i2_min + a, lj.n2) * lj.n1;
    }
  }
  __device__ __forceinline__ R operator()(int c, int a, int b) const {
    return d[cs * c + (col[a] + row[b])];
  }
};
template <typename R, int PITCH>
struct LamWindow {  // an LDS copy: w at the particle's first slot, cstride per component
  const R* w;
  int      cstride;
  __device__ __forceinline__ R operator()(int c, int a, int b) const {
    return w[c * cstride + (b * PITCH + a)];
  }
};

template <int O, typename R, class LamAt>
__device__ __forceinline__ void deposit_vjp_rev(const ParticleView& P, int64_t p, LamAt&& lam_at,
                                                R charge, R inv_dt, R* acc) {
  constexpr int M = O + 2;
  R   iS1[M], fS1[M], iS2[M], fS2[M], diS1[M], dfS1[M], diS2[M], dfS2[M];
  int i1_min, d1, i2_min, d2;
  for_deposit_d<O, R>(P.i1p[p], static_cast<R>(P.dx1p[p]), P.i1[p], static_cast<R>(P.dx1[p]),
                      i1_min, d1, iS1, fS1, diS1, dfS1);
  for_deposit_d<O, R>(P.i2p[p], static_cast<R>(P.dx2p[p]), P.i2[p], static_cast<R>(P.dx2[p]),
                      i2_min, d2, iS2, fS2, diS2, dfS2);
  const R u1 = P.ux1[p], u2 = P.ux2[p], u3 = P.ux3[p];
  const R gam   = sqrt(R(1) + (u1 * u1 + u2 * u2 + u3 * u3));
  const R vz    = u3 / gam;
  const R coeff = static_cast<R>(P.w[p]) * charge;
  const R c0    = -(coeff * inv_dt) * R(0.5);
  const R third = static_cast<R>(0.333333f);
  // lam(c, a, b): the current adjoint at slot (i1_min + a, i2_min + b) of
  // component c, from the accessor lam_at picks for this particle's slots
  lam_at(i1_min, i2_min, [&](const auto& lam) {
  // adjoints of the shape values
  R gi1[M], gf1[M], gi2[M], gf2[M];
#pragma unroll
  for (int j = 0; j < M; ++j) gi1[j] = gf1[j] = gi2[j] = gf2[j] = R(0);
  // jx1(a, b) = sum_{a' <= a} c0 (fS1 - iS1)[a'] (fS2 + iS2)[b], a < d1, b <= d2
#pragma unroll
  for (int b = 0; b < M; ++b) {
    if (b > d2) continue;
    const R B = fS2[b] + iS2[b];
    R       L = R(0), gB = R(0);
#pragma unroll
    for (int a = M - 1; a >= 0; --a) {
      if (a < d1) L += lam(0, a, b);
      const R gA = c0 * B * L;
      gf1[a] += gA;
      gi1[a] -= gA;
      gB += (fS1[a] - iS1[a]) * L;
    }
    gf2[b] += c0 * gB;
    gi2[b] += c0 * gB;
  }
  // jx2(a, b) = sum_{b' <= b} c0 (fS1 + iS1)[a] (fS2 - iS2)[b'], a <= d1, b < d2
#pragma unroll
  for (int a = 0; a < M; ++a) {
    if (a > d1) continue;
    const R A = fS1[a] + iS1[a];
    R       L = R(0), gA = R(0);
#pragma unroll
    for (int b = M - 1; b >= 0; --b) {
      if (b < d2) L += lam(1, a, b);
      const R gB = c0 * A * L;
      gf2[b] += gB;
      gi2[b] -= gB;
      gA += (fS2[b] - iS2[b]) * L;
    }
    gf1[a] += c0 * gA;
    gi1[a] += c0 * gA;
  }
  // jx3(a, b) = coeff vz third (fS2[b] (iS1[a] / 2 + fS1[a]) + iS2[b] (fS1[a] / 2 + iS1[a]))
  const R q3 = coeff * vz * third;
  R       K3 = R(0);  // d<lam, jx3> / d(coeff vz)
#pragma unroll
  for (int a = 0; a < M; ++a)
#pragma unroll
    for (int b = 0; b < M; ++b) {
      if (a > d1 || b > d2) continue;
      const R l  = lam(2, a, b);
      const R p1 = R(0.5) * iS1[a] + fS1[a], p2 = R(0.5) * fS1[a] + iS1[a];
      K3 += l * (fS2[b] * p1 + iS2[b] * p2);
      const R k = q3 * l;
      gf2[b] += k * p1;
      gi2[b] += k * p2;
      gf1[a] += k * (fS2[b] + R(0.5) * iS2[b]);
      gi1[a] += k * (R(0.5) * fS2[b] + iS2[b]);
    }
  K3 *= third;
  R s0 = R(0), s1 = R(0), s2 = R(0), s3 = R(0);
#pragma unroll
  for (int j = 0; j < M; ++j) {
    s0 += gi1[j] * diS1[j];
    s1 += gi2[j] * diS2[j];
    s2 += gf1[j] * dfS1[j];
    s3 += gf2[j] * dfS2[j];
  }
  acc[0] = s0;
  acc[1] = s1;
  acc[2] = s2;
  acc[3] = s3;
  // v_z = u3 / gamma
  const R g3 = gam * gam * gam, kz = coeff * K3;
  acc[4] = kz * (-u3 * u1 / g3);
  acc[5] = kz * (-u3 * u2 / g3);
  acc[6] = kz * (R(1) / gam - u3 * u3 / g3);
  });
}

template <int O, typename R>
__global__ __launch_bounds__(256) void k_deposit_vjp(ParticleView P, const int64_t* ids,
                                                     AdjField<R> lj, R charge, R inv_dt,
                                                     PAdj<R> post, PAdj<R> pre, uint32_t* err) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  const int64_t q = adj_index(ids, p, min(pre.size, post.size), err);
  if (q < 0) return;
  if (P.tag[p] != NTT_TAG_ALIVE) {
    pre.lx1[q] = pre.lx2[q] = R(0);
    return;
  }
  R acc[7];
  deposit_vjp_acc<O, R>(P, p, lj, charge, inv_dt, acc);
  pre.lx1[q] = acc[0];
  pre.lx2[q] = acc[1];
  post.lx1[q] += acc[2];
  post.lx2[q] += acc[3];
  post.lu1[q] += acc[4];
  post.lu2[q] += acc[5];
  post.lu3[q] += acc[6];
}

// Reverse sweep of Boris (sr.hpp:337-368) and the position update of the 2D
// push: given the adjoint lo5 of (dx1', dx2', u1', u2', u3'), writes to out[0..8]
// the adjoint of (u1..3, e1..3, b1..3), e and b being the gathered fields
// before the in-plane dxm scaling.  A few times the forward arithmetic; equal
// to the dual-number form to round-off.
template <typename R>
__device__ __forceinline__ void boris_vjp_rev(R coeff, R dt, R dxm, const R* u, const R* e,
                                              const R* b, const R* lo5, R* out) {
  auto cross = [](const R* x, const R* y, R* z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  // forward, keeping the intermediates
  const R es[3] = { e[0] * dxm, e[1] * dxm, e[2] }, bs[3] = { b[0] * dxm, b[1] * dxm, b[2] };
  R e0[3], u0[3], b0[3], v[3], u1[3], up[3], t[3];
  for (int c = 0; c < 3; ++c) {
    e0[c] = coeff * es[c];
    u0[c] = u[c] + e0[c];
  }
  // one division per reciprocal: in double each one is a long instruction chain
  const R idxm = R(1) / dxm;
  const R ig0  = R(1) / sqrt(R(1) + (u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]));
  const R C1   = coeff * ig0;
  for (int c = 0; c < 3; ++c) b0[c] = bs[c] * C1;
  const R C2 = R(2) / (R(1) + (b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]));
  cross(u0, b0, t);
  for (int c = 0; c < 3; ++c) {
    v[c]  = u0[c] + t[c];
    u1[c] = v[c] * C2;
  }
  cross(u1, b0, t);
  for (int c = 0; c < 3; ++c) up[c] = u0[c] + t[c] + e0[c];
  const R igp  = R(1) / sqrt(R(1) + (up[0] * up[0] + up[1] * up[1] + up[2] * up[2]));
  const R dtie = dt * igp;
  // reverse
  R lup[3] = { lo5[2], lo5[3], lo5[4] };
  // X_k = up_k / dxm * dtie (k = 1, 2), dtie = dt / gamma(up)
  lup[0] += lo5[0] * dtie * idxm;
  lup[1] += lo5[1] * dtie * idxm;
  const R ldtie = (lo5[0] * up[0] + lo5[1] * up[1]) * idxm;
  const R kg    = -ldtie * dt * (igp * igp * igp);
  for (int c = 0; c < 3; ++c) lup[c] += kg * up[c];
  // up = u0 + u1 x b0 + e0
  R lu0[3], le0[3], lu1[3], lb0[3];
  cross(b0, lup, lu1);              // d<l, u1 x b0>/d u1 = b0 x l
  cross(lup, u1, lb0);              // d<l, u1 x b0>/d b0 = l x u1
  for (int c = 0; c < 3; ++c) {
    lu0[c] = lup[c];
    le0[c] = lup[c];
  }
  // u1 = v C2, v = u0 + u0 x b0
  R lv[3];
  const R lC2 = lu1[0] * v[0] + lu1[1] * v[1] + lu1[2] * v[2];
  for (int c = 0; c < 3; ++c) lv[c] = lu1[c] * C2;
  cross(b0, lv, t);                 // d<lv, u0 x b0>/d u0 = b0 x lv
  for (int c = 0; c < 3; ++c) lu0[c] += lv[c] + t[c];
  cross(lv, u0, t);                 // d<lv, u0 x b0>/d b0 = lv x u0
  for (int c = 0; c < 3; ++c) lb0[c] += t[c];
  // C2 = 2 / (1 + |b0|^2): dC2/db0 = -C2^2 b0
  for (int c = 0; c < 3; ++c) lb0[c] += lC2 * (-C2 * C2 * b0[c]);
  // b0 = bs C1, C1 = coeff / gamma(u0)
  R lbs[3];
  const R lC1 = lb0[0] * bs[0] + lb0[1] * bs[1] + lb0[2] * bs[2];
  for (int c = 0; c < 3; ++c) lbs[c] = lb0[c] * C1;
  const R k1 = -lC1 * coeff * (ig0 * ig0 * ig0);
  for (int c = 0; c < 3; ++c) lu0[c] += k1 * u0[c];
  // u0 = u + e0, e0 = coeff es
  for (int c = 0; c < 3; ++c) le0[c] += lu0[c];
  out[0] = lu0[0];
  out[1] = lu0[1];
  out[2] = lu0[2];
  out[3] = coeff * le0[0] * dxm;
  out[4] = coeff * le0[1] * dxm;
  out[5] = coeff * le0[2];
  out[6] = lbs[0] * dxm;
  out[7] = lbs[1] * dxm;
  out[8] = lbs[2];
}

// ---- push VJP -------------------------------------------------------------------
// Forward of one particle (sr_move, 2D, Boris): e, b gathered at (i, dx),
// in-plane components times dxm, Boris, dx' = dx + (u'/dxm) dt/gamma'.
// Inputs of the Jacobian: dx1, dx2, u1..3, e1..3, b1..3 (11); outputs dx1',
// dx2', u1'..3' (5).
//
// The adjoint of the gather is a scatter of 6 x (O+1)^2 values per particle.
// A workgroup takes a chunk of kVjpChunk consecutive particles; when their
// cells (with the stencil reach [-1, +2]) fit a kVW x kVW window, in the
// plain frame or in the frame rotated by n/2 (a chunk straddling the periodic
// boundary), the scatter goes to an LDS window of doubles (ds_add_f64, about
// 20x the throughput of the 32-bit float LDS atomic on this chip), which is
// then added to the field adjoint with one global atomic per non-zero entry.
// Chunks that do not fit scatter with global atomics.
// Particles per thread and window side: two particles per thread with 28 x 28
// windows (the field adjoint in doubles plus copies of the current adjoint and
// of E/B: 66 KB of LDS in float, two workgroups per CU) measured faster than
// one particle per thread, and than 24 x 24 windows at three workgroups per CU
// (2D order 2, 128 x 128, 262,144 particles per species).
#ifndef NTT_VJP_PER_THREAD
#define NTT_VJP_PER_THREAD 2
#endif
constexpr int kVjpThreads = 256, kVjpWaves = kVjpThreads / 64, kVjpPerThread = NTT_VJP_PER_THREAD;
constexpr int kVjpChunk = kVjpThreads * kVjpPerThread;
#ifndef NTT_VJP_WINDOW
#define NTT_VJP_WINDOW 28
#endif
constexpr int kVW = NTT_VJP_WINDOW;

__device__ __forceinline__ int vjp_wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int vjp_wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int vjp_rotated(int c, int n) {
  const int h = n >> 1;
  const int r = c + h;
  return r >= n ? r - n : r;
}

// DEP = true fuses the deposit VJP of the same step in front: Q holds the
// particles after the push (same order as P), lj the current adjoint; the
// deposit's adjoint then feeds the push's in registers, post is only read and
// pre.lx is written (not accumulated).
template <int O, typename R, bool DEP>
__global__ __launch_bounds__(kVjpThreads) void k_push_vjp(ParticleView P, const int64_t* ids,
                                                          EmView F, int ngh, R ndth, R dt, R dxm,
                                                          PAdj<R> post, PAdj<R> pre,
                                                          AdjField<R> lem, uint32_t* err,
                                                          ParticleView Q, AdjField<R> lj,
                                                          R charge, R inv_dt) {
  __shared__ double win[6 * kVW * kVW];
  // with DEP, an LDS copy of the current adjoint over the chunk's deposit reach,
  // and of E/B over the gather's reach (the cells the scatter writes)
  __shared__ R      ljw[DEP ? 3 * kVW * kVW : 1];
  __shared__ float  emw[DEP ? 6 * kVW * kVW : 1];
  __shared__ int    red[8][kVjpWaves];
  const int     tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int64_t beg = static_cast<int64_t>(blockIdx.x) * kVjpChunk;
  const int64_t end = min(beg + kVjpChunk, P.n);
  const int     n1 = F.N1 - 2 * ngh, n2 = F.N2 - 2 * ngh;
  // dead particles, and particles outside the active cells (which the
  // forward push rejects), carry no adjoint
  auto valid = [&](int64_t p) {
    return P.tag[p] == NTT_TAG_ALIVE && P.i1[p] >= 0 && P.i1[p] < n1 && P.i2[p] >= 0 &&
           P.i2[p] < n2 && P.dx1[p] >= 0.0f && P.dx1[p] <= 1.0f && P.dx2[p] >= 0.0f &&
           P.dx2[p] <= 1.0f;
  };

  // ---- pass 1: the chunk's cell box in the plain and the rotated frame
  int lo[2] = { INT_MAX, INT_MAX }, hi[2] = { INT_MIN, INT_MIN };
  int rlo[2] = { INT_MAX, INT_MAX }, rhi[2] = { INT_MIN, INT_MIN };
  for (int64_t p = beg + tid; p < end; p += kVjpThreads) {
    if (!valid(p)) continue;
    const int c[2] = { P.i1[p], P.i2[p] }, n[2] = { n1, n2 };
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      const int r = vjp_rotated(c[d], n[d]);
      lo[d]  = min(lo[d], c[d]);
      hi[d]  = max(hi[d], c[d]);
      rlo[d] = min(rlo[d], r);
      rhi[d] = max(rhi[d], r);
    }
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int l = vjp_wave_min(lo[d]), h = vjp_wave_max(hi[d]);
    const int rl = vjp_wave_min(rlo[d]), rh = vjp_wave_max(rhi[d]);
    if (lane == 0) {
      red[4 * d][wv]     = l;
      red[4 * d + 1][wv] = h;
      red[4 * d + 2][wv] = rl;
      red[4 * d + 3][wv] = rh;
    }
  }
  __syncthreads();
  bool fits = true, fitsJ = DEP, rot[2];
  int  wlo[2], ext[2], wloJ[2], extJ[2];
  {
    const int n[2] = { n1, n2 };
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      int l = red[4 * d][0], h = red[4 * d + 1][0], rl = red[4 * d + 2][0], rh = red[4 * d + 3][0];
#pragma unroll
      for (int w = 1; w < kVjpWaves; ++w) {
        l  = min(l, red[4 * d][w]);
        h  = max(h, red[4 * d + 1][w]);
        rl = min(rl, red[4 * d + 2][w]);
        rh = max(rh, red[4 * d + 3][w]);
      }
      if (l > h) fits = fitsJ = false;  // no valid particle
      rot[d] = (rh - rl) < (h - l);
      // unwrapped cell coordinates: in the rotated frame, cells c >= n - n/2
      // are taken as c - n
      const int ul = rot[d] ? rl - (n[d] >> 1) : l, uh = rot[d] ? rh - (n[d] >> 1) : h;
      wlo[d] = ul - 1;
      ext[d] = uh - ul + 4;
      if (ext[d] > kVW) fits = false;
      // the deposit of a particle that moved at most one cell reaches
      // [cell - 2, cell + 3]
      wloJ[d] = ul - 2;
      extJ[d] = uh - ul + 6;
      if (extJ[d] > kVW) fitsJ = false;
    }
  }
  if (fits) {
    for (int e = tid; e < 6 * ext[0] * ext[1]; e += kVjpThreads) {
      const int x = e % ext[0], r = e / ext[0];
      win[r * kVW + x] = 0.0;  // r = c * ext[1] + y -> row (c, y) of the window
    }
  }
  if constexpr (DEP) {
    if (fitsJ) {
      const int64_t cs = static_cast<int64_t>(lj.n1) * lj.n2;
      for (int e = tid; e < 3 * extJ[0] * extJ[1]; e += kVjpThreads) {
        const int x = e % extJ[0], r = e / extJ[0];  // r = c * extJ[1] + y
        const int c = r / extJ[1], y = r - c * extJ[1];
        const int gx = AdjField<R>::wrap(wloJ[0] + x, lj.n1);
        const int gy = AdjField<R>::wrap(wloJ[1] + y, lj.n2);
        ljw[r * kVW + x] = lj.d[cs * c + (gx + static_cast<int64_t>(lj.n1) * gy)];
      }
    }
    if (fits) {
      for (int e = tid; e < 6 * ext[0] * ext[1]; e += kVjpThreads) {
        const int x = e % ext[0], r = e / ext[0];  // r = c * ext[1] + y
        const int c = r / ext[1], y = r - c * ext[1];
        // ghost-extended index of the periodic image of the window cell
        const int gx = AdjField<R>::wrap(wlo[0] + x, n1) + ngh;
        const int gy = AdjField<R>::wrap(wlo[1] + y, n2) + ngh;
        emw[r * kVW + x] = F.d[F.cs * c + (gx + static_cast<int64_t>(F.N1) * gy)];
      }
    }
  }
  __syncthreads();

  // ---- pass 2: per-particle VJP
  using D9 = Dual<9, R>;
  using D2 = Dual<2, R>;
  for (int64_t p = beg + tid; p < end; p += kVjpThreads) {
    const int64_t q = adj_index(ids, p, min(pre.size, post.size), err);
    if (q < 0) continue;
    R a7[7] = { R(0), R(0), R(0), R(0), R(0), R(0), R(0) };
    if constexpr (DEP) {
      // the current adjoint from the LDS copy when all of the particle's slots
      // fall inside it, else from HBM.  Q's cells are in its own periodic
      // frame, P's previous cell plus (Q.i_prev - P.i); the window is in P's
      // unwrapped frame.
      auto lam_at = [&](int i1m, int i2m, auto&& body) {
        constexpr int M = O + 2;
        const int sh1 = (rot[0] && P.i1[p] >= n1 - (n1 >> 1)) ? -n1 : 0;
        const int sh2 = (rot[1] && P.i2[p] >= n2 - (n2 >> 1)) ? -n2 : 0;
        const int x0 = i1m - (Q.i1p[p] - P.i1[p]) + sh1 - wloJ[0];
        const int y0 = i2m - (Q.i2p[p] - P.i2[p]) + sh2 - wloJ[1];
        if (fitsJ && x0 >= 0 && x0 + M <= extJ[0] && y0 >= 0 && y0 + M <= extJ[1])
          body(LamWindow<R, kVW>{ ljw + (y0 * kVW + x0), extJ[1] * kVW });
        else
          body(LamGlobal<R, M>(lj, i1m, i2m));
      };
      if (Q.tag[p] == NTT_TAG_ALIVE) deposit_vjp_rev<O, R>(Q, p, lam_at, charge, inv_dt, a7);
    }
    if (!valid(p)) {
      pre.lu1[q] = pre.lu2[q] = pre.lu3[q] = R(0);
      if constexpr (DEP) {
        pre.lx1[q] = a7[0];
        pre.lx2[q] = a7[1];
      }
      continue;
    }
    const int i = P.i1[p] + ngh, j = P.i2[p] + ngh;
    // cell offset of this particle's unwrapped frame (0 or -n)
    const int sh1 = (rot[0] && P.i1[p] >= n1 - (n1 >> 1)) ? -n1 : 0;
    const int sh2 = (rot[1] && P.i2[p] >= n2 - (n2 >> 1)) ? -n2 : 0;
    // gathered E, B (gv) and their derivatives in dx1 (g1) and dx2 (g2); the
    // shapes (values) are kept for the scatter
    R   gv[6], g1[6], g2[6];
    R   S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1];
    int ip, id, jp, jd;
    if constexpr (DEP) {
      // separable stencils: e = sum_b B_b sum_a A_a f, with the derivatives
      // taken through A (dx1) and B (dx2) only
      R T1p[O + 1], T1d[O + 1], T2p[O + 1], T2d[O + 1];
      shape_d<O, R>(i, static_cast<R>(P.dx1[p]), ip, S1p, T1p);
      shape_d<O, R>(j, static_cast<R>(P.dx2[p]), jp, S2p, T2p);
      shape_sd<O, R>(i, static_cast<R>(P.dx1[p]), id, S1d, T1d);
      shape_sd<O, R>(j, static_cast<R>(P.dx2[p]), jd, S2d, T2d);
      // f(c, imin, jmin) -> row access f_row(a, b) at (imin + a, jmin + b)
      auto gather = [&](const auto& rows, int c, int imin, int jmin, const R* A, const R* dA,
                        const R* B, const R* dB) {
        const auto f = rows(c, imin, jmin);
        R e = R(0), e1 = R(0), e2 = R(0);
#pragma unroll
        for (int b = 0; b < O + 1; b++) {
          R r = R(0), r1 = R(0);
#pragma unroll
          for (int a = 0; a < O + 1; a++) {
            const R v = static_cast<R>(f(a, b));
            r += A[a] * v;
            r1 += dA[a] * v;
          }
          e += r * B[b];
          e1 += r1 * B[b];
          e2 += r * dB[b];
        }
        gv[c] = e;
        g1[c] = e1;
        g2[c] = e2;
      };
      auto gather6 = [&](const auto& rows) {
        gather(rows, 0, id, jp, S1d, T1d, S2p, T2p);
        gather(rows, 1, ip, jd, S1p, T1p, S2d, T2d);
        gather(rows, 2, ip, jp, S1p, T1p, S2p, T2p);
        gather(rows, 3, ip, jd, S1p, T1p, S2d, T2d);
        gather(rows, 4, id, jp, S1d, T1d, S2p, T2p);
        gather(rows, 5, id, jd, S1d, T1d, S2d, T2d);
      };
      if (fits) {  // workgroup-uniform: the LDS copy of the window
        gather6([&](int c, int imin, int jmin) {
          const float* w = emw + ((c * ext[1] + (jmin - ngh + sh2 - wlo[1])) * kVW +
                                  (imin - ngh + sh1 - wlo[0]));
          return [w](int a, int b) { return w[b * kVW + a]; };
        });
      } else {
        gather6([&](int c, int imin, int jmin) {
          const float* g = F.d + F.cs * c + (imin + static_cast<int64_t>(F.N1) * jmin);
          const int    N1 = F.N1;
          return [g, N1](int a, int b) { return g[a + N1 * b]; };
        });
      }
    } else {
      // forward-mode reference: dual numbers over (dx1, dx2)
      const D2 x1 = D2::var(static_cast<R>(P.dx1[p]), 0), x2 = D2::var(static_cast<R>(P.dx2[p]), 1);
      D2 D1p[O + 1], D1d[O + 1], D2p[O + 1], D2d[O + 1];
      shape_t<false, O, D2, R>(i, x1, ip, D1p);
      shape_t<false, O, D2, R>(j, x2, jp, D2p);
      shape_t<true, O, D2, R>(i, x1, id, D1d);
      shape_t<true, O, D2, R>(j, x2, jd, D2d);
      auto gather = [&](int c, int imin, int jmin, const D2* A, const D2* B) {
        D2 e(R(0));
        for (int b = 0; b < O + 1; b++) {
          D2 c0(R(0));
          for (int a = 0; a < O + 1; a++) c0 = c0 + A[a] * static_cast<R>(F(c, imin + a, jmin + b));
          e = e + c0 * B[b];
        }
        gv[c] = e.v;
        g1[c] = e.d[0];
        g2[c] = e.d[1];
      };
      gather(0, id, jp, D1d, D2p);
      gather(1, ip, jd, D1p, D2d);
      gather(2, ip, jp, D1p, D2p);
      gather(3, ip, jd, D1p, D2d);
      gather(4, id, jp, D1d, D2p);
      gather(5, id, jd, D1d, D2d);
      for (int a = 0; a < O + 1; a++) {
        S1p[a] = D1p[a].v;
        S1d[a] = D1d[a].v;
        S2p[a] = D2p[a].v;
        S2d[a] = D2d[a].v;
      }
    }
    const R lo5[5] = { post.lx1[q] + a7[2], post.lx2[q] + a7[3], post.lu1[q] + a7[4],
                       post.lu2[q] + a7[5], post.lu3[q] + a7[6] };
    // lin = (d/d dx1, d/d dx2, d/du1..3, d/de1..3, d/db1..3) of <lo5, (dx', u')>.
    // The offsets enter the update only through the gathered fields (the
    // gather derivatives above) and additively (d dx'/d dx = 1).
    R lin[11];
    lin[0] = lo5[0];
    lin[1] = lo5[1];
    const R uin[3] = { static_cast<R>(P.ux1[p]), static_cast<R>(P.ux2[p]),
                       static_cast<R>(P.ux3[p]) };
    const R ein[3] = { gv[0], gv[1], gv[2] }, bin[3] = { gv[3], gv[4], gv[5] };
    if constexpr (DEP) {
      boris_vjp_rev<R>(ndth, dt, dxm, uin, ein, bin, lo5, lin + 2);
    } else {
      // forward-mode reference: dual numbers over the 9 inputs u1..3, e1..3, b1..3
      D9 u[3] = { D9::var(uin[0], 0), D9::var(uin[1], 1), D9::var(uin[2], 2) };
      D9 e[3], b[3];
      for (int c = 0; c < 3; ++c) {
        e[c] = D9::var(ein[c], 3 + c);
        b[c] = D9::var(bin[c], 6 + c);
      }
      for (int c = 0; c < 2; ++c) {
        e[c] = e[c] * dxm;
        b[c] = b[c] * dxm;
      }
      boris_t<D9, R>(ndth, u, e, b);
      const D9 dtie = dt / tsqrt(R(1) + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
      const D9 X1 = (u[0] / D9(dxm)) * dtie;   // dx1' - dx1
      const D9 X2 = (u[1] / D9(dxm)) * dtie;
      for (int k = 0; k < 9; ++k)
        lin[2 + k] = lo5[0] * X1.d[k] + lo5[1] * X2.d[k] + lo5[2] * u[0].d[k] +
                     lo5[3] * u[1].d[k] + lo5[4] * u[2].d[k];
    }
    // adjoint of the gather: the position derivative of the gathered values,
    // and a scatter into the field adjoint
    R lx1 = lin[0], lx2 = lin[1];
    for (int c = 0; c < 6; ++c) {
      const R l = lin[5 + c];
      lx1 += l * g1[c];
      lx2 += l * g2[c];
    }
    if (fits) {  // workgroup-uniform
      auto scatter = [&](int c, int imin, int jmin, const R* A, const R* B, R l) {
        if (l == R(0)) return;
        // window row (c, y) holds cells x = 0..ext[0]-1 of component c
        double* w = win + ((c * ext[1] + (jmin - ngh + sh2 - wlo[1])) * kVW +
                           (imin - ngh + sh1 - wlo[0]));
#pragma unroll
        for (int bb = 0; bb < O + 1; bb++) {
          const R lb = l * B[bb];
#pragma unroll
          for (int a = 0; a < O + 1; a++)
            atomicAdd(w + (bb * kVW + a), static_cast<double>(lb * A[a]));
        }
      };
      scatter(0, id, jp, S1d, S2p, lin[5]);
      scatter(1, ip, jd, S1p, S2d, lin[6]);
      scatter(2, ip, jp, S1p, S2p, lin[7]);
      scatter(3, ip, jd, S1p, S2d, lin[8]);
      scatter(4, id, jp, S1d, S2p, lin[9]);
      scatter(5, id, jd, S1d, S2d, lin[10]);
    } else {
      auto scatter = [&](int c, int imin, int jmin, const R* A, const R* B, R l) {
        if (l == R(0)) return;
        for (int bb = 0; bb < O + 1; bb++) {
          const R lb = l * B[bb];
          for (int a = 0; a < O + 1; a++)
            lem.add(c, imin + a - ngh, jmin + bb - ngh, lb * A[a]);
        }
      };
      scatter(0, id, jp, S1d, S2p, lin[5]);
      scatter(1, ip, jd, S1p, S2d, lin[6]);
      scatter(2, ip, jp, S1p, S2p, lin[7]);
      scatter(3, ip, jd, S1p, S2d, lin[8]);
      scatter(4, id, jp, S1d, S2p, lin[9]);
      scatter(5, id, jd, S1d, S2d, lin[10]);
    }
    if constexpr (DEP) {
      pre.lx1[q] = a7[0] + lx1;
      pre.lx2[q] = a7[1] + lx2;
    } else {
      pre.lx1[q] += lx1;
      pre.lx2[q] += lx2;
    }
    pre.lu1[q] = lin[2];
    pre.lu2[q] = lin[3];
    pre.lu3[q] = lin[4];
  }
  if (!fits) return;
  __syncthreads();
  // ---- flush: one global atomic per non-zero window entry
  for (int e = tid; e < 6 * ext[0] * ext[1]; e += kVjpThreads) {
    const int    x = e % ext[0], r = e / ext[0];
    const double v = win[r * kVW + x];
    if (v == 0.0) continue;
    const int c = r / ext[1], y = r - c * ext[1];
    lem.add(c, wlo[0] + x, wlo[1] + y, static_cast<R>(v));
  }
}

// ---- field VJPs (2D, periodic) -------------------------------------------------
// Faraday (B += M E with B1 -= c1 dy E3, B2 += c1 dx E3,
// B3 += c2 (dy E1 - dx E2), forward differences): lam_E += M^T lam_B.
template <typename R>
__global__ void k_faraday_vjp(AdjField<R> lb, AdjField<R> le, R c1, R c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= le.n1) return;
  const R e1 = c2 * (lb.get(2, i, j - 1) - lb.get(2, i, j));
  const R e2 = c2 * (lb.get(2, i, j) - lb.get(2, i - 1, j));
  const R e3 = c1 * (lb.get(0, i, j) - lb.get(0, i, j - 1) + lb.get(1, i - 1, j) - lb.get(1, i, j));
  le.d[le.at(0, i, j)] += e1;
  le.d[le.at(1, i, j)] += e2;
  le.d[le.at(2, i, j)] += e3;
}

// Ampere (E1 += c1 (B3 - B3(j-1)), E2 += c1 (B3(i-1) - B3),
// E3 += c2 (B1(j-1) - B1 + B2 - B2(i-1))): lam_B += N^T lam_E.
template <typename R>
__global__ void k_ampere_vjp(AdjField<R> le, AdjField<R> lb, R c1, R c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= lb.n1) return;
  const R b3 = c1 * (le.get(0, i, j) - le.get(0, i, j + 1) + le.get(1, i + 1, j) - le.get(1, i, j));
  const R b1 = c2 * (le.get(2, i, j + 1) - le.get(2, i, j));
  const R b2 = c2 * (le.get(2, i, j) - le.get(2, i + 1, j));
  lb.d[lb.at(0, i, j)] += b1;
  lb.d[lb.at(1, i, j)] += b2;
  lb.d[lb.at(2, i, j)] += b3;
}

// the 2D binomial filter (digital_filter.hpp:182-194) is symmetric: its
// adjoint is the same stencil
template <typename R>
__global__ void k_filter_vjp(AdjField<R> src, AdjField<R> dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= src.n1) return;
  for (int c = 0; c < 3; ++c) {
    auto b = [&](int x, int y) { return src.get(c, x, y); };
    dst.d[dst.at(c, i, j)] =
      R(0.25) * b(i, j) + R(0.125) * (b(i - 1, j) + b(i + 1, j) + b(i, j - 1) + b(i, j + 1)) +
      R(0.0625) * (b(i - 1, j - 1) + b(i + 1, j + 1) + b(i - 1, j + 1) + b(i + 1, j - 1));
  }
}

template <typename R>
PAdj<R> padj(const ntt_particle_adjoint_t* a) {
  return PAdj<R>{ static_cast<R*>(a->lx1), static_cast<R*>(a->lx2), static_cast<R*>(a->lu1),
                  static_cast<R*>(a->lu2), static_cast<R*>(a->lu3), a->size };
}

// every array present (positions only when lx_only), and room for every
// particle when they are indexed by position
bool padj_ok(const ntt_particle_adjoint_t* a, const ntt_particles_t* p, bool lx_only = false) {
  if (!a || !a->lx1 || !a->lx2 || a->size < 0) return false;
  if (!lx_only && (!a->lu1 || !a->lu2 || !a->lu3)) return false;
  return p->pld_i != nullptr || a->size >= p->npart;
}

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_sr_push_vjp(const ntt_sr_pusher_ctx_t* ctx, const ntt_particles_t* pre,
                                      const ntt_field_t* em, const ntt_minkowski_t* metric,
                                      int32_t shape_order, int32_t precision,
                                      const ntt_particle_adjoint_t* post_adj,
                                      ntt_particle_adjoint_t* pre_adj, void* lam_em,
                                      void* stream) {
  if (!ctx || !pre || !em || !metric || !lam_em || !padj_ok(post_adj, pre) ||
      !padj_ok(pre_adj, pre))
    return NTT_ERR_ARG;
  if (em->dim != 2 || metric->dim != 2) return NTT_ERR_UNSUPPORTED;
  if (shape_order != 1 && shape_order != 2) return NTT_ERR_UNSUPPORTED;
  if (ctx->pusher_flags != NTT_PUSHER_BORIS) return NTT_ERR_UNSUPPORTED;
  if (precision != 32 && precision != 64) return NTT_ERR_ARG;
  ParticleView P;
  ntt_status s = make_particle_view(pre, 2, true, &P);
  if (s != NTT_OK) return s;
  FieldView F;
  s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  if (P.n == 0) return NTT_OK;
  const EmView     ev{ F.d, F.N1, F.N2, F.cstride };
  const int64_t*   ids = static_cast<const int64_t*>(pre->pld_i);
  uint32_t*        err = device_error_word();
  hipStream_t      st  = as_stream(stream);
  const unsigned   grid = grid_for(P.n, kVjpChunk);
  // sr.hpp:111: normalized_dt_half = 1/2 (q/m) omegaB0 dt
  const float ndth = HALF * (ctx->charge / ctx->mass) * ctx->omegaB0 * ctx->dt;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    AdjField<R> L{ static_cast<R*>(lam_em), F.n1, F.n2 };
    const auto  po = padj<R>(post_adj), pr = padj<R>(pre_adj);
    const AdjField<R> none{ nullptr, 1, 1 };
    if (shape_order == 1)
      k_push_vjp<1, R, false><<<grid, kVjpThreads, 0, st>>>(P, ids, ev, F.ngh, ndth, ctx->dt,
                                                            metric->dx, po, pr, L, err, P, none,
                                                            R(0), R(0));
    else
      k_push_vjp<2, R, false><<<grid, kVjpThreads, 0, st>>>(P, ids, ev, F.ngh, ndth, ctx->dt,
                                                            metric->dx, po, pr, L, err, P, none,
                                                            R(0), R(0));
  };
  if (precision == 64) run(double{});
  else run(float{});
  return check_launch();
}

extern "C" ntt_status ntt_deposit_vjp(const ntt_particles_t* post, const int32_t n[2],
                                      float charge, float dt, int32_t shape_order,
                                      int32_t precision, const void* lam_j,
                                      ntt_particle_adjoint_t* post_adj,
                                      ntt_particle_adjoint_t* pre_adj, void* stream) {
  if (!post || !n || !lam_j || !padj_ok(post_adj, post) || !padj_ok(pre_adj, post, true))
    return NTT_ERR_ARG;
  if (shape_order != 1 && shape_order != 2) return NTT_ERR_UNSUPPORTED;
  if (precision != 32 && precision != 64) return NTT_ERR_ARG;
  if (n[0] < 1 || n[1] < 1 || !(dt > 0.0f)) return NTT_ERR_ARG;
  ParticleView P;
  ntt_status s = make_particle_view(post, 2, true, &P);
  if (s != NTT_OK) return s;
  if (P.n == 0) return NTT_OK;
  const int64_t* ids  = static_cast<const int64_t*>(post->pld_i);
  uint32_t*      err  = device_error_word();
  hipStream_t    st   = as_stream(stream);
  const unsigned grid = grid_for(P.n, 256);
  auto run = [&](auto tag) {
    using R = decltype(tag);
    AdjField<R> L{ const_cast<R*>(static_cast<const R*>(lam_j)), n[0], n[1] };
    const R     inv_dt = R(1) / static_cast<R>(dt);
    // pre_adj's momentum arrays are not touched (they may be null)
    PAdj<R> pr{ static_cast<R*>(pre_adj->lx1), static_cast<R*>(pre_adj->lx2), nullptr, nullptr,
                nullptr, pre_adj->size };
    if (shape_order == 1)
      k_deposit_vjp<1, R><<<grid, 256, 0, st>>>(P, ids, L, charge, inv_dt, padj<R>(post_adj), pr,
                                                err);
    else
      k_deposit_vjp<2, R><<<grid, 256, 0, st>>>(P, ids, L, charge, inv_dt, padj<R>(post_adj), pr,
                                                err);
  };
  if (precision == 64) run(double{});
  else run(float{});
  return check_launch();
}

extern "C" ntt_status ntt_sr_particle_vjp(const ntt_sr_pusher_ctx_t* ctx,
                                          const ntt_particles_t* pre, const ntt_particles_t* post,
                                          const ntt_field_t* em, const ntt_minkowski_t* metric,
                                          int32_t shape_order, int32_t precision,
                                          const void* lam_j,
                                          const ntt_particle_adjoint_t* post_adj,
                                          ntt_particle_adjoint_t* pre_adj, void* lam_em,
                                          void* stream) {
  if (!ctx || !pre || !post || !em || !metric || !lam_j || !lam_em || !padj_ok(post_adj, pre) ||
      !padj_ok(pre_adj, pre))
    return NTT_ERR_ARG;
  if (pre->npart != post->npart) return NTT_ERR_ARG;
  if (em->dim != 2 || metric->dim != 2) return NTT_ERR_UNSUPPORTED;
  if (shape_order != 1 && shape_order != 2) return NTT_ERR_UNSUPPORTED;
  if (ctx->pusher_flags != NTT_PUSHER_BORIS) return NTT_ERR_UNSUPPORTED;
  if (precision != 32 && precision != 64) return NTT_ERR_ARG;
  if (!(ctx->dt > 0.0f)) return NTT_ERR_ARG;
  ParticleView P, Q;
  ntt_status s = make_particle_view(pre, 2, true, &P);
  if (s != NTT_OK) return s;
  s = make_particle_view(post, 2, true, &Q);
  if (s != NTT_OK) return s;
  FieldView F;
  s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  if (P.n == 0) return NTT_OK;
  const EmView   ev{ F.d, F.N1, F.N2, F.cstride };
  const int64_t* ids  = static_cast<const int64_t*>(pre->pld_i);
  uint32_t*      err  = device_error_word();
  hipStream_t    st   = as_stream(stream);
  const unsigned grid = grid_for(P.n, kVjpChunk);
  const float    ndth = HALF * (ctx->charge / ctx->mass) * ctx->omegaB0 * ctx->dt;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    AdjField<R> L{ static_cast<R*>(lam_em), F.n1, F.n2 };
    AdjField<R> LJ{ const_cast<R*>(static_cast<const R*>(lam_j)), F.n1, F.n2 };
    const auto  po = padj<R>(post_adj), pr = padj<R>(pre_adj);
    const R     inv_dt = R(1) / static_cast<R>(ctx->dt);
    if (shape_order == 1)
      k_push_vjp<1, R, true><<<grid, kVjpThreads, 0, st>>>(P, ids, ev, F.ngh, ndth, ctx->dt,
                                                           metric->dx, po, pr, L, err, Q, LJ,
                                                           static_cast<R>(ctx->charge), inv_dt);
    else
      k_push_vjp<2, R, true><<<grid, kVjpThreads, 0, st>>>(P, ids, ev, F.ngh, ndth, ctx->dt,
                                                           metric->dx, po, pr, L, err, Q, LJ,
                                                           static_cast<R>(ctx->charge), inv_dt);
  };
  if (precision == 64) run(double{});
  else run(float{});
  return check_launch();
}

extern "C" ntt_status ntt_yee_vjp(int32_t op, const int32_t n[2], float c1, float c2,
                                  int32_t precision, const void* lam_in, void* lam_out,
                                  void* stream) {
  if (!n || !lam_in || !lam_out || n[0] < 1 || n[1] < 1 || lam_in == lam_out) return NTT_ERR_ARG;
  if (precision != 32 && precision != 64) return NTT_ERR_ARG;
  if (op != NTT_VJP_FARADAY && op != NTT_VJP_AMPERE && op != NTT_VJP_FILTER) return NTT_ERR_ARG;
  hipStream_t st = as_stream(stream);
  const dim3 grid((n[0] + 255) / 256, n[1]);
  auto run = [&](auto tag) {
    using R = decltype(tag);
    AdjField<R> in{ const_cast<R*>(static_cast<const R*>(lam_in)), n[0], n[1] };
    AdjField<R> out{ static_cast<R*>(lam_out), n[0], n[1] };
    if (op == NTT_VJP_FARADAY) k_faraday_vjp<R><<<grid, 256, 0, st>>>(in, out, c1, c2);
    else if (op == NTT_VJP_AMPERE) k_ampere_vjp<R><<<grid, 256, 0, st>>>(in, out, c1, c2);
    else k_filter_vjp<R><<<grid, 256, 0, st>>>(in, out);
  };
  if (precision == 64) run(double{});
  else run(float{});
  return check_launch();
}
