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
  __device__ __forceinline__ int64_t at(int c, int i, int j) const {
    i = ((i % n1) + n1) % n1;
    j = ((j % n2) + n2) % n2;
    return static_cast<int64_t>(i) + static_cast<int64_t>(n1) * j +
           static_cast<int64_t>(n1) * n2 * c;
  }
  __device__ __forceinline__ R get(int c, int i, int j) const { return d[at(c, i, j)]; }
  __device__ __forceinline__ void add(int c, int i, int j, R v) const { atomicAdd(d + at(c, i, j), v); }
};

template <typename R>
struct PAdj {  // per-particle adjoints
  R *lx1, *lx2, *lu1, *lu2, *lu3;
};

// ---- push VJP -------------------------------------------------------------------
// Forward of one particle (sr_move, 2D, Boris): e, b gathered at (i, dx),
// in-plane components times dxm, Boris, dx' = dx + (u'/dxm) dt/gamma'.
// Inputs of the Jacobian: dx1, dx2, u1..3, e1..3, b1..3 (11); outputs dx1',
// dx2', u1'..3' (5).
template <int O, typename R>
__global__ __launch_bounds__(256) void k_push_vjp(ParticleView P, EmView F, int ngh, R ndth,
                                                  R dt, R dxm, PAdj<R> post, PAdj<R> pre,
                                                  AdjField<R> lem) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  // dead particles, and particles outside the active cells (which the
  // forward push rejects), carry no adjoint
  const int n1 = F.N1 - 2 * ngh, n2 = F.N2 - 2 * ngh;
  if (P.tag[p] != NTT_TAG_ALIVE || P.i1[p] < 0 || P.i1[p] >= n1 || P.i2[p] < 0 ||
      P.i2[p] >= n2 || !(P.dx1[p] >= 0.0f && P.dx1[p] <= 1.0f) ||
      !(P.dx2[p] >= 0.0f && P.dx2[p] <= 1.0f)) {
    pre.lu1[p] = pre.lu2[p] = pre.lu3[p] = R(0);
    return;
  }
  using D11 = Dual<11, R>;
  using D2  = Dual<2, R>;
  const int  i = P.i1[p] + ngh, j = P.i2[p] + ngh;
  const D2   x1 = D2::var(static_cast<R>(P.dx1[p]), 0), x2 = D2::var(static_cast<R>(P.dx2[p]), 1);
  // gather with position derivatives: e_c, b_c as functions of (dx1, dx2)
  D2  S1p[O + 1], S1d[O + 1], S2p[O + 1], S2d[O + 1];
  int ip, id, jp, jd;
  shape_t<false, O, D2, R>(i, x1, ip, S1p);
  shape_t<false, O, D2, R>(j, x2, jp, S2p);
  shape_t<true, O, D2, R>(i, x1, id, S1d);
  shape_t<true, O, D2, R>(j, x2, jd, S2d);
  auto gather = [&](int c, int imin, int jmin, const D2* A, const D2* B) -> D2 {
    D2 e(R(0));
    for (int b = 0; b < O + 1; b++) {
      D2 c0(R(0));
      for (int a = 0; a < O + 1; a++) c0 = c0 + A[a] * static_cast<R>(F(c, imin + a, jmin + b));
      e = e + c0 * B[b];
    }
    return e;
  };
  D2 g[6];
  g[0] = gather(0, id, jp, S1d, S2p);
  g[1] = gather(1, ip, jd, S1p, S2d);
  g[2] = gather(2, ip, jp, S1p, S2p);
  g[3] = gather(3, ip, jd, S1p, S2d);
  g[4] = gather(4, id, jp, S1d, S2p);
  g[5] = gather(5, id, jd, S1d, S2d);
  // Boris + position update over the 11 inputs
  D11 u[3] = { D11::var(static_cast<R>(P.ux1[p]), 2), D11::var(static_cast<R>(P.ux2[p]), 3),
               D11::var(static_cast<R>(P.ux3[p]), 4) };
  D11 e[3], b[3];
  for (int c = 0; c < 3; ++c) {
    e[c] = D11::var(g[c].v, 5 + c);
    b[c] = D11::var(g[3 + c].v, 8 + c);
  }
  for (int c = 0; c < 2; ++c) {
    e[c] = e[c] * dxm;
    b[c] = b[c] * dxm;
  }
  boris_t<D11, R>(ndth, u, e, b);
  const D11 dtie = dt / tsqrt(R(1) + (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  const D11 X1 = D11::var(static_cast<R>(P.dx1[p]), 0) + (u[0] / D11(dxm)) * dtie;
  const D11 X2 = D11::var(static_cast<R>(P.dx2[p]), 1) + (u[1] / D11(dxm)) * dtie;
  const R lo[5] = { post.lx1[p], post.lx2[p], post.lu1[p], post.lu2[p], post.lu3[p] };
  R lin[11];
  for (int k = 0; k < 11; ++k)
    lin[k] = lo[0] * X1.d[k] + lo[1] * X2.d[k] + lo[2] * u[0].d[k] + lo[3] * u[1].d[k] +
             lo[4] * u[2].d[k];
  // adjoint of the gather: scatter into the field adjoint, and the
  // position derivative of the gathered values
  R lx1 = lin[0], lx2 = lin[1];
  for (int c = 0; c < 6; ++c) {
    const R l = lin[5 + c];
    lx1 += l * g[c].d[0];
    lx2 += l * g[c].d[1];
  }
  auto scatter = [&](int c, int imin, int jmin, const D2* A, const D2* B, R l) {
    if (l == R(0)) return;
    for (int bb = 0; bb < O + 1; bb++)
      for (int a = 0; a < O + 1; a++)
        lem.add(c, imin + a - ngh, jmin + bb - ngh, l * A[a].v * B[bb].v);
  };
  scatter(0, id, jp, S1d, S2p, lin[5]);
  scatter(1, ip, jd, S1p, S2d, lin[6]);
  scatter(2, ip, jp, S1p, S2p, lin[7]);
  scatter(3, ip, jd, S1p, S2d, lin[8]);
  scatter(4, id, jp, S1d, S2p, lin[9]);
  scatter(5, id, jd, S1d, S2d, lin[10]);
  pre.lx1[p] += lx1;
  pre.lx2[p] += lx2;
  pre.lu1[p] = lin[2];
  pre.lu2[p] = lin[3];
  pre.lu3[p] = lin[4];
}

// ---- deposit VJP ----------------------------------------------------------------
// Esirkepov 2D (currents_deposit.hpp:476-555) over the 7 inputs dx_prev1,2,
// dx1,2, u1..3 (u enters through the out-of-plane velocity); the adjoint of
// each current slot is read from the current adjoint at its periodic image.
template <int O, typename R>
__global__ __launch_bounds__(256) void k_deposit_vjp(ParticleView P, AdjField<R> lj, R charge,
                                                     R inv_dt, PAdj<R> post, R* lxp1, R* lxp2) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  if (P.tag[p] != NTT_TAG_ALIVE) {
    lxp1[p] = lxp2[p] = R(0);
    return;
  }
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
  R acc[7] = { R(0), R(0), R(0), R(0), R(0), R(0), R(0) };
  auto take = [&](int c, int a, int b, const D7& v) {
    const R l = lj.get(c, i1_min + a, i2_min + b);
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
  lxp1[p] = acc[0];
  lxp2[p] = acc[1];
  post.lx1[p] += acc[2];
  post.lx2[p] += acc[3];
  post.lu1[p] += acc[4];
  post.lu2[p] += acc[5];
  post.lu3[p] += acc[6];
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
                  static_cast<R*>(a->lu2), static_cast<R*>(a->lu3) };
}

bool padj_ok(const ntt_particle_adjoint_t* a) {
  return a && a->lx1 && a->lx2 && a->lu1 && a->lu2 && a->lu3;
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
  if (!ctx || !pre || !em || !metric || !padj_ok(post_adj) || !padj_ok(pre_adj) || !lam_em)
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
  const EmView ev{ F.d, F.N1, F.N2, F.cstride };
  hipStream_t st = as_stream(stream);
  const unsigned grid = grid_for(P.n, 256);
  // sr.hpp:111: normalized_dt_half = 1/2 (q/m) omegaB0 dt
  const float ndth = HALF * (ctx->charge / ctx->mass) * ctx->omegaB0 * ctx->dt;
  if (precision == 64) {
    AdjField<double> L{ static_cast<double*>(lam_em), F.n1, F.n2 };
    if (shape_order == 1)
      k_push_vjp<1, double><<<grid, 256, 0, st>>>(P, ev, F.ngh, ndth, ctx->dt, metric->dx,
                                                  padj<double>(post_adj), padj<double>(pre_adj), L);
    else
      k_push_vjp<2, double><<<grid, 256, 0, st>>>(P, ev, F.ngh, ndth, ctx->dt, metric->dx,
                                                  padj<double>(post_adj), padj<double>(pre_adj), L);
  } else {
    AdjField<float> L{ static_cast<float*>(lam_em), F.n1, F.n2 };
    if (shape_order == 1)
      k_push_vjp<1, float><<<grid, 256, 0, st>>>(P, ev, F.ngh, ndth, ctx->dt, metric->dx,
                                                 padj<float>(post_adj), padj<float>(pre_adj), L);
    else
      k_push_vjp<2, float><<<grid, 256, 0, st>>>(P, ev, F.ngh, ndth, ctx->dt, metric->dx,
                                                 padj<float>(post_adj), padj<float>(pre_adj), L);
  }
  return check_launch();
}

extern "C" ntt_status ntt_deposit_vjp(const ntt_particles_t* post, const int32_t n[2],
                                      float charge, float dt, int32_t shape_order,
                                      int32_t precision, const void* lam_j,
                                      ntt_particle_adjoint_t* post_adj, void* lx_prev1,
                                      void* lx_prev2, void* stream) {
  if (!post || !n || !lam_j || !padj_ok(post_adj) || !lx_prev1 || !lx_prev2) return NTT_ERR_ARG;
  if (shape_order != 1 && shape_order != 2) return NTT_ERR_UNSUPPORTED;
  if (precision != 32 && precision != 64) return NTT_ERR_ARG;
  if (n[0] < 1 || n[1] < 1 || !(dt > 0.0f)) return NTT_ERR_ARG;
  ParticleView P;
  ntt_status s = make_particle_view(post, 2, true, &P);
  if (s != NTT_OK) return s;
  if (P.n == 0) return NTT_OK;
  hipStream_t st = as_stream(stream);
  const unsigned grid = grid_for(P.n, 256);
  if (precision == 64) {
    AdjField<double> L{ const_cast<double*>(static_cast<const double*>(lam_j)), n[0], n[1] };
    auto pa = padj<double>(post_adj);
    if (shape_order == 1)
      k_deposit_vjp<1, double><<<grid, 256, 0, st>>>(P, L, charge, 1.0 / dt, pa,
                                                     static_cast<double*>(lx_prev1),
                                                     static_cast<double*>(lx_prev2));
    else
      k_deposit_vjp<2, double><<<grid, 256, 0, st>>>(P, L, charge, 1.0 / dt, pa,
                                                     static_cast<double*>(lx_prev1),
                                                     static_cast<double*>(lx_prev2));
  } else {
    AdjField<float> L{ const_cast<float*>(static_cast<const float*>(lam_j)), n[0], n[1] };
    auto pa = padj<float>(post_adj);
    if (shape_order == 1)
      k_deposit_vjp<1, float><<<grid, 256, 0, st>>>(P, L, charge, ONE / dt, pa,
                                                    static_cast<float*>(lx_prev1),
                                                    static_cast<float*>(lx_prev2));
    else
      k_deposit_vjp<2, float><<<grid, 256, 0, st>>>(P, L, charge, ONE / dt, pa,
                                                    static_cast<float*>(lx_prev1),
                                                    static_cast<float*>(lx_prev2));
  }
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
