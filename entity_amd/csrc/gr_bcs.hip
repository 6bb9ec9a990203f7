// gr_bcs.hip -- what the full GRPIC step adds around the GR field solver and
// the GR pusher, for a 2D axisymmetric (r, theta) box on gfx950: the field
// boundaries (match layer, horizon, axis), the absorbing layer of the
// current, the current filter of spherical grids, and the GR current deposit.
//
// Reference:
// * kernel::bc::MatchBoundaries_kernel<GRPIC, M, FS, in::x1>, 2D
//   (src/kernels/fields_bcs.hpp:176-345), with its range from
//   grpic::MatchFieldsIn (src/engines/grpic/fields_bcs.h:41-138);
// * kernel::bc::gr::HorizonBoundaries_kernel (fields_bcs.hpp:1188-1240);
// * kernel::bc::AxisBoundaries_kernel (fields_bcs.hpp:825-867);
// * kernel::bc::gr::AbsorbCurrents_kernel (fields_bcs.hpp:1243-1280);
// * kernel::DigitalFilter_kernel<2D, non-Cartesian> (digital_filter.hpp:
//   196-266) over the range of grpic::CurrentsFilter (currents.h:80-110);
// * kernel::DepositCurrents_kernel<GRPIC, M> (currents_deposit.hpp:108-555),
//   always order 1 in the GRPIC engine (src/engines/grpic/currents.h:48).
//
// Design: every kernel is one thread per cell (per column for the axis and
// horizon rows, per particle for the deposit).  The match layer evaluates
// the radius in FP64 and blends in FP64, rounding once to the FP32 field.
// The target of the match layer is an array with the layout of em holding
// the problem's field at every component's own staggered point: the
// reference evaluates its field setter there (fields_bcs.hpp:183-330); a
// caller fills the array once from the same function.  The deposit's
// out-of-plane velocity uses the metric in the pusher's precision (FP64 by
// default) and the in-plane Esirkepov current of the SR deposit, summed in an
// LDS window of doubles per workgroup and added to J with FP32 atomics.
#include "gr_metric.hpp"

#include <climits>

namespace ntt {
namespace {

using R = double;

struct BFld {
  float*  d;
  int32_t N1, N2;
  int64_t cs;
  __device__ __forceinline__ float& at(int i, int j, int c) const {
    return d[static_cast<int64_t>(i) + static_cast<int64_t>(N1) * j + cs * c];
  }
};
BFld bfld(const FieldView& v) { return BFld{ v.d, v.N1, v.N2, v.cstride }; }

constexpr int kB = 256;
constexpr uint32_t kTagD = 0x7u, kTagB = 0x38u;

__device__ __forceinline__ bool cell2(int lo1, int hi1, int lo2, int hi2, int& i1, int& i2) {
  const int w = hi1 - lo1;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w <= 0 || t >= static_cast<int64_t>(w) * (hi2 - lo2)) return false;
  i1 = lo1 + static_cast<int>(t % w);
  i2 = lo2 + static_cast<int>(t / w);
  return true;
}
unsigned blocks2(int lo1, int hi1, int lo2, int hi2) {
  const int64_t n = static_cast<int64_t>(hi1 - lo1) * (hi2 - lo2);
  return static_cast<unsigned>(n > 0 ? (n + kB - 1) / kB : 1);
}

// Match layer: F = s F + (1 - s) T, s = tanh(4 |r - r_edge| / ds), with r at
// each component's x1 position; D^3 on and past an axis row is only damped
// (fields_bcs.hpp:296-310).
struct MatchArgs {
  R        r_edge, four_inv_ds;
  int      i1_lo, i1_hi, G, N2;
  uint32_t comps;
  int      setD, setB, axis_min, axis_max;
};

template <int KIND>
__global__ __launch_bounds__(kB) void k_match(BFld F, BFld T, KSMetric<R, KIND> M, MatchArgs a) {
  int i1, i2;
  if (!cell2(a.i1_lo, a.i1_hi, 0, a.N2, i1, i2)) return;
  const R x1 = static_cast<R>(i1 - a.G);
  auto shape = [&](R x) { return tanh(fabs(M.r_of(x) - a.r_edge) * a.four_inv_ds); };
  auto has = [&](int c) { return ((a.comps >> c) & 1u) != 0; };
  auto blend = [&](int c, R s) {
    F.at(i1, i2, c) = static_cast<float>(s * F.at(i1, i2, c) + (R(1) - s) * T.at(i1, i2, c));
  };
  if ((has(0) && a.setD) || (has(4) && a.setB)) {   // (i1 + 1/2, i2)
    const R s = shape(x1 + R(0.5));
    if (has(0) && a.setD) blend(0, s);
    if (has(4) && a.setB) blend(4, s);
  }
  if ((has(1) && a.setD) || (has(3) && a.setB) || (has(2) && a.setD)) {   // x1 = i1
    const R s = shape(x1);
    if (has(1) && a.setD) blend(1, s);
    if (has(3) && a.setB) blend(3, s);
    if (has(2) && a.setD) {
      R v = s * F.at(i1, i2, 2);
      if ((!a.axis_min || i2 > a.G) && (!a.axis_max || i2 < a.N2 - a.G))
        v += (R(1) - s) * T.at(i1, i2, 2);
      F.at(i1, i2, 2) = static_cast<float>(v);
    }
  }
  if (has(5) && a.setB) blend(5, shape(x1 + R(0.5)));   // (i1 + 1/2, i2 + 1/2)
}

// Horizon: x1 rows 0 .. 2 + nfilter take row G + 1 + nfilter, x2 rows
// [G, G + n2]; one thread per x2 row.
__global__ __launch_bounds__(kB) void k_horizon(BFld F, int G, int n2, int nfilter, int setD,
                                                int setB) {
  const int i2 = G + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i2 > G + n2) return;
  const int src = G + 1 + nfilter;
  for (int c = 0; c < 6; ++c) {
    if ((c < 3 && !setD) || (c >= 3 && !setB)) continue;
    const float v = F.at(src, i2, c);
    for (int i = 0; i <= 2 + nfilter; ++i) F.at(i, i2, c) = v;
  }
}

// Axis rows (one thread per x1 column, ghosts included): the theta
// components are odd across the axis, the others even; B^theta and D^phi
// vanish on it.
__global__ __launch_bounds__(kB) void k_axis(BFld F, int e, int upper, int setD, int setB) {
  const int i1 = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i1 >= F.N1) return;
  if (!upper) {
    if (setD) {
      F.at(i1, e - 1, 1) = -F.at(i1, e, 1);
      F.at(i1, e, 2)     = 0.0f;
      F.at(i1, e - 1, 2) = F.at(i1, e + 1, 2);
    }
    if (setB) {
      F.at(i1, e - 1, 3) = F.at(i1, e, 3);
      F.at(i1, e, 4)     = 0.0f;
      F.at(i1, e - 1, 4) = -F.at(i1, e + 1, 4);
      F.at(i1, e - 1, 5) = F.at(i1, e, 5);
    }
  } else {
    if (setD) {
      F.at(i1, e, 1)     = -F.at(i1, e - 1, 1);
      F.at(i1, e, 2)     = 0.0f;
      F.at(i1, e + 1, 2) = F.at(i1, e - 1, 2);
    }
    if (setB) {
      F.at(i1, e, 3)     = F.at(i1, e - 1, 3);
      F.at(i1, e, 4)     = 0.0f;
      F.at(i1, e + 1, 4) = -F.at(i1, e - 1, 4);
      F.at(i1, e, 5)     = F.at(i1, e - 1, 5);
    }
  }
}

// Absorbing layer of the current: J *= tanh(|r(i1) - r_edge| / (ds / 4)).
template <int KIND>
__global__ __launch_bounds__(kB) void k_absorb_j(BFld J, KSMetric<R, KIND> M, R r_edge,
                                                 R inv_q, int i1_lo, int i1_hi, int G) {
  int i1, i2;
  if (!cell2(i1_lo, i1_hi, 0, J.N2, i1, i2)) return;
  const R f = tanh(fabs(M.r_of(static_cast<R>(i1 - G)) - r_edge) * inv_q);
  for (int c = 0; c < 3; ++c) J.at(i1, i2, c) = static_cast<float>(f * J.at(i1, i2, c));
}

// Current filter of a spherical grid: the 9-point binomial stencil off the
// axis; on the two rows next to each axis the r-filtered values combine in
// theta as the axis parities require (digital_filter.hpp:196-266).
__global__ __launch_bounds__(kB) void k_filter_sph(BFld A, BFld B, int G, int n1, int n2,
                                                   int axis_min, int axis_max) {
  int i1, i2;
  if (!cell2(G, G + n1, G, G + n2 + 1, i1, i2)) return;
  const float H = 0.5f, Q = 0.25f, E8 = 0.125f, S16 = 0.0625f;
  auto f1 = [&](int c, int j) {
    return H * B.at(i1, j, c) + Q * (B.at(i1 - 1, j, c) + B.at(i1 + 1, j, c));
  };
  const int jmin = G, jmax = G + n2;
  if (axis_min && i2 == jmin) {
    A.at(i1, i2, 0) = H * f1(0, i2) + H * f1(0, i2 + 1);
    A.at(i1, i2, 2) = 0.0f;
    A.at(i1, i2, 1) = Q * (f1(1, i2) + f1(1, i2 + 1));
  } else if (axis_min && i2 == jmin + 1) {
    A.at(i1, i2, 0) = H * f1(0, i2) + Q * (f1(0, i2 + 1) + f1(0, i2 - 1));
    A.at(i1, i2, 2) = H * f1(2, i2) + Q * f1(2, i2 + 1);
    A.at(i1, i2, 1) = H * f1(1, i2) + Q * (f1(1, i2 - 1) + f1(1, i2 + 1));
  } else if (axis_max && i2 == jmax - 1) {
    A.at(i1, i2, 0) = H * f1(0, i2) + Q * (f1(0, i2 - 1) + f1(0, i2 + 1));
    A.at(i1, i2, 2) = H * f1(2, i2) + Q * f1(2, i2 - 1);
    A.at(i1, i2, 1) = Q * (f1(1, i2) + f1(1, i2 - 1));
  } else if (axis_max && i2 == jmax) {
    A.at(i1, i2, 0) = H * f1(0, i2) + H * f1(0, i2 - 1);
    A.at(i1, i2, 2) = 0.0f;
  } else {
    for (int c = 0; c < 3; ++c)
      A.at(i1, i2, c) = Q * B.at(i1, i2, c) +
                        E8 * (B.at(i1 - 1, i2, c) + B.at(i1 + 1, i2, c) + B.at(i1, i2 - 1, c) +
                              B.at(i1, i2 + 1, c)) +
                        S16 * (B.at(i1 - 1, i2 - 1, c) + B.at(i1 + 1, i2 + 1, c) +
                               B.at(i1 - 1, i2 + 1, c) + B.at(i1 + 1, i2 - 1, c));
  }
}

// J sink for the deposit: FP32 global atomics into the ghost-extended array
struct AtomicJ {
  BFld J;
  int  i0 = 0, j0 = 0;
  __device__ __forceinline__ void set_origin(int i, int j, int) {
    i0 = i;
    j0 = j;
  }
  __device__ __forceinline__ void add_rel(int a, int b, int, int c, float v) {
    unsafeAtomicAdd(&J.at(i0 + a, j0 + b, c), v);
  }
  __device__ __forceinline__ void add(int i, int j, int, int c, float v) {
    unsafeAtomicAdd(&J.at(i, j, c), v);
  }
};

// J sink into an LDS window of doubles (ds_add_f64, about 20x the throughput
// of the FP32 LDS atomic on gfx950), origin (w1, w2) in ghost-extended cells
constexpr int kGW1 = 64, kGW2 = 16, kGWC = kGW1 * kGW2;
struct LdsJ2 {
  double* w;
  int     w1, w2;
  int     base = 0;
  __device__ __forceinline__ void set_origin(int i, int j, int) { base = (i - w1) + kGW1 * (j - w2); }
  __device__ __forceinline__ void add_rel(int a, int b, int, int c, float v) {
    atomicAdd(w + c * kGWC + base + a + kGW1 * b, static_cast<double>(v));
  }
  __device__ __forceinline__ void add(int i, int j, int, int c, float v) {
    atomicAdd(w + c * kGWC + (i - w1) + kGW1 * (j - w2), static_cast<double>(v));
  }
};

struct GRDepArgs {
  ParticleView P;
  BFld         J;
  int          G;
  float        charge, inv_dt;
  uint32_t*    err;
};

// workgroup-wide min (MAX = false) or max of v; every thread calls it
template <bool MAX>
__device__ __forceinline__ int blk_reduce(int v, int* red) {
  auto op = [](int x, int y) { return MAX ? max(x, y) : min(x, y); };
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
#pragma unroll
  for (int w = 1; w < kB / 64; ++w) r = op(r, red[w]);
  __syncthreads();
  return r;
}

// DepositCurrents_kernel<GRPIC, M> at order 1: theta clamped as the pusher
// does, v^i = h^ij u_j, v^3 alpha / sqrt(1 + u_i v^i) as the out-of-plane
// velocity (zero if not finite), then the 2D Esirkepov deposit.  A workgroup
// takes kB consecutive particles; when their deposit stencils fit a
// kGW1 x kGW2 window (particles kept in cell order, as a fresh or re-sorted
// plasma is, mostly do), the deposit goes to an LDS window of doubles and
// is added to J with one FP32 atomic per non-zero entry, else straight to J:
// the FP32 atomics of neighbouring particles no longer queue on the same
// addresses.  Rounding: the window sums in FP64, so only the order of the
// final FP32 additions is free, as it is for the direct atomics.
template <int KIND, typename RR>
__global__ __launch_bounds__(kB) void k_gr_deposit(GRDepArgs a, KSMetric<RR, KIND> M) {
  __shared__ double jw[3 * kGWC];
  __shared__ int    red[kB / 64];
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  bool live = p < a.P.n && a.P.tag[p] != NTT_TAG_DEAD;
  PState s;
  int lo[2] = { INT_MAX, INT_MAX }, hi[2] = { INT_MIN, INT_MIN };
  if (live) {
    s.i[0] = a.P.i1[p];   s.ip[0] = a.P.i1p[p];
    s.dx[0] = a.P.dx1[p]; s.dxp[0] = a.P.dx1p[p];
    s.i[1] = a.P.i2[p];   s.ip[1] = a.P.i2p[p];
    s.dx[1] = a.P.dx2[p]; s.dxp[1] = a.P.dx2p[p];
    for (int d = 0; d < 2; ++d) {
      deposit_extent<1>(s.i[d], s.dx[d], s.ip[d], s.dxp[d], lo[d], hi[d]);
      if (lo[d] + a.G < 0 || hi[d] + a.G >= (d == 0 ? a.J.N1 : a.J.N2)) {
        raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
        live = false;
      }
    }
    if (!live) lo[0] = lo[1] = INT_MAX, hi[0] = hi[1] = INT_MIN;
  }
  // the workgroup's box (ghost-extended cells) and whether it fits the window
  const int b1 = blk_reduce<false>(lo[0], red), b2 = blk_reduce<false>(lo[1], red);
  const int e1 = blk_reduce<true>(hi[0], red), e2 = blk_reduce<true>(hi[1], red);
  if (b1 > e1) return;   // no live particle (uniform across the workgroup)
  const bool fits = e1 - b1 < kGW1 && e2 - b2 < kGW2;
  const int w1 = b1 + a.G, w2 = b2 + a.G;
  if (fits) {
    for (int e = threadIdx.x; e < 3 * kGWC; e += kB) jw[e] = 0.0;
    __syncthreads();
  }
  if (live) {
    const RR u1 = a.P.ux1[p], u2 = a.P.ux2[p], u3 = a.P.ux3[p];
    const RR x1 = static_cast<RR>(s.i[0]) + static_cast<RR>(s.dx[0]);
    RR x2 = static_cast<RR>(s.i[1]) + static_cast<RR>(s.dx[1]);
    const RR th = M.theta_ph(x2);
    if (th < M.k.small_angle) x2 = M.k.th_small_cd;
    else if (static_cast<double>(th) >= M.k.large_angle) x2 = M.k.th_large_cd;
    MPt<RR> m;
    M.eval(x1, x2, m);
    const RR v1 = u1 * m.h11 + u3 * m.h13, v2 = u2 * m.h22, v3 = u1 * m.h13 + u3 * m.h33;
    const RR inv_energy = m.alpha / sqrt(RR(1) + u1 * v1 + u2 * v2 + u3 * v3);
    RR vz = v3;
    if (isnan(vz) || isinf(vz)) vz = RR(0);
    const float vzf = static_cast<float>(vz * inv_energy), coeff = a.P.w[p] * a.charge;
    if (fits) {
      LdsJ2 J{ jw, w1, w2 };
      deposit_esirkepov_2d<1>(J, a.G, s, vzf, coeff, a.inv_dt);
    } else {
      AtomicJ J{ a.J };
      deposit_esirkepov_2d<1>(J, a.G, s, vzf, coeff, a.inv_dt);
    }
  }
  if (!fits) return;
  __syncthreads();
  const int n1 = e1 - b1 + 1, n2 = e2 - b2 + 1;
  for (int e = threadIdx.x; e < 3 * n1 * n2; e += kB) {
    const int c = e / (n1 * n2), r = e - c * (n1 * n2);
    const int y = r / n1, x = r - y * n1;
    const double v = jw[c * kGWC + x + kGW1 * y];
    if (v != 0.0) unsafeAtomicAdd(&a.J.at(w1 + x, w2 + y, c), static_cast<float>(v));
  }
}

template <typename RR>
void launch_deposit(int kind, const ntt_ks_metric_t* metric, const GRDepArgs& a, hipStream_t st,
                    ntt_status* s) {
  KSConst<RR> k;
  *s = make_const<RR>(metric, &k);
  if (*s != NTT_OK) return;
  const unsigned g = grid_for(a.P.n, kB);
  if (kind == NTT_METRIC_KERR_SCHILD)
    k_gr_deposit<NTT_METRIC_KERR_SCHILD, RR><<<g, kB, 0, st>>>(a, KSMetric<RR, NTT_METRIC_KERR_SCHILD>{ k });
  else if (kind == NTT_METRIC_QKERR_SCHILD)
    k_gr_deposit<NTT_METRIC_QKERR_SCHILD, RR><<<g, kB, 0, st>>>(a, KSMetric<RR, NTT_METRIC_QKERR_SCHILD>{ k });
  else
    k_gr_deposit<NTT_METRIC_KERR_SCHILD_0, RR><<<g, kB, 0, st>>>(a, KSMetric<RR, NTT_METRIC_KERR_SCHILD_0>{ k });
}

bool metric_kind_ok(const ntt_ks_metric_t* m) {
  return m != nullptr && (m->kind == NTT_METRIC_KERR_SCHILD || m->kind == NTT_METRIC_QKERR_SCHILD ||
                          m->kind == NTT_METRIC_KERR_SCHILD_0);
}

// a 2D field with the reference's two ghost layers, matching the metric's grid
ntt_status view2(const ntt_field_t* f, int ncomp, const ntt_ks_metric_t* metric, FieldView* v) {
  ntt_status s = make_field_view(f, ncomp, v);
  if (s != NTT_OK) return s;
  if (f->dim != 2) return NTT_ERR_UNSUPPORTED;
  if (v->ngh != 2) return NTT_ERR_ARG;
  if (metric != nullptr && (v->n1 != metric->nx1 || v->n2 != metric->nx2)) return NTT_ERR_ARG;
  return NTT_OK;
}

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_gr_match_boundaries(ntt_field_t* em, const ntt_field_t* target,
                                              const ntt_ks_metric_t* metric,
                                              const ntt_flds_bc_t* bc, const ntt_gr_match_t* match,
                                              void* stream) {
  if (!metric_kind_ok(metric) || match == nullptr) return metric ? NTT_ERR_UNSUPPORTED : NTT_ERR_ARG;
  FieldView F, T;
  ntt_status s = view2(em, 6, metric, &F);
  if (s != NTT_OK) return s;
  if ((s = view2(target, 6, metric, &T)) != NTT_OK) return s;
  if (!(match->ds > 0.0) || (match->comps & ~0x3Fu) != 0) return NTT_ERR_ARG;
  if (match->i1_lo < 0 || match->i1_hi > F.N1 || match->i1_lo > match->i1_hi) return NTT_ERR_ARG;
  if (match->i1_lo == match->i1_hi) return NTT_OK;
  KSConst<R> k;
  if ((s = make_const<R>(metric, &k)) != NTT_OK) return s;
  MatchArgs a;
  a.r_edge = match->r_edge;
  a.four_inv_ds = R(4) / match->ds;
  a.i1_lo = match->i1_lo;
  a.i1_hi = match->i1_hi;
  a.G = F.ngh;
  a.N2 = F.N2;
  a.comps = match->comps;
  a.setD = (match->tags & kTagD) != 0;
  a.setB = (match->tags & kTagB) != 0;
  a.axis_min = bc != nullptr && bc->bc[1][0] == NTT_FLDS_BC_AXIS;
  a.axis_max = bc != nullptr && bc->bc[1][1] == NTT_FLDS_BC_AXIS;
  const unsigned g = blocks2(a.i1_lo, a.i1_hi, 0, a.N2);
  hipStream_t st = as_stream(stream);
  if (metric->kind == NTT_METRIC_KERR_SCHILD)
    k_match<NTT_METRIC_KERR_SCHILD><<<g, kB, 0, st>>>(bfld(F), bfld(T), KSMetric<R, NTT_METRIC_KERR_SCHILD>{ k }, a);
  else if (metric->kind == NTT_METRIC_QKERR_SCHILD)
    k_match<NTT_METRIC_QKERR_SCHILD><<<g, kB, 0, st>>>(bfld(F), bfld(T), KSMetric<R, NTT_METRIC_QKERR_SCHILD>{ k }, a);
  else
    k_match<NTT_METRIC_KERR_SCHILD_0><<<g, kB, 0, st>>>(bfld(F), bfld(T), KSMetric<R, NTT_METRIC_KERR_SCHILD_0>{ k }, a);
  return check_launch();
}

extern "C" ntt_status ntt_gr_horizon_boundaries(ntt_field_t* em, uint32_t tags, int32_t nfilter,
                                                void* stream) {
  FieldView F;
  ntt_status s = view2(em, 6, nullptr, &F);
  if (s != NTT_OK) return s;
  // the copied row G + 1 + nfilter must lie inside the x1 extent
  if (nfilter < 0 || F.ngh + 1 + nfilter >= F.N1) return NTT_ERR_ARG;
  const int setD = (tags & kTagD) != 0, setB = (tags & kTagB) != 0;
  if (!setD && !setB) return NTT_OK;
  k_horizon<<<grid_for(F.n2 + 1, kB), kB, 0, as_stream(stream)>>>(bfld(F), F.ngh, F.n2, nfilter,
                                                                   setD, setB);
  return check_launch();
}

extern "C" ntt_status ntt_gr_axis_boundaries(ntt_field_t* em, int32_t upper, uint32_t tags,
                                             void* stream) {
  FieldView F;
  ntt_status s = view2(em, 6, nullptr, &F);
  if (s != NTT_OK) return s;
  const int setD = (tags & kTagD) != 0, setB = (tags & kTagB) != 0;
  if (!setD && !setB) return NTT_OK;
  const int e = upper ? F.ngh + F.n2 : F.ngh;
  k_axis<<<grid_for(F.N1, kB), kB, 0, as_stream(stream)>>>(bfld(F), e, upper ? 1 : 0, setD, setB);
  return check_launch();
}

extern "C" ntt_status ntt_gr_absorb_currents(ntt_field_t* cur, const ntt_ks_metric_t* metric,
                                             double r_edge, double ds, int32_t i1_lo,
                                             int32_t i1_hi, void* stream) {
  if (!metric_kind_ok(metric)) return metric ? NTT_ERR_UNSUPPORTED : NTT_ERR_ARG;
  FieldView J;
  ntt_status s = view2(cur, 3, metric, &J);
  if (s != NTT_OK) return s;
  if (!(ds > 0.0) || i1_lo < 0 || i1_hi > J.N1 || i1_lo > i1_hi) return NTT_ERR_ARG;
  if (i1_lo == i1_hi) return NTT_OK;
  KSConst<R> k;
  if ((s = make_const<R>(metric, &k)) != NTT_OK) return s;
  const unsigned g = blocks2(i1_lo, i1_hi, 0, J.N2);
  const R inv_q = R(4) / ds;
  hipStream_t st = as_stream(stream);
  if (metric->kind == NTT_METRIC_KERR_SCHILD)
    k_absorb_j<NTT_METRIC_KERR_SCHILD><<<g, kB, 0, st>>>(bfld(J), KSMetric<R, NTT_METRIC_KERR_SCHILD>{ k }, r_edge, inv_q, i1_lo, i1_hi, J.ngh);
  else if (metric->kind == NTT_METRIC_QKERR_SCHILD)
    k_absorb_j<NTT_METRIC_QKERR_SCHILD><<<g, kB, 0, st>>>(bfld(J), KSMetric<R, NTT_METRIC_QKERR_SCHILD>{ k }, r_edge, inv_q, i1_lo, i1_hi, J.ngh);
  else
    k_absorb_j<NTT_METRIC_KERR_SCHILD_0><<<g, kB, 0, st>>>(bfld(J), KSMetric<R, NTT_METRIC_KERR_SCHILD_0>{ k }, r_edge, inv_q, i1_lo, i1_hi, J.ngh);
  return check_launch();
}

extern "C" ntt_status ntt_gr_digital_filter(ntt_field_t* cur, const ntt_field_t* buff,
                                            const ntt_flds_bc_t* bc, void* stream) {
  FieldView A, B;
  ntt_status s = view2(cur, 3, nullptr, &A);
  if (s != NTT_OK) return s;
  if ((s = view2(buff, 3, nullptr, &B)) != NTT_OK) return s;
  if (A.n1 != B.n1 || A.n2 != B.n2 || A.d == B.d) return NTT_ERR_ARG;
  const int amin = bc != nullptr && bc->bc[1][0] == NTT_FLDS_BC_AXIS;
  const int amax = bc != nullptr && bc->bc[1][1] == NTT_FLDS_BC_AXIS;
  const int G = A.ngh;
  k_filter_sph<<<blocks2(G, G + A.n1, G, G + A.n2 + 1), kB, 0, as_stream(stream)>>>(
    bfld(A), bfld(B), G, A.n1, A.n2, amin, amax);
  return check_launch();
}

extern "C" ntt_status ntt_gr_deposit_currents(const ntt_particles_t* prtls, ntt_field_t* cur,
                                              const ntt_ks_metric_t* metric, float charge,
                                              float dt, int32_t precision, void* stream) {
  if (!metric_kind_ok(metric)) return metric ? NTT_ERR_UNSUPPORTED : NTT_ERR_ARG;
  if (precision != 0 && precision != 32 && precision != 64) return NTT_ERR_ARG;
  if (!(dt > 0.0f)) return NTT_ERR_ARG;
  FieldView J;
  ntt_status s = view2(cur, 3, metric, &J);
  if (s != NTT_OK) return s;
  GRDepArgs a;
  if ((s = make_particle_view(prtls, 2, true, &a.P)) != NTT_OK) return s;
  if (a.P.i1p == nullptr || a.P.dx1p == nullptr || a.P.i2p == nullptr || a.P.dx2p == nullptr)
    return NTT_ERR_ARG;
  if (a.P.n == 0) return NTT_OK;
  a.J = bfld(J);
  a.G = J.ngh;
  a.charge = charge;
  a.inv_dt = 1.0f / dt;
  a.err = device_error_word();
  if (precision == 32) launch_deposit<float>(metric->kind, metric, a, as_stream(stream), &s);
  else launch_deposit<double>(metric->kind, metric, a, as_stream(stream), &s);
  if (s != NTT_OK) return s;
  return check_launch();
}
