// gr_fields.hip -- the GR field solver of a 2D axisymmetric (r, theta) box over
// the Kerr-Schild family of metrics, on gfx950: auxiliary E and H, Faraday,
// Ampere, the current source and the time averages of the GRPIC step.
//
// Reference: kernel::gr::Faraday_kernel (src/kernels/faraday_gr.hpp:32-106),
// Ampere_kernel and CurrentsAmpere_kernel (src/kernels/ampere_gr.hpp:33-181),
// ComputeAuxE_kernel, ComputeAuxH_kernel, TimeAverageDB_kernel and
// TimeAverageJ_kernel (src/kernels/aux_fields_gr.hpp:31-306), with the launch
// ranges of src/engines/grpic/fieldsolvers.h:53-266.
//
// Design: one thread per cell.  The metric is evaluated in FP64 at the
// staggered points a cell's update needs (the lapse, shift, lower metric and
// sqrt(det h) share one evaluation per point), on FP32 field storage.  The
// work is FP64 VALU with a few field reads per cell: these kernels run once
// or twice per step and are small next to the particle push.
//
// Ampere on the axis row reads D of the row above (D^1 on the axis takes
// the value next to it).  When Din and Dout are the same array the reference
// lets the two rows race; here the axis rows run in a launch of their own
// before the other rows, so they read the value from before the update.
#include "gr_metric.hpp"

namespace ntt {
namespace {

using R = double;

struct GRFld {
  float*  d;
  int32_t N1, N2;
  int64_t cs;
  __device__ __forceinline__ float& at(int i, int j, int c) const {
    return d[static_cast<int64_t>(i) + static_cast<int64_t>(N1) * j + cs * c];
  }
};

GRFld grfld(const FieldView& v) { return GRFld{ v.d, v.N1, v.N2, v.cstride }; }

// cell index range [lo, hi) per dimension (ghost-extended indices)
struct Range2 {
  int i1lo, i1hi, i2lo, i2hi;
};

__device__ __forceinline__ bool cell_of(const Range2& r, int& i1, int& i2) {
  const int w = r.i1hi - r.i1lo;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = static_cast<int64_t>(w) * (r.i2hi - r.i2lo);
  if (t >= n) return false;
  i1 = r.i1lo + static_cast<int>(t % w);
  i2 = r.i2lo + static_cast<int>(t / w);
  return true;
}

constexpr int kB = 256;
unsigned blocks_of(const Range2& r) {
  const int64_t n = static_cast<int64_t>(r.i1hi - r.i1lo) * (r.i2hi - r.i2lo);
  return static_cast<unsigned>((n + kB - 1) / kB);
}

// ComputeAuxE_kernel (aux_fields_gr.hpp:31-125): E from D and B
template <int KIND>
__global__ __launch_bounds__(kB) void k_aux_e(GRFld Df, GRFld Bf, GRFld Ef, KSMetric<R, KIND> M,
                                              Range2 rg, int G) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  const R x1 = i1 - G, x2 = i2 - G, H = 0.5;
  const auto p_m0 = M.fpoint(x1 - H, x2), p_p0 = M.fpoint(x1 + H, x2), p_00 = M.fpoint(x1, x2);
  const auto p_0p = M.fpoint(x1, x2 + H), p_mp = M.fpoint(x1 - H, x2 + H);
  const auto p_pp = M.fpoint(x1 + H, x2 + H), p_10 = M.fpoint(x1 + 1.0, x2);
  R ww = 2.0 * p_00.sdht;
  const R D1_half = (p_m0.sdht * p_m0.h_13 * Df.at(i1 - 1, i2, 0) +
                     p_p0.sdht * p_p0.h_13 * Df.at(i1, i2, 0)) / ww;
  const R B2_half = (p_m0.sdht * p_m0.sdh * p_m0.beta1 * Bf.at(i1 - 1, i2, 4) +
                     p_p0.sdht * p_p0.sdh * p_p0.beta1 * Bf.at(i1, i2, 4)) / ww;
  ww = 2.0 * p_0p.sdht;
  const R B3_half = (p_mp.sdht * p_mp.sdh * p_mp.beta1 * Bf.at(i1 - 1, i2, 5) +
                     p_pp.sdht * p_pp.sdh * p_pp.beta1 * Bf.at(i1, i2, 5)) / ww;
  ww = 2.0 * p_p0.sdht;
  const R D3_half = (p_00.sdht * p_00.h_13 * Df.at(i1, i2, 2) +
                     p_10.sdht * p_10.h_13 * Df.at(i1 + 1, i2, 2)) / ww;
  const R D1_cov = p_p0.h_11 * Df.at(i1, i2, 0) + D3_half;
  const R D2_cov = p_0p.h_22 * Df.at(i1, i2, 1);
  const R D3_cov = p_00.h_33 * Df.at(i1, i2, 2) + D1_half;
  Ef.at(i1, i2, 0) = static_cast<float>(p_p0.alpha * D1_cov);
  Ef.at(i1, i2, 1) = static_cast<float>(p_0p.alpha * D2_cov - B3_half);
  Ef.at(i1, i2, 2) = static_cast<float>(p_00.alpha * D3_cov + B2_half);
}

// ComputeAuxH_kernel (aux_fields_gr.hpp:136-234): H from D and B
template <int KIND>
__global__ __launch_bounds__(kB) void k_aux_h(GRFld Df, GRFld Bf, GRFld Hf, KSMetric<R, KIND> M,
                                              Range2 rg, int G) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  const R x1 = i1 - G, x2 = i2 - G, H = 0.5;
  const auto p_0p = M.fpoint(x1, x2 + H), p_1p = M.fpoint(x1 + 1.0, x2 + H);
  const auto p_pp = M.fpoint(x1 + H, x2 + H), p_00 = M.fpoint(x1, x2);
  const auto p_10 = M.fpoint(x1 + 1.0, x2), p_p0 = M.fpoint(x1 + H, x2);
  const auto p_mp = M.fpoint(x1 - H, x2 + H);
  R ww = 2.0 * p_pp.sdht;
  const R B1_half = (p_0p.sdht * p_0p.h_13 * Bf.at(i1, i2, 3) +
                     p_1p.sdht * p_1p.h_13 * Bf.at(i1 + 1, i2, 3)) / ww;
  const R D2_half = (p_0p.sdht * p_0p.sdh * p_0p.beta1 * Df.at(i1, i2, 1) +
                     p_1p.sdht * p_1p.sdh * p_1p.beta1 * Df.at(i1 + 1, i2, 1)) / ww;
  ww = 2.0 * p_p0.sdht;
  const R D3_half = (p_00.sdht * p_00.sdh * p_00.beta1 * Df.at(i1, i2, 2) +
                     p_10.sdht * p_10.sdh * p_10.beta1 * Df.at(i1 + 1, i2, 2)) / ww;
  ww = 2.0 * p_0p.sdht;
  const R B3_half = (p_mp.sdht * p_mp.h_13 * Bf.at(i1 - 1, i2, 5) +
                     p_pp.sdht * p_pp.h_13 * Bf.at(i1, i2, 5)) / ww;
  const R B1_cov = p_0p.h_11 * Bf.at(i1, i2, 3) + B3_half;
  const R B2_cov = p_p0.h_22 * Bf.at(i1, i2, 4);
  const R B3_cov = p_pp.h_33 * Bf.at(i1, i2, 5) + B1_half;
  Hf.at(i1, i2, 3) = static_cast<float>(p_0p.alpha * B1_cov);
  Hf.at(i1, i2, 4) = static_cast<float>(p_p0.alpha * B2_cov + D3_half);
  Hf.at(i1, i2, 5) = static_cast<float>(p_pp.alpha * B3_cov - D2_half);
}

// Faraday_kernel (faraday_gr.hpp:32-106): Bout = Bin - coeff curl E
template <int KIND>
__global__ __launch_bounds__(kB) void k_faraday(GRFld Bin, GRFld Bout, GRFld E,
                                                KSMetric<R, KIND> M, R coeff, Range2 rg, int G,
                                                int axis_i2min) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  const R x1 = i1 - G, x2 = i2 - G, H = 0.5;
  const R inv_0p = 1.0 / M.fpoint(x1, x2 + H).sdh;
  const R inv_pp = 1.0 / M.fpoint(x1 + H, x2 + H).sdh;
  const float b1 = Bin.at(i1, i2, 3), b2 = Bin.at(i1, i2, 4), b3 = Bin.at(i1, i2, 5);
  Bout.at(i1, i2, 3) = static_cast<float>(
    b1 + coeff * inv_0p * (R(E.at(i1, i2, 2)) - R(E.at(i1, i2 + 1, 2))));
  if (i2 == G && axis_i2min) {
    Bout.at(i1, i2, 4) = 0.0f;
  } else {
    const R inv_p0 = 1.0 / M.fpoint(x1 + H, x2).sdh;
    Bout.at(i1, i2, 4) = static_cast<float>(
      b2 + coeff * inv_p0 * (R(E.at(i1 + 1, i2, 2)) - R(E.at(i1, i2, 2))));
  }
  Bout.at(i1, i2, 5) = static_cast<float>(
    b3 + coeff * inv_pp *
           (R(E.at(i1, i2 + 1, 0)) - R(E.at(i1, i2, 0)) + R(E.at(i1, i2, 1)) -
            R(E.at(i1 + 1, i2, 1))));
}

// Ampere_kernel (ampere_gr.hpp:33-100): Dout = Din + coeff curl H
template <int KIND>
__global__ __launch_bounds__(kB) void k_ampere(GRFld Din, GRFld Dout, GRFld Hf,
                                               KSMetric<R, KIND> M, R coeff, Range2 rg, int G,
                                               int i2max, int axis_i2min, int axis_i2max) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  const R x1 = i1 - G, x2 = i2 - G, H = 0.5;
  if (i2 == G && axis_i2min) {
    // theta = 0
    const R inv_0p = 1.0 / M.fpoint(x1, H).sdh;
    const float d1 = Din.at(i1, i2 + 1, 0);
    const float d2 = Din.at(i1, i2, 1);
    Dout.at(i1, i2, 0) = d1;
    Dout.at(i1, i2, 1) = static_cast<float>(
      d2 + coeff * inv_0p * (R(Hf.at(i1 - 1, i2, 5)) - R(Hf.at(i1, i2, 5))));
  } else if (i2 == i2max && axis_i2max) {
    // theta = pi
    Dout.at(i1, i2, 0) = Din.at(i1, i2 - 1, 0);
  } else {
    const R inv_0p = 1.0 / M.fpoint(x1, x2 + H).sdh;
    const R inv_00 = 1.0 / M.fpoint(x1, x2).sdh;
    const R inv_p0 = 1.0 / M.fpoint(x1 + H, x2).sdh;
    const float d1 = Din.at(i1, i2, 0), d2 = Din.at(i1, i2, 1), d3 = Din.at(i1, i2, 2);
    Dout.at(i1, i2, 0) = static_cast<float>(
      d1 + coeff * inv_p0 * (R(Hf.at(i1, i2, 5)) - R(Hf.at(i1, i2 - 1, 5))));
    Dout.at(i1, i2, 1) = static_cast<float>(
      d2 + coeff * inv_0p * (R(Hf.at(i1 - 1, i2, 5)) - R(Hf.at(i1, i2, 5))));
    Dout.at(i1, i2, 2) = static_cast<float>(
      d3 + coeff * inv_00 *
             ((R(Hf.at(i1, i2 - 1, 3)) - R(Hf.at(i1, i2, 3))) +
              (R(Hf.at(i1, i2, 4)) - R(Hf.at(i1 - 1, i2, 4)))));
  }
}

// CurrentsAmpere_kernel (ampere_gr.hpp:105-178): D += coeff J / sqrt(det h)
template <int KIND>
__global__ __launch_bounds__(kB) void k_currents_ampere(GRFld Df, GRFld J, KSMetric<R, KIND> M,
                                                        R coeff, Range2 rg, int G, int i2max,
                                                        int axis_i2min, int axis_i2max) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  const R x1 = i1 - G, x2 = i2 - G, H = 0.5;
  if (i2 == G && axis_i2min) {
    const R inv_0p = 1.0 / M.fpoint(x1, x2 + H).sdh;
    Df.at(i1, i2, 0) = static_cast<float>(
      Df.at(i1, i2, 0) + R(J.at(i1, i2, 0)) * 0.5 * coeff / M.polar_area(x1 + H));
    Df.at(i1, i2, 1) = static_cast<float>(Df.at(i1, i2, 1) + R(J.at(i1, i2, 1)) * coeff * inv_0p);
  } else if (i2 == i2max && axis_i2max) {
    Df.at(i1, i2, 0) = static_cast<float>(
      Df.at(i1, i2, 0) + R(J.at(i1, i2, 0)) * 0.5 * coeff / M.polar_area(x1 + H));
  } else {
    const R inv_0p = 1.0 / M.fpoint(x1, x2 + H).sdh;
    const R inv_00 = 1.0 / M.fpoint(x1, x2).sdh;
    const R inv_p0 = 1.0 / M.fpoint(x1 + H, x2).sdh;
    Df.at(i1, i2, 0) = static_cast<float>(Df.at(i1, i2, 0) + R(J.at(i1, i2, 0)) * coeff * inv_p0);
    Df.at(i1, i2, 1) = static_cast<float>(Df.at(i1, i2, 1) + R(J.at(i1, i2, 1)) * coeff * inv_0p);
    Df.at(i1, i2, 2) = static_cast<float>(Df.at(i1, i2, 2) + R(J.at(i1, i2, 2)) * coeff * inv_00);
  }
}

// TimeAverageDB_kernel / TimeAverageJ_kernel (aux_fields_gr.hpp:240-303):
// dst = (dst + src) / 2 over components [c0, c1)
__global__ __launch_bounds__(kB) void k_time_average(GRFld dst, GRFld src, Range2 rg, int c0,
                                                     int c1) {
  int i1, i2;
  if (!cell_of(rg, i1, i2)) return;
  for (int c = c0; c < c1; ++c)
    dst.at(i1, i2, c) = 0.5f * (dst.at(i1, i2, c) + src.at(i1, i2, c));
}

struct GRSetup {
  FieldView             a, b, c;
  KSConst<R>            k;
  int                   kind, G, n1, n2;
  bool                  axis_min, axis_max;
};

ntt_status setup(const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc, GRSetup* S) {
  if (metric == nullptr) return NTT_ERR_ARG;
  if (metric->kind != NTT_METRIC_KERR_SCHILD && metric->kind != NTT_METRIC_QKERR_SCHILD &&
      metric->kind != NTT_METRIC_KERR_SCHILD_0)
    return NTT_ERR_UNSUPPORTED;
  ntt_status s = make_const<R>(metric, &S->k);
  if (s != NTT_OK) return s;
  S->kind     = metric->kind;
  S->axis_min = bc != nullptr && bc->bc[1][0] == NTT_FLDS_BC_AXIS;
  S->axis_max = bc != nullptr && bc->bc[1][1] == NTT_FLDS_BC_AXIS;
  return NTT_OK;
}

ntt_status view2d(const ntt_field_t* f, int ncomp, const ntt_ks_metric_t* metric, FieldView* v) {
  ntt_status s = make_field_view(f, ncomp, v);
  if (s != NTT_OK) return s;
  if (f->dim != 2) return NTT_ERR_UNSUPPORTED;
  if (v->ngh < 2) return NTT_ERR_ARG;
  if (metric != nullptr && (v->n1 != metric->nx1 || v->n2 != metric->nx2)) return NTT_ERR_ARG;
  return NTT_OK;
}

bool same(const FieldView& a, const FieldView& b) {
  return a.n1 == b.n1 && a.n2 == b.n2 && a.ngh == b.ngh;
}

template <template <int> class Launch, class... Args>
void by_kind(int kind, const KSConst<R>& k, Args&&... args) {
  if (kind == NTT_METRIC_KERR_SCHILD)
    Launch<NTT_METRIC_KERR_SCHILD>::run(KSMetric<R, NTT_METRIC_KERR_SCHILD>{ k }, args...);
  else if (kind == NTT_METRIC_QKERR_SCHILD)
    Launch<NTT_METRIC_QKERR_SCHILD>::run(KSMetric<R, NTT_METRIC_QKERR_SCHILD>{ k }, args...);
  else
    Launch<NTT_METRIC_KERR_SCHILD_0>::run(KSMetric<R, NTT_METRIC_KERR_SCHILD_0>{ k }, args...);
}

template <int KIND> struct LAuxE {
  static void run(KSMetric<R, KIND> M, GRFld D, GRFld B, GRFld E, Range2 rg, int G,
                  hipStream_t st) {
    k_aux_e<KIND><<<blocks_of(rg), kB, 0, st>>>(D, B, E, M, rg, G);
  }
};
template <int KIND> struct LAuxH {
  static void run(KSMetric<R, KIND> M, GRFld D, GRFld B, GRFld H, Range2 rg, int G,
                  hipStream_t st) {
    k_aux_h<KIND><<<blocks_of(rg), kB, 0, st>>>(D, B, H, M, rg, G);
  }
};
template <int KIND> struct LFaraday {
  static void run(KSMetric<R, KIND> M, GRFld Bin, GRFld Bout, GRFld E, R coeff, Range2 rg, int G,
                  int axis_min, hipStream_t st) {
    k_faraday<KIND><<<blocks_of(rg), kB, 0, st>>>(Bin, Bout, E, M, coeff, rg, G, axis_min);
  }
};
template <int KIND> struct LAmpere {
  static void run(KSMetric<R, KIND> M, GRFld Din, GRFld Dout, GRFld H, R coeff, Range2 rg, int G,
                  int i2max, int amin, int amax, hipStream_t st) {
    k_ampere<KIND><<<blocks_of(rg), kB, 0, st>>>(Din, Dout, H, M, coeff, rg, G, i2max, amin,
                                                 amax);
  }
};
template <int KIND> struct LCurAmpere {
  static void run(KSMetric<R, KIND> M, GRFld D, GRFld J, R coeff, Range2 rg, int G, int i2max,
                  int amin, int amax, hipStream_t st) {
    k_currents_ampere<KIND><<<blocks_of(rg), kB, 0, st>>>(D, J, M, coeff, rg, G, i2max, amin,
                                                          amax);
  }
};

// range_with_axis_BCs (fieldsolvers.h:53-75): one more cell below in x1, and
// one more row in x2 when the upper theta boundary is the axis
Range2 aux_range(const FieldView& v, bool axis_max) {
  const int G = v.ngh;
  return Range2{ G - 1, G + v.n1, G, G + v.n2 + (axis_max ? 1 : 0) };
}

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_gr_aux_e(const ntt_field_t* D, const ntt_field_t* B, ntt_field_t* E,
                                   const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc,
                                   void* stream) {
  GRSetup S;
  ntt_status s = setup(metric, bc, &S);
  if (s == NTT_OK) s = view2d(D, 6, metric, &S.a);
  if (s == NTT_OK) s = view2d(B, 6, metric, &S.b);
  if (s == NTT_OK) s = view2d(E, 6, metric, &S.c);
  if (s != NTT_OK) return s;
  if (!same(S.a, S.b) || !same(S.a, S.c) || S.c.d == S.a.d || S.c.d == S.b.d) return NTT_ERR_ARG;
  const Range2 rg = aux_range(S.a, S.axis_max);
  by_kind<LAuxE>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), rg, S.a.ngh, as_stream(stream));
  return check_launch();
}

extern "C" ntt_status ntt_gr_aux_h(const ntt_field_t* D, const ntt_field_t* B, ntt_field_t* H,
                                   const ntt_ks_metric_t* metric, const ntt_flds_bc_t* bc,
                                   void* stream) {
  GRSetup S;
  ntt_status s = setup(metric, bc, &S);
  if (s == NTT_OK) s = view2d(D, 6, metric, &S.a);
  if (s == NTT_OK) s = view2d(B, 6, metric, &S.b);
  if (s == NTT_OK) s = view2d(H, 6, metric, &S.c);
  if (s != NTT_OK) return s;
  if (!same(S.a, S.b) || !same(S.a, S.c) || S.c.d == S.a.d || S.c.d == S.b.d) return NTT_ERR_ARG;
  const Range2 rg = aux_range(S.a, S.axis_max);
  by_kind<LAuxH>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), rg, S.a.ngh, as_stream(stream));
  return check_launch();
}

extern "C" ntt_status ntt_gr_faraday(const ntt_field_t* Bin, ntt_field_t* Bout,
                                     const ntt_field_t* E, const ntt_ks_metric_t* metric,
                                     float coeff, const ntt_flds_bc_t* bc, void* stream) {
  GRSetup S;
  ntt_status s = setup(metric, bc, &S);
  if (s == NTT_OK) s = view2d(Bin, 6, metric, &S.a);
  if (s == NTT_OK) s = view2d(Bout, 6, metric, &S.b);
  if (s == NTT_OK) s = view2d(E, 6, metric, &S.c);
  if (s != NTT_OK) return s;
  if (!same(S.a, S.b) || !same(S.a, S.c) || S.c.d == S.b.d) return NTT_ERR_ARG;
  const int    G  = S.a.ngh;
  const Range2 rg{ G, G + S.a.n1, G, G + S.a.n2 };
  by_kind<LFaraday>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), static_cast<R>(coeff), rg,
                    G, static_cast<int>(S.axis_min), as_stream(stream));
  return check_launch();
}

extern "C" ntt_status ntt_gr_ampere(const ntt_field_t* Din, ntt_field_t* Dout,
                                    const ntt_field_t* H, const ntt_ks_metric_t* metric,
                                    float coeff, const ntt_flds_bc_t* bc, void* stream) {
  GRSetup S;
  ntt_status s = setup(metric, bc, &S);
  if (s == NTT_OK) s = view2d(Din, 6, metric, &S.a);
  if (s == NTT_OK) s = view2d(Dout, 6, metric, &S.b);
  if (s == NTT_OK) s = view2d(H, 6, metric, &S.c);
  if (s != NTT_OK) return s;
  if (!same(S.a, S.b) || !same(S.a, S.c) || S.c.d == S.b.d) return NTT_ERR_ARG;
  const int G = S.a.ngh, i2max = G + S.a.n2;
  hipStream_t st = as_stream(stream);
  // fieldsolvers.h:168-170: x2 over [i2min, i2max] (one row past the last
  // active row); the axis rows first, then the rest
  if (S.axis_min)
    by_kind<LAmpere>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), static_cast<R>(coeff),
                     Range2{ G, G + S.a.n1, G, G + 1 }, G, i2max, 1, 0, st);
  if (S.axis_max)
    by_kind<LAmpere>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), static_cast<R>(coeff),
                     Range2{ G, G + S.a.n1, i2max, i2max + 1 }, G, i2max, 0, 1, st);
  const Range2 rest{ G, G + S.a.n1, G + (S.axis_min ? 1 : 0), i2max + (S.axis_max ? 0 : 1) };
  if (rest.i2hi > rest.i2lo)
    by_kind<LAmpere>(S.kind, S.k, grfld(S.a), grfld(S.b), grfld(S.c), static_cast<R>(coeff), rest,
                     G, i2max, 0, 0, st);
  return check_launch();
}

extern "C" ntt_status ntt_gr_currents_ampere(ntt_field_t* D, const ntt_field_t* J,
                                             const ntt_ks_metric_t* metric, float coeff,
                                             const ntt_flds_bc_t* bc, void* stream) {
  GRSetup S;
  ntt_status s = setup(metric, bc, &S);
  if (s == NTT_OK) s = view2d(D, 6, metric, &S.a);
  if (s == NTT_OK) s = view2d(J, 3, metric, &S.b);
  if (s != NTT_OK) return s;
  if (!same(S.a, S.b)) return NTT_ERR_ARG;
  const int G = S.a.ngh, i2max = G + S.a.n2;
  by_kind<LCurAmpere>(S.kind, S.k, grfld(S.a), grfld(S.b), static_cast<R>(coeff),
                      Range2{ G, G + S.a.n1, G, i2max + 1 }, G, i2max,
                      static_cast<int>(S.axis_min), static_cast<int>(S.axis_max),
                      as_stream(stream));
  return check_launch();
}

namespace {
ntt_status time_average(ntt_field_t* dst, const ntt_field_t* src, int ncomp, void* stream) {
  FieldView a, b;
  ntt_status s = view2d(dst, ncomp, nullptr, &a);
  if (s == NTT_OK) s = view2d(src, ncomp, nullptr, &b);
  if (s != NTT_OK) return s;
  if (!same(a, b) || a.d == b.d) return NTT_ERR_ARG;
  const int    G = a.ngh;
  const Range2 rg{ G, G + a.n1, G, G + a.n2 };
  k_time_average<<<blocks_of(rg), kB, 0, as_stream(stream)>>>(grfld(a), grfld(b), rg, 0, ncomp);
  return check_launch();
}
}  // namespace

extern "C" ntt_status ntt_gr_time_average_db(const ntt_field_t* BD, ntt_field_t* BD0,
                                             void* stream) {
  return time_average(BD0, BD, 6, stream);
}

extern "C" ntt_status ntt_gr_time_average_j(ntt_field_t* J, const ntt_field_t* J0, void* stream) {
  return time_average(J, J0, 3, stream);
}
