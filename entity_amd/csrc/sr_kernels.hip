// sr_kernels.hip -- kernel-level SR pusher and current deposit over an
// arbitrary (unsorted) particle range, reading E/B from HBM and accumulating J
// at the reference's ghost-extended indices: per workgroup in an LDS window of
// doubles when its particles are spatially close (then one FP32 atomic per
// touched entry), else with FP32 atomics straight into J.
//
// These are the drop-ins for kernel::sr::Pusher_kernel and
// kernel::DepositCurrents_kernel; the fused, tile-windowed push+deposit of
// the engine step lives in push_deposit.hip.
#include "sr_particle.hpp"

#include <climits>

namespace ntt {
namespace {

constexpr int kBlock = 256;

struct PushArgs {
  ParticleView P;
  GlobalEM     F;
  int32_t      ngh, N[3];
  float        ndth, dt, dxm;
  PrtlBCs      bc;
  uint32_t*    err;
};

template <int D, int O, int ALG>
__global__ __launch_bounds__(kBlock) void k_sr_push(PushArgs a) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= a.P.n) return;
  const int16_t tag = a.P.tag[p];
  if (tag != NTT_TAG_ALIVE) {
    if (tag != NTT_TAG_DEAD) raise_device_error(a.err, NTT_DEVERR_INVALID_TAG);
    return;
  }
  PState s;
  s.tag  = tag;
  s.i[0] = a.P.i1[p];
  s.dx[0] = a.P.dx1[p];
  s.i[1] = a.P.i2[p];
  s.dx[1] = a.P.dx2[p];
  if constexpr (D == 3) {
    s.i[2]  = a.P.i3[p];
    s.dx[2] = a.P.dx3[p];
  }
  s.u[0] = a.P.ux1[p];
  s.u[1] = a.P.ux2[p];
  s.u[2] = a.P.ux3[p];
  if constexpr (ALG != ALG_PHOTON) {
    int lo, hi;
    interp_reach<O>(lo, hi);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int g = s.i[d] + a.ngh;
      // offsets outside [0, 1] (or NaN) would move the order-0 stencil
      if (g + lo < 0 || g + hi >= a.N[d] || !(s.dx[d] >= ZERO && s.dx[d] <= ONE)) {
        raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
        return;
      }
    }
  }
  sr_update<D, O, ALG>(a.F, a.ngh, a.ndth, a.dt, a.dxm, a.bc, s);
  a.P.i1p[p] = s.ip[0];
  a.P.dx1p[p] = s.dxp[0];
  a.P.i1[p]  = s.i[0];
  a.P.dx1[p] = s.dx[0];
  a.P.i2p[p] = s.ip[1];
  a.P.dx2p[p] = s.dxp[1];
  a.P.i2[p]  = s.i[1];
  a.P.dx2[p] = s.dx[1];
  if constexpr (D == 3) {
    a.P.i3p[p]  = s.ip[2];
    a.P.dx3p[p] = s.dxp[2];
    a.P.i3[p]   = s.i[2];
    a.P.dx3[p]  = s.dx[2];
  }
  a.P.ux1[p] = s.u[0];
  a.P.ux2[p] = s.u[1];
  a.P.ux3[p] = s.u[2];
  if (s.tag != tag) a.P.tag[p] = s.tag;
}

// J sink: hardware float atomics into the ghost-extended current array
struct GlobalJ {
  float*  d;
  int32_t N1, N2;
  int64_t cstride;
  int     i0 = 0, j0 = 0, k0 = 0;
  __device__ __forceinline__ void set_origin(int i, int j, int k) {
    i0 = i;
    j0 = j;
    k0 = k;
  }
  __device__ __forceinline__ void add_rel(int a, int b, int g, int c, float v) {
    add(i0 + a, j0 + b, k0 + g, c, v);
  }
  __device__ __forceinline__ void add(int i, int j, int k, int c, float v) {
    unsafeAtomicAdd(d + static_cast<int64_t>(i) +
                      static_cast<int64_t>(N1) * (static_cast<int64_t>(j) + static_cast<int64_t>(N2) * k) +
                      cstride * c,
                    v);
  }
};

struct DepositArgs {
  ParticleView P;
  GlobalJ      J;
  int32_t      ngh, N[3];
  float        charge, inv_dt;
  uint32_t*    err;
};

// LDS window of the deposit (cells per dimension): a workgroup's 256
// consecutive particles, when kept in spatial order (the reference sorts
// them periodically), deposit into it with FP64 LDS atomics; the window is
// then added to J with one FP32 atomic per non-zero entry.
template <int D> struct DepWin;
template <> struct DepWin<2> { static constexpr int W1 = 64, W2 = 16, W3 = 1; };
template <> struct DepWin<3> { static constexpr int W1 = 14, W2 = 10, W3 = 10; };

template <int D>
struct LdsJd {
  static constexpr int W1 = DepWin<D>::W1, W2 = DepWin<D>::W2, WC = W1 * W2 * DepWin<D>::W3;
  double* w;
  int     o1, o2, o3;
  int     base = 0;
  __device__ __forceinline__ int index(int i, int j, int k) const {
    return (i - o1) + W1 * ((j - o2) + W2 * (k - o3));
  }
  __device__ __forceinline__ void set_origin(int i, int j, int k) { base = index(i, j, k); }
  __device__ __forceinline__ void add_rel(int a, int b, int g, int c, float v) {
    atomicAdd(w + c * WC + base + a + W1 * (b + W2 * g), static_cast<double>(v));
  }
  __device__ __forceinline__ void add(int i, int j, int k, int c, float v) {
    atomicAdd(w + c * WC + index(i, j, k), static_cast<double>(v));
  }
};

template <bool MAX>
__device__ __forceinline__ int block_reduce(int v, int* red) {
  auto op = [](int x, int y) { return MAX ? max(x, y) : min(x, y); };
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = red[0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) r = op(r, red[w]);
  __syncthreads();
  return r;
}

// particles per thread of the deposit: a workgroup takes kBlock * kDepPPT
// consecutive particles, so the window's clearing and flush are paid once
// per 1024 particles
constexpr int kDepPPT = 4;

template <int D, int O>
__device__ __forceinline__ bool deposit_state(const DepositArgs& a, int64_t p, PState& s,
                                              int* lo, int* hi) {
  if (p >= a.P.n || a.P.tag[p] == NTT_TAG_DEAD) return false;
  s.i[0] = a.P.i1[p];   s.ip[0] = a.P.i1p[p];
  s.dx[0] = a.P.dx1[p]; s.dxp[0] = a.P.dx1p[p];
  s.i[1] = a.P.i2[p];   s.ip[1] = a.P.i2p[p];
  s.dx[1] = a.P.dx2[p]; s.dxp[1] = a.P.dx2p[p];
  if constexpr (D == 3) {
    s.i[2] = a.P.i3[p];   s.ip[2] = a.P.i3p[p];
    s.dx[2] = a.P.dx3[p]; s.dxp[2] = a.P.dx3p[p];
  }
  // particles whose stencil leaves the ghost-extended array are refused
#pragma unroll
  for (int d = 0; d < D; ++d) {
    deposit_extent<O>(s.i[d], s.dx[d], s.ip[d], s.dxp[d], lo[d], hi[d]);
    if (lo[d] + a.ngh < 0 || hi[d] + a.ngh >= a.N[d]) return false;
  }
  return true;
}

template <int D, int O>
__global__ __launch_bounds__(kBlock) void k_deposit(DepositArgs a) {
  constexpr int W1 = DepWin<D>::W1, W2 = DepWin<D>::W2, W3 = DepWin<D>::W3;
  __shared__ double jw[3 * W1 * W2 * W3];
  __shared__ int    red[kBlock / 64];
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * (kBlock * kDepPPT) + threadIdx.x;
  // the workgroup's box of deposit cells, and whether it fits the window
  int blo[3] = { INT_MAX, INT_MAX, INT_MAX }, bhi[3] = { INT_MIN, INT_MIN, INT_MIN };
#pragma unroll 1
  for (int q = 0; q < kDepPPT; ++q) {
    const int64_t p = p0 + static_cast<int64_t>(q) * kBlock;
    PState s;
    int lo[3], hi[3];
    if (deposit_state<D, O>(a, p, s, lo, hi)) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        blo[d] = min(blo[d], lo[d]);
        bhi[d] = max(bhi[d], hi[d]);
      }
    } else if (p < a.P.n && a.P.tag[p] != NTT_TAG_DEAD) {
      raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
    }
  }
  bool fits = true;
  int  b[3] = { 0, 0, 0 }, e[3] = { 0, 0, 0 };
  const int Wd[3] = { W1, W2, W3 };
#pragma unroll
  for (int d = 0; d < D; ++d) {
    b[d] = block_reduce<false>(blo[d], red);
    e[d] = block_reduce<true>(bhi[d], red);
    if (b[d] > e[d]) return;   // no live particle (uniform across the workgroup)
    if (e[d] - b[d] >= Wd[d]) fits = false;
  }
  const int o1 = b[0] + a.ngh, o2 = b[1] + a.ngh, o3 = (D == 3) ? b[2] + a.ngh : 0;
  if (fits) {
    for (int k = threadIdx.x; k < 3 * W1 * W2 * W3; k += kBlock) jw[k] = 0.0;
    __syncthreads();
  }
#pragma unroll 1
  for (int q = 0; q < kDepPPT; ++q) {
    const int64_t p = p0 + static_cast<int64_t>(q) * kBlock;
    PState s;
    int lo[3], hi[3];
    if (!deposit_state<D, O>(a, p, s, lo, hi)) continue;
    if constexpr (D == 2) {
      s.u[0] = a.P.ux1[p];
      s.u[1] = a.P.ux2[p];
      s.u[2] = a.P.ux3[p];
    }
    const float coeff = a.P.w[p] * a.charge;
    if (fits) {
      LdsJd<D> J{ jw, o1, o2, o3 };
      deposit_particle<D, O>(J, a.ngh, s, coeff, a.inv_dt);
    } else {
      GlobalJ J = a.J;
      deposit_particle<D, O>(J, a.ngh, s, coeff, a.inv_dt);
    }
  }
  if (!fits) return;
  __syncthreads();
  const int n1 = e[0] - b[0] + 1, n2 = e[1] - b[1] + 1, n3 = (D == 3) ? e[2] - b[2] + 1 : 1;
  const int nb = n1 * n2 * n3;
  for (int k = threadIdx.x; k < 3 * nb; k += kBlock) {
    const int c = k / nb, r = k - c * nb;
    const int z = r / (n1 * n2), r2 = r - z * (n1 * n2);
    const int y = r2 / n1, x = r2 - y * n1;
    const double v = jw[c * (W1 * W2 * W3) + x + W1 * (y + W2 * z)];
    if (v != 0.0) a.J.add(o1 + x, o2 + y, o3 + z, c, static_cast<float>(v));
  }
}

template <int D, int O, int ALG>
void launch_push(const PushArgs& a, hipStream_t st) {
  k_sr_push<D, O, ALG><<<grid_for(a.P.n, kBlock), kBlock, 0, st>>>(a);
}

template <int D, int ALG>
void launch_push_o(int order, const PushArgs& a, hipStream_t st) {
  if (order == 0) launch_push<D, 0, ALG>(a, st);
  else if (order == 1) launch_push<D, 1, ALG>(a, st);
  else launch_push<D, 2, ALG>(a, st);
}

}  // namespace

// Shared validation of pusher inputs, also used by the engine step.
ntt_status prepare_push(const ntt_sr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                        const ntt_particles_t* prtls, const ntt_field_t* em,
                        const ntt_minkowski_t* metric, int32_t shape_order, int* alg_out,
                        ParticleView* pv, FieldView* fv, PrtlBCs* bcs, float* ndth) {
  if (ctx == nullptr || bc == nullptr || metric == nullptr) return NTT_ERR_ARG;
  if (shape_order < 0 || shape_order > 2) return NTT_ERR_UNSUPPORTED;
  ntt_status s = make_field_view(em, 6, fv);
  if (s != NTT_OK) return s;
  const int D = em->dim;
  if (metric->dim != D || !(metric->dx > 0.0f)) return NTT_ERR_ARG;
  if (fv->ngh < 2) return NTT_ERR_ARG;
  const uint8_t fl = ctx->pusher_flags;
  if (fl == NTT_PUSHER_NONE) return NTT_ERR_ARG;  // "No particle pusher specified"
  if (fl & NTT_PUSHER_GCA) return NTT_ERR_UNSUPPORTED;
  int alg;
  if (fl == NTT_PUSHER_PHOTON) alg = ALG_PHOTON;
  else if (fl & NTT_PUSHER_BORIS) alg = ALG_BORIS;
  else if (fl & NTT_PUSHER_VAY) alg = ALG_VAY;
  else return NTT_ERR_ARG;
  if (alg != ALG_PHOTON && !(ctx->mass != 0.0f)) return NTT_ERR_ARG;
  s = make_particle_view(prtls, D, true, pv);
  if (s != NTT_OK) return s;
  const int ni[3] = { ctx->ni1, ctx->ni2, ctx->ni3 };
  const int na[3] = { fv->n1, fv->n2, fv->n3 };
  for (int d = 0; d < D; ++d) {
    if (ni[d] != na[d]) return NTT_ERR_ARG;
    for (int e = 0; e < 2; ++e) {
      const uint8_t b = bc->bc[d][e];
      const bool ok = b == NTT_PRTL_BC_PERIODIC || b == NTT_PRTL_BC_ABSORB ||
                      b == NTT_PRTL_BC_ATMOSPHERE || b == NTT_PRTL_BC_REFLECT ||
                      (d == 1 && b == NTT_PRTL_BC_AXIS);
      if (!ok) return NTT_ERR_UNSUPPORTED;
    }
    bcs->lo[d] = bc->bc[d][0];
    bcs->hi[d] = bc->bc[d][1];
    bcs->ni[d] = ni[d];
  }
  for (int d = D; d < 3; ++d) {
    bcs->lo[d] = bcs->hi[d] = 0;
    bcs->ni[d] = 1;
  }
  // normalized_dt_half (sr.hpp:111): HALF * (charge / mass) * omegaB0 * dt
  *ndth = HALF * (ctx->charge / ctx->mass) * ctx->omegaB0 * ctx->dt;
  *alg_out = alg;
  return NTT_OK;
}

ntt_status launch_sr_push(const ParticleView& pv, const FieldView& fv, const PrtlBCs& bcs,
                          int D, int order, int alg, float ndth, float dt, float dxm,
                          hipStream_t st) {
  if (pv.n == 0) return NTT_OK;
  uint32_t* err = device_error_word();
  if (err == nullptr) return NTT_ERR_HIP;
  PushArgs a;
  a.P = pv;
  a.F = GlobalEM{ fv.d, fv.N1, fv.N2, fv.cstride };
  a.ngh = fv.ngh;
  a.N[0] = fv.N1; a.N[1] = fv.N2; a.N[2] = fv.N3;
  a.ndth = ndth; a.dt = dt; a.dxm = dxm;
  a.bc = bcs;
  a.err = err;
  if (D == 2) {
    if (alg == ALG_BORIS) launch_push_o<2, ALG_BORIS>(order, a, st);
    else if (alg == ALG_VAY) launch_push_o<2, ALG_VAY>(order, a, st);
    else launch_push<2, 0, ALG_PHOTON>(a, st);
  } else {
    if (alg == ALG_BORIS) launch_push_o<3, ALG_BORIS>(order, a, st);
    else if (alg == ALG_VAY) launch_push_o<3, ALG_VAY>(order, a, st);
    else launch_push<3, 0, ALG_PHOTON>(a, st);
  }
  return check_launch();
}

ntt_status launch_deposit(const ParticleView& pv, const FieldView& J, int D, int order,
                          float charge, float dt, hipStream_t st) {
  if (pv.n == 0) return NTT_OK;
  uint32_t* err = device_error_word();
  if (err == nullptr) return NTT_ERR_HIP;
  DepositArgs a;
  a.P = pv;
  a.J = GlobalJ{ J.d, J.N1, J.N2, J.cstride };
  a.ngh = J.ngh;
  a.N[0] = J.N1; a.N[1] = J.N2; a.N[2] = J.N3;
  a.charge = charge;
  a.inv_dt = ONE / dt;
  a.err = err;
  const unsigned g = grid_for(pv.n, kBlock * kDepPPT);
  if (D == 2) {
    if (order == 0) k_deposit<2, 0><<<g, kBlock, 0, st>>>(a);
    else if (order == 1) k_deposit<2, 1><<<g, kBlock, 0, st>>>(a);
    else k_deposit<2, 2><<<g, kBlock, 0, st>>>(a);
  } else {
    if (order == 0) k_deposit<3, 0><<<g, kBlock, 0, st>>>(a);
    else if (order == 1) k_deposit<3, 1><<<g, kBlock, 0, st>>>(a);
    else k_deposit<3, 2><<<g, kBlock, 0, st>>>(a);
  }
  return check_launch();
}

}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_sr_push(const ntt_sr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                                  const ntt_particles_t* prtls, const ntt_field_t* em,
                                  const ntt_minkowski_t* metric, int32_t shape_order,
                                  void* stream) {
  ParticleView pv;
  FieldView    fv;
  PrtlBCs      bcs;
  float        ndth;
  int          alg;
  ntt_status s = prepare_push(ctx, bc, prtls, em, metric, shape_order, &alg, &pv, &fv, &bcs, &ndth);
  if (s != NTT_OK) return s;
  return launch_sr_push(pv, fv, bcs, em->dim, shape_order, alg, ndth, ctx->dt, metric->dx,
                        as_stream(stream));
}

extern "C" ntt_status ntt_deposit_currents(const ntt_particles_t* prtls, ntt_field_t* cur,
                                           const ntt_minkowski_t* metric, float charge, float dt,
                                           int32_t shape_order, void* stream) {
  if (metric == nullptr) return NTT_ERR_ARG;
  if (shape_order < 0 || shape_order > 2) return NTT_ERR_UNSUPPORTED;
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (metric->dim != cur->dim || !(dt > 0.0f)) return NTT_ERR_ARG;
  if (J.ngh < 2) return NTT_ERR_ARG;
  ParticleView pv;
  s = make_particle_view(prtls, cur->dim, cur->dim == 2, &pv);
  if (s != NTT_OK) return s;
  return launch_deposit(pv, J, cur->dim, shape_order, charge, dt, as_stream(stream));
}
