// sr_kernels.hip -- kernel-level SR pusher and current deposit over an
// arbitrary (unsorted) particle range, reading E/B from HBM and accumulating J
// with hardware float atomics at the reference's ghost-extended indices.
//
// These are the drop-ins for kernel::sr::Pusher_kernel and
// kernel::DepositCurrents_kernel; the fused, tile-windowed push+deposit of
// the engine step lives in push_deposit.hip.
#include "sr_particle.hpp"

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

template <int D, int O>
__global__ __launch_bounds__(kBlock) void k_deposit(DepositArgs a) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= a.P.n) return;
  if (a.P.tag[p] == NTT_TAG_DEAD) return;
  PState s;
  s.i[0] = a.P.i1[p];   s.ip[0] = a.P.i1p[p];
  s.dx[0] = a.P.dx1[p]; s.dxp[0] = a.P.dx1p[p];
  s.i[1] = a.P.i2[p];   s.ip[1] = a.P.i2p[p];
  s.dx[1] = a.P.dx2[p]; s.dxp[1] = a.P.dx2p[p];
  if constexpr (D == 3) {
    s.i[2] = a.P.i3[p];   s.ip[2] = a.P.i3p[p];
    s.dx[2] = a.P.dx3[p]; s.dxp[2] = a.P.dx3p[p];
  }
  if constexpr (D == 2) {
    s.u[0] = a.P.ux1[p];
    s.u[1] = a.P.ux2[p];
    s.u[2] = a.P.ux3[p];
  }
  // refuse particles whose stencil leaves the ghost-extended array
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int lo, hi;
    deposit_extent<O>(s.i[d], s.dx[d], s.ip[d], s.dxp[d], lo, hi);
    if (lo + a.ngh < 0 || hi + a.ngh >= a.N[d]) {
      raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
      return;
    }
  }
  const float coeff = a.P.w[p] * a.charge;
  deposit_particle<D, O>(a.J, a.ngh, s, coeff, a.inv_dt);
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
  const unsigned g = grid_for(pv.n, kBlock);
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
