// engine.hip -- host-side driver of one SRPIC step for a periodic box (the
// counterpart of SRPICEngine::step_forward, src/engines/srpic/srpic.hpp:65-188,
// with the drivers of src/engines/srpic/{fieldsolvers.h, currents.h,
// particle_pusher.h}).  All launches are enqueued on one stream.
#include <cmath>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace ntt;

namespace {

// launch-timing record of the fused push+deposit kernels
struct Profile {
  std::mutex                                  mu;
  bool                                        on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare;
  double                                      ms = 0.0;
  int64_t                                     launches = 0;
};
Profile g_prof;

bool prof_begin(hipStream_t st, std::pair<hipEvent_t, hipEvent_t>* ev) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  if (!g_prof.on) return false;
  if (!g_prof.spare.empty()) {
    *ev = g_prof.spare.back();
    g_prof.spare.pop_back();
  } else {
    if (hipEventCreate(&ev->first) != hipSuccess) return false;
    if (hipEventCreate(&ev->second) != hipSuccess) {
      (void)hipEventDestroy(ev->first);
      return false;
    }
  }
  if (hipEventRecord(ev->first, st) != hipSuccess) return false;
  return true;
}

void prof_end(hipStream_t st, const std::pair<hipEvent_t, hipEvent_t>& ev) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  if (hipEventRecord(ev.second, st) == hipSuccess) g_prof.pending.push_back(ev);
}

bool same_grid(const FieldView& a, const FieldView& b) {
  return a.n1 == b.n1 && a.n2 == b.n2 && a.n3 == b.n3 && a.ngh == b.ngh && a.N3 == b.N3;
}

bool fd_nonzero(const ntt_fd_coeffs_t& f) {
  return f.delta_x != 0 || f.delta_y != 0 || f.delta_z != 0 || f.beta_xy != 0 ||
         f.beta_yx != 0 || f.beta_xz != 0 || f.beta_zx != 0 || f.beta_yz != 0 || f.beta_zy != 0;
}

}  // namespace

extern "C" ntt_status ntt_faraday_mink(ntt_field_t*, float, float, const ntt_fd_coeffs_t*, void*);

extern "C" ntt_status ntt_sr_step_workspace_size(const ntt_field_t* cur, uint64_t* bytes) {
  if (bytes == nullptr) return NTT_ERR_ARG;
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  *bytes = static_cast<uint64_t>(J.cstride) * 3 * sizeof(unsigned long long);
  return NTT_OK;
}

extern "C" ntt_status ntt_sr_step(ntt_field_t* em, ntt_field_t* cur, ntt_field_t* buff,
                                  ntt_species_t* species, int32_t nspecies,
                                  const ntt_sr_step_params_t* prm, const ntt_minkowski_t* metric,
                                  void* workspace, uint64_t workspace_bytes, void* stream) {
  if (prm == nullptr || metric == nullptr || (nspecies > 0 && species == nullptr) || nspecies < 0)
    return NTT_ERR_ARG;
  FieldView E, J, Bf;
  ntt_status s = make_field_view(em, 6, &E);
  if (s != NTT_OK) return s;
  s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  s = make_field_view(buff, 3, &Bf);
  if (s != NTT_OK) return s;
  if (!same_grid(E, J) || !same_grid(E, Bf) || J.d == Bf.d) return NTT_ERR_ARG;
  const uint64_t ws_need = static_cast<uint64_t>(J.cstride) * 3 * sizeof(unsigned long long);
  if (prm != nullptr && prm->deposit_enable && (workspace == nullptr || workspace_bytes < ws_need))
    return NTT_ERR_CAPACITY;
  unsigned long long* acc = static_cast<unsigned long long*>(workspace);
  const int D = em->dim;
  if (metric->dim != D || !(metric->dx > 0.0f) || !(prm->dt > 0.0f)) return NTT_ERR_ARG;
  if (E.ngh != 2) return NTT_ERR_UNSUPPORTED;
  const int na[3] = { E.n1, E.n2, E.n3 };
  for (int d = 0; d < D; ++d) {
    if (na[d] < 2 * E.ngh) return NTT_ERR_ARG;
    if (prm->prtl_bc.bc[d][0] != NTT_PRTL_BC_PERIODIC || prm->prtl_bc.bc[d][1] != NTT_PRTL_BC_PERIODIC)
      return NTT_ERR_UNSUPPORTED;
  }
  if (prm->shape_order < 0 || prm->shape_order > 2 || prm->n_filters < 0) return NTT_ERR_ARG;
  hipStream_t st = as_stream(stream);
  const float dt = prm->dt;
  // dx = sqrt(h_11) as the drivers compute it (fieldsolvers.h:50,110)
  const float dxh = sqrtf(metric->dx * metric->dx);
  const bool  ext = fd_nonzero(prm->fd);

  // validate every species before enqueueing anything
  struct Prep {
    ParticleView pv;
    PrtlBCs      bcs;
    float        ndth;
    int          alg;
    bool         run, dep;
  };
  Prep prep[64];
  if (nspecies > 64) return NTT_ERR_UNSUPPORTED;
  for (int sp = 0; sp < nspecies; ++sp) {
    Prep& p = prep[sp];
    p.run = species[sp].pusher_flags != NTT_PUSHER_NONE && species[sp].prtls.npart > 0;
    p.dep = p.run && prm->deposit_enable && species[sp].charge != 0.0f;
    if (!p.run) continue;
    ntt_sr_pusher_ctx_t ctx{ species[sp].mass, species[sp].charge, dt, prm->omegaB0,
                             E.n1, D >= 2 ? E.n2 : 1, D >= 3 ? E.n3 : 1,
                             species[sp].pusher_flags };
    FieldView fv;
    s = prepare_push(&ctx, &prm->prtl_bc, &species[sp].prtls, em, metric, prm->shape_order,
                     &p.alg, &p.pv, &fv, &p.bcs, &p.ndth);
    if (s != NTT_OK) return s;
  }

  // srpic.hpp:71-82: fill E and B ghosts at step 0
  if (prm->step == 0) launch_ghost_fill(E, 0, 6, st);

  auto faraday_half = [&]() -> ntt_status {
    const float dT = HALF * prm->correction * dt;  // fieldsolvers.h:44-47
    float c1, c2;
    if (D == 2) { c1 = dT / (dxh * dxh); c2 = dT; }
    else { c1 = dT / dxh; c2 = ZERO; }
    if (!ext) {
      launch_faraday_periodic(E, c1, c2, st);
      return check_launch();
    }
    ntt_status r = ntt_faraday_mink(em, c1, c2, &prm->fd, stream);
    if (r != NTT_OK) return r;
    launch_ghost_fill(E, 3, 6, st);
    return check_launch();
  };

  if (prm->fieldsolver_enable) {
    s = faraday_half();
    if (s != NTT_OK) return s;
  }
  // currents.h:68: J = 0 -- here the fixed-point accumulator (a fill kernel
  // rather than a memset, so that the step can be captured in a graph)
  if (prm->deposit_enable) launch_zero_u64(acc, J.cstride * 3, st);
  // fixed-point scale of the accumulator, common to every species
  float qmax = 0.0f;
  for (int sp = 0; sp < nspecies; ++sp)
    if (prep[sp].dep) qmax = fmaxf(qmax, fabsf(species[sp].charge));
  const float wmax   = prm->weight_max > 0.0f ? prm->weight_max : 1.0f;
  const float fscale = fixed_point_scale(qmax * wmax, dt);
  for (int sp = 0; sp < nspecies; ++sp) {
    const Prep& p = prep[sp];
    if (!p.run) continue;
    const int ppt = prm->particles_per_thread > 0 ? prm->particles_per_thread : 4;
    std::pair<hipEvent_t, hipEvent_t> ev;
    const bool timed = prof_begin(st, &ev);
    s = launch_push_deposit_periodic(p.pv, E, J, acc, D, prm->shape_order, p.alg, p.ndth, dt,
                                     metric->dx, species[sp].charge, p.dep, fscale,
                                     species[sp].tile_offsets, species[sp].ntiles, ppt, st);
    if (timed) prof_end(st, ev);
    if (s != NTT_OK) return s;
  }
  if (prm->deposit_enable) {
    // folded current (active cells + periodic ghosts) from the accumulator
    launch_j_from_fixed(J, acc, 1.0 / static_cast<double>(fscale), st);
    // currents.h:89-119: n filter passes (ping-pong between cur and buff)
    if (prm->n_filters > 0) {
      const FieldView* src = &J;
      const FieldView* dst = &Bf;
      for (int f = 0; f < prm->n_filters; ++f) {
        launch_filter_periodic(*dst, *src, st);
        const FieldView* t = src;
        src = dst;
        dst = t;
      }
      if (src != &J) launch_copy(J.d, Bf.d, J.cstride * 3, st);
    }
  }
  if (prm->fieldsolver_enable) {
    s = faraday_half();
    if (s != NTT_OK) return s;
    const float dT = ONE * prm->correction * dt;
    float c1, c2;
    if (D == 2) { c1 = dT / (dxh * dxh); c2 = dT; }
    else { c1 = dT / dxh; c2 = ZERO; }
    // fieldsolvers.h:157: coeff = -dt * q0 / (B0 * V0)
    const float coeff = -dt * prm->q0 / (prm->B0 * prm->V0);
    launch_ampere_currents_periodic(E, J, c1, c2, coeff, prm->ppc0, prm->deposit_enable != 0, st);
  }
  return check_launch();
}

extern "C" ntt_status ntt_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  g_prof.on = on != 0;
  return NTT_OK;
}

extern "C" ntt_status ntt_profile_read(double* ms_total, int64_t* launches, int32_t reset) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  for (auto& ev : g_prof.pending) {
    if (hipEventSynchronize(ev.second) != hipSuccess) return NTT_ERR_HIP;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev.first, ev.second) != hipSuccess) return NTT_ERR_HIP;
    g_prof.ms += ms;
    g_prof.launches += 1;
    g_prof.spare.push_back(ev);
  }
  g_prof.pending.clear();
  if (ms_total) *ms_total = g_prof.ms;
  if (launches) *launches = g_prof.launches;
  if (reset) {
    g_prof.ms = 0.0;
    g_prof.launches = 0;
  }
  return NTT_OK;
}
