// engine.hip -- host-side driver of one SRPIC step for a periodic box (the
// counterpart of SRPICEngine::step_forward, src/engines/srpic/srpic.hpp:65-188,
// with the drivers of src/engines/srpic/{fieldsolvers.h, currents.h,
// particle_pusher.h}).  All launches are enqueued on one stream.
#include <algorithm>
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

namespace {

// Workspace of ntt_sr_step: the fixed-point current accumulator (the layout
// of `cur`, 64-bit entries), then, in 3D, the chunk flags of the split push
// and deposit launches for up to one chunk per cell plus the tail (a sort
// tile holds at least one cell), then the push's per-chunk box records for
// tiles of 16 cells or more (smaller tiles run without them): one set shared
// by the species, or one set per species when the workspace holds them
// (ntt_sr_step_workspace_size_species), which lets each species' push start
// from the boxes its previous push recorded.
struct StepWorkspace {
  uint64_t acc_bytes, flag_bytes, rec_bytes;
  int64_t  flag_chunks, rec_chunks;
};
StepWorkspace step_workspace(const FieldView& J, int dim) {
  StepWorkspace w{};
  w.acc_bytes = static_cast<uint64_t>(J.cstride) * 3 * sizeof(unsigned long long);
  if (dim == 3) {
    const int64_t ncells = static_cast<int64_t>(J.n1) * J.n2 * J.n3;
    w.flag_chunks = ncells + 1;
    w.flag_bytes  = static_cast<uint64_t>(split_flag_bytes(w.flag_chunks));
    w.rec_chunks  = (ncells / 16 + 1 + 15) / 16 * 16;
    w.rec_bytes   = static_cast<uint64_t>(split_record_bytes(w.rec_chunks));
  }
  return w;
}

}  // namespace

extern "C" ntt_status ntt_sr_step_workspace_size_species(const ntt_field_t* cur,
                                                         int32_t nspecies, uint64_t* bytes) {
  if (bytes == nullptr || nspecies < 0) return NTT_ERR_ARG;
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  const StepWorkspace w = step_workspace(J, cur->dim);
  *bytes = w.acc_bytes + w.flag_bytes + w.rec_bytes * static_cast<uint64_t>(std::max(nspecies, 1));
  return NTT_OK;
}

extern "C" ntt_status ntt_sr_step_workspace_size(const ntt_field_t* cur, uint64_t* bytes) {
  return ntt_sr_step_workspace_size_species(cur, 1, bytes);
}

namespace {

// Everything one SRPIC step of a periodic box needs, validated and derived
// from the step's arguments before anything is enqueued.
struct Prep {
  ParticleView pv;
  PrtlBCs      bcs;
  float        ndth;
  int          alg;
  bool         run, dep;
};

struct StepPlan {
  FieldView           E, J, Bf;
  unsigned long long* acc;
  uint32_t*           flags;             // split-path chunk flags (null: fused kernel)
  int64_t             flag_chunks;
  void*               recs;              // split-path box records (null: none)
  int64_t             rec_chunks;
  uint64_t            rec_stride;        // bytes between species' record sets (0: shared)
  int                 D;
  float               dt, dxh;
  bool                ext;               // extended Faraday stencil
  std::vector<Prep>   prep;
  float               fscale;
  float               far_c1, far_c2;    // Faraday(1/2)
  float               amp_c1, amp_c2, amp_coeff;
};

ntt_status plan_step(ntt_field_t* em, ntt_field_t* cur, ntt_field_t* buff,
                     ntt_species_t* species, int32_t nspecies, const ntt_sr_step_params_t* prm,
                     const ntt_minkowski_t* metric, void* workspace, uint64_t workspace_bytes,
                     StepPlan* P) {
  if (prm == nullptr || metric == nullptr || (nspecies > 0 && species == nullptr) || nspecies < 0)
    return NTT_ERR_ARG;
  ntt_status s = make_field_view(em, 6, &P->E);
  if (s != NTT_OK) return s;
  s = make_field_view(cur, 3, &P->J);
  if (s != NTT_OK) return s;
  s = make_field_view(buff, 3, &P->Bf);
  if (s != NTT_OK) return s;
  const FieldView &E = P->E, &J = P->J, &Bf = P->Bf;
  if (!same_grid(E, J) || !same_grid(E, Bf) || J.d == Bf.d) return NTT_ERR_ARG;
  const int D = em->dim;
  // the accumulator is required; the flags after it are used when the
  // workspace holds them (ntt_sr_step_workspace_size counts them)
  const StepWorkspace wsz = step_workspace(J, D);
  if (prm->deposit_enable && (workspace == nullptr || workspace_bytes < wsz.acc_bytes))
    return NTT_ERR_CAPACITY;
  P->acc = static_cast<unsigned long long*>(workspace);
  P->flags = nullptr;
  P->flag_chunks = 0;
  P->recs = nullptr;
  P->rec_chunks = 0;
  P->rec_stride = 0;
  if (workspace != nullptr && wsz.flag_bytes > 0 &&
      workspace_bytes >= wsz.acc_bytes + wsz.flag_bytes) {
    P->flags = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + wsz.acc_bytes);
    P->flag_chunks = wsz.flag_chunks;
    const uint64_t base = wsz.acc_bytes + wsz.flag_bytes;
    if (workspace_bytes >= base + wsz.rec_bytes) {
      P->recs = static_cast<char*>(workspace) + base;
      P->rec_chunks = wsz.rec_chunks;
      // one record set per species (the push reuses its own boxes), or one
      // set shared by the species (every push runs its box pass)
      if (nspecies == 1 ||
          workspace_bytes >= base + wsz.rec_bytes * static_cast<uint64_t>(nspecies))
        P->rec_stride = wsz.rec_bytes;
    }
  }
  P->D = D;
  if (metric->dim != D || !(metric->dx > 0.0f) || !(prm->dt > 0.0f)) return NTT_ERR_ARG;
  if (E.ngh != 2) return NTT_ERR_UNSUPPORTED;
  const int na[3] = { E.n1, E.n2, E.n3 };
  for (int d = 0; d < D; ++d) {
    if (na[d] < 2 * E.ngh) return NTT_ERR_ARG;
    if (prm->prtl_bc.bc[d][0] != NTT_PRTL_BC_PERIODIC || prm->prtl_bc.bc[d][1] != NTT_PRTL_BC_PERIODIC)
      return NTT_ERR_UNSUPPORTED;
  }
  if (prm->shape_order < 0 || prm->shape_order > 2 || prm->n_filters < 0) return NTT_ERR_ARG;
  const float dt = prm->dt;
  P->dt = dt;
  // dx = sqrt(h_11) as the drivers compute it (fieldsolvers.h:50,110)
  const float dxh = sqrtf(metric->dx * metric->dx);
  P->dxh = dxh;
  P->ext = fd_nonzero(prm->fd);
  P->prep.assign(nspecies, Prep{});
  for (int sp = 0; sp < nspecies; ++sp) {
    Prep& p = P->prep[sp];
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
  // fixed-point scale of the accumulator, common to every species
  float qmax = 0.0f;
  for (int sp = 0; sp < nspecies; ++sp)
    if (P->prep[sp].dep) qmax = fmaxf(qmax, fabsf(species[sp].charge));
  const float wmax = prm->weight_max > 0.0f ? prm->weight_max : 1.0f;
  P->fscale = fixed_point_scale(qmax * wmax, dt);
  // fieldsolvers.h:44-47: Faraday(1/2) coefficients
  const float dT = HALF * prm->correction * dt;
  if (D == 2) { P->far_c1 = dT / (dxh * dxh); P->far_c2 = dT; }
  else { P->far_c1 = dT / dxh; P->far_c2 = ZERO; }
  // fieldsolvers.h:108-117, 157: Ampere and the current term
  const float dTa = ONE * prm->correction * dt;
  if (D == 2) { P->amp_c1 = dTa / (dxh * dxh); P->amp_c2 = dTa; }
  else { P->amp_c1 = dTa / dxh; P->amp_c2 = ZERO; }
  P->amp_coeff = -dt * prm->q0 / (prm->B0 * prm->V0);
  return NTT_OK;
}

}  // namespace

namespace {

// One SRPIC step (ntt_sr_step); `timed` lets the fused launches be bracketed
// by the profiling events when profiling is on.
ntt_status sr_step(ntt_field_t* em, ntt_field_t* cur, ntt_field_t* buff, ntt_species_t* species,
                   int32_t nspecies, const ntt_sr_step_params_t* prm,
                   const ntt_minkowski_t* metric, void* workspace, uint64_t workspace_bytes,
                   void* stream, bool timed_ok) {
  StepPlan P;
  ntt_status s = plan_step(em, cur, buff, species, nspecies, prm, metric, workspace,
                           workspace_bytes, &P);
  if (s != NTT_OK) return s;
  hipStream_t st = as_stream(stream);
  const FieldView &E = P.E, &J = P.J, &Bf = P.Bf;

  // srpic.hpp:71-82: fill E and B ghosts at step 0
  if (prm->step == 0) launch_ghost_fill(E, 0, 6, st);

  auto faraday_half = [&]() -> ntt_status {
    if (!P.ext) {
      launch_faraday_periodic(E, P.far_c1, P.far_c2, st);
      return check_launch();
    }
    ntt_status r = ntt_faraday_mink(em, P.far_c1, P.far_c2, &prm->fd, stream);
    if (r != NTT_OK) return r;
    launch_ghost_fill(E, 3, 6, st);
    return check_launch();
  };

  if (prm->fieldsolver_enable) {
    s = faraday_half();
    if (s != NTT_OK) return s;
  }
  // the fields the push gathers, for a caller that records them
  if (prm->em_half_out != nullptr) launch_copy(prm->em_half_out, E.d, E.cstride * 6, st);
  // currents.h:68: J = 0 -- here the fixed-point accumulator (a fill kernel
  // rather than a memset, so that the step can be captured in a graph)
  if (prm->deposit_enable) launch_zero_u64(P.acc, J.cstride * 3, st);
  for (int sp = 0; sp < nspecies; ++sp) {
    const Prep& p = P.prep[sp];
    if (!p.run) continue;
    const int ppt = prm->particles_per_thread > 0 ? prm->particles_per_thread : 4;
    std::pair<hipEvent_t, hipEvent_t> ev;
    const bool timed = timed_ok && prof_begin(st, &ev);
    void* recs = (P.recs != nullptr && P.rec_stride > 0)
                   ? static_cast<char*>(P.recs) + P.rec_stride * static_cast<uint64_t>(sp)
                   : P.recs;
    s = launch_push_deposit_periodic(p.pv, E, J, P.acc, P.D, prm->shape_order, p.alg, p.ndth,
                                     P.dt, metric->dx, species[sp].charge, p.dep, P.fscale,
                                     species[sp].tile_offsets, species[sp].ntiles, ppt,
                                     P.flags, P.flag_chunks, recs, P.rec_chunks,
                                     P.rec_stride > 0, st);
    if (timed) prof_end(st, ev);
    if (s != NTT_OK) return s;
  }
  if (prm->deposit_enable) {
    // folded current (active cells + periodic ghosts) from the accumulator
    launch_j_from_fixed(J, P.acc, 1.0 / static_cast<double>(P.fscale), st);
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
    launch_ampere_currents_periodic(E, J, P.amp_c1, P.amp_c2, P.amp_coeff, prm->ppc0,
                                    prm->deposit_enable != 0, st);
  }
  return check_launch();
}

// The fixed-point current accumulator of ntt_sr_step_static: one array per
// device (each GPU holds its own copy of this module's globals), used through
// the address hipGetSymbolAddress returns on that device.  Every call on a
// device shares it; the step clears it before depositing.
constexpr uint64_t kStaticScratchBytes = NTT_SR_STEP_STATIC_MAX_BYTES;
static_assert(kStaticScratchBytes % 256 == 0, "scratch size");
__device__ __align__(256) unsigned long long scratch[kStaticScratchBytes / 8];

constexpr int kScratchDevices = 64;
std::mutex    g_scratch_mu;
void*         g_scratch_ptr[kScratchDevices] = {};

// the current device's address of `scratch` (looked up once per device)
ntt_status static_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return NTT_ERR_HIP;
  if (dev < 0 || dev >= kScratchDevices) return NTT_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr)
      return NTT_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NTT_OK;
}

}  // namespace

extern "C" ntt_status ntt_sr_step(ntt_field_t* em, ntt_field_t* cur, ntt_field_t* buff,
                                  ntt_species_t* species, int32_t nspecies,
                                  const ntt_sr_step_params_t* prm, const ntt_minkowski_t* metric,
                                  void* workspace, uint64_t workspace_bytes, void* stream) {
  return sr_step(em, cur, buff, species, nspecies, prm, metric, workspace, workspace_bytes,
                 stream, true);
}

extern "C" ntt_status ntt_sr_step_static(ntt_field_t* em, ntt_field_t* cur, ntt_field_t* buff,
                                         ntt_species_t* species, int32_t nspecies,
                                         const ntt_sr_step_params_t* prm,
                                         const ntt_minkowski_t* metric, void* stream) {
  // the accumulator must fit; the split path's flags and records are used
  // when they fit after it
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (step_workspace(J, cur->dim).acc_bytes > kStaticScratchBytes) return NTT_ERR_CAPACITY;
  void* ws = nullptr;
  s = static_scratch(&ws);
  if (s != NTT_OK) return s;
  return sr_step(em, cur, buff, species, nspecies, prm, metric, ws, kStaticScratchBytes, stream,
                 false);
}

// ---- ensembles of independent periodic boxes ------------------------------------
struct ntt_ensemble {
  struct Group {
    int      D, order, alg;
    void*    table;    // device FusedArgs[nent]
    int64_t* chunk0;   // device, nent + 1
    int      nent;
    int64_t  nchunks;
  };
  struct Seg {
    void*   dev;
    int     n;
    int64_t blocks;
  };
  std::vector<FieldView> ghost_E;   // step 0: E/B ghost fill per box
  Seg                    far, zero, jfix, amp, copy;
  // batched tile re-sort of every (box, species) whose box names a tile
  Seg                    sort;      // device SortSeg[n]
  int64_t*               sort_pbase = nullptr;
  int64_t*               sort_kbase = nullptr;
  int64_t*               sort_bbase = nullptr;
  int64_t                sort_ptot = 0, sort_ktot = 0, sort_blocks = 0;
  uint32_t               sort_live = 0;  // bit a: SortSeg::arr[a] non-null in some segment
  void*                  sort_ws = nullptr;
  std::vector<Seg>       filter_pass;
  std::vector<Group>     groups;
  std::vector<void*>     allocations;
  int                    device = 0;
};

namespace {

ntt_status upload(ntt_ensemble* e, const void* host, size_t bytes, void** dev) {
  *dev = nullptr;
  if (bytes == 0) return NTT_OK;
  if (hipMalloc(dev, bytes) != hipSuccess) return NTT_ERR_HIP;
  e->allocations.push_back(*dev);
  if (hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice) != hipSuccess) return NTT_ERR_HIP;
  return NTT_OK;
}

template <class T>
ntt_status upload_segs(ntt_ensemble* e, std::vector<T>& v, int64_t blocks, ntt_ensemble::Seg* s) {
  s->n = static_cast<int>(v.size());
  s->blocks = blocks;
  return upload(e, v.data(), v.size() * sizeof(T), &s->dev);
}

void destroy(ntt_ensemble* e) {
  for (void* p : e->allocations) (void)hipFree(p);
  delete e;
}

}  // namespace

extern "C" ntt_status ntt_ensemble_create(const ntt_box_t* boxes, int32_t nboxes,
                                          ntt_ensemble_t* out) {
  if (out == nullptr || nboxes < 0 || (nboxes > 0 && boxes == nullptr)) return NTT_ERR_ARG;
  *out = nullptr;
  ntt_ensemble* e = new ntt_ensemble();
  (void)hipGetDevice(&e->device);
  std::vector<FieldSeg> far, jfix, amp;
  std::vector<FlatSeg>  zero, copy;
  std::vector<std::vector<FieldSeg>> passes;
  int64_t nb_far = 0, nb_jfix = 0, nb_amp = 0, nb_zero = 0, nb_copy = 0;
  std::vector<int64_t> nb_pass;
  struct GroupHost {
    int D, order, alg;
    std::vector<char>    args;
    std::vector<int64_t> chunk0;
  };
  std::vector<GroupHost> groups;
  std::vector<SortSeg>   sorts;
  std::vector<int64_t>   pbase{ 0 }, kbase{ 0 }, bbase{ 0 };
  const size_t asz = fused_args_size();
  ntt_status s = NTT_OK;
  for (int b = 0; b < nboxes && s == NTT_OK; ++b) {
    const ntt_box_t& B = boxes[b];
    ntt_field_t em = B.em, cur = B.cur, buff = B.buff;
    StepPlan P;
    s = plan_step(&em, &cur, &buff, B.species, B.nspecies, &B.prm, &B.metric, B.workspace,
                  B.workspace_bytes, &P);
    if (s != NTT_OK) break;
    if (P.ext) { s = NTT_ERR_UNSUPPORTED; break; }  // extended Faraday: use ntt_sr_step
    const ntt_sr_step_params_t& prm = B.prm;
    e->ghost_E.push_back(P.E);
    FieldSeg g{};
    int32_t gx, gy;
    const int64_t nblk = field_seg_blocks(P.E, &gx, &gy);
    g.gx = gx;
    g.gy = gy;
    if (prm.fieldsolver_enable) {
      g.F = P.E;
      g.c1 = P.far_c1;
      g.c2 = P.far_c2;
      g.blk0 = nb_far;
      far.push_back(g);
      nb_far += nblk;
      FieldSeg a = g;
      a.J = P.J;
      a.c1 = P.amp_c1;
      a.c2 = P.amp_c2;
      a.coeff = P.amp_coeff;
      a.ppc0 = prm.ppc0;
      a.with_j = prm.deposit_enable != 0;
      a.blk0 = nb_amp;
      amp.push_back(a);
      nb_amp += nblk;
    }
    if (prm.deposit_enable) {
      const int64_t n = P.J.cstride * 3;
      zero.push_back(FlatSeg{ P.acc, nullptr, n, nb_zero });
      nb_zero += flat_seg_blocks(n);
      FieldSeg j{};
      j.gx = gx;
      j.gy = gy;
      j.F = P.J;
      j.acc = P.acc;
      j.inv_scale = 1.0 / static_cast<double>(P.fscale);
      j.blk0 = nb_jfix;
      jfix.push_back(j);
      nb_jfix += nblk;
      const FieldView* src = &P.J;
      const FieldView* dst = &P.Bf;
      for (int f = 0; f < prm.n_filters; ++f) {
        if (static_cast<int>(passes.size()) <= f) {
          passes.emplace_back();
          nb_pass.push_back(0);
        }
        FieldSeg q{};
        q.gx = gx;
        q.gy = gy;
        q.F = *dst;
        q.J = *src;
        q.blk0 = nb_pass[f];
        passes[f].push_back(q);
        nb_pass[f] += nblk;
        std::swap(src, dst);
      }
      if (src != &P.J) {
        copy.push_back(FlatSeg{ P.J.d, P.Bf.d, n, nb_copy });
        nb_copy += flat_seg_blocks(n);
      }
    }
    if (B.tile[0] > 0) {
      const int32_t n3[3]    = { B.em.n[0], B.em.n[1], P.D == 3 ? B.em.n[2] : 1 };
      const int32_t tile3[3] = { B.tile[0], B.tile[1], P.D == 3 ? B.tile[2] : 1 };
      for (int sp = 0; sp < B.nspecies && s == NTT_OK; ++sp) {
        SortSeg g{};
        // ntt_ensemble_sort rewrites the offsets the species hands to the step
        s = make_sort_seg(&B.species[sp].prtls, P.D, n3, tile3,
                          const_cast<uint32_t*>(B.species[sp].tile_offsets), &g);
        if (s != NTT_OK) break;
        // the step's push rewrites the prev arrays of a pushed species before
        // anything reads them: the re-sort leaves them out
        if (B.species[sp].pusher_flags != NTT_PUSHER_NONE)
          for (int a : { 3, 4, 5, 9, 10, 11 }) g.arr[a] = nullptr;  // i*_prev, dx*_prev
        // the step's chunking reads these offsets: they must cover every tile
        if (B.species[sp].tile_offsets != nullptr &&
            B.species[sp].ntiles != static_cast<int64_t>(g.dead_key)) {
          s = NTT_ERR_ARG;
          break;
        }
        sorts.push_back(g);
        pbase.push_back(pbase.back() + g.n);
        kbase.push_back(kbase.back() + static_cast<int64_t>(g.dead_key) + 1);
        bbase.push_back(bbase.back() + sort_blocks(g.n));
      }
      if (s != NTT_OK) break;
    }
    const int ppt = prm.particles_per_thread > 0 ? prm.particles_per_thread : 4;
    for (int sp = 0; sp < B.nspecies && s == NTT_OK; ++sp) {
      const Prep& p = P.prep[sp];
      if (!p.run) continue;
      const int order = (p.alg == ALG_PHOTON) ? 0 : prm.shape_order;
      GroupHost* grp = nullptr;
      for (auto& gh : groups)
        if (gh.D == P.D && gh.order == order && gh.alg == p.alg) grp = &gh;
      if (grp == nullptr) {
        groups.push_back(GroupHost{ P.D, order, p.alg, {}, { 0 } });
        grp = &groups.back();
      }
      std::vector<char> a(asz);
      int64_t           nch = 0;
      s = build_fused_args(a.data(), &nch, p.pv, P.E, P.J, P.acc, P.D, p.ndth, P.dt, B.metric.dx,
                           B.species[sp].charge, p.dep, P.fscale, B.species[sp].tile_offsets,
                           B.species[sp].ntiles, ppt);
      if (s != NTT_OK) break;
      if (nch == 0) continue;
      grp->args.insert(grp->args.end(), a.begin(), a.end());
      grp->chunk0.push_back(grp->chunk0.back() + nch);
    }
  }
  if (s == NTT_OK) s = upload_segs(e, far, nb_far, &e->far);
  if (s == NTT_OK) s = upload_segs(e, amp, nb_amp, &e->amp);
  if (s == NTT_OK) s = upload_segs(e, jfix, nb_jfix, &e->jfix);
  if (s == NTT_OK) s = upload_segs(e, zero, nb_zero, &e->zero);
  if (s == NTT_OK) s = upload_segs(e, copy, nb_copy, &e->copy);
  for (size_t f = 0; f < passes.size() && s == NTT_OK; ++f) {
    ntt_ensemble::Seg seg;
    s = upload_segs(e, passes[f], nb_pass[f], &seg);
    e->filter_pass.push_back(seg);
  }
  if (s == NTT_OK && !sorts.empty()) {
    s = upload_segs(e, sorts, 0, &e->sort);
    void* pb = nullptr;
    void* kb = nullptr;
    if (s == NTT_OK) s = upload(e, pbase.data(), pbase.size() * sizeof(int64_t), &pb);
    void* bb = nullptr;
    if (s == NTT_OK) s = upload(e, kbase.data(), kbase.size() * sizeof(int64_t), &kb);
    if (s == NTT_OK) s = upload(e, bbase.data(), bbase.size() * sizeof(int64_t), &bb);
    e->sort_pbase  = static_cast<int64_t*>(pb);
    e->sort_kbase  = static_cast<int64_t*>(kb);
    e->sort_bbase  = static_cast<int64_t*>(bb);
    e->sort_ptot   = pbase.back();
    e->sort_ktot   = kbase.back();
    e->sort_blocks = bbase.back();
    for (const SortSeg& g : sorts)
      for (int a = 0; a < kSortArrays; ++a)
        if (g.arr[a] != nullptr) e->sort_live |= 1u << a;
    if (s == NTT_OK) {
      const int64_t bytes = sort_batched_bytes(e->sort_ptot, e->sort_ktot);
      if (hipMalloc(&e->sort_ws, bytes) != hipSuccess) s = NTT_ERR_HIP;
      else e->allocations.push_back(e->sort_ws);
    }
  }
  for (auto& gh : groups) {
    if (s != NTT_OK) break;
    ntt_ensemble::Group g{ gh.D, gh.order, gh.alg, nullptr, nullptr,
                           static_cast<int>(gh.chunk0.size()) - 1, gh.chunk0.back() };
    if (g.nent == 0) continue;
    s = upload(e, gh.args.data(), gh.args.size(), &g.table);
    if (s == NTT_OK) {
      void* c0 = nullptr;
      s = upload(e, gh.chunk0.data(), gh.chunk0.size() * sizeof(int64_t), &c0);
      g.chunk0 = static_cast<int64_t*>(c0);
    }
    e->groups.push_back(g);
  }
  if (s != NTT_OK) {
    destroy(e);
    return s;
  }
  *out = e;
  return NTT_OK;
}

extern "C" ntt_status ntt_ensemble_step(ntt_ensemble_t e, int64_t step, void* stream) {
  if (e == nullptr) return NTT_ERR_ARG;
  hipStream_t st = as_stream(stream);
  if (step == 0)
    for (const FieldView& E : e->ghost_E) launch_ghost_fill(E, 0, 6, st);
  auto fields = [&](int op, const ntt_ensemble::Seg& sg) {
    launch_fields_batched(op, static_cast<const FieldSeg*>(sg.dev), sg.n, sg.blocks, st);
  };
  auto flat = [&](int op, const ntt_ensemble::Seg& sg) {
    launch_flat_batched(op, static_cast<const FlatSeg*>(sg.dev), sg.n, sg.blocks, st);
  };
  fields(SEG_FARADAY, e->far);
  flat(FLAT_ZERO_U64, e->zero);
  for (const auto& g : e->groups) {
    ntt_status s = launch_push_deposit_batched(g.D, g.order, g.alg, g.table, g.chunk0, g.nent,
                                               g.nchunks, st);
    if (s != NTT_OK) return s;
  }
  fields(SEG_JFIX, e->jfix);
  for (const auto& p : e->filter_pass) fields(SEG_FILTER, p);
  flat(FLAT_COPY_F32, e->copy);
  fields(SEG_FARADAY, e->far);
  fields(SEG_AMPERE, e->amp);
  return check_launch();
}

extern "C" ntt_status ntt_ensemble_sort(ntt_ensemble_t e, void* stream) {
  if (e == nullptr) return NTT_ERR_ARG;
  if (e->sort.n == 0) return NTT_OK;
  ntt_status s = sort_batched(static_cast<const SortSeg*>(e->sort.dev), e->sort_pbase,
                              e->sort_kbase, e->sort_bbase, e->sort_blocks, e->sort.n,
                              e->sort_ptot, e->sort_ktot, e->sort_live, e->sort_ws,
                              as_stream(stream));
  if (s != NTT_OK) return s;
  return check_launch();
}

extern "C" ntt_status ntt_ensemble_destroy(ntt_ensemble_t e) {
  if (e == nullptr) return NTT_OK;
  destroy(e);
  return NTT_OK;
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
