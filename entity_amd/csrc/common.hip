// common.hip -- ABI bookkeeping: status strings, argument validation, the
// device error word.
#include "ntt_common.hpp"

namespace ntt {

namespace {
__device__ uint32_t g_device_errors = 0;
constexpr int kMaxDevices = 64;
uint32_t* g_error_word_ptr[kMaxDevices] = {};
}  // namespace

uint32_t* device_error_word() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  if (g_error_word_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_device_errors)) != hipSuccess) return nullptr;
    g_error_word_ptr[dev] = static_cast<uint32_t*>(p);
  }
  return g_error_word_ptr[dev];
}

ntt_status make_field_view(const ntt_field_t* f, int want_ncomp, FieldView* out) {
  if (f == nullptr || f->data == nullptr) return NTT_ERR_ARG;
  if (f->dim != 2 && f->dim != 3) return NTT_ERR_UNSUPPORTED;
  if (f->ngh < 1) return NTT_ERR_ARG;
  if (want_ncomp > 0 && f->ncomp != want_ncomp) return NTT_ERR_ARG;
  for (int d = 0; d < f->dim; ++d)
    if (f->n[d] < 1) return NTT_ERR_ARG;
  FieldView v;
  v.d   = f->data;
  v.ngh = f->ngh;
  v.nc  = f->ncomp;
  v.n1  = f->n[0];
  v.n2  = f->n[1];
  v.n3  = f->dim >= 3 ? f->n[2] : 1;
  v.N1  = v.n1 + 2 * f->ngh;
  v.N2  = v.n2 + 2 * f->ngh;
  v.N3  = f->dim >= 3 ? v.n3 + 2 * f->ngh : 1;
  v.cstride = static_cast<int64_t>(v.N1) * v.N2 * v.N3;
  *out = v;
  return NTT_OK;
}

ntt_status make_particle_view(const ntt_particles_t* p, int dim, bool need_u, ParticleView* out) {
  if (p == nullptr) return NTT_ERR_ARG;
  if (p->npart < 0 || (p->maxnpart > 0 && p->npart > p->maxnpart)) return NTT_ERR_ARG;
  ParticleView v{};
  v.n = p->npart;
  if (v.n == 0) {
    *out = v;
    return NTT_OK;
  }
  if (!p->i1 || !p->i1_prev || !p->dx1 || !p->dx1_prev || !p->tag || !p->weight)
    return NTT_ERR_ARG;
  if (dim >= 2 && (!p->i2 || !p->i2_prev || !p->dx2 || !p->dx2_prev)) return NTT_ERR_ARG;
  if (dim >= 3 && (!p->i3 || !p->i3_prev || !p->dx3 || !p->dx3_prev)) return NTT_ERR_ARG;
  if (need_u && (!p->ux1 || !p->ux2 || !p->ux3)) return NTT_ERR_ARG;
  v.i1 = p->i1; v.i2 = p->i2; v.i3 = p->i3;
  v.i1p = p->i1_prev; v.i2p = p->i2_prev; v.i3p = p->i3_prev;
  v.dx1 = p->dx1; v.dx2 = p->dx2; v.dx3 = p->dx3;
  v.dx1p = p->dx1_prev; v.dx2p = p->dx2_prev; v.dx3p = p->dx3_prev;
  v.ux1 = p->ux1; v.ux2 = p->ux2; v.ux3 = p->ux3;
  v.w = p->weight;
  v.tag = p->tag;
  *out = v;
  return NTT_OK;
}

ntt_status check_launch(void) {
  return hipGetLastError() == hipSuccess ? NTT_OK : NTT_ERR_HIP;
}

}  // namespace ntt

extern "C" int32_t ntt_abi_version(void) { return NTT_ABI_VERSION; }

extern "C" const char* ntt_status_string(ntt_status s) {
  switch (s) {
    case NTT_OK: return "ok";
    case NTT_ERR_ARG: return "invalid argument";
    case NTT_ERR_HIP: return "HIP runtime error";
    case NTT_ERR_DEVICE: return "device error flag set";
    case NTT_ERR_UNSUPPORTED: return "unsupported configuration";
    case NTT_ERR_CAPACITY: return "capacity or workspace too small";
  }
  return "unknown status";
}

extern "C" ntt_status ntt_get_device_errors(uint32_t* flags) {
  if (flags == nullptr) return NTT_ERR_ARG;
  if (hipDeviceSynchronize() != hipSuccess) return NTT_ERR_HIP;
  uint32_t* w = ntt::device_error_word();
  if (w == nullptr) return NTT_ERR_HIP;
  uint32_t v = 0, zero = 0;
  if (hipMemcpy(&v, w, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return NTT_ERR_HIP;
  if (hipMemcpy(w, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return NTT_ERR_HIP;
  *flags = v;
  return NTT_OK;
}
