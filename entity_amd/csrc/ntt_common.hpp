// ntt_common.hpp -- shared device/host helpers of libntt_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ntt.h"

namespace ntt {

// single-precision literal constants of the reference
// (src/global/utils/numeric.h:31-48; THIRD is 0.333333f, not 1/3)
constexpr float ONE = 1.0f, TWO = 2.0f, THREE = 3.0f, FOUR = 4.0f;
constexpr float ZERO = 0.0f, HALF = 0.5f, THIRD = 0.333333f;
constexpr float THREE_FOURTHS = 0.75f, THREE_HALFS = 1.5f;
constexpr float INV_2 = 0.5f, INV_4 = 0.25f, INV_8 = 0.125f;
constexpr float INV_16 = 0.0625f, INV_32 = 0.03125f, INV_64 = 0.015625f;

// Device error word (bits NTT_DEVERR_*): one per device, defined in
// common.hip.  Kernels receive its address as an argument.
uint32_t* device_error_word();

__device__ __forceinline__ void raise_device_error(uint32_t* word, uint32_t bit) {
  atomicOr(word, bit);
}

// field view: LayoutLeft, component slowest (Kokkos default on HIP)
struct FieldView {
  float*  d;
  int32_t n1, n2, n3;  // active cells (1 for unused dims)
  int32_t N1, N2, N3;  // extents including ghosts (1 for unused dims)
  int32_t ngh, nc;
  int64_t cstride;     // N1 * N2 * N3

  __host__ __device__ __forceinline__ int64_t idx(int i, int j, int k, int c) const {
    return static_cast<int64_t>(i) +
           static_cast<int64_t>(N1) * (static_cast<int64_t>(j) +
                                       static_cast<int64_t>(N2) * static_cast<int64_t>(k)) +
           cstride * c;
  }
  __device__ __forceinline__ float& operator()(int i, int j, int k, int c) const {
    return d[idx(i, j, k, c)];
  }
};

struct ParticleView {
  int32_t *i1, *i2, *i3, *i1p, *i2p, *i3p;
  float *dx1, *dx2, *dx3, *dx1p, *dx2p, *dx3p;
  float *ux1, *ux2, *ux3, *w;
  int16_t* tag;
  int64_t  n;
};

// host helpers --------------------------------------------------------------
ntt_status make_field_view(const ntt_field_t* f, int want_ncomp, FieldView* out);
ntt_status make_particle_view(const ntt_particles_t* p, int dim, bool need_u, ParticleView* out);
ntt_status check_launch(void);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace ntt
