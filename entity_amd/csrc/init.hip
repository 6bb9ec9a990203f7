// init.hip -- seeded synthetic plasma generator (uniform density, drifting
// relativistic thermal distribution) written straight into the SoA.
//
// Sampling follows the reference archetypes (src/archetypes/energy_dist.h):
// JuttnerSinge (:79-135) draws a Gaussian in u (Box-Muller) for
// theta < 0.5 and uses the Sobol rejection method otherwise; a drift along a
// principal axis is applied with the flipping method (:136-170).  Random
// numbers come from a counter-based hash of (seed, stream, particle, draw),
// so the result does not depend on the launch shape.  Particle p of a species
// lives in tile p / (ppc * cells_per_tile), and consecutive particles of a
// tile visit its cells round-robin; positions depend on the position seed
// only, so two species sharing it are co-located pairwise, as the reference's
// uniform injector places them (src/kernels/injectors.hpp:198-218).
#include "internal.hpp"

namespace ntt {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finalizer
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]
__device__ __forceinline__ float urand(uint64_t seed, uint64_t stream, uint64_t p, uint32_t draw) {
  const uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull ^ mix64(p * 8 + draw)));
  return (static_cast<float>(h >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

struct InitArgs {
  ParticleView P;
  int32_t  dim, n[3], tile[3], nt[3];
  int32_t  ppc, cpt;
  uint64_t seed, pos_seed, stream;
  float    theta, weight;
  float    drift[3];
};

__device__ void sample_u(const InitArgs& a, int64_t p, float* v) {
  uint32_t draw = 0;
  auto r = [&]() { return urand(a.seed, a.stream, static_cast<uint64_t>(p), draw++); };
  const float th = a.theta;
  if (th <= 0.0f) {
    v[0] = v[1] = v[2] = 0.0f;
  } else if (th < 0.5f) {
    for (int d = 0; d < 3; ++d) {
      const float x1 = sqrtf(-2.0f * logf(r()));
      const float x2 = 6.283185307179586f * r();
      v[d] = x1 * cosf(x2) * sqrtf(th);
    }
  } else {
    float u = 1.0f, eta = r();
    for (int it = 0; it < 256 && eta * eta <= u * u + 1.0f; ++it) {
      const float x1 = r() * r() * r();
      u   = -th * logf(x1);
      eta = -th * logf(x1 * r());
    }
    const float x1 = r(), x2 = r();
    v[0] = u * (2.0f * x1 - 1.0f);
    v[2] = 2.0f * u * sqrtf(x1 * (1.0f - x1));
    v[1] = v[2] * cosf(6.283185307179586f * x2);
    v[2] = v[2] * sinf(6.283185307179586f * x2);
  }
  for (int d = 0; d < 3; ++d) {
    const float ud = a.drift[d];
    if (ud == 0.0f) continue;
    const float beta  = ud / sqrtf(1.0f + ud * ud);
    const float gamma = sqrtf(1.0f + v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (-beta * v[d] > gamma * r()) v[d] = -v[d];
    v[d] = sqrtf(1.0f + ud * ud) * (v[d] + beta * gamma);
  }
}

__global__ __launch_bounds__(256) void k_init(InitArgs a, int64_t npart) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (p >= npart) return;
  const int64_t per_tile = static_cast<int64_t>(a.ppc) * a.cpt;
  const int64_t t        = p / per_tile;
  const int64_t q        = p - t * per_tile;
  const int     lc       = static_cast<int>(q % a.cpt);
  // tile coordinates (x fastest) and the cell inside the tile
  int tc[3], lcc[3];
  int64_t rem = t;
  for (int d = 0; d < a.dim; ++d) {
    tc[d] = static_cast<int>(rem % a.nt[d]);
    rem /= a.nt[d];
  }
  int lrem = lc;
  for (int d = 0; d < a.dim; ++d) {
    lcc[d] = lrem % a.tile[d];
    lrem /= a.tile[d];
  }
  int cell[3] = { 0, 0, 0 };
  for (int d = 0; d < a.dim; ++d) cell[d] = tc[d] * a.tile[d] + lcc[d];
  float dx[3];
  for (int d = 0; d < a.dim; ++d) {
    float x = urand(a.pos_seed, 0x5eedull, static_cast<uint64_t>(p), 7 - d) - 1e-7f;
    dx[d] = fminf(fmaxf(x, 0.0f), 0.99999994f);
  }
  float v[3];
  sample_u(a, p, v);
  a.P.i1[p] = cell[0];
  a.P.i1p[p] = cell[0];
  a.P.dx1[p] = dx[0];
  a.P.dx1p[p] = dx[0];
  a.P.i2[p] = cell[1];
  a.P.i2p[p] = cell[1];
  a.P.dx2[p] = dx[1];
  a.P.dx2p[p] = dx[1];
  if (a.dim == 3) {
    a.P.i3[p] = cell[2];
    a.P.i3p[p] = cell[2];
    a.P.dx3[p] = dx[2];
    a.P.dx3p[p] = dx[2];
  }
  a.P.ux1[p] = v[0];
  a.P.ux2[p] = v[1];
  a.P.ux3[p] = v[2];
  a.P.w[p] = a.weight;
  a.P.tag[p] = NTT_TAG_ALIVE;
}

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_init_uniform_plasma(ntt_particles_t* prtls, int32_t dim,
                                              const int32_t n[3],
                                              const ntt_plasma_init_t* init, void* stream) {
  if (prtls == nullptr || n == nullptr || init == nullptr) return NTT_ERR_ARG;
  if (dim != 2 && dim != 3) return NTT_ERR_UNSUPPORTED;
  if (init->ppc < 1 || init->temperature < 0.0f) return NTT_ERR_ARG;
  InitArgs a{};
  a.dim = dim;
  int64_t ncells = 1;
  a.cpt = 1;
  for (int d = 0; d < dim; ++d) {
    if (n[d] < 1 || init->tile[d] < 1 || n[d] % init->tile[d] != 0) return NTT_ERR_ARG;
    a.n[d] = n[d];
    a.tile[d] = init->tile[d];
    a.nt[d] = n[d] / init->tile[d];
    ncells *= n[d];
    a.cpt *= init->tile[d];
  }
  const int64_t npart = ncells * init->ppc;
  if (prtls->maxnpart > 0 && npart > prtls->maxnpart) return NTT_ERR_CAPACITY;
  ntt_particles_t tmp = *prtls;
  tmp.npart = npart;
  ntt_status s = make_particle_view(&tmp, dim, true, &a.P);
  if (s != NTT_OK) return s;
  a.ppc = init->ppc;
  a.seed = init->seed;
  a.pos_seed = init->position_seed;
  a.stream = static_cast<uint64_t>(init->species_id);
  a.theta = init->temperature;
  a.weight = init->weight;
  for (int d = 0; d < 3; ++d) a.drift[d] = init->drift_u[d];
  int nz = 0;
  for (int d = 0; d < 3; ++d) nz += (a.drift[d] != 0.0f);
  if (nz > 1) return NTT_ERR_UNSUPPORTED;  // drift along one principal axis
  k_init<<<grid_for(npart, 256), 256, 0, as_stream(stream)>>>(a, npart);
  s = check_launch();
  if (s == NTT_OK) prtls->npart = npart;
  return s;
}
