// push_deposit.hip -- the hot path of one SR step for a periodic box: the
// particle push (interpolation, Boris/Vay, position update, periodic wrap)
// fused with the charge-conserving current deposit.
//
// Design (gfx950):
// * Work is cut into chunks of consecutive particles: the particle range of
//   one spatial tile when the tile offsets of the last sort are given, else
//   fixed chunks of 256 * ppt particles.  One workgroup = one chunk.  Chunks
//   are mapped so that each XCD (workgroups are dealt round-robin over the 8
//   XCDs) walks a contiguous range of chunks, whose windows overlap in its L2.
// * The workgroup reduces the cell bounding box of its live particles, in
//   the plain frame and in the frame rotated by half the box per dimension
//   (a tile whose particles straddle the periodic boundary is compact in the
//   rotated one), and keeps the smaller.  If box + [-1, +2] cells (the reach
//   of the order-0..2 gather and of the deposit of a particle that moves at
//   most one cell) fits the LDS window, the E/B values of that box (periodic
//   images of the active cells) are staged into LDS once, every particle
//   gathers from LDS, and every deposit is accumulated into an LDS current
//   window as 64-bit fixed point (units of 2^-32) with integer LDS atomics:
//   on gfx950 a 64-bit integer LDS atomic costs ~7 clocks per wavefront
//   instruction against ~190 for the float one.  The window is then added to
//   an HBM fixed-point accumulator with one 64-bit integer atomic per
//   non-zero entry.  Integer sums do not depend on the order of the
//   additions, so the deposited current is bit-reproducible.  Particles
//   whose exact deposit stencil leaves the window add their current straight
//   to the HBM accumulator.
// * Chunks whose box does not fit (unsorted input) run the same code against
//   HBM (gather from the ghost-extended em array, global atomics).
// * Deposits use the pre-wrap cell indices and are folded onto the periodic
//   image at the flush, which produces the folded current of
//   Metadomain::SynchronizeFields directly (metadomain_comm.cpp:408-562).
//
// Reference: src/kernels/pushers/sr.hpp:117-814, src/kernels/currents_deposit.hpp:108-761.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "sr_particle.hpp"

namespace ntt {

// LDS window extents (cells).  3D: an 8 x 4 x 4 tile plus the [-1, +2] reach
// and one cell of drift on each side needs 14 x 10 x 10 (67.8 KB with the
// current window, two workgroups per CU).  Against 4^3 tiles in a 10^3 window
// it halves the workgroups, and with them the per-workgroup latency chain
// (box, staging, flush) and the partial last pass of tiles just over the
// mean count: 38.6 against 45.6 ms per launch at 512 x 256 x 256, order 2.
#ifndef NTT_PD_W3D
#define NTT_PD_W3D 14, 10, 10
#endif
template <int D> struct Win;
template <> struct Win<3> {
  static constexpr int W[3] = { NTT_PD_W3D };
  static constexpr int W1 = W[0], W2 = W[1], W3 = W[2];
};
#ifndef NTT_PD_W2D
#define NTT_PD_W2D 32, 16
#endif
template <> struct Win<2> {
  static constexpr int W[2] = { NTT_PD_W2D };
  static constexpr int W1 = W[0], W2 = W[1], W3 = 1;
};

namespace {

// Workgroup size and occupancy target of the fused kernel.  The LDS windows
// (67.8 KB in 3D) admit two workgroups per CU; with 512 threads and at most
// 128 VGPRs two workgroups give 4 waves per SIMD.
#ifndef NTT_PD_THREADS
#define NTT_PD_THREADS 512
#endif
#ifndef NTT_PD_WAVES_PER_EU
#define NTT_PD_WAVES_PER_EU 4
#endif
constexpr int kThreads = NTT_PD_THREADS;
constexpr int kWaves   = kThreads / 64;

// LDS window accessor for em (6 components)
template <int D>
struct LdsEM {
  const float* w;  // [6][W1*W2*W3]
  int o1, o2, o3;  // ghost-extended origin of the window
  static constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, W3 = Win<D>::W3;
  static constexpr int WC = W1 * W2 * W3;
  __device__ __forceinline__ float operator()(int c, int i, int j, int k) const {
    if constexpr (D == 2)
      return w[c * WC + (i - o1) + W1 * (j - o2)];
    else
      return w[c * WC + (i - o1) + W1 * ((j - o2) + W2 * (k - o3))];
  }
};

// 64-bit fixed point of a value x already multiplied by the accumulator's
// scale (the deposit coefficient carries the factor, an exact power of two).
// General form: x = h * 2^32 + r with h = rint(x / 2^32) and |r| <= 2^31; r is
// exact in float (a multiple of ulp(x) no larger than |x|), so the result is
// x truncated toward zero to an integer (error below one unit).
__device__ __forceinline__ unsigned long long to_fixed(float x) {
  const float   h  = rintf(x * 2.3283064365386963e-10f);
  const float   r  = fminf(fmaf(h, -4294967296.0f, x), 2147483520.0f);  // < 2^31
  const int32_t lo = static_cast<int32_t>(r);
  const int32_t hi = static_cast<int32_t>(h) + (lo >> 31);
  return (static_cast<unsigned long long>(static_cast<uint32_t>(hi)) << 32) |
         static_cast<uint32_t>(lo);
}
// Short form for |x| < 2^31: the 32-bit truncation, sign-extended.
__device__ __forceinline__ unsigned long long to_fixed_small(float x) {
  const int32_t lo = static_cast<int32_t>(x);
  return (static_cast<unsigned long long>(static_cast<uint32_t>(lo >> 31)) << 32) |
         static_cast<uint32_t>(lo);
}

// LDS current window sink (3 components), 64-bit integer LDS atomics
template <int D>
struct LdsJ {
  unsigned long long* w;
  int                 o1, o2, o3;
  int                 base = 0;
  static constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, W3 = Win<D>::W3;
  static constexpr int WC = W1 * W2 * W3;
  __device__ __forceinline__ int index(int i, int j, int k) const {
    if constexpr (D == 2) return (i - o1) + W1 * (j - o2);
    else return (i - o1) + W1 * ((j - o2) + W2 * (k - o3));
  }
  __device__ __forceinline__ void set_origin(int i, int j, int k) { base = index(i, j, k); }
  // values are below 2^31 units (checked per particle before the deposit)
  __device__ __forceinline__ void add_rel(int a, int b, int g, int c, float v) {
    atomicAdd(w + base + (c * WC + a + W1 * (b + W2 * g)), to_fixed_small(v));
  }
  __device__ __forceinline__ void add(int i, int j, int k, int c, float v) {
    atomicAdd(w + c * WC + index(i, j, k), to_fixed_small(v));
  }
};

// HBM fixed-point current accumulator (layout of the ghost-extended current
// array), folding ghost-extended indices onto the periodic image
struct WrapJ {
  unsigned long long* d;
  int32_t             n1, n2, n3, G, N1, N2;
  int64_t             cstride;
  __device__ __forceinline__ static int wrap(int x, int n) {
    while (x < 0) x += n;
    while (x >= n) x -= n;
    return x;
  }
  // linear index of ghost-extended (a, b, g) in component c (no wrap)
  __device__ __forceinline__ int64_t linear(int a, int b, int g, int c) const {
    return static_cast<int64_t>(a) +
           static_cast<int64_t>(N1) * (static_cast<int64_t>(b) + static_cast<int64_t>(N2) * g) +
           cstride * c;
  }
  __device__ __forceinline__ int64_t index(int i, int j, int k, int c) const {
    const int a = wrap(i - G, n1) + G;
    const int b = wrap(j - G, n2) + G;
    const int g = (n3 > 1) ? wrap(k - G, n3) + G : 0;
    return linear(a, b, g, c);
  }
  int i0 = 0, j0 = 0, k0 = 0;
  __device__ __forceinline__ void set_origin(int i, int j, int k) {
    i0 = i;
    j0 = j;
    k0 = k;
  }
  __device__ __forceinline__ void add_rel(int a, int b, int g, int c, float v) {
    add(i0 + a, j0 + b, k0 + g, c, v);
  }
  __device__ __forceinline__ void add(int i, int j, int k, int c, float v) {
    atomicAdd(d + index(i, j, k, c), to_fixed(v));
  }
  __device__ __forceinline__ void add_fixed(int i, int j, int k, int c, unsigned long long v) {
    atomicAdd(d + index(i, j, k, c), v);
  }
  __device__ __forceinline__ void add_fixed_at(int64_t idx, unsigned long long v) const {
    atomicAdd(d + idx, v);
  }
};

struct FusedArgs {
  ParticleView    P;
  GlobalEM        F;     // ghost-extended em (periodic ghosts filled)
  WrapJ           J;     // active current, folded
  int32_t         ngh;
  int32_t         n[3];
  float           ndth, dt, dxm, charge, inv_dt;
  float           fscale;      // fixed-point units per unit of current (2^k)
  float           small_coeff; // |coeff * fscale| up to this: LDS short path
  float           big_coeff;   // |coeff * fscale| up to this: general conversion
  int64_t         nchunks;
  int32_t         ppt;   // particles per thread of a fixed chunk
  const uint32_t* toff;  // tile offsets (nchunks entries: ntiles + 1) or null
  uint32_t*       err;
};

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

template <int D>
__device__ __forceinline__ void load_state(const ParticleView& P, int64_t p, PState& s) {
  s.i[0]  = P.i1[p];
  s.i[1]  = P.i2[p];
  s.i[2]  = (D == 3) ? P.i3[p] : 0;
  s.dx[0] = P.dx1[p];
  s.dx[1] = P.dx2[p];
  if constexpr (D == 3) s.dx[2] = P.dx3[p];
  s.u[0] = P.ux1[p];
  s.u[1] = P.ux2[p];
  s.u[2] = P.ux3[p];
}

template <int D>
__device__ __forceinline__ void store_state(const ParticleView& P, int64_t p, const PState& s) {
  P.i1p[p]  = s.ip[0];
  P.dx1p[p] = s.dxp[0];
  P.i1[p]   = s.i[0];
  P.dx1[p]  = s.dx[0];
  P.i2p[p]  = s.ip[1];
  P.dx2p[p] = s.dxp[1];
  P.i2[p]   = s.i[1];
  P.dx2[p]  = s.dx[1];
  if constexpr (D == 3) {
    P.i3p[p]  = s.ip[2];
    P.dx3p[p] = s.dxp[2];
    P.i3[p]   = s.i[2];
    P.dx3[p]  = s.dx[2];
  }
  P.ux1[p] = s.u[0];
  P.ux2[p] = s.u[1];
  P.ux3[p] = s.u[2];
}

// Periodic frame of one dimension for a chunk.  A chunk whose particles sit
// on both sides of the periodic boundary (e.g. cells n-1 and 0) is compact
// in the frame rotated by h = n/2: cells c >= n - h are taken as c - n.  The
// push and the deposit then run on these unwrapped cell indices, and the
// periodic wrap that follows the push maps them back into [0, n) with the
// same i / i_prev shifts as for a particle that crossed the boundary.
__device__ __forceinline__ int rotated(int c, int n) {
  const int h = n >> 1;
  const int r = c + h;
  return r >= n ? r - n : r;
}
__device__ __forceinline__ int wrap_any(int x, int n) {
  while (x < 0) x += n;
  while (x >= n) x -= n;
  return x;
}

// XCD-aware chunk order: blockIdx.x % 8 is the XCD a workgroup runs on; each
// XCD gets a contiguous run of chunks, whose windows overlap in its L2.
// Returns -1 for the padding blocks of the last run.
__device__ __forceinline__ int64_t xcd_chunk(int64_t nb) {
  const int64_t bid  = blockIdx.x;
  const int64_t per  = (nb + 7) / 8;
  const int64_t xcd  = bid % 8;
  const int64_t slot = bid / 8;
  const int64_t full = nb - 8 * (per - 1);  // XCDs owning `per` chunks
  if (slot >= ((xcd < full) ? per : per - 1)) return -1;
  return (xcd < full) ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
}

// Push and deposit of one chunk (one workgroup).
template <int D, int O, int ALG>
__device__ __forceinline__ void push_deposit_chunk(const FusedArgs& a, int64_t chunk) {
  constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, W3 = Win<D>::W3;
  constexpr int WC = W1 * W2 * W3;
  constexpr int WM = (W1 > W2 ? W1 : W2) > W3 ? (W1 > W2 ? W1 : W2) : W3;
  __shared__ float              em_w[6 * WC];
  __shared__ unsigned long long j_w[3 * WC];
  __shared__ int                red[4 * 3 + 1][kWaves];
  // ghost-extended index (em) and active index (current) of every window
  // row/column, per dimension, after the periodic wrap
  __shared__ int gtab[3][WM];

  const int tid    = threadIdx.x;
  const int G      = a.ngh;
  const bool do_dep = a.charge != ZERO;
  int64_t    pbeg, pend;
  if (a.toff != nullptr) {
    // chunk ntiles is the tail [toff[ntiles], n): the dead particles a sort
    // put last, and any particles appended after that sort
    pbeg = a.toff[chunk];
    pend = (chunk + 1 < a.nchunks) ? static_cast<int64_t>(a.toff[chunk + 1]) : a.P.n;
  } else {
    pbeg = chunk * (static_cast<int64_t>(kThreads) * a.ppt);
    pend = pbeg + static_cast<int64_t>(kThreads) * a.ppt;
  }
  pend = min(pend, a.P.n);

  // ---- pass 1: bounding box of the live particles' cells, in the plain and
  // in the rotated frame; cells outside [0, n) send the chunk to the HBM
  // path, which reports them (pass 2 reads i and tag again; the re-read hits
  // in L2)
  int lo[3] = { INT_MAX, INT_MAX, INT_MAX }, hi[3] = { INT_MIN, INT_MIN, INT_MIN };
  int rlo[3] = { INT_MAX, INT_MAX, INT_MAX }, rhi[3] = { INT_MIN, INT_MIN, INT_MIN };
  int bad = 0;
  for (int64_t p = pbeg + tid; p < pend; p += kThreads) {
    const int16_t tag = a.P.tag[p];
    if (tag == NTT_TAG_ALIVE) {
      int c[3] = { a.P.i1[p], a.P.i2[p], 0 };
      if constexpr (D == 3) c[2] = a.P.i3[p];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (c[d] < 0 || c[d] >= a.n[d]) bad = 1;
        const int r = rotated(c[d], a.n[d]);
        lo[d]  = min(lo[d], c[d]);
        hi[d]  = max(hi[d], c[d]);
        rlo[d] = min(rlo[d], r);
        rhi[d] = max(rhi[d], r);
      }
    } else if (tag != NTT_TAG_DEAD) {
      raise_device_error(a.err, NTT_DEVERR_INVALID_TAG);
    }
  }
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const int l = wave_min(lo[d]), h = wave_max(hi[d]);
    const int rl = wave_min(rlo[d]), rh = wave_max(rhi[d]);
    if (lane == 0) {
      red[4 * d][wv]     = l;
      red[4 * d + 1][wv] = h;
      red[4 * d + 2][wv] = rl;
      red[4 * d + 3][wv] = rh;
    }
  }
  {
    const int b = wave_max(bad);
    if (lane == 0) red[12][wv] = b;
  }
  __syncthreads();
  int  ulo[3] = { 0, 0, 0 }, uhi[3] = { 0, 0, 0 };
  bool rot[3] = { false, false, false };
  bool any = true, fits = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int l = red[4 * d][0], h = red[4 * d + 1][0];
    int rl = red[4 * d + 2][0], rh = red[4 * d + 3][0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) {
      l  = min(l, red[4 * d][w]);
      h  = max(h, red[4 * d + 1][w]);
      rl = min(rl, red[4 * d + 2][w]);
      rh = max(rh, red[4 * d + 3][w]);
    }
    if (l > h) any = false;
    rot[d] = (rh - rl) < (h - l);
    const int hshift = a.n[d] >> 1;
    ulo[d] = rot[d] ? rl - hshift : l;
    uhi[d] = rot[d] ? rh - hshift : h;
  }
  if (!any) return;  // no live particle in this chunk (uniform across the block)
#pragma unroll
  for (int w = 0; w < kWaves; ++w)
    if (red[12][w]) fits = false;
  const int Wd[3] = { W1, W2, W3 };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (uhi[d] - ulo[d] + 4 > Wd[d]) fits = false;

  if (!fits) {
    // ---- HBM path -------------------------------------------------------------
#pragma unroll 1
    for (int64_t p = pbeg + tid; p < pend; p += kThreads) {
      if (a.P.tag[p] != NTT_TAG_ALIVE) continue;
      PState s;
      s.tag = NTT_TAG_ALIVE;
      load_state<D>(a.P, p, s);
      bool ok = true;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int g = s.i[d] + G;
        if (g - 1 < 0 || g + 2 >= a.n[d] + 2 * G || !(s.dx[d] >= ZERO && s.dx[d] <= ONE)) ok = false;
      }
      if (!ok) {
        raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
        continue;
      }
      sr_move<D, O, ALG>(a.F, G, a.ndth, a.dt, a.dxm, s);
      if (do_dep) {
        const float coeff = a.P.w[p] * a.charge * a.fscale;
        if (fabsf(coeff) <= a.big_coeff) {
          WrapJ js = a.J;
          deposit_particle<D, O>(js, G, s, coeff, a.inv_dt);
        } else {
          raise_device_error(a.err, NTT_DEVERR_CURRENT_RANGE);
        }
      }
      sr_wrap_periodic<D>(a.n, s);
      store_state<D>(a.P, p, s);
    }
    return;
  }

  // ---- window path ----------------------------------------------------------
  // window origin (ghost-extended, unwrapped frame): box low corner - 1
  const int o1 = ulo[0] - 1 + G, o2 = ulo[1] - 1 + G, o3 = (D == 3) ? ulo[2] - 1 + G : 0;
  const int e1 = uhi[0] - ulo[0] + 4, e2 = uhi[1] - ulo[1] + 4;
  const int e3 = (D == 3) ? uhi[2] - ulo[2] + 4 : 1;
  {
    const int oo[3] = { o1, o2, o3 }, ee[3] = { e1, e2, e3 };
    for (int idx = tid; idx < D * WM; idx += kThreads) {
      const int d = idx / WM, l = idx - d * WM;
      if (l < ee[d]) gtab[d][l] = wrap_any(oo[d] + l - G, a.n[d]);
    }
  }
  __syncthreads();
  // stage E, B of the box (periodic images of the active cells) and zero the
  // current window; rows of constant width W1 keep the index math to
  // constant divisions
  const int rows = e3 * W2 * W1;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    for (int idx = tid; idx < rows; idx += kThreads) {
      const int l3 = idx / (W1 * W2);
      const int r  = idx - l3 * (W1 * W2);
      const int l2 = r / W1, l1 = r - l2 * W1;
      if (l1 < e1 && l2 < e2) {
        const int g3 = (D == 3) ? gtab[2][l3] + G : 0;
        em_w[c * WC + idx] = a.F(c, gtab[0][l1] + G, gtab[1][l2] + G, g3);
      }
    }
  }
  for (int idx = tid; idx < 3 * WC; idx += kThreads) j_w[idx] = 0ull;
  __syncthreads();

  const LdsEM<D> F{ em_w, o1, o2, o3 };
  LdsJ<D>        jw{ j_w, o1, o2, o3 };
#pragma unroll 1
  for (int64_t p = pbeg + tid; p < pend; p += kThreads) {
    if (a.P.tag[p] != NTT_TAG_ALIVE) continue;
    PState s;
    s.tag = NTT_TAG_ALIVE;
    load_state<D>(a.P, p, s);
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (!(s.dx[d] >= ZERO && s.dx[d] <= ONE)) ok = false;
      if (rot[d] && s.i[d] >= a.n[d] - (a.n[d] >> 1)) s.i[d] -= a.n[d];
    }
    if (!ok) {
      raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
      continue;
    }
    const float w = a.P.w[p];
    sr_move<D, O, ALG>(F, G, a.ndth, a.dt, a.dxm, s);
    // exact deposit stencil inside the window?
    bool inwin = true;
    const int oo[3] = { o1, o2, o3 }, ee[3] = { e1, e2, e3 };
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int l, h;
      deposit_extent<O>(s.i[d], s.dx[d], s.ip[d], s.dxp[d], l, h);
      if (l + G < oo[d] || h + G >= oo[d] + ee[d]) inwin = false;
    }
    if (do_dep) {
      // coefficient times the fixed-point scale (exact): the sinks take
      // values in fixed-point units; the LDS sink's short conversion needs
      // every contribution below 2^31 units
      const float coeff = w * a.charge * a.fscale;
      if (inwin && fabsf(coeff) <= a.small_coeff) {
        deposit_particle<D, O>(jw, G, s, coeff, a.inv_dt);
      } else if (fabsf(coeff) <= a.big_coeff) {
        WrapJ js = a.J;
        deposit_particle<D, O>(js, G, s, coeff, a.inv_dt);
      } else {
        raise_device_error(a.err, NTT_DEVERR_CURRENT_RANGE);
      }
    }
    sr_wrap_periodic<D>(a.n, s);
    store_state<D>(a.P, p, s);
  }
  if (!do_dep) return;
  __syncthreads();
  // flush the current window onto the periodic image of the HBM accumulator
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    for (int idx = tid; idx < rows; idx += kThreads) {
      const int l3 = idx / (W1 * W2);
      const int r  = idx - l3 * (W1 * W2);
      const int l2 = r / W1, l1 = r - l2 * W1;
      if (l1 < e1 && l2 < e2) {
        const unsigned long long v = j_w[c * WC + idx];
        if (v != 0ull) {
          const int g3 = (D == 3) ? gtab[2][l3] + G : 0;
          a.J.add_fixed_at(a.J.linear(gtab[0][l1] + G, gtab[1][l2] + G, g3, c), v);
        }
      }
    }
  }
}

template <int D, int O, int ALG>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NTT_PD_WAVES_PER_EU)))
void k_push_deposit(FusedArgs a) {
  // With tile offsets the last chunk is the tail past the last tile (dead or
  // appended particles, usually none).  It takes the last block: the tiles
  // keep their XCD-aware mapping, and a box whose tiles fill the GPU exactly
  // once (two workgroups per CU) does not send a real tile to a second round
  // behind them.
  int64_t chunk;
  if (a.toff != nullptr) {
    const int64_t ntiles = a.nchunks - 1;
    chunk = (static_cast<int64_t>(blockIdx.x) == ntiles) ? ntiles : xcd_chunk(ntiles);
  } else {
    chunk = xcd_chunk(a.nchunks);
  }
  if (chunk < 0) return;
  push_deposit_chunk<D, O, ALG>(a, chunk);
}

// The same over the boxes of an ensemble: entry e owns the global chunks
// [chunk0[e], chunk0[e + 1]).
template <int D, int O, int ALG>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NTT_PD_WAVES_PER_EU)))
void k_push_deposit_batched(const FusedArgs* __restrict__ table,
                            const int64_t* __restrict__ chunk0, int nent, int64_t nchunks) {
  const int64_t g = xcd_chunk(nchunks);
  if (g < 0) return;
  int lo = 0, hi = nent - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunk0[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  // the entry is uniform across the workgroup: read its arguments in place
  // (scalar loads) rather than copying the block into registers
  const int e = __builtin_amdgcn_readfirstlane(lo);
  push_deposit_chunk<D, O, ALG>(table[e], g - chunk0[e]);
}

template <int D, int O, int ALG>
void launch_fused(const FusedArgs& a, hipStream_t st) {
  if (a.nchunks == 0) return;
  k_push_deposit<D, O, ALG><<<static_cast<unsigned>(a.nchunks), kThreads, 0, st>>>(a);
}

template <int D, int ALG>
void launch_fused_o(int order, const FusedArgs& a, hipStream_t st) {
  if (order == 0) launch_fused<D, 0, ALG>(a, st);
  else if (order == 1) launch_fused<D, 1, ALG>(a, st);
  else launch_fused<D, 2, ALG>(a, st);
}

}  // namespace

// Every deposit contribution of a particle is bounded by |charge * weight| *
// max(1/dt, 1) (Esirkepov: prefix sums of shape differences and the
// transverse weights are at most 1; zig-zag: displacements below one cell;
// out-of-plane: |v| < 1).  Contributions of particles with |coeff * scale|
// up to small_coeff therefore stay below 2^30 units (a factor 2 of margin
// below the 2^31 of the short conversion); the general conversion holds
// contributions below 2^62 units, i.e. |coeff * scale| below 2^32 small_coeff.
float fixed_point_small_coeff(float dt) { return 1073741824.0f / fmaxf(ONE / dt, ONE); }

float fixed_point_scale(float qw_max, float dt) {
  if (!(qw_max > 0.0f) || !std::isfinite(qw_max)) return 4294967296.0f;
  int k = 0;
  std::frexp(fixed_point_small_coeff(dt) / qw_max, &k);  // ratio in [2^(k-1), 2^k)
  k -= 1;
  k = std::max(-100, std::min(k, 100));
  return std::ldexp(1.0f, k);
}

size_t fused_args_size() { return sizeof(FusedArgs); }

// Arguments of the fused push + deposit of one species over a periodic box
// (the kernel argument block, written to `out`).
ntt_status build_fused_args(void* out, int64_t* nchunks, const ParticleView& pv,
                            const FieldView& em, const FieldView& cur, unsigned long long* acc,
                            int D, float ndth, float dt, float dxm, float charge, bool deposit,
                            float fscale, const uint32_t* tile_offsets, int64_t ntiles,
                            int ppt) {
  uint32_t* err = device_error_word();
  if (err == nullptr) return NTT_ERR_HIP;
  FusedArgs a;
  a.P = pv;
  a.F = GlobalEM{ em.d, em.N1, em.N2, em.cstride };
  a.J = WrapJ{ acc, cur.n1, cur.n2, (D == 3) ? cur.n3 : 1, cur.ngh, cur.N1, cur.N2, cur.cstride };
  a.ngh = em.ngh;
  a.n[0] = em.n1;
  a.n[1] = em.n2;
  a.n[2] = (D == 3) ? em.n3 : 1;
  a.ndth = ndth;
  a.dt = dt;
  a.dxm = dxm;
  a.charge = deposit ? charge : ZERO;
  a.inv_dt = ONE / dt;
  a.fscale = fscale;
  a.small_coeff = fixed_point_small_coeff(dt);
  a.big_coeff = a.small_coeff * 4294967296.0f;
  a.err = err;
  a.toff = tile_offsets;
  a.ppt = ppt < 1 ? 1 : ppt;
  if (pv.n == 0) {
    a.nchunks = 0;
  } else if (tile_offsets != nullptr) {
    a.nchunks = ntiles + 1;  // the tiles, then the tail past the last tile
  } else {
    const int64_t per = static_cast<int64_t>(kThreads) * a.ppt;
    a.nchunks = (pv.n + per - 1) / per;
  }
  if (a.nchunks > 0x7FFFFFFF) return NTT_ERR_UNSUPPORTED;
  *nchunks = a.nchunks;
  std::memcpy(out, &a, sizeof(a));
  return NTT_OK;
}

// Fused push + deposit of one species over a periodic box.  em must have its
// periodic ghosts filled; the folded current is added, as 64-bit fixed point
// in units of 1/fscale, to the active cells of `acc`, an integer array with
// the layout of the current field `cur` (the caller zeroes it first).
ntt_status launch_push_deposit_periodic(const ParticleView& pv, const FieldView& em,
                                        const FieldView& cur, unsigned long long* acc,
                                        int D, int order, int alg,
                                        float ndth, float dt, float dxm, float charge,
                                        bool deposit, float fscale,
                                        const uint32_t* tile_offsets,
                                        int64_t ntiles, int ppt, hipStream_t st) {
  if (pv.n == 0) return NTT_OK;
  FusedArgs a;
  int64_t   nch = 0;
  ntt_status s = build_fused_args(&a, &nch, pv, em, cur, acc, D, ndth, dt, dxm, charge, deposit,
                                  fscale, tile_offsets, ntiles, ppt);
  if (s != NTT_OK) return s;
  if (D == 2) {
    if (alg == ALG_BORIS) launch_fused_o<2, ALG_BORIS>(order, a, st);
    else if (alg == ALG_VAY) launch_fused_o<2, ALG_VAY>(order, a, st);
    else launch_fused<2, 0, ALG_PHOTON>(a, st);
  } else {
    if (alg == ALG_BORIS) launch_fused_o<3, ALG_BORIS>(order, a, st);
    else if (alg == ALG_VAY) launch_fused_o<3, ALG_VAY>(order, a, st);
    else launch_fused<3, 0, ALG_PHOTON>(a, st);
  }
  return check_launch();
}

namespace {
template <int D, int O, int ALG>
void launch_batched_t(const void* table, const int64_t* chunk0, int nent, int64_t nchunks,
                      hipStream_t st) {
  k_push_deposit_batched<D, O, ALG><<<static_cast<unsigned>(nchunks), kThreads, 0, st>>>(
    static_cast<const FusedArgs*>(table), chunk0, nent, nchunks);
}
template <int D, int ALG>
void launch_batched_o(int order, const void* table, const int64_t* chunk0, int nent,
                      int64_t nchunks, hipStream_t st) {
  if (order == 0) launch_batched_t<D, 0, ALG>(table, chunk0, nent, nchunks, st);
  else if (order == 1) launch_batched_t<D, 1, ALG>(table, chunk0, nent, nchunks, st);
  else launch_batched_t<D, 2, ALG>(table, chunk0, nent, nchunks, st);
}
}  // namespace

// One launch over the table entries (all of one dimension, shape order and
// pusher algorithm); entry e owns global chunks [chunk0[e], chunk0[e + 1]).
ntt_status launch_push_deposit_batched(int D, int order, int alg, const void* table,
                                       const int64_t* chunk0, int nent, int64_t nchunks,
                                       hipStream_t st) {
  if (nent == 0 || nchunks == 0) return NTT_OK;
  if (nchunks > 0x7FFFFFFF) return NTT_ERR_UNSUPPORTED;
  if (D == 2) {
    if (alg == ALG_BORIS) launch_batched_o<2, ALG_BORIS>(order, table, chunk0, nent, nchunks, st);
    else if (alg == ALG_VAY) launch_batched_o<2, ALG_VAY>(order, table, chunk0, nent, nchunks, st);
    else launch_batched_t<2, 0, ALG_PHOTON>(table, chunk0, nent, nchunks, st);
  } else {
    if (alg == ALG_BORIS) launch_batched_o<3, ALG_BORIS>(order, table, chunk0, nent, nchunks, st);
    else if (alg == ALG_VAY) launch_batched_o<3, ALG_VAY>(order, table, chunk0, nent, nchunks, st);
    else launch_batched_t<3, 0, ALG_PHOTON>(table, chunk0, nent, nchunks, st);
  }
  return check_launch();
}

}  // namespace ntt

// LDS window extents (cells per dimension) of the fused push + deposit kernel.
extern "C" ntt_status ntt_fused_window(int32_t dim, int32_t w[3]) {
  if (w == nullptr) return NTT_ERR_ARG;
  if (dim == 2) {
    w[0] = ntt::Win<2>::W1;
    w[1] = ntt::Win<2>::W2;
    w[2] = 1;
  } else if (dim == 3) {
    w[0] = ntt::Win<3>::W1;
    w[1] = ntt::Win<3>::W2;
    w[2] = ntt::Win<3>::W3;
  } else {
    return NTT_ERR_ARG;
  }
  return NTT_OK;
}
