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
  uint32_t*       flags; // split path: 2 bits per chunk (kFlagPush, kFlagDeposit), or null
  uint4*          recs;  // split path: per chunk, the box the push leaves (or null)
  int32_t         reuse; // split path: recs holds the records of this species' last push
};

// A chunk's box record, written by the split path's push for its deposit:
// the box of every live particle's cells before and after the push, in the
// push's frame (x, y, z: lo | hi << 16 as 16-bit integers; w: bit d set when
// dimension d is rotated, bit 30 when the box is in a frame where each cell
// has one representative (the next push can reuse it), bit 31 when the record
// is valid).
__device__ __forceinline__ uint4 pack_box(const int* lo, const int* hi, const bool* rot,
                                          bool valid) {
  uint4 r;
  r.x = (static_cast<uint32_t>(lo[0]) & 0xFFFFu) | (static_cast<uint32_t>(hi[0]) << 16);
  r.y = (static_cast<uint32_t>(lo[1]) & 0xFFFFu) | (static_cast<uint32_t>(hi[1]) << 16);
  r.z = (static_cast<uint32_t>(lo[2]) & 0xFFFFu) | (static_cast<uint32_t>(hi[2]) << 16);
  r.w = (rot[0] ? 1u : 0u) | (rot[1] ? 2u : 0u) | (rot[2] ? 4u : 0u) | (valid ? 0x80000000u : 0u);
  return r;
}

// Chunk flags of the split path (push and deposit in separate launches): a
// chunk whose box does not fit the push window is left to the fallback launch
// whole (kFlagPush); one pushed from the window whose deposit cannot run from
// the current window is left to it for the deposit only (kFlagDeposit).
constexpr uint32_t kFlagPush = 1u, kFlagDeposit = 2u;
// i1_prev of a particle that the window push left for the fallback launch
// (its cell lay outside a reused box): the deposit launch skips it, the
// fallback pushes and deposits it
constexpr int32_t kUnpushed = INT32_MIN;
constexpr int64_t split_flag_words(int64_t nchunks) { return (nchunks + 15) / 16; }
__device__ __forceinline__ uint32_t chunk_flags(const uint32_t* f, int64_t c) {
  return (f[c >> 4] >> (static_cast<uint32_t>(c & 15) * 2u)) & 3u;
}
__device__ __forceinline__ void set_chunk_flag(uint32_t* f, int64_t c, uint32_t bit) {
  atomicOr(f + (c >> 4), bit << (static_cast<uint32_t>(c & 15) * 2u));
}

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

// A chunk's particle arrays from its first particle on, addressed by a 32-bit
// byte offset from a base that is uniform over the workgroup: each access is
// then the scalar base plus one vector offset shared by every array, with no
// 64-bit address arithmetic per array.  Chunks of more than kMaxRel
// particles are left to the 64-bit paths.
constexpr int64_t kMaxRel = int64_t{ 1 } << 28;
template <class T>
__device__ __forceinline__ T* rel(T* base, uint32_t k) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + k * static_cast<uint32_t>(sizeof(T)));
}
__device__ __forceinline__ ParticleView rebase(const ParticleView& P, int64_t p0) {
  ParticleView R = P;
  R.i1 += p0;  R.i2 += p0;  R.i3 += p0;
  R.i1p += p0; R.i2p += p0; R.i3p += p0;
  R.dx1 += p0;  R.dx2 += p0;  R.dx3 += p0;
  R.dx1p += p0; R.dx2p += p0; R.dx3p += p0;
  R.ux1 += p0; R.ux2 += p0; R.ux3 += p0;
  R.w += p0;
  R.tag += p0;
  return R;
}
template <int D>
__device__ __forceinline__ void load_state_rel(const ParticleView& R, uint32_t k, PState& s) {
  s.i[0]  = *rel(R.i1, k);
  s.i[1]  = *rel(R.i2, k);
  s.i[2]  = (D == 3) ? *rel(R.i3, k) : 0;
  s.dx[0] = *rel(R.dx1, k);
  s.dx[1] = *rel(R.dx2, k);
  if constexpr (D == 3) s.dx[2] = *rel(R.dx3, k);
  s.u[0] = *rel(R.ux1, k);
  s.u[1] = *rel(R.ux2, k);
  s.u[2] = *rel(R.ux3, k);
}
template <int D>
__device__ __forceinline__ void store_state_rel(const ParticleView& R, uint32_t k,
                                                const PState& s) {
  *rel(R.i1p, k)  = s.ip[0];
  *rel(R.dx1p, k) = s.dxp[0];
  *rel(R.i1, k)   = s.i[0];
  *rel(R.dx1, k)  = s.dx[0];
  *rel(R.i2p, k)  = s.ip[1];
  *rel(R.dx2p, k) = s.dxp[1];
  *rel(R.i2, k)   = s.i[1];
  *rel(R.dx2, k)  = s.dx[1];
  if constexpr (D == 3) {
    *rel(R.i3p, k)  = s.ip[2];
    *rel(R.dx3p, k) = s.dxp[2];
    *rel(R.i3, k)   = s.i[2];
    *rel(R.dx3, k)  = s.dx[2];
  }
  *rel(R.ux1, k) = s.u[0];
  *rel(R.ux2, k) = s.u[1];
  *rel(R.ux3, k) = s.u[2];
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
// the cell before the push in [0, n): the periodic wrap that follows the push
// shifts i_prev by the same n as i, so it lies in [-1, n]
__device__ __forceinline__ int wrap_prev(int q, int n) {
  return q < 0 ? q + n : (q >= n ? q - n : q);
}

// Moves a box [lo, hi] of unwrapped cells (one dimension) into a frame in
// which each of its cells has one representative: the plain frame [0, n) or
// the rotated one [-(n/2), n - n/2), shifted by a multiple of n.  A cell c in
// [0, n) then maps into the box as the frame maps it (rotated: c - n for c >=
// n - n/2).  False when the box fits neither (wider than the frames allow).
__device__ __forceinline__ bool canonical_box(int& lo, int& hi, bool& rot, int n) {
  const int h = n >> 1;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const bool r = (k < 3) ? rot : !rot;
    const int  s = (k % 3 == 0) ? 0 : ((k % 3 == 1) ? -n : n);
    const int  l = lo + s, u = hi + s;
    if (r ? (l >= -h && u < n - h) : (l >= 0 && u < n)) {
      lo  = l;
      hi  = u;
      rot = r;
      return true;
    }
  }
  return false;
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

// Particle range [pbeg, pend) of a chunk: one tile of the last sort (the last
// chunk is the tail [toff[ntiles], n): the dead particles a sort put last,
// and any particles appended after that sort), else a fixed run of particles.
__device__ __forceinline__ void chunk_range(const FusedArgs& a, int64_t chunk, int64_t& pbeg,
                                            int64_t& pend) {
  if (a.toff != nullptr) {
    pbeg = a.toff[chunk];
    pend = (chunk + 1 < a.nchunks) ? static_cast<int64_t>(a.toff[chunk + 1]) : a.P.n;
  } else {
    pbeg = chunk * (static_cast<int64_t>(kThreads) * a.ppt);
    pend = pbeg + static_cast<int64_t>(kThreads) * a.ppt;
  }
  pend = min(pend, a.P.n);
}

// Cell box of a chunk's live particles and the frame it is taken in.
struct ChunkBox {
  int  lo[3], hi[3];  // box in the chosen frame (cells, without ghosts)
  bool rot[3];        // frame rotated by n/2 in that dimension
  bool any;           // some particle is alive
  bool fits;          // box + [-1, +2] fits the LDS window (and nothing below disqualifies it)
};

// Pass 1 of a window kernel: the box of the live particles' cells, in the
// plain and in the rotated frame, keeping the smaller.  Cells outside [0, n)
// disqualify the window (the HBM path reports them).  PREV (particles
// already pushed): the box spans the cells before and after the push, placed
// as the push placed them: the frame is taken at the cell before the push
// (i_prev unwrapped into [0, n)) and the cell after it is i_prev's plus the
// move (the periodic wrap shifts i and i_prev together).  `red` is a shared
// [13][kWaves] scratch.
template <int D, bool PREV>
__device__ __forceinline__ ChunkBox chunk_box(const FusedArgs& a, int64_t pbeg, int64_t pend,
                                              int (*red)[kWaves]) {
  const int tid = threadIdx.x;
  int lo[3] = { INT_MAX, INT_MAX, INT_MAX }, hi[3] = { INT_MIN, INT_MIN, INT_MIN };
  int rlo[3] = { INT_MAX, INT_MAX, INT_MAX }, rhi[3] = { INT_MIN, INT_MIN, INT_MIN };
  int bad = 0;
  // two particles per iteration, branch-free: all their loads are in flight
  // together, and a particle that is not alive (or past the end: it reads
  // the last particle again) leaves the box unchanged
  for (int64_t p0 = pbeg + tid; p0 < pend; p0 += 2 * kThreads) {
    int16_t tag[2];
    bool    in[2];
    int     c[2][3], q[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int64_t p = min(p0 + u * kThreads, pend - 1);
      in[u]   = p0 + u * kThreads < pend;
      tag[u]  = a.P.tag[p];
      c[u][0] = a.P.i1[p];
      c[u][1] = a.P.i2[p];
      c[u][2] = (D == 3) ? a.P.i3[p] : 0;
      q[u][0] = q[u][1] = q[u][2] = 0;
      if constexpr (PREV) {
        q[u][0] = a.P.i1p[p];
        q[u][1] = a.P.i2p[p];
        if constexpr (D == 3) q[u][2] = a.P.i3p[p];
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const bool alive = in[u] && tag[u] == NTT_TAG_ALIVE && (!PREV || q[u][0] != kUnpushed);
      if (!PREV && in[u] && !alive && tag[u] != NTT_TAG_DEAD)
        raise_device_error(a.err, NTT_DEVERR_INVALID_TAG);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        int cd = c[u][d], dl = 0, dh = 0;
        if constexpr (PREV) {
          const int dd = cd - q[u][d];  // the move
          cd = wrap_prev(q[u][d], a.n[d]);
          dl = min(dd, 0);
          dh = max(dd, 0);
        }
        const int r = rotated(cd, a.n[d]);
        bad |= (alive && static_cast<uint32_t>(cd) >= static_cast<uint32_t>(a.n[d])) ? 1 : 0;
        lo[d]  = min(lo[d], alive ? cd + dl : INT_MAX);
        hi[d]  = max(hi[d], alive ? cd + dh : INT_MIN);
        rlo[d] = min(rlo[d], alive ? r + dl : INT_MAX);
        rhi[d] = max(rhi[d], alive ? r + dh : INT_MIN);
      }
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
  ChunkBox bx;
  bx.any  = true;
  bx.fits = true;
  constexpr int Wd[3] = { Win<D>::W1, Win<D>::W2, Win<D>::W3 };
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    bx.lo[d]  = 0;
    bx.hi[d]  = 0;
    bx.rot[d] = false;
  }
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
    if (l > h) bx.any = false;
    bx.rot[d] = (rh - rl) < (h - l);
    const int hshift = a.n[d] >> 1;
    bx.lo[d] = bx.rot[d] ? rl - hshift : l;
    bx.hi[d] = bx.rot[d] ? rh - hshift : h;
    if (bx.hi[d] - bx.lo[d] + 4 > Wd[d]) bx.fits = false;
  }
#pragma unroll
  for (int w = 0; w < kWaves; ++w)
    if (red[12][w]) bx.fits = false;
  return bx;
}

// LDS window of a chunk: ghost-extended origin of box - 1 (unwrapped frame)
// and extent box + 4 per dimension, with the wrapped active index of every
// window row/column per dimension in `gtab` (each row has WM entries).
template <int D>
struct WinGeom {
  int o[3], e[3];
};
template <int D>
__device__ __forceinline__ WinGeom<D> window_geometry(const FusedArgs& a, const ChunkBox& bx,
                                                      int* gtab, int WM) {
  const int G = a.ngh;
  WinGeom<D> g;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g.o[d] = (d < D) ? bx.lo[d] - 1 + G : 0;
    g.e[d] = (d < D) ? bx.hi[d] - bx.lo[d] + 4 : 1;
  }
  // one dimension at a time: constant indices keep g in registers
#pragma unroll
  for (int d = 0; d < D; ++d)
    for (int l = threadIdx.x; l < g.e[d]; l += kThreads)
      gtab[d * WM + l] = wrap_any(g.o[d] + l - G, a.n[d]);
  return g;
}

// Stages E, B of the window (periodic images of the active cells) into LDS;
// rows of constant width W1 keep the index math to constant divisions, and
// each window cell's source offset is computed once for its six components,
// whose loads are then in flight together.
template <int D>
__device__ __forceinline__ void stage_em(const FusedArgs& a, const WinGeom<D>& g,
                                         const int* gtab, int WM, float* em_w) {
  constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, WC = W1 * W2 * Win<D>::W3;
  const int     G    = a.ngh;
  const int     rows = g.e[2] * W2 * W1;
  const int64_t cs   = a.F.cstride;
  for (int idx = threadIdx.x; idx < rows; idx += kThreads) {
    const int l3 = idx / (W1 * W2);
    const int r  = idx - l3 * (W1 * W2);
    const int l2 = r / W1, l1 = r - l2 * W1;
    if (l1 < g.e[0] && l2 < g.e[1]) {
      const int    g3  = (D == 3) ? gtab[2 * WM + l3] + G : 0;
      const float* src = a.F.d + (static_cast<int64_t>(gtab[l1] + G) +
                                  static_cast<int64_t>(a.F.N1) *
                                    (static_cast<int64_t>(gtab[WM + l2] + G) +
                                     static_cast<int64_t>(a.F.N2) * g3));
      float v[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) v[c] = src[c * cs];
#pragma unroll
      for (int c = 0; c < 6; ++c) em_w[c * WC + idx] = v[c];
    }
  }
}

// Adds the current window to the periodic image of the HBM accumulator, one
// 64-bit integer atomic per non-zero entry (the target offset computed once
// per window cell for its three components).
template <int D>
__device__ __forceinline__ void flush_j(const FusedArgs& a, const WinGeom<D>& g, const int* gtab,
                                        int WM, const unsigned long long* j_w) {
  constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, WC = W1 * W2 * Win<D>::W3;
  const int G = a.ngh;
  const int rows = g.e[2] * W2 * W1;
  for (int idx = threadIdx.x; idx < rows; idx += kThreads) {
    const int l3 = idx / (W1 * W2);
    const int r  = idx - l3 * (W1 * W2);
    const int l2 = r / W1, l1 = r - l2 * W1;
    if (l1 < g.e[0] && l2 < g.e[1]) {
      unsigned long long v[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = j_w[c * WC + idx];
      if ((v[0] | v[1] | v[2]) != 0ull) {
        const int     g3  = (D == 3) ? gtab[2 * WM + l3] + G : 0;
        const int64_t lin = a.J.linear(gtab[l1] + G, gtab[WM + l2] + G, g3, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          if (v[c] != 0ull) a.J.add_fixed_at(lin + a.J.cstride * c, v[c]);
      }
    }
  }
}

// Push (and deposit) of a chunk's particles straight from HBM: gather from
// the ghost-extended em array, deposit with global atomics (periodic fold per
// index).  For chunks whose box does not fit the window; MARKED: only the
// particles the window push left (i1_prev == kUnpushed).
template <int D, int O, int ALG, bool MARKED = false>
__device__ __forceinline__ void hbm_push_deposit(const FusedArgs& a, int64_t pbeg, int64_t pend) {
  const int  G      = a.ngh;
  const bool do_dep = a.charge != ZERO;
#pragma unroll 1
  for (int64_t p = pbeg + threadIdx.x; p < pend; p += kThreads) {
    if (a.P.tag[p] != NTT_TAG_ALIVE) continue;
    if (MARKED && a.P.i1p[p] != kUnpushed) continue;
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
}

// Deposit of a chunk's pushed particles straight to HBM (global atomics,
// periodic fold per index), from the state the push stored.  Particles whose
// offset is outside [0, 1] were skipped (and reported) by the push, and those
// it left for the fallback (kUnpushed) are not pushed yet.
template <int D, int O>
__device__ __forceinline__ void hbm_deposit(const FusedArgs& a, int64_t pbeg, int64_t pend) {
  const int G = a.ngh;
#pragma unroll 1
  for (int64_t p = pbeg + threadIdx.x; p < pend; p += kThreads) {
    if (a.P.tag[p] != NTT_TAG_ALIVE || a.P.i1p[p] == kUnpushed) continue;
    PState s;
    s.tag   = NTT_TAG_ALIVE;
    s.i[0]  = a.P.i1[p];
    s.i[1]  = a.P.i2[p];
    s.i[2]  = (D == 3) ? a.P.i3[p] : 0;
    s.ip[0] = a.P.i1p[p];
    s.ip[1] = a.P.i2p[p];
    s.ip[2] = (D == 3) ? a.P.i3p[p] : 0;
    s.dx[0] = a.P.dx1[p];
    s.dx[1] = a.P.dx2[p];
    s.dx[2] = (D == 3) ? a.P.dx3[p] : ZERO;
    s.dxp[0] = a.P.dx1p[p];
    s.dxp[1] = a.P.dx2p[p];
    s.dxp[2] = (D == 3) ? a.P.dx3p[p] : ZERO;
    s.u[0] = s.u[1] = s.u[2] = ZERO;
    if constexpr (D == 2) {
      s.u[0] = a.P.ux1[p];
      s.u[1] = a.P.ux2[p];
      s.u[2] = a.P.ux3[p];
    }
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d)
      if (!(s.dx[d] >= ZERO && s.dx[d] <= ONE)) ok = false;
    if (!ok) continue;
    const float coeff = a.P.w[p] * a.charge * a.fscale;
    if (fabsf(coeff) <= a.big_coeff) {
      WrapJ js = a.J;
      deposit_particle<D, O>(js, G, s, coeff, a.inv_dt);
    } else {
      raise_device_error(a.err, NTT_DEVERR_CURRENT_RANGE);
    }
  }
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
  __shared__ int                gtab[3 * WM];

  const int  tid    = threadIdx.x;
  const int  G      = a.ngh;
  const bool do_dep = a.charge != ZERO;
  int64_t    pbeg, pend;
  chunk_range(a, chunk, pbeg, pend);
  const ChunkBox bx = chunk_box<D, false>(a, pbeg, pend, red);
  if (!bx.any) return;  // no live particle in this chunk (uniform across the block)
  if (!bx.fits) {
    hbm_push_deposit<D, O, ALG>(a, pbeg, pend);
    return;
  }

  // ---- window path ----------------------------------------------------------
  const WinGeom<D> g = window_geometry<D>(a, bx, gtab, WM);
  __syncthreads();
  stage_em<D>(a, g, gtab, WM, em_w);
  for (int idx = tid; idx < 3 * WC; idx += kThreads) j_w[idx] = 0ull;
  __syncthreads();

  const LdsEM<D> F{ em_w, g.o[0], g.o[1], g.o[2] };
  LdsJ<D>        jw{ j_w, g.o[0], g.o[1], g.o[2] };
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
      if (bx.rot[d] && s.i[d] >= a.n[d] - (a.n[d] >> 1)) s.i[d] -= a.n[d];
    }
    if (!ok) {
      raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
      continue;
    }
    const float w = a.P.w[p];
    sr_move<D, O, ALG>(F, G, a.ndth, a.dt, a.dxm, s);
    // exact deposit stencil inside the window?
    bool inwin = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      int l, h;
      deposit_extent<O>(s.i[d], s.dx[d], s.ip[d], s.dxp[d], l, h);
      if (l + G < g.o[d] || h + G >= g.o[d] + g.e[d]) inwin = false;
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
  flush_j<D>(a, g, gtab, WM, j_w);
}

// Chunk of a workgroup with tile offsets: the tiles in XCD-aware order, and
// the tail past the last tile (dead or appended particles, usually none) in
// the last block, so that a box whose tiles fill the GPU exactly once (two
// workgroups per CU) does not send a real tile to a second round behind it.
__device__ __forceinline__ int64_t tile_chunk(const FusedArgs& a) {
  if (a.toff != nullptr) {
    const int64_t ntiles = a.nchunks - 1;
    return (static_cast<int64_t>(blockIdx.x) == ntiles) ? ntiles : xcd_chunk(ntiles);
  }
  return xcd_chunk(a.nchunks);
}

template <int D, int O, int ALG>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NTT_PD_WAVES_PER_EU)))
void k_push_deposit(FusedArgs a) {
  const int64_t chunk = tile_chunk(a);
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

// ---- split path: push and deposit in separate launches ----------------------
// Each kernel keeps one LDS window and the registers of its half of the work,
// so the push runs at eight waves per SIMD (64 registers) and the deposit at
// six (its LDS window and list admit three workgroups per CU), where the fused
// kernel runs at four.
#ifndef NTT_SPLIT_PUSH_WAVES
#define NTT_SPLIT_PUSH_WAVES 8
#endif
#ifndef NTT_SPLIT_DEP_WAVES
#define NTT_SPLIT_DEP_WAVES 6
#endif

// Push of one chunk from its E/B window; the next particle's state is loaded
// while the current one is pushed.  A chunk whose box does not fit is left
// to the fallback launch (kFlagPush).
template <int D, int O, int ALG>
__device__ __forceinline__ void push_chunk(const FusedArgs& a, int64_t chunk) {
  constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, W3 = Win<D>::W3;
  constexpr int WM = (W1 > W2 ? W1 : W2) > W3 ? (W1 > W2 ? W1 : W2) : W3;
  __shared__ float    em_w[6 * W1 * W2 * W3];
  __shared__ int      red[4 * 3 + 1][kWaves];
  __shared__ int      gtab[3 * WM];
  __shared__ uint32_t moved, skipped;
  const int G = a.ngh;
  int64_t   pbeg, pend;
  chunk_range(a, chunk, pbeg, pend);
  // the box: this species' record of the previous push when it fits the
  // window (a superset of the cells now: the cells before and after that
  // push), else the box pass.  A particle outside a reused box is left to
  // the fallback launch, so a stale record costs time, never correctness.
  ChunkBox bx;
  bool     reuse = false;
  if (a.reuse && pend - pbeg <= kMaxRel) {
    const uint4 r = a.recs[chunk];
    if ((r.w >> 30) == 3u) {
      const uint32_t xs[3] = { r.x, r.y, r.z };
      bx.any  = true;
      bx.fits = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bx.lo[d]  = static_cast<int16_t>(xs[d] & 0xFFFFu);
        bx.hi[d]  = static_cast<int16_t>(xs[d] >> 16);
        bx.rot[d] = (r.w >> d) & 1u;
        if (d < D && (bx.hi[d] - bx.lo[d] + 4 > Win<D>::W[d] || bx.lo[d] < -a.n[d] ||
                      bx.hi[d] >= 2 * a.n[d]))
          bx.fits = false;
      }
      reuse = bx.fits;
    }
  }
  if (!reuse) bx = chunk_box<D, false>(a, pbeg, pend, red);
  const bool fits = bx.fits && pend - pbeg <= kMaxRel;
  if (!bx.any || !fits) {
    if (threadIdx.x == 0) {
      if (a.recs != nullptr) a.recs[chunk] = pack_box(bx.lo, bx.hi, bx.rot, false);
      if (!fits) set_chunk_flag(a.flags, chunk, kFlagPush);
    }
    return;
  }
  if (threadIdx.x == 0) {
    moved   = 0u;
    skipped = 0u;
  }
  const WinGeom<D> g = window_geometry<D>(a, bx, gtab, WM);
  __syncthreads();
  stage_em<D>(a, g, gtab, WM, em_w);
  __syncthreads();
  const LdsEM<D>     F{ em_w, g.o[0], g.o[1], g.o[2] };
  const ParticleView R   = rebase(a.P, pbeg);
  const uint32_t     cnt = static_cast<uint32_t>(pend - pbeg);
  int16_t ntag = NTT_TAG_DEAD;
  PState  nx;
  if (threadIdx.x < cnt) {
    ntag = *rel(R.tag, threadIdx.x);
    load_state_rel<D>(R, threadIdx.x, nx);
  }
#pragma unroll 1
  for (uint32_t k = threadIdx.x; k < cnt; k += kThreads) {
    const int16_t tag = ntag;
    PState        s   = nx;
    {
      const uint32_t q = k + kThreads;
      if (q < cnt) {
        ntag = *rel(R.tag, q);
        load_state_rel<D>(R, q, nx);
      }
    }
    if (tag != NTT_TAG_ALIVE) {
      if (reuse && tag != NTT_TAG_DEAD) raise_device_error(a.err, NTT_DEVERR_INVALID_TAG);
      continue;
    }
    s.tag   = NTT_TAG_ALIVE;
    bool ok = true, in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (!(s.dx[d] >= ZERO && s.dx[d] <= ONE)) ok = false;
      // a cell outside [0, n) (the box pass sends its chunk to the HBM
      // path, which reports it) is left to the fallback with a reused box
      if (static_cast<uint32_t>(s.i[d]) >= static_cast<uint32_t>(a.n[d])) in = false;
      if (bx.rot[d] && s.i[d] >= a.n[d] - (a.n[d] >> 1)) s.i[d] -= a.n[d];
    }
    if (!ok) {
      raise_device_error(a.err, NTT_DEVERR_OUT_OF_BOUNDS);
      continue;
    }
    if (reuse) {
#pragma unroll
      for (int d = 0; d < D; ++d)
        if (s.i[d] < bx.lo[d] || s.i[d] > bx.hi[d]) in = false;
      if (!in) {
        *rel(R.i1p, k) = kUnpushed;
        skipped        = 1u;
        continue;
      }
    }
    sr_move<D, O, ALG>(F, G, a.ndth, a.dt, a.dxm, s);
    // bit 2d: the particle left the box below in dimension d, 2d + 1: above
    // (rare: one LDS atomic, no state carried between particles)
    uint32_t ext = 0u;
#pragma unroll
    for (int d = 0; d < D; ++d)
      ext |= ((s.i[d] < bx.lo[d]) ? 1u : 0u) << (2 * d) | ((s.i[d] > bx.hi[d]) ? 2u : 0u) << (2 * d);
    if (ext != 0u) atomicOr(&moved, ext);
    sr_wrap_periodic<D>(a.n, s);
    store_state_rel<D>(R, k, s);
  }
  if (a.recs == nullptr) return;
  // the box of the cells before and after the push (a particle moves at most
  // one cell), for the deposit launch; a chunk with particles left outside
  // the box goes to the fallback launch for those (kFlagPush, valid record)
  __syncthreads();
  if (threadIdx.x == 0) {
    if (skipped != 0u) set_chunk_flag(a.flags, chunk, kFlagPush);
    int  lo[3], hi[3];
    bool rot[3], canon = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d]  = bx.lo[d] - static_cast<int>((moved >> (2 * d)) & 1u);
      hi[d]  = bx.hi[d] + static_cast<int>((moved >> (2 * d + 1)) & 1u);
      rot[d] = bx.rot[d];
      if (d < D && !canonical_box(lo[d], hi[d], rot[d], a.n[d])) canon = false;
    }
    uint4 r = pack_box(lo, hi, rot, true);
    if (canon) r.w |= 0x40000000u;
    a.recs[chunk] = r;
  }
}

// Deposit of one pushed chunk into an LDS current window.  The window spans
// the cells of every live particle before and after the push (pass 1 reads
// tag, i and i_prev), so every stencil fits.  The particles are then taken
// kThreads at a time, one per thread: each is loaded once, reduced to a
// 32-byte record (cells relative to the window, offsets before and after
// the push, coefficient) and written to an LDS list, the particles whose
// stencil stays along every direction from the front and the others from
// the back; after a barrier every thread deposits one record of the list,
// so that most wavefronts run the fixed 3 (O + 1)^2 O-slot stencil without
// per-slot branches.  A chunk whose box does not fit, or with a particle too
// heavy for the short fixed-point conversion, is left to the fallback launch
// (kFlagDeposit, the window discarded), as is one whose push was (kFlagPush).
// Particles with an offset outside [0, 1] were skipped (and reported) by the
// push and are skipped here.
struct DepRecord {
  uint4 a, b;  // a: packed cells, dx1..3; b: dxp1..3, coefficient
};
// Order of the deposit list: the particles of one bank class (stencil base
// slot mod 32) take ranks 0, 1, ... within their class (ranks from the last
// level of a class on share it).  Classes: the stay group (kRanks levels),
// then the particles whose stencil shifts along x only, y only, z only, and
// along several directions (kCrossRanks levels each), so that the wavefronts
// of the general stencil mostly share one shape and skip the same slots.
constexpr int kRanks      = 32;
constexpr int kCrossRanks = 8;
constexpr int kGroups     = kRanks + 4 * kCrossRanks;
static_assert(kGroups == 64, "one wavefront scans the groups");

template <int D, int O>
__device__ __forceinline__ void deposit_chunk(const FusedArgs& a, int64_t chunk) {
  static_assert(D == 3, "split deposit: 3D");
  constexpr int W1 = Win<D>::W1, W2 = Win<D>::W2, W3 = Win<D>::W3;
  constexpr int WC = W1 * W2 * W3;
  constexpr int WM = (W1 > W2 ? W1 : W2) > W3 ? (W1 > W2 ? W1 : W2) : W3;
  static_assert(WM <= 32, "window cells are packed in 5 bits");
  __shared__ unsigned long long j_w[3 * WC];
  __shared__ int                red[4 * 3 + 1][kWaves];
  __shared__ int                gtab[3 * WM];
  __shared__ DepRecord          list[kThreads];
  // per segment: particles per (class, bank class), per group, the groups'
  // fill pointers, the counts
  __shared__ uint32_t           ctr[5][32];
  __shared__ uint32_t           gsz[kGroups];
  __shared__ uint32_t           gpos[kGroups];
  __shared__ uint32_t           nstay_s, nall_s, heavy;
  const int tid = threadIdx.x, lane = tid & 63;
  const int G   = a.ngh;
  const uint4 r = (a.recs != nullptr) ? a.recs[chunk] : make_uint4(0u, 0u, 0u, 0u);
  // a chunk the push left whole (no valid record) is the fallback's; one it
  // left in part is deposited here without the particles it left
  if ((chunk_flags(a.flags, chunk) & kFlagPush) && !(r.w >> 31)) return;
  int64_t   pbeg, pend;
  chunk_range(a, chunk, pbeg, pend);
  // the box from the push's record when it has one that fits, else pass 1
  ChunkBox bx;
  bool     have = false;
  if (a.recs != nullptr) {
    if (r.w >> 31) {
      const uint32_t xs[3] = { r.x, r.y, r.z };
      bx.any  = true;
      bx.fits = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        bx.lo[d]  = static_cast<int16_t>(xs[d] & 0xFFFFu);
        bx.hi[d]  = static_cast<int16_t>(xs[d] >> 16);
        bx.rot[d] = (r.w >> d) & 1u;
        if (bx.hi[d] - bx.lo[d] + 4 > Win<D>::W[d]) bx.fits = false;
      }
      have = bx.fits;
    }
  }
  if (!have) bx = chunk_box<D, true>(a, pbeg, pend, red);
  if (!bx.any) return;
  if (!bx.fits || pend - pbeg > kMaxRel) {
    if (tid == 0) set_chunk_flag(a.flags, chunk, kFlagDeposit);
    return;
  }
  const WinGeom<D> g = window_geometry<D>(a, bx, gtab, WM);
  for (int idx = tid; idx < 3 * WC; idx += kThreads) j_w[idx] = 0ull;
  if (tid == 0) heavy = 0u;
  LdsJ<D>            jw{ j_w, g.o[0], g.o[1], g.o[2] };
  const ParticleView R   = rebase(a.P, pbeg);
  const uint32_t     cnt = static_cast<uint32_t>(pend - pbeg);
#pragma unroll 1
  for (uint32_t sb = 0; sb < cnt; sb += kThreads) {
    // (the previous segment's list and counts are read before the barrier
    // after the next classification, and the group counters are not read
    // by its deposits)
    for (int k = tid; k < 160 + kGroups; k += kThreads) {
      if (k < 160) ctr[k >> 5][k & 31] = 0u;
      else gsz[k - 160] = 0u;
    }
    __syncthreads();
    // load one particle per thread; its record, class and the bank class of
    // its stencil base (the window's 64-bit slot of the lowest cell, mod 32)
    const uint32_t k    = sb + tid;
    bool           live = false, stay = true;
    uint32_t       shift = 0u;  // bit d: the stencil shifts along d
    DepRecord      r;
    int            grp = 0;
    if (k < cnt && *rel(R.tag, k) == NTT_TAG_ALIVE) {
      const int   c[3]  = { *rel(R.i1, k), *rel(R.i2, k), *rel(R.i3, k) };
      const int   cp[3] = { *rel(R.i1p, k), *rel(R.i2p, k), *rel(R.i3p, k) };
      const float x[3]  = { *rel(R.dx1, k), *rel(R.dx2, k), *rel(R.dx3, k) };
      const float xp[3] = { *rel(R.dx1p, k), *rel(R.dx2p, k), *rel(R.dx3p, k) };
      const float coeff = *rel(R.w, k) * a.charge * a.fscale;
      live = cp[0] != kUnpushed;
      uint32_t cells = 0u;
      int      slot = 0, stride = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        if (!(x[d] >= ZERO && x[d] <= ONE)) live = false;
        if (!deposit_stays<O>(c[d], x[d], cp[d], xp[d])) shift |= 1u << d;
        // the push's frame: taken at the cell before the push, unwrapped
        const int q  = wrap_prev(cp[d], a.n[d]);
        const int fp = (bx.rot[d] && q >= a.n[d] - (a.n[d] >> 1)) ? q - a.n[d] : q;
        const int f  = fp + c[d] - cp[d];
        int lo, hi;
        deposit_extent<O>(f, x[d], fp, xp[d], lo, hi);
        slot += (lo + G - g.o[d]) * stride;
        stride *= Win<D>::W[d];
        cells |= (static_cast<uint32_t>(f + G - g.o[d]) |
                  (static_cast<uint32_t>(cp[d] - c[d] + 1) << 5)) << (7 * d);
      }
      if (live && !(fabsf(coeff) <= a.small_coeff)) {
        atomicAdd(&heavy, 1u);
        live = false;
      }
      r.a = make_uint4(cells, __float_as_uint(x[0]), __float_as_uint(x[1]),
                       __float_as_uint(x[2]));
      r.b = make_uint4(__float_as_uint(xp[0]), __float_as_uint(xp[1]), __float_as_uint(xp[2]),
                       __float_as_uint(coeff));
      stay = shift == 0u;
      if (live) {
        // 0: stay; 1, 2, 3: shifts along x, y, z only; 4: along several
        const int cls = stay ? 0 : (shift == 1u ? 1 : (shift == 2u ? 2 : (shift == 4u ? 3 : 4)));
        const uint32_t rank = atomicAdd(&ctr[cls][slot & 31], 1u);
        grp = stay ? static_cast<int>(min(rank, static_cast<uint32_t>(kRanks - 1)))
                   : kRanks + (cls - 1) * kCrossRanks +
                       static_cast<int>(min(rank, static_cast<uint32_t>(kCrossRanks - 1)));
        atomicAdd(&gsz[grp], 1u);
      }
    }
    __syncthreads();
    // group starts: exclusive scan over the groups, stay before cross
    if (tid < kGroups) {
      const uint32_t v = gsz[tid];
      uint32_t       incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      gpos[tid] = incl - v;
      if (tid == kRanks - 1) nstay_s = incl;
      if (tid == 63) nall_s = incl;
    }
    __syncthreads();
    // list order: group by group, each group in the order its particles
    // arrived.  A rank group holds at most one particle per bank class, so a
    // wavefront over two full groups puts two lanes on each bank pair, as
    // consecutive addresses do (a few lanes may repeat an address).  Measured
    // on 512 x 256 x 256, order 2, per launch: with a box pass, 18.2 ms
    // against 23.8 ms in particle order, and 19.1 ms when ranked by stencil
    // slot first (no repeated addresses, one more atomic); with the box from
    // the push's record, 16.1 ms, and 17.2 ms with each group placed in
    // bank-class order
    if (live) list[atomicAdd(&gpos[grp], 1u)] = r;
    __syncthreads();
    const int ns = static_cast<int>(nstay_s), nall = static_cast<int>(nall_s);
    if (tid < nall) {
      const bool      st = tid < ns;
      const DepRecord e  = list[tid];
      PState          s;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const uint32_t v = e.a.x >> (7 * d);
        s.i[d]  = static_cast<int>(v & 31u) + g.o[d] - G;
        s.ip[d] = s.i[d] + static_cast<int>((v >> 5) & 3u) - 1;
      }
      s.dx[0]  = __uint_as_float(e.a.y);
      s.dx[1]  = __uint_as_float(e.a.z);
      s.dx[2]  = __uint_as_float(e.a.w);
      s.dxp[0] = __uint_as_float(e.b.x);
      s.dxp[1] = __uint_as_float(e.b.y);
      s.dxp[2] = __uint_as_float(e.b.z);
      const float coeff = __uint_as_float(e.b.w);
      if (st) deposit_esirkepov_3d_stay<O>(jw, G, s, coeff, a.inv_dt);
      else deposit_particle<D, O>(jw, G, s, coeff, a.inv_dt);
    }
  }
  __syncthreads();
  if (heavy != 0u) {
    // a heavy particle: the fallback deposits the whole chunk from HBM
    if (tid == 0) set_chunk_flag(a.flags, chunk, kFlagDeposit);
    return;
  }
  flush_j<D>(a, g, gtab, WM, j_w);
}

template <int D, int O, int ALG>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NTT_SPLIT_PUSH_WAVES)))
void k_push_window(FusedArgs a) {
  const int64_t chunk = tile_chunk(a);
  if (chunk < 0) return;
  push_chunk<D, O, ALG>(a, chunk);
}

template <int D, int O>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NTT_SPLIT_DEP_WAVES)))
void k_deposit_window(FusedArgs a) {
  const int64_t chunk = tile_chunk(a);
  if (chunk < 0) return;
  deposit_chunk<D, O>(a, chunk);
}

// The chunks the window kernels left: push and deposit from HBM (kFlagPush:
// the whole chunk, or with a valid record only the particles the push left
// outside its reused box), or the deposit of the pushed particles
// (kFlagDeposit).  One workgroup per flag word (16 chunks), which exits at
// once when the word is clear.
template <int D, int O, int ALG>
__global__ __launch_bounds__(kThreads) void k_split_fallback(FusedArgs a) {
  const uint32_t word = a.flags[blockIdx.x];
  if (word == 0u) return;
#pragma unroll 1
  for (int k = 0; k < 16; ++k) {
    const uint32_t f = (word >> (2 * k)) & 3u;
    if (f == 0u) continue;
    const int64_t chunk = static_cast<int64_t>(blockIdx.x) * 16 + k;
    int64_t pbeg, pend;
    chunk_range(a, chunk, pbeg, pend);
    if (!(f & kFlagPush)) {
      if (a.charge != ZERO) hbm_deposit<D, O>(a, pbeg, pend);
    } else if (a.recs == nullptr || !(a.recs[chunk].w >> 31)) {
      hbm_push_deposit<D, O, ALG>(a, pbeg, pend);  // the whole chunk
    } else {
      // the particles the window push left outside a reused box; the others
      // too when the deposit launch left the chunk (kFlagDeposit)
      if ((f & kFlagDeposit) && a.charge != ZERO) hbm_deposit<D, O>(a, pbeg, pend);
      hbm_push_deposit<D, O, ALG, true>(a, pbeg, pend);
    }
  }
}

__global__ void k_zero_words(uint32_t* __restrict__ w, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0u;
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

// Split path of a 3D species (orders 1, 2, massive pushers): clear the chunk
// flags, push from the E/B windows, deposit into the current windows, then
// run what the windows left from HBM.
template <int O, int ALG>
void launch_split(const FusedArgs& a, hipStream_t st) {
  const int64_t  nwords = split_flag_words(a.nchunks);
  const unsigned nb     = static_cast<unsigned>(a.nchunks);
  k_zero_words<<<static_cast<unsigned>((nwords + 255) / 256), 256, 0, st>>>(a.flags, nwords);
  k_push_window<3, O, ALG><<<nb, kThreads, 0, st>>>(a);
  if (a.charge != ZERO) k_deposit_window<3, O><<<nb, kThreads, 0, st>>>(a);
  k_split_fallback<3, O, ALG><<<static_cast<unsigned>(nwords), kThreads, 0, st>>>(a);
}

template <int ALG>
void launch_split_o(int order, const FusedArgs& a, hipStream_t st) {
  if (order == 1) launch_split<1, ALG>(a, st);
  else launch_split<2, ALG>(a, st);
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

int64_t split_flag_bytes(int64_t max_chunks) {
  return ((split_flag_words(max_chunks) * 4 + 255) / 256) * 256;
}

int64_t split_record_bytes(int64_t max_chunks) { return max_chunks * static_cast<int64_t>(sizeof(uint4)); }

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
  a.flags = nullptr;
  a.recs = nullptr;
  a.reuse = 0;
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
                                        int64_t ntiles, int ppt, uint32_t* flags,
                                        int64_t flag_chunks, void* recs, int64_t rec_chunks,
                                        bool reuse_boxes, hipStream_t st) {
  if (pv.n == 0) return NTT_OK;
  FusedArgs a;
  int64_t   nch = 0;
  ntt_status s = build_fused_args(&a, &nch, pv, em, cur, acc, D, ndth, dt, dxm, charge, deposit,
                                  fscale, tile_offsets, ntiles, ppt);
  if (s != NTT_OK) return s;
  if (D == 3 && (order == 1 || order == 2) && (alg == ALG_BORIS || alg == ALG_VAY) &&
      tile_offsets != nullptr && flags != nullptr && nch > 0 && nch <= flag_chunks) {
    a.flags = flags;
    // box records: 16-bit box corners (|cells| below 2^14), one per chunk
    if (recs != nullptr && nch <= rec_chunks && em.n1 < 16384 && em.n2 < 16384 && em.n3 < 16384) {
      a.recs  = static_cast<uint4*>(recs);
      a.reuse = reuse_boxes ? 1 : 0;
    }
    if (alg == ALG_BORIS) launch_split_o<ALG_BORIS>(order, a, st);
    else launch_split_o<ALG_VAY>(order, a, st);
    return check_launch();
  }
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
