// sort.hip -- spatial sort of a species by tile (counting sort), which also
// compacts dead particles to the tail.
//
// Replaces Particles::SortSpatially (BinSort by cell key, bins unsorted
// inside) and Particles::RemoveDead (dead particles moved past the live ones,
// npart reduced), src/framework/containers/particles_sort.cpp:103-253, with
// the key of sort::PositionToTileIndex (src/global/utils/sorting.h:65-137):
// tiles of tile[0] x tile[1] (x tile[2]) cells, numbered x fastest; dead
// particles get the key one past the last tile.
//
// Steps: key + ranked histogram -> exclusive scan of the histogram -> each
// particle's destination (prefix of its key + its rank) -> the SoA arrays
// moved there: in place, four per pass through a scratch buffer; out of place
// (into a second set of arrays), up to twelve per pass, each destination
// computed by the pass itself.  All scratch lives in the caller's workspace.
#include "internal.hpp"

namespace ntt {
namespace {

constexpr int kB = 256;
// permutations by destination (coalesced reads, writes scattered inside a
// tile): measured faster than by source at 3D order 2, 1.07e9 particles
#ifndef NTT_SORT_SCATTER
#define NTT_SORT_SCATTER 1
#endif
constexpr int kScanItems = 4;                 // per thread in the block scan
constexpr int kScanTile  = kB * kScanItems;   // elements per scan block

struct TileKey {
  int32_t n[3], t[3], nt[3];
  int32_t sh[3];  // log2 t[d] when t[d] is a power of two, else -1
  int32_t dim;
  uint32_t dead_key;
  // cell / tile edge, unsigned (the cell is clamped to [0, n) first): a
  // shift for power-of-two edges (a branch uniform over the launch)
  __device__ __forceinline__ uint32_t tdiv(int c, int d) const {
    const uint32_t x = static_cast<uint32_t>(c);
    return sh[d] >= 0 ? x >> sh[d] : x / static_cast<uint32_t>(t[d]);
  }
  __device__ __forceinline__ uint32_t operator()(int i1, int i2, int i3) const {
    // particles outside [0, n) are clamped to the edge tile
    i1 = min(max(i1, 0), n[0] - 1);
    i2 = min(max(i2, 0), n[1] - 1);
    const uint32_t a = tdiv(i1, 0), b = tdiv(i2, 1);
    if (dim == 2) return a + static_cast<uint32_t>(nt[0]) * b;
    i3 = min(max(i3, 0), n[2] - 1);
    const uint32_t c = tdiv(i3, 2);
    return a + static_cast<uint32_t>(nt[0]) * (b + static_cast<uint32_t>(nt[1]) * c);
  }
};

__global__ void k_zero_u32(uint32_t* d, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB)
    d[i] = 0;
}

// Wavefront-aggregated atomic increment: the lanes of a wavefront that
// share a key are served by one atomic of their leader (a nearly sorted
// species has a handful of distinct keys per wavefront).  Returns the old
// value plus the lane's rank among the lanes with the same key.
__device__ __forceinline__ uint32_t aggregated_increment(uint32_t* counters, uint32_t key,
                                                         bool active) {
  uint32_t result = 0;
  uint64_t pending = __ballot(active);
  while (pending != 0ull) {
    const int      leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
    const uint32_t lkey   = __shfl(key, leader);
    const uint64_t same   = __ballot(active && key == lkey) & pending;
    const uint32_t rank =
      __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(same >> 32),
                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(same), 0u));
    uint32_t base = 0;
    if ((threadIdx.x & 63) == leader)
      base = atomicAdd(counters + lkey, static_cast<uint32_t>(__popcll(same)));
    base = __shfl(base, leader);
    if ((same >> (threadIdx.x & 63)) & 1ull) result = base + rank;
    pending &= ~same;
  }
  return result;
}

// Ranked histogram: every particle gets its key and its rank among the
// particles of that key, so that the scatter needs no atomics (destination =
// exclusive prefix of its key + rank).  A workgroup takes kB * R consecutive
// particles.  Keys near the workgroup's first key are counted in LDS, and the
// workgroup then reserves each non-empty bin's range with one global atomic;
// other keys use wavefront-aggregated global atomics.  This keeps the atomics
// on the few counters of a small species from queuing behind each other.
// The LDS bins: with zs = 0, kHist keys from the first key minus kHist / 4 (a
// nearly sorted species moves few particles to lower tiles); with zs > 0 (3D
// tiles, zs = tiles per z layer), three windows of kHist / 4 keys centred on
// the first key and on the keys one tile layer below and above it, where the
// particles that crossed a z tile boundary land.
constexpr int kHist = 1024;
constexpr int kWin  = kHist / 4;

// base, n: the workgroup's first particle and the particle count (in the
// numbering of key_of); keys and rank are written at gofs + p.
template <int R, class KeyOf>
__device__ __forceinline__ void ranked_keys(int64_t base, int64_t n, int64_t gofs, KeyOf key_of,
                                            uint32_t* keys, uint32_t* rank, uint32_t* counts,
                                            uint32_t* hist, uint32_t* kref_s, uint32_t zs = 0u) {
  for (int b = threadIdx.x; b < kHist; b += kB) hist[b] = 0;
  uint32_t k[R];
  bool     act[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t p = base + r * kB + threadIdx.x;
    act[r]          = p < n;
    k[r]            = act[r] ? key_of(p) : 0u;
    if (act[r]) keys[gofs + p] = k[r];
  }
  if (threadIdx.x == 0)
    *kref_s = (zs == 0u) ? (k[0] >= kHist / 4 ? k[0] - kHist / 4 : 0u) : k[0] - kWin / 2;
  __syncthreads();
  const uint32_t kref = *kref_s;
  // bin of a key (kHist: none); with zs, window w in {0, 1, 2} holds the keys
  // kref + (w - 1) zs + [0, kWin) at bins w kWin + [0, kWin) (modulo 2^32)
  auto bin_of = [&](uint32_t key) -> uint32_t {
    if (zs == 0u) return key - kref < static_cast<uint32_t>(kHist) ? key - kref : kHist;
    const uint32_t d = key - kref;
    if (d < static_cast<uint32_t>(kWin)) return kWin + d;
    if (d + zs < static_cast<uint32_t>(kWin)) return d + zs;
    if (d - zs < static_cast<uint32_t>(kWin)) return 2 * kWin + (d - zs);
    return kHist;
  };
  uint32_t loc[R], bin[R];
  bool     win[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    bin[r]   = act[r] ? bin_of(k[r]) : static_cast<uint32_t>(kHist);
    win[r]   = bin[r] < static_cast<uint32_t>(kHist);
    loc[r]   = win[r] ? atomicAdd(hist + bin[r], 1u) : 0u;
    const uint32_t g = aggregated_increment(counts, k[r], act[r] && !win[r]);
    if (act[r] && !win[r]) rank[gofs + base + r * kB + threadIdx.x] = g;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kHist; b += kB) {
    const uint32_t c = hist[b];
    if (c != 0u) {
      const uint32_t key =
        (zs == 0u) ? kref + b : kref + (static_cast<uint32_t>(b / kWin) - 1u) * zs + b % kWin;
      hist[b] = atomicAdd(counts + key, c);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (win[r]) rank[gofs + base + r * kB + threadIdx.x] = hist[bin[r]] + loc[r];
}

constexpr int kKeyItems = 8;  // particles per thread of the single-species key pass

__global__ __launch_bounds__(kB) void k_keys(const int32_t* __restrict__ i1,
                                             const int32_t* __restrict__ i2,
                                             const int32_t* __restrict__ i3,
                                             const int16_t* __restrict__ tag, int64_t n,
                                             TileKey K, uint32_t* __restrict__ keys,
                                             uint32_t* __restrict__ rank,
                                             uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[kHist];
  __shared__ uint32_t kref;
  auto key_of = [&](int64_t p) -> uint32_t {
    return tag[p] == NTT_TAG_DEAD ? K.dead_key : K(i1[p], i2[p], K.dim == 3 ? i3[p] : 0);
  };
  const uint32_t zs = (K.dim == 3) ? static_cast<uint32_t>(K.nt[0]) * K.nt[1] : 0u;
  ranked_keys<kKeyItems>(static_cast<int64_t>(blockIdx.x) * (kB * kKeyItems), n, 0, key_of, keys,
                         rank, counts, hist, &kref, zs);
}

// block-level exclusive scan of kScanTile elements; writes block totals
__global__ __launch_bounds__(kB) void k_scan_block(uint32_t* __restrict__ data, int64_t n,
                                                   uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s[kB];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + threadIdx.x * kScanItems + k;
    v[k] = (i < n) ? data[i] : 0u;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  // Hillis-Steele inclusive scan of the per-thread sums
  for (int off = 1; off < kB; off <<= 1) {
    const uint32_t t = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = s[threadIdx.x] - sum;  // exclusive prefix of this thread
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + threadIdx.x * kScanItems + k;
    if (i < n) data[i] = run;
    run += v[k];
  }
  if (threadIdx.x == kB - 1) block_sums[blockIdx.x] = s[kB - 1];
}

__global__ __launch_bounds__(kB) void k_scan_add(uint32_t* __restrict__ data, int64_t n,
                                                 const uint32_t* __restrict__ block_prefix) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  const uint32_t add = block_prefix[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + threadIdx.x * kScanItems + k;
    if (i < n) data[i] += add;
  }
}

// Order inside a tile.  Ranks follow the old order: the tile's own particles,
// then the ones that moved in, which arrive in runs from a boundary layer of a
// few cells.  A wavefront over such a run hits the same cells with most of
// its lanes, and the push's LDS current atomics serialise on them.  With
// NTT_SORT_INTERLEAVE the tile's ranks are laid out as a q x 61 array and its
// slots filled column by column: neighbouring slots take ranks 61 apart,
// mixing the runs and every cell cycle of the old order (a bijection of
// [0, count)).  Measured: the 2D order-1 push 10% faster, the 3D order-2 push
// unchanged, and the permutation by destination, whose writes no longer run
// in order inside a tile, so much slower that every step lost; off.
#ifndef NTT_SORT_INTERLEAVE
#define NTT_SORT_INTERLEAVE 0
#endif
constexpr uint32_t kInterleave = 61;

__device__ __forceinline__ uint32_t interleaved_slot(uint32_t r, uint32_t count) {
  const uint32_t q = count / kInterleave, m = count - q * kInterleave;
  const uint32_t a = r / kInterleave, b = r - a * kInterleave;
  return b * q + min(b, m) + a;  // column b holds q + (b < m) ranks
}

// destination = exclusive prefix of the key + slot of the rank; inverse writes
// it over the key (keys[p] = destination), else perm[destination] = p.
// prefix has nkeys entries; the last key's run ends at n.
__global__ __launch_bounds__(kB) void k_scatter(uint32_t* __restrict__ keys, int64_t n,
                                                const uint32_t* __restrict__ prefix,
                                                int64_t nkeys, const uint32_t* __restrict__ rank,
                                                uint32_t* __restrict__ perm, int inverse) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (p >= n) return;
  const uint32_t k     = keys[p];
  const uint32_t start = prefix[k];
  uint32_t       slot  = rank[p];
  if (NTT_SORT_INTERLEAVE) {
    const uint32_t end = k + 1 < nkeys ? prefix[k + 1] : static_cast<uint32_t>(n);
    slot = interleaved_slot(slot, end - start);
  }
  const uint32_t pos = start + slot;
  if (inverse) keys[p] = pos;
  else perm[pos] = static_cast<uint32_t>(p);
}

// up to kBatch particle arrays permuted per pass, each staged in a slot of
// its own element size per particle of the scratch buffer (slot k starts at
// off[k] * n bytes, off[k] a multiple of the element size; at most 8 * kBatch
// bytes per particle in all)
constexpr int kBatch = 4;
struct ArrayBatch {
  void*   arr[kBatch];
  int32_t bytes[kBatch];  // element size: 2, 4 or 8
  int32_t off[kBatch];    // sum of the element sizes of the slots before
  int32_t count;
};

template <typename T>
__device__ __forceinline__ void gather_one(void* dst, const void* src, int64_t to, int64_t from) {
  static_cast<T*>(dst)[to] = static_cast<const T*>(src)[from];
}

// stage the permuted arrays: slot[p] = arr[perm[p]], or by destination
// (by_dest) slot[inv[p]] = arr[p] with coalesced reads
__global__ __launch_bounds__(kB) void k_gather_batch(ArrayBatch b, char* __restrict__ tmp,
                                                     const uint32_t* __restrict__ perm,
                                                     int64_t n, int by_dest) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (p >= n) return;
  const int64_t q  = perm[p];
  const int64_t to = by_dest ? q : p, from = by_dest ? p : q;
#pragma unroll
  for (int k = 0; k < kBatch; ++k) {
    if (k >= b.count) break;
    void* slot = tmp + n * b.off[k];
    if (b.bytes[k] == 2) gather_one<int16_t>(slot, b.arr[k], to, from);
    else if (b.bytes[k] == 4) gather_one<int32_t>(slot, b.arr[k], to, from);
    else gather_one<int64_t>(slot, b.arr[k], to, from);
  }
}

__global__ __launch_bounds__(kB) void k_copy_batch(ArrayBatch b, const char* __restrict__ tmp,
                                                   int64_t n) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int k = 0; k < kBatch; ++k) {
    if (k >= b.count) break;
    const void* slot = tmp + n * b.off[k];
    if (b.bytes[k] == 2) gather_one<int16_t>(b.arr[k], slot, p, p);
    else if (b.bytes[k] == 4) gather_one<int32_t>(b.arr[k], slot, p, p);
    else gather_one<int64_t>(b.arr[k], slot, p, p);
  }
}

// out of place: dst[k][p] = src[k][perm[p]] for up to kDirect arrays per pass
// (every array of a 3D species without its prev arrays in one pass)
constexpr int kDirect = 12;
struct ArrayPairs {
  const void* src[kDirect];
  void*       dst[kDirect];
  int32_t     bytes[kDirect];
  int32_t     count;
};

// the same by destination: dst[k][inv[p]] = src[k][p] (coalesced reads),
// kPermItems particles per thread, so that each array's loads of a thread are
// in flight together
constexpr int kPermItems = 8;
template <typename T>
__device__ __forceinline__ void scatter_items(void* dst, const void* src, const int64_t* p,
                                              const uint32_t* q, const bool* act) {
  T v[kPermItems];
#pragma unroll
  for (int r = 0; r < kPermItems; ++r)
    if (act[r]) v[r] = static_cast<const T*>(src)[p[r]];
#pragma unroll
  for (int r = 0; r < kPermItems; ++r)
    if (act[r]) static_cast<T*>(dst)[q[r]] = v[r];
}
// The destination is inv[p], or, with prefix given (ranks), the exclusive
// prefix of the particle's key plus its rank -- k_scatter's arithmetic done in
// place, so that no pass writes the inverse permutation first.
__global__ __launch_bounds__(kB) void k_permute_scatter(ArrayPairs b,
                                                        const uint32_t* __restrict__ inv,
                                                        const uint32_t* __restrict__ prefix,
                                                        const uint32_t* __restrict__ rank,
                                                        int64_t n) {
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * (kB * kPermItems) + threadIdx.x;
  int64_t       p[kPermItems];
  uint32_t      q[kPermItems];
  bool          act[kPermItems];
#pragma unroll
  for (int r = 0; r < kPermItems; ++r) {
    p[r]   = p0 + r * kB;
    act[r] = p[r] < n;
    q[r]   = 0u;
    if (act[r]) q[r] = (prefix != nullptr) ? prefix[inv[p[r]]] + rank[p[r]] : inv[p[r]];
  }
#pragma unroll
  for (int k = 0; k < kDirect; ++k) {
    if (k >= b.count) break;
    if (b.bytes[k] == 2) scatter_items<int16_t>(b.dst[k], b.src[k], p, q, act);
    else if (b.bytes[k] == 4) scatter_items<int32_t>(b.dst[k], b.src[k], p, q, act);
    else scatter_items<int64_t>(b.dst[k], b.src[k], p, q, act);
  }
}

__global__ __launch_bounds__(kB) void k_permute_direct(ArrayPairs b,
                                                       const uint32_t* __restrict__ perm,
                                                       int64_t n) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (p >= n) return;
  const int64_t q = perm[p];
#pragma unroll
  for (int k = 0; k < kDirect; ++k) {
    if (k >= b.count) break;
    if (b.bytes[k] == 2) gather_one<int16_t>(b.dst[k], b.src[k], p, q);
    else if (b.bytes[k] == 4) gather_one<int32_t>(b.dst[k], b.src[k], p, q);
    else gather_one<int64_t>(b.dst[k], b.src[k], p, q);
  }
}

// exclusive scan of n uint32 in place; tmp holds the block sums of every level
ntt_status scan_exclusive(uint32_t* data, int64_t n, uint32_t* tmp, hipStream_t st) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  k_scan_block<<<static_cast<unsigned>(nb), kB, 0, st>>>(data, n, tmp);
  if (nb > 1) {
    ntt_status s = scan_exclusive(tmp, nb, tmp + nb, st);
    if (s != NTT_OK) return s;
    k_scan_add<<<static_cast<unsigned>(nb), kB, 0, st>>>(data, n, tmp);
  }
  return check_launch();
}

int64_t scan_tmp_elems(int64_t n) {
  int64_t total = 0;
  while (n > 1) {
    n = (n + kScanTile - 1) / kScanTile;
    total += n;
  }
  return total + 1;
}

struct Layout {
  int64_t keys, perm, rank, counts, scan, tmp, total;  // byte offsets
  int64_t nkeys;
};

int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

ntt_status make_layout(int64_t npart, int32_t dim, const int32_t* n, const int32_t* tile,
                       TileKey* K, Layout* L, bool staging = true) {
  if (dim != 2 && dim != 3) return NTT_ERR_UNSUPPORTED;
  if (npart < 0 || npart > 0xFFFFFFFFll) return NTT_ERR_ARG;
  int64_t ntiles = 1;
  for (int d = 0; d < dim; ++d) {
    if (n[d] < 1 || tile[d] < 1) return NTT_ERR_ARG;
    K->n[d]  = n[d];
    K->t[d]  = tile[d];
    K->nt[d] = (n[d] + tile[d] - 1) / tile[d];
    ntiles *= K->nt[d];
  }
  for (int d = dim; d < 3; ++d) { K->n[d] = 1; K->t[d] = 1; K->nt[d] = 1; }
  for (int d = 0; d < 3; ++d) {
    const int32_t t = K->t[d];
    K->sh[d] = (t & (t - 1)) == 0 ? __builtin_ctz(static_cast<unsigned>(t)) : -1;
  }
  if (ntiles >= 0xFFFFFFFFll) return NTT_ERR_ARG;
  K->dim      = dim;
  K->dead_key = static_cast<uint32_t>(ntiles);
  L->nkeys    = ntiles + 1;
  int64_t off = 0;
  L->keys   = off; off = align256(off + 4 * npart);
  L->perm   = off; off = align256(off + 4 * npart);
  // ranks share the permutation's buffer when the permutation is kept in the
  // keys (by destination)
  L->rank   = L->perm;
  if (!NTT_SORT_SCATTER) { L->rank = off; off = align256(off + 4 * npart); }
  L->counts = off; off = align256(off + 4 * L->nkeys);
  L->scan   = off; off = align256(off + 4 * scan_tmp_elems(L->nkeys));
  L->tmp    = off; off = align256(off + (staging ? 8 * kBatch * npart : 0));
  L->total  = off;
  return NTT_OK;
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kB - 1) / kB); }

// applies perm to every listed array (dst[p] = src[perm[p]]), kBatch at a time
void permute_all(void* const* arrs, const int32_t* bytes, int narr, char* tmp,
                 const uint32_t* perm, int64_t n, hipStream_t st, int by_dest) {
  ArrayBatch b{};
  for (int k = 0; k <= narr; ++k) {
    if (k < narr && arrs[k] != nullptr) {
      b.arr[b.count]   = arrs[k];
      b.bytes[b.count] = bytes[k];
      // slots start at multiples of their element size (aligned accesses)
      const int end    = b.count ? b.off[b.count - 1] + b.bytes[b.count - 1] : 0;
      b.off[b.count]   = (end + bytes[k] - 1) / bytes[k] * bytes[k];
      ++b.count;
    }
    if (b.count == kBatch || (k == narr && b.count > 0)) {
      k_gather_batch<<<blocks_for(n), kB, 0, st>>>(b, tmp, perm, n, by_dest);
      k_copy_batch<<<blocks_for(n), kB, 0, st>>>(b, tmp, n);
      b = ArrayBatch{};
    }
  }
}

// ---- batched: the segments of many boxes in one pass ---------------------------

// segment holding global index g of a prefix array base[0..nseg] (base[0] = 0)
__device__ __forceinline__ int find_seg(const int64_t* __restrict__ base, int nseg, int64_t g) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t seg_key(const SortSeg& s, int64_t p) {
  if (static_cast<const int16_t*>(s.arr[17])[p] == NTT_TAG_DEAD) return s.dead_key;
  int c[3] = { static_cast<const int32_t*>(s.arr[0])[p], static_cast<const int32_t*>(s.arr[1])[p],
               s.dim == 3 ? static_cast<const int32_t*>(s.arr[2])[p] : 0 };
  uint32_t key = 0, mul = 1;
  for (int d = 0; d < s.dim; ++d) {
    const int x = min(max(c[d], 0), s.ext[d] - 1);
    key += mul * static_cast<uint32_t>(x / s.tile[d]);
    mul *= static_cast<uint32_t>(s.nt[d]);
  }
  return key;
}

// Every workgroup works inside one segment (bbase: first block of each
// segment), so the segment index and its table entry are wave-uniform.  A
// workgroup takes kB * kSegItems consecutive particles of its segment, thread
// t the particles p + r * kB (r < kSegItems): the segment lookup, a chain of
// dependent loads, is paid once per 1024 particles rather than per 256.
constexpr int kSegItems = 4;
struct SegBlock {
  int     s;
  int64_t p;  // the thread's first particle index inside the segment
};
__device__ __forceinline__ SegBlock seg_block(const int64_t* __restrict__ bbase, int nseg) {
  const int s = __builtin_amdgcn_readfirstlane(find_seg(bbase, nseg, blockIdx.x));
  return SegBlock{ s,
                   (static_cast<int64_t>(blockIdx.x) - bbase[s]) * (kB * kSegItems) + threadIdx.x };
}

__global__ __launch_bounds__(kB) void k_keys_batched(const SortSeg* __restrict__ segs,
                                                     const int64_t* __restrict__ pbase,
                                                     const int64_t* __restrict__ kbase,
                                                     const int64_t* __restrict__ bbase, int nseg,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ rank,
                                                     uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[kHist];
  __shared__ uint32_t kref;
  const SegBlock b    = seg_block(bbase, nseg);
  const SortSeg& sg   = segs[b.s];
  const uint32_t koff = static_cast<uint32_t>(kbase[b.s]);
  auto key_of = [&](int64_t p) -> uint32_t { return koff + seg_key(sg, p); };
  // the workgroups follow the segments' block table, kSegItems particles per thread
  ranked_keys<kSegItems>(b.p - threadIdx.x, sg.n, pbase[b.s], key_of, keys, rank, counts, hist,
                         &kref);
}

// tile offsets of every segment: the global exclusive prefix minus pbase
__global__ __launch_bounds__(kB) void k_offsets_batched(const SortSeg* __restrict__ segs,
                                                        const int64_t* __restrict__ pbase,
                                                        const int64_t* __restrict__ kbase,
                                                        int nseg, int64_t ktot,
                                                        const uint32_t* __restrict__ prefix) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  if (k >= ktot) return;
  const int s = find_seg(kbase, nseg, k);
  uint32_t* off = segs[s].tile_offsets;
  if (off) off[k - kbase[s]] = static_cast<uint32_t>(prefix[k] - pbase[s]);
}

// the element size of each of the kSortArrays arrays
__device__ __forceinline__ int arr_bytes(int a) { return a == 17 ? 2 : (a == 18 ? 8 : 4); }

__device__ __forceinline__ void move_elem(void* dst, int64_t to, const void* src, int64_t from,
                                          int bytes) {
  if (bytes == 2) static_cast<int16_t*>(dst)[to] = static_cast<const int16_t*>(src)[from];
  else if (bytes == 4) static_cast<int32_t*>(dst)[to] = static_cast<const int32_t*>(src)[from];
  else static_cast<int64_t*>(dst)[to] = static_cast<const int64_t*>(src)[from];
}

// the arrays (indices into SortSeg::arr) one pass of the batched sort moves:
// up to kBatch of those that are non-null in some segment
struct ArrList {
  int32_t a[kBatch];
  int32_t count;
};

// arrays L.a[0 .. L.count) of every segment: tmp slot k <- src[perm]
// by_dest: perm holds each particle's global destination, and the slot is
// written there (coalesced reads); otherwise the slot gathers from perm
__global__ __launch_bounds__(kB) void k_gather_batched(const SortSeg* __restrict__ segs,
                                                       const int64_t* __restrict__ pbase,
                                                       const int64_t* __restrict__ bbase, int nseg,
                                                       int64_t ptot, ArrList L,
                                                       const uint32_t* __restrict__ perm,
                                                       char* __restrict__ tmp, int by_dest) {
  const SegBlock b  = seg_block(bbase, nseg);
  const SortSeg& sg = segs[b.s];
  const int64_t  pb = pbase[b.s];
  // g: slot index (global), q: source index (in the segment)
  int64_t g[kSegItems], q[kSegItems];
  bool    act[kSegItems];
#pragma unroll
  for (int r = 0; r < kSegItems; ++r) {
    const int64_t p = b.p + r * kB;
    act[r]          = p < sg.n;
    const int64_t d = act[r] ? static_cast<int64_t>(perm[pb + p]) : 0;
    g[r]            = by_dest ? d : pb + p;
    q[r]            = by_dest ? p : d - pb;
  }
  int64_t off = 0;  // slot offset in bytes per particle (the element sizes before)
#pragma unroll
  for (int k = 0; k < kBatch; ++k) {
    if (k >= L.count) break;
    const int a  = L.a[k];
    const int sz = arr_bytes(a);
    off = (off + sz - 1) / sz * sz;  // aligned slot
    const void* src = sg.arr[a];
    if (src != nullptr) {
#pragma unroll
      for (int r = 0; r < kSegItems; ++r)
        if (act[r]) move_elem(tmp + ptot * off, g[r], src, q[r], sz);
    }
    off += sz;
  }
}

__global__ __launch_bounds__(kB) void k_copy_batched(const SortSeg* __restrict__ segs,
                                                     const int64_t* __restrict__ pbase,
                                                     const int64_t* __restrict__ bbase, int nseg,
                                                     int64_t ptot, ArrList L,
                                                     const char* __restrict__ tmp) {
  const SegBlock b  = seg_block(bbase, nseg);
  const SortSeg& sg = segs[b.s];
  const int64_t  pb = pbase[b.s];
  int64_t off = 0;
#pragma unroll
  for (int k = 0; k < kBatch; ++k) {
    if (k >= L.count) break;
    const int a  = L.a[k];
    const int sz = arr_bytes(a);
    off = (off + sz - 1) / sz * sz;
    void* dst = sg.arr[a];
    if (dst != nullptr) {
#pragma unroll
      for (int r = 0; r < kSegItems; ++r) {
        const int64_t p = b.p + r * kB;
        if (p < sg.n) move_elem(dst, p, tmp + ptot * off, pb + p, sz);
      }
    }
    off += sz;
  }
}

}  // namespace

ntt_status make_sort_seg(const ntt_particles_t* p, int32_t dim, const int32_t n[3],
                         const int32_t tile[3], uint32_t* tile_offsets, SortSeg* out) {
  ParticleView pv;
  ntt_status s = make_particle_view(p, dim, true, &pv);
  if (s != NTT_OK) return s;
  TileKey K;
  Layout  L;
  s = make_layout(pv.n, dim, n, tile, &K, &L);
  if (s != NTT_OK) return s;
  SortSeg g{};
  void* arrs[kSortArrays] = {
    p->i1, p->i2, dim == 3 ? p->i3 : nullptr,
    p->i1_prev, p->i2_prev, dim == 3 ? p->i3_prev : nullptr,
    p->dx1, p->dx2, dim == 3 ? p->dx3 : nullptr,
    p->dx1_prev, p->dx2_prev, dim == 3 ? p->dx3_prev : nullptr,
    p->ux1, p->ux2, p->ux3, p->weight, p->phi, p->tag, p->pld_i
  };
  for (int a = 0; a < kSortArrays; ++a) g.arr[a] = arrs[a];
  g.n   = pv.n;
  g.dim = dim;
  for (int d = 0; d < 3; ++d) {
    g.ext[d]  = K.n[d];
    g.tile[d] = K.t[d];
    g.nt[d]   = K.nt[d];
  }
  g.dead_key     = K.dead_key;
  g.tile_offsets = tile_offsets;
  *out = g;
  return NTT_OK;
}

int64_t sort_blocks(int64_t n) { return (n + kB * kSegItems - 1) / (kB * kSegItems); }

int64_t sort_batched_bytes(int64_t ptot, int64_t ktot) {
  int64_t off = 0;
  off = align256(off + 4 * ptot);                    // keys
  off = align256(off + 4 * ptot);                    // perm (and ranks)
  if (!NTT_SORT_SCATTER) off = align256(off + 4 * ptot);  // ranks
  off = align256(off + 4 * ktot);                    // counts
  off = align256(off + 4 * scan_tmp_elems(ktot));    // scan
  off = align256(off + 8 * kBatch * ptot);           // staging
  return off;
}

ntt_status sort_batched(const SortSeg* segs, const int64_t* pbase, const int64_t* kbase,
                        const int64_t* bbase, int64_t nblocks, int nseg, int64_t ptot,
                        int64_t ktot, uint32_t live, void* ws, hipStream_t st) {
  if (ptot == 0 || nseg == 0 || nblocks == 0) return NTT_OK;
  if (nblocks > 0x7FFFFFFFll) return NTT_ERR_ARG;
  const unsigned nb = static_cast<unsigned>(nblocks);
  if (ptot > 0xFFFFFFFFll || ktot > 0xFFFFFFFFll) return NTT_ERR_ARG;
  char*     w    = static_cast<char*>(ws);
  int64_t   off  = 0;
  uint32_t* keys = reinterpret_cast<uint32_t*>(w + off); off = align256(off + 4 * ptot);
  uint32_t* perm = reinterpret_cast<uint32_t*>(w + off); off = align256(off + 4 * ptot);
  uint32_t* rank = perm;
  if (!NTT_SORT_SCATTER) { rank = reinterpret_cast<uint32_t*>(w + off); off = align256(off + 4 * ptot); }
  uint32_t* cnt  = reinterpret_cast<uint32_t*>(w + off); off = align256(off + 4 * ktot);
  uint32_t* scn  = reinterpret_cast<uint32_t*>(w + off); off = align256(off + 4 * scan_tmp_elems(ktot));
  char*     tmp  = w + off;
  k_zero_u32<<<blocks_for(ktot) > 4096 ? 4096 : blocks_for(ktot), kB, 0, st>>>(cnt, ktot);
  k_keys_batched<<<nb, kB, 0, st>>>(segs, pbase, kbase, bbase, nseg, keys, rank, cnt);
  ntt_status s = scan_exclusive(cnt, ktot, scn, st);
  if (s != NTT_OK) return s;
  k_scatter<<<blocks_for(ptot), kB, 0, st>>>(keys, ptot, cnt, ktot, rank, perm, NTT_SORT_SCATTER);
  k_offsets_batched<<<blocks_for(ktot), kB, 0, st>>>(segs, pbase, kbase, nseg, ktot, cnt);
  // the live arrays, kBatch per pass (a 2D species without its prev arrays
  // has 9 of the 19: three passes)
  ArrList L{};
  for (int a = 0; a < kSortArrays; ++a) {
    if ((live >> a) & 1u) L.a[L.count++] = a;
    if (L.count == kBatch || (a == kSortArrays - 1 && L.count > 0)) {
      k_gather_batched<<<nb, kB, 0, st>>>(segs, pbase, bbase, nseg, ptot, L,
                                          NTT_SORT_SCATTER ? keys : perm, tmp, NTT_SORT_SCATTER);
      k_copy_batched<<<nb, kB, 0, st>>>(segs, pbase, bbase, nseg, ptot, L, tmp);
      L.count = 0;
    }
  }
  return check_launch();
}

}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_sort_workspace_size(int64_t maxnpart, int32_t dim, const int32_t n[3],
                                              const int32_t tile[3], uint64_t* bytes) {
  if (n == nullptr || tile == nullptr || bytes == nullptr) return NTT_ERR_ARG;
  TileKey K;
  Layout  L;
  ntt_status s = make_layout(maxnpart, dim, n, tile, &K, &L);
  if (s != NTT_OK) return s;
  *bytes = static_cast<uint64_t>(L.total);
  return NTT_OK;
}

extern "C" ntt_status ntt_sort_ntiles(int32_t dim, const int32_t n[3], const int32_t tile[3],
                                      int64_t* ntiles) {
  if (n == nullptr || tile == nullptr || ntiles == nullptr) return NTT_ERR_ARG;
  TileKey K;
  Layout  L;
  ntt_status s = make_layout(0, dim, n, tile, &K, &L);
  if (s != NTT_OK) return s;
  *ntiles = L.nkeys - 1;
  return NTT_OK;
}

namespace {

// every SoA array of a species, in the order of SortSeg (coordinates beyond
// dim are not used)
void species_arrays(const ntt_particles_t* p, int32_t dim, void* arrs[kSortArrays]) {
  void* a[kSortArrays] = {
    p->i1, p->i2, dim == 3 ? p->i3 : nullptr,
    p->i1_prev, p->i2_prev, dim == 3 ? p->i3_prev : nullptr,
    p->dx1, p->dx2, dim == 3 ? p->dx3 : nullptr,
    p->dx1_prev, p->dx2_prev, dim == 3 ? p->dx3_prev : nullptr,
    p->ux1, p->ux2, p->ux3, p->weight, p->phi, p->tag, p->pld_i
  };
  for (int k = 0; k < kSortArrays; ++k) arrs[k] = a[k];
}
constexpr int32_t kArrBytes[kSortArrays] = { 4, 4, 4, 4, 4, 4, 4, 4, 4, 4,
                                            4, 4, 4, 4, 4, 4, 4, 2, 8 };

// ntt_sort_particles (dst == nullptr: in place, through the staging buffer)
// and ntt_sort_particles_to (dst: every array gathered straight into dst)
ntt_status sort_species(ntt_particles_t* src, ntt_particles_t* dst, int32_t dim,
                        const int32_t n[3], const int32_t tile[3], void* workspace,
                        uint64_t workspace_bytes, uint32_t* tile_offsets,
                        int64_t* npart_alive, void* stream) {
  if (src == nullptr || n == nullptr || tile == nullptr) return NTT_ERR_ARG;
  // the key needs the cell indices and the tags; every other array is
  // permuted when given (NULL arrays are left as they are)
  if (src->npart < 0 || (src->maxnpart > 0 && src->npart > src->maxnpart)) return NTT_ERR_ARG;
  ParticleView pv{};
  pv.n = src->npart;
  if (pv.n > 0 && (!src->i1 || !src->tag || (dim >= 2 && !src->i2) || (dim >= 3 && !src->i3)))
    return NTT_ERR_ARG;
  pv.i1  = src->i1;
  pv.i2  = src->i2;
  pv.i3  = src->i3;
  pv.tag = src->tag;
  ntt_status s = NTT_OK;
  void* sa[kSortArrays];
  void* da[kSortArrays];
  species_arrays(src, dim, sa);
  if (dst != nullptr) {
    if (dst->maxnpart > 0 && dst->maxnpart < src->npart) return NTT_ERR_CAPACITY;
    species_arrays(dst, dim, da);
    for (int k = 0; k < kSortArrays; ++k) {
      if (sa[k] == nullptr) continue;
      if (da[k] == nullptr) return NTT_ERR_ARG;
      for (int j = 0; j < kSortArrays; ++j)
        if (sa[j] != nullptr && da[k] == sa[j]) return NTT_ERR_ARG;  // no aliasing
    }
  }
  TileKey K;
  Layout  L;
  s = make_layout(pv.n, dim, n, tile, &K, &L, dst == nullptr);
  if (s != NTT_OK) return s;
  if (pv.n == 0) {
    if (tile_offsets != nullptr) {
      k_zero_u32<<<blocks_for(L.nkeys) > 4096 ? 4096 : blocks_for(L.nkeys), kB, 0,
                   as_stream(stream)>>>(tile_offsets, L.nkeys);
      s = check_launch();
      if (s != NTT_OK) return s;
    }
    if (npart_alive) *npart_alive = 0;
    if (dst != nullptr) dst->npart = 0;  // the sorted (empty) species
    return NTT_OK;
  }
  if (workspace == nullptr || static_cast<int64_t>(workspace_bytes) < L.total)
    return NTT_ERR_CAPACITY;
  hipStream_t st   = as_stream(stream);
  char*       ws   = static_cast<char*>(workspace);
  uint32_t*   keys = reinterpret_cast<uint32_t*>(ws + L.keys);
  uint32_t*   perm = reinterpret_cast<uint32_t*>(ws + L.perm);
  uint32_t*   cnt  = reinterpret_cast<uint32_t*>(ws + L.counts);
  uint32_t*   rank = reinterpret_cast<uint32_t*>(ws + L.rank);
  uint32_t*   scn  = reinterpret_cast<uint32_t*>(ws + L.scan);
  void*       tmp  = ws + L.tmp;
  const int64_t np = pv.n;

  k_zero_u32<<<blocks_for(L.nkeys) > 4096 ? 4096 : blocks_for(L.nkeys), kB, 0, st>>>(cnt, L.nkeys);
  k_keys<<<static_cast<unsigned>((np + kB * kKeyItems - 1) / (kB * kKeyItems)), kB, 0, st>>>(
    pv.i1, pv.i2, pv.i3, pv.tag, np, K, keys, rank, cnt);
  s = scan_exclusive(cnt, L.nkeys, scn, st);
  if (s != NTT_OK) return s;
  // out of place by destination: the permute computes each destination itself
  const bool fused_dest = dst != nullptr && NTT_SORT_SCATTER && !NTT_SORT_INTERLEAVE;
  if (!fused_dest)
    k_scatter<<<blocks_for(np), kB, 0, st>>>(keys, np, cnt, L.nkeys, rank, perm, NTT_SORT_SCATTER);
  if (tile_offsets != nullptr &&
      hipMemcpyAsync(tile_offsets, cnt, 4 * L.nkeys, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return NTT_ERR_HIP;

  if (dst == nullptr) {
    permute_all(sa, kArrBytes, kSortArrays, static_cast<char*>(tmp),
                NTT_SORT_SCATTER ? keys : perm, np, st, NTT_SORT_SCATTER);
  } else {
    ArrayPairs b{};
    for (int k = 0; k <= kSortArrays; ++k) {
      if (k < kSortArrays && sa[k] != nullptr) {
        b.src[b.count]   = sa[k];
        b.dst[b.count]   = da[k];
        b.bytes[b.count] = kArrBytes[k];
        ++b.count;
      }
      if (b.count == kDirect || (k == kSortArrays && b.count > 0)) {
        if (NTT_SORT_SCATTER)
          k_permute_scatter<<<blocks_for((np + kPermItems - 1) / kPermItems), kB, 0, st>>>(
            b, keys, fused_dest ? cnt : nullptr, fused_dest ? rank : nullptr, np);
        else k_permute_direct<<<blocks_for(np), kB, 0, st>>>(b, perm, np);
        b = ArrayPairs{};
      }
    }
    dst->npart = src->npart;
  }
  s = check_launch();
  if (s != NTT_OK) return s;
  if (npart_alive != nullptr) {
    // live particles = exclusive prefix at the dead key
    uint32_t alive = 0;
    if (hipMemcpyAsync(&alive, cnt + (L.nkeys - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess)
      return NTT_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return NTT_ERR_HIP;
    *npart_alive = alive;
  }
  return NTT_OK;
}

}  // namespace

extern "C" ntt_status ntt_sort_particles(ntt_particles_t* prtls, int32_t dim, const int32_t n[3],
                                         const int32_t tile[3], void* workspace,
                                         uint64_t workspace_bytes, uint32_t* tile_offsets,
                                         int64_t* npart_alive, void* stream) {
  return sort_species(prtls, nullptr, dim, n, tile, workspace, workspace_bytes, tile_offsets,
                      npart_alive, stream);
}

extern "C" ntt_status ntt_sort_particles_to(const ntt_particles_t* src, ntt_particles_t* dst,
                                            int32_t dim, const int32_t n[3],
                                            const int32_t tile[3], void* workspace,
                                            uint64_t workspace_bytes, uint32_t* tile_offsets,
                                            int64_t* npart_alive, void* stream) {
  if (dst == nullptr) return NTT_ERR_ARG;
  return sort_species(const_cast<ntt_particles_t*>(src), dst, dim, n, tile, workspace,
                      workspace_bytes, tile_offsets, npart_alive, stream);
}

extern "C" ntt_status ntt_sort_workspace_size_to(int64_t maxnpart, int32_t dim,
                                                 const int32_t n[3], const int32_t tile[3],
                                                 uint64_t* bytes) {
  if (n == nullptr || tile == nullptr || bytes == nullptr) return NTT_ERR_ARG;
  TileKey K;
  Layout  L;
  ntt_status s = make_layout(maxnpart, dim, n, tile, &K, &L, false);
  if (s != NTT_OK) return s;
  *bytes = static_cast<uint64_t>(L.total);
  return NTT_OK;
}
