// internal.hpp -- launchers shared between the translation units of libntt_hip.
#pragma once

#include "sr_particle.hpp"

namespace ntt {

// sr_kernels.hip
ntt_status prepare_push(const ntt_sr_pusher_ctx_t* ctx, const ntt_prtl_bc_t* bc,
                        const ntt_particles_t* prtls, const ntt_field_t* em,
                        const ntt_minkowski_t* metric, int32_t shape_order, int* alg_out,
                        ParticleView* pv, FieldView* fv, PrtlBCs* bcs, float* ndth);
ntt_status launch_sr_push(const ParticleView& pv, const FieldView& fv, const PrtlBCs& bcs, int D,
                          int order, int alg, float ndth, float dt, float dxm, hipStream_t st);
ntt_status launch_deposit(const ParticleView& pv, const FieldView& J, int D, int order,
                          float charge, float dt, hipStream_t st);

// push_deposit.hip
ntt_status launch_push_deposit_periodic(const ParticleView& pv, const FieldView& em,
                                        const FieldView& cur, unsigned long long* acc,
                                        int D, int order, int alg,
                                        float ndth, float dt, float dxm, float charge,
                                        bool deposit, float fscale,
                                        const uint32_t* tile_offsets,
                                        int64_t ntiles, int ppt, uint32_t* flags,
                                        int64_t flag_chunks, void* recs, int64_t rec_chunks,
                                        bool reuse_boxes, hipStream_t st);
// bytes of the chunk flags of the split (push, then deposit) launches for up
// to max_chunks chunks; `flags` above holds flag_chunks of them (null: the
// fused kernel runs)
int64_t split_flag_bytes(int64_t max_chunks);
// bytes of the per-chunk box records the split push leaves for its deposit
// (16 B per chunk, a multiple of 256 B for a multiple of 16 chunks); `recs`
// above holds rec_chunks of them (null, or too few: the deposit finds each
// chunk's box itself)
int64_t split_record_bytes(int64_t max_chunks);
// the same, as a kernel argument block written to host memory (for the
// batched launch over the boxes of an ensemble)
size_t     fused_args_size();
ntt_status build_fused_args(void* out, int64_t* nchunks, const ParticleView& pv,
                            const FieldView& em, const FieldView& cur, unsigned long long* acc,
                            int D, float ndth, float dt, float dxm, float charge, bool deposit,
                            float fscale, const uint32_t* tile_offsets, int64_t ntiles, int ppt);
ntt_status launch_push_deposit_batched(int D, int order, int alg, const void* table,
                                       const int64_t* chunk0, int nent, int64_t nchunks,
                                       hipStream_t st);
// fixed-point scale 2^k of the current accumulator: the largest power of two
// for which a particle with |charge * weight| <= qw_max keeps every deposit
// contribution below 2^30 units
float fixed_point_scale(float qw_max, float dt);
float fixed_point_small_coeff(float dt);

// fields.hip
void launch_faraday_periodic(const FieldView& F, float c1, float c2, hipStream_t st);
void launch_ampere_currents_periodic(const FieldView& F, const FieldView& J, float c1, float c2,
                                     float coeff, float ppc0, bool with_j, hipStream_t st);
void launch_filter_periodic(const FieldView& dst, const FieldView& src, hipStream_t st);
void launch_ghost_fill(const FieldView& F, int c0, int c1, hipStream_t st);
void launch_fold(const FieldView& J, hipStream_t st);
void launch_zero(float* d, int64_t n, hipStream_t st);
void launch_copy(float* dst, const float* src, int64_t n, hipStream_t st);
void launch_zero_u64(unsigned long long* d, int64_t n, hipStream_t st);
// J (active cells and their periodic ghost copies) = acc * inv_scale
void launch_j_from_fixed(const FieldView& J, const unsigned long long* acc, double inv_scale,
                         hipStream_t st);

// ---- batched field launches over the boxes of an ensemble (fields.hip) ----
struct FieldSeg {
  FieldView                 F;          // em (Faraday, Ampere), dst (filter), cur (J from acc)
  FieldView                 J;          // cur (Ampere), src (filter)
  const unsigned long long* acc;        // J from the fixed-point accumulator
  double                    inv_scale;
  float                     c1, c2, coeff, ppc0;
  int32_t                   with_j;
  int32_t                   gx, gy;     // blocks of this box's active grid in x, y
  int64_t                   blk0;       // first block of this box in the launch
};
enum { SEG_FARADAY = 0, SEG_AMPERE = 1, SEG_FILTER = 2, SEG_JFIX = 3 };
// blocks a box needs in a batched field launch
int64_t field_seg_blocks(const FieldView& F, int32_t* gx, int32_t* gy);
void    launch_fields_batched(int op, const FieldSeg* segs, int nseg, int64_t nblocks,
                              hipStream_t st);
// sort.hip: one tile sort over the species of many boxes at once.  Every
// (box, species) is a segment: its particles occupy [pbase, pbase + n) of a
// global index space and its keys [kbase, kbase + ntiles + 1) of a global key
// space, so one counting sort over the union sorts every segment in place.
constexpr int kSortArrays = 19;
struct SortSeg {
  void*     arr[kSortArrays];  // i1..3, i*_prev, dx1..3, dx*_prev, u1..3, weight, phi, tag, pld_i
  int64_t   n;
  int32_t   dim, ext[3], tile[3], nt[3];
  uint32_t  dead_key;          // = ntiles
  uint32_t* tile_offsets;      // ntiles + 1 entries, written (may be null)
};
ntt_status make_sort_seg(const ntt_particles_t* p, int32_t dim, const int32_t n[3],
                         const int32_t tile[3], uint32_t* tile_offsets, SortSeg* out);
// workspace of sort_batched for ptot particles and ktot keys in all
int64_t    sort_batched_bytes(int64_t ptot, int64_t ktot);
// workgroups a segment of n particles takes
int64_t    sort_blocks(int64_t n);
// segs, pbase, kbase, bbase: device arrays (nseg, then nseg + 1 prefix sums
// of particles, keys and sort_blocks); nblocks = bbase[nseg].  live: bit a set
// when arr[a] is non-null in some segment (only those arrays are moved).
ntt_status sort_batched(const SortSeg* segs, const int64_t* pbase, const int64_t* kbase,
                        const int64_t* bbase, int64_t nblocks, int nseg, int64_t ptot,
                        int64_t ktot, uint32_t live, void* ws, hipStream_t st);

struct FlatSeg {
  void*       dst;
  const void* src;
  int64_t     n;      // elements
  int64_t     blk0;
};
enum { FLAT_ZERO_U64 = 0, FLAT_COPY_F32 = 1 };
int64_t flat_seg_blocks(int64_t n);
void    launch_flat_batched(int op, const FlatSeg* segs, int nseg, int64_t nblocks,
                            hipStream_t st);

}  // namespace ntt
