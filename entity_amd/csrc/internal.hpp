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
                                        int64_t ntiles, int ppt, hipStream_t st);
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

}  // namespace ntt
