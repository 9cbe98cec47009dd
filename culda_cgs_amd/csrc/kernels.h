// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include "layout.h"

namespace culda {

hipError_t launch_init_z(const InitArgs &a, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_add_u32(uint32_t *p, uint32_t v, hipStream_t s);
hipError_t launch_zero_rows(const ModelDev *models, const int2 *rows, int32_t n, hipStream_t s);
hipError_t launch_phi_rebuild(const PhiArgs &a, hipStream_t s);
hipError_t launch_den(const ModelDev *models, int32_t num_models, hipStream_t s);
hipError_t launch_phi_head(const int32_t *nwk, const float *den, float *head, int32_t num_words, int32_t k,
                           float beta, hipStream_t s);
hipError_t launch_theta_rebuild(const ThetaArgs &a, hipStream_t s);
// grid is used in persistent mode (a.queue != null) only
hipError_t launch_sample(const SampleArgs &a, int32_t grid, hipStream_t s);
// workgroups of the persistent sampler resident per CU (0 if unknown)
int32_t sample_blocks_per_cu();

// mh.hip
hipError_t launch_heldout(const HeldoutArgs &a, hipStream_t s);
hipError_t launch_ll_reduce(const LLArgs &a, hipStream_t s);
hipError_t launch_joint_ll_words(const int32_t *nwk, int32_t num_words, int32_t k, double lg_beta, float beta,
                                 double *partial, int32_t grid, hipStream_t s);
hipError_t launch_joint_ll_docs(const DocMeta *meta, const uint32_t *theta, int32_t num_docs,
                                uint32_t count_bits, float alpha, double *partial, int32_t grid, hipStream_t s);
hipError_t launch_copy(const float4 *src, float4 *dst, int64_t n, hipStream_t s);

}  // namespace culda
