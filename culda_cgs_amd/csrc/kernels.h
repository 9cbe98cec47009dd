// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include "layout.h"

namespace culda {

hipError_t launch_init_z(const InitArgs &a, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_zero_rows(int32_t *nwk, const int32_t *words, int32_t n, int32_t k, hipStream_t s);
hipError_t launch_phi_rebuild(const PhiArgs &a, int32_t grid, hipStream_t s);
hipError_t launch_den(const int32_t *nk, float *den, int32_t k, float beta, int32_t num_words, hipStream_t s);
hipError_t launch_phi_head(const int32_t *nwk, const float *den, float *head, int32_t num_words, int32_t k,
                           float beta, hipStream_t s);
hipError_t launch_theta_rebuild(const ThetaArgs &a, int32_t grid, hipStream_t s);
hipError_t launch_sample(const SampleArgs &a, hipStream_t s);
constexpr int kLLGroups = 256;
hipError_t launch_ll_reduce(const double *partial, int64_t n, double *scratch, double *out, hipStream_t s);
hipError_t launch_joint_ll_words(const int32_t *nwk, int32_t num_words, int32_t k, double lg_beta, float beta,
                                 double *partial, int32_t grid, hipStream_t s);
hipError_t launch_joint_ll_docs(const DocMeta *meta, const uint32_t *theta, int32_t num_docs,
                                uint32_t count_bits, float alpha, double *partial, int32_t grid, hipStream_t s);
hipError_t launch_copy(const float4 *src, float4 *dst, int64_t n, hipStream_t s);

}  // namespace culda
