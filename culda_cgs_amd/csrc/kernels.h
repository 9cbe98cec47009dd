// kernels.h -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include "layout.h"

namespace culda {

hipError_t launch_init_z(const InitArgs &a, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
hipError_t launch_add_u32(uint32_t *p, uint32_t v, hipStream_t s);
// fill[0, n_fill) = 0 and the n listed (model, word) n_wk rows zeroed, one
// launch; with ll, that launch also sums the LL packets (stage 1)
hipError_t launch_phi_clear(uint32_t *fill, int64_t n_fill, const ModelDev *models, const int2 *rows, int32_t n,
                            hipStream_t s, const LLArgs *ll = nullptr);
hipError_t launch_phi_rebuild(const PhiArgs &a, hipStream_t s);
// the word-slot rebuild p and the rare tiles' count pass r (r.mode 0) in one launch
hipError_t launch_phi_update(const PhiArgs &p, const RareArgs &r, hipStream_t s);
hipError_t launch_phi_rare(const RareArgs &a, hipStream_t s);
// one launch: theta rebuild t (a one-unit model, the single-chunk kernel),
// phi slots p and rare tiles r (r.mode 0)
hipError_t launch_count_update(const PhiArgs &p, const RareArgs &r, const ThetaArgs &t, hipStream_t s);
// head denominators; with ll also the LL stage 2, with advance (one model)
// the sweep counter then moves on by one
hipError_t launch_den(const ModelDev *models, int32_t num_models, hipStream_t s, const LLArgs *ll = nullptr,
                      uint32_t *advance = nullptr);
hipError_t launch_phi_head(const int32_t *nwk, const float *den, float *head, int32_t num_words, int32_t k,
                           float beta, hipStream_t s);
// alone: nothing else runs beside it (a single-chunk model), see
// theta_rebuild_alone_kernel; else at most max_groups workgroups (0: one
// per packet) take the packets in turn
hipError_t launch_theta_rebuild(const ThetaArgs &a, bool alone, int32_t max_groups, hipStream_t s);
hipError_t launch_theta_rebuild_slim(const ThetaArgs &a, int32_t max_groups, hipStream_t s);
// grid is used in persistent mode (a.queue != null) only
hipError_t launch_sample(const SampleArgs &a, int32_t grid, hipStream_t s);
hipError_t launch_copy_u16(const uint16_t *src, uint16_t *dst, int64_t n, hipStream_t s);
// workgroups of the persistent sampler resident per CU (0 if unknown)
int32_t sample_blocks_per_cu(int32_t rs_stride);
int32_t phi_blocks_per_cu();  // resident phi-rebuild workgroups per CU (0 if unknown)
int32_t sample_rs_stride(int64_t max_entries);
size_t sample_dyn_lds(int32_t rs_stride);

// mh.hip
hipError_t launch_heldout(const HeldoutArgs &a, hipStream_t s);
// A stored [m][k], B stored [n][k]; 128 x 128 tiles when m and n allow
hipError_t launch_gemm_bf16(const GemmArgs &g, hipStream_t s, int32_t splits = 1);
hipError_t launch_sketch(const DocMeta *meta, const uint16_t *theta, int32_t num_docs, uint16_t *out,
                         hipStream_t s);
hipError_t launch_gather_train(const uint16_t *sketches, const uint16_t *zs, const uint32_t *tok_doc,
                               const uint32_t *tok_word, const uint16_t *emb16, uint64_t mul, uint64_t add, uint64_t T,
                               uint64_t N, int64_t D, int64_t p0, int32_t rows, int32_t rows_pad, uint16_t *X,
                               int32_t *tgt, int32_t *wrow, hipStream_t s);
hipError_t launch_gather_query(const uint16_t *sketch, const uint32_t *docs, const uint32_t *words,
                               const uint16_t *emb16, int64_t n, int64_t rows_pad, uint16_t *X, hipStream_t s);
// loss_fx (optional): += sum of the row losses in 2^-24 units
hipError_t launch_softmax_ce(const float *logits, int32_t K, const int32_t *tgt, float inv_rows, uint16_t *G,
                             float *loss, unsigned long long *loss_fx, int32_t rows_pad, hipStream_t s);
// part: colsum_parts(rows) * cols floats of scratch
int32_t colsum_parts(int32_t rows);
hipError_t launch_colsum_bf16(const uint16_t *x, int64_t ld, int32_t rows, int32_t cols, float *part, float *out,
                              hipStream_t s);
// part: splits * m * n floats of scratch; c = sum of the split products, added in order
hipError_t launch_gemm_bf16_splitk(const GemmArgs &g, int32_t splits, float *part, hipStream_t s);
hipError_t launch_emb_grad(const float *dxe, const int32_t *wrow, int32_t rows, long long *acc, hipStream_t s);
// p16t (optional): BF16 copy of the [rows][n / rows] matrix, transposed
hipError_t launch_adam(float *p, float *m, float *v, const float *g, long long *g64, uint16_t *p16, uint16_t *p16t,
                       int64_t rows, int64_t n, float step, float b1, float b2, float eps, float sqrt_bc2,
                       hipStream_t s);
hipError_t launch_to_bf16(const float *x, uint16_t *y, uint16_t *yt, int64_t rows, int64_t n, hipStream_t s);
// dst[c][r] = src[r][c]; rows, cols multiples of 64
hipError_t launch_transpose_bf16(const uint16_t *src, int64_t rows, int64_t cols, int64_t ld_src, uint16_t *dst,
                                 int64_t ld_dst, hipStream_t s);
hipError_t launch_mh(const MHArgs &a, hipStream_t s);
hipError_t launch_ll_reduce(const LLArgs &a, hipStream_t s);
hipError_t launch_joint_ll_words(const int32_t *nwk, int32_t num_words, int32_t k, double lg_beta, float beta,
                                 double *partial, int32_t grid, hipStream_t s);
hipError_t launch_joint_ll_docs(const DocMeta *meta, const uint16_t *theta, int32_t num_docs, float alpha,
                                double *partial, int32_t grid, hipStream_t s);
hipError_t launch_copy(const float4 *src, float4 *dst, int64_t n, hipStream_t s);
// out[0] += sum entries, out[1] += sum n_d * entries over a unit's docs
hipError_t launch_meta_stats(const DocMeta *meta, int32_t num_docs, unsigned long long *out, hipStream_t s);

// Allocation-free joint LL: a static device array of kJointGrid * (1 +
// kJointUnitsMax) partials plus the result; joint_scratch_init (once per
// device, that device current) finds its address, joint_scratch_ptr
// returns it (null before the lookup).
hipError_t joint_scratch_init(int device);
double *joint_scratch_ptr(int device);
hipError_t launch_joint_final(const double *partial, int32_t num_partials, const int32_t *nk, int32_t k, double vb,
                              const UnitDev *units, int32_t unit_begin, int32_t unit_end, double ka, double *out,
                              hipStream_t s);

}  // namespace culda
