// context.h -- device-resident model state behind the C ABI.
//
// Plays the roles of the reference's DocChunk device buffers
// (model/doc_chunk.h:43-57), ModelPhiGPU (model/model_phi_gpu.h:17-44) and
// ModelThetaChunk (model/model_theta_chunk.h:20-59), owned by one context
// instead of by model objects that each call cudaMalloc.
#pragma once
#include <hip/hip_runtime_api.h>

#include <vector>

#include "corpus.h"
#include "layout.h"

namespace culda {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct ChunkState {
    int32_t doc_start = 0, doc_end = 0;
    int64_t num_tokens = 0;
    int64_t theta_cap = 0;
    // device
    uint32_t *tokens = nullptr;
    uint16_t *z = nullptr;
    int64_t *word_idx = nullptr;
    uint32_t *rev_off = nullptr;
    uint32_t *rev_idx = nullptr;
    DocMeta *meta = nullptr;
    uint32_t *theta = nullptr;
    SlotDesc *slots = nullptr;
    int32_t num_slots = 0;
    PhiSlot *phi_slots = nullptr;  // this chunk alone (chunks-in-turn mode)
    int32_t num_phi_slots = 0;
    int32_t *split_words = nullptr;
    int32_t num_split_words = 0;
    int64_t ll_offset = 0;  // first partial of this chunk in ll_partial
    // host copies needed for sizing / downloads
    std::vector<int64_t> word_idx_h;
    std::vector<uint32_t> meta_start_h;
};

struct Context {
    culda_params prm{};
    hipStream_t stream = nullptr;   // sampler, LL, phi
    hipStream_t stream2 = nullptr;  // theta rebuilds, overlapped with the next chunk's sweep
    std::vector<hipEvent_t> ev_sampled;  // per chunk: its sweep finished
    hipEvent_t ev_theta_done = nullptr;
    int32_t num_words = 0, num_docs = 0;
    int64_t num_tokens = 0;
    uint32_t count_bits = 0;  // packed theta entry: count field width
    bool uploaded = false, theta_ready = false, phi_ready = false, sampled = false;
    std::vector<int64_t> doc_len;  // global
    std::vector<ChunkState> chunks;
    // global device state
    int32_t *nwk = nullptr;
    int32_t *nk = nullptr;
    float *den = nullptr;
    double *ll_partial = nullptr;
    int64_t ll_partials = 0;
    double *ll_out = nullptr;
    double *ll_scratch = nullptr;
    unsigned long long *stats = nullptr;
    ChunkPtrs *chunk_ptrs = nullptr;
    PhiSlot *phi_slots_all = nullptr;
    int32_t num_phi_slots_all = 0;
    int32_t *split_words_all = nullptr;
    int32_t num_split_words_all = 0;
    int32_t cu_count = 0;
    std::vector<void *> allocations;
    // profiling
    hipEvent_t ev[16] = {};
    double sampler_ms = 0, phi_ms = 0, theta_ms = 0, other_ms = 0;
    int32_t sampler_launches = 0;
    double last_sampler_bytes = 0;
    int64_t sum_doc_nnz = 0, sum_token_nnz = 0;
    int64_t last_iter_token_nnz = 0;
    int32_t last_sweep = 0;

    ~Context();
    void *dalloc(size_t bytes);
    void upload(const Corpus &c);
    void init_z();
    void rebuild_phi(int32_t chunk, bool accumulate);
    void phi_head();
    void rebuild_theta(int32_t chunk, hipStream_t s = nullptr);
    void sample(int32_t chunk, int32_t sweep);
    double loglik();
    void iterate(int32_t sweep, double *ll, double *seconds);
    void rebuild_all();
    void read_stats();
    void sync();
};

void hip_check(hipError_t e, const char *what);

}  // namespace culda
