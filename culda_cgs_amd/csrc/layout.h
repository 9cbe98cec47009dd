// layout.h -- HBM layout shared by the host context and the gfx950 kernels.
//
// Per chunk (all chunks stay resident):
//   tokens[N_c]   u32  chunk-local doc id of each token, word-major (the
//                      reference's wordTokens, doc_chunk.h:39, rebased)
//   z[N_c]        u16  topic of each token, word-major (wordTopics)
//   word_idx[V+1] i64  chunk-local token offsets of each word (wordIndices)
//   rev_off[D_c+1] u32 doc-major offsets = cumulative doc lengths
//   rev_idx[N_c]  u32  word-major position of each token, doc-major
//                      (docRevIndices / docRevIdx, doc_chunk.cpp:225-255)
//   meta[D_c]     DocMeta: theta slab start, nnz, n_d
//   theta[cap_c]  u32  packed ((4 * topic) << count_bits) | count,
//                      count_bits = 30 - ceil(log2 K) (20 at K=1024), so
//                      e >> count_bits is the byte offset of phi_topic in
//                      LDS; ascending topic within a doc; slab capacity of doc d is
//                      roundup16(min(n_d, K)) (model_theta.cpp:49-56 uses
//                      roundup32 for 32-lane warps)
// Global: n_wk[V*K] i32 (row per word), n_k[K] i32, den[K] f32.
#pragma once
#include <stdint.h>

namespace culda {

constexpr int kWave = 64;       // CDNA wavefront
constexpr int kBlock = 16;      // FP32 prefix block = one DPP row
constexpr int kMaxK = 1024;     // Q block totals scanned by one wave
constexpr int kWavesPerGroup = 4;   // rebuild kernels
constexpr int kSampleWaves = 1;     // sampler: one wave per workgroup

// Sampler work unit: tokens [begin, end) of one word, one wavefront.
// gdelta turns a chunk-local token index into the chunking-invariant
// global id: g = i + gdelta.
struct SlotDesc {
    int32_t word;
    int32_t pad;
    int64_t begin;
    int64_t end;
    int64_t gdelta;
};

struct DocMeta {
    uint32_t start;  // theta slab start (entries)
    uint32_t nnz;    // nonzero topics after the last rebuild
    uint32_t nd;     // doc length
    uint32_t pad;
};

// Phi rebuild work unit.  chunk < 0: every token of `word` in the launch's
// chunk range, written as a whole row.  chunk >= 0: tokens [begin, end) of
// chunk `chunk`, added to a pre-zeroed row with atomics (hot words).
struct PhiSlot {
    int32_t word;
    int32_t chunk;
    int64_t begin;
    int64_t end;
};

struct ChunkPtrs {
    const uint16_t *z;
    const int64_t *word_idx;
};

struct SampleArgs {
    const SlotDesc *slots;
    const uint32_t *tokens;
    uint16_t *z;
    const DocMeta *meta;
    const uint32_t *theta;
    const int32_t *nwk;
    const float *den;
    double *ll_partial;
    uint64_t seed;
    double k_alpha;  // (double)K * (double)alpha
    float alpha;
    float beta;
    int32_t num_slots;
    int32_t k;
    uint32_t count_bits;
    uint32_t sweep;
};

struct PhiArgs {
    const PhiSlot *slots;
    const ChunkPtrs *chunks;
    int32_t *nwk;
    int32_t *nk;
    int32_t num_slots;
    int32_t k;
    int32_t chunk_begin;
    int32_t chunk_end;
    int32_t accumulate;
};

struct ThetaArgs {
    const uint32_t *rev_off;
    const uint32_t *rev_idx;
    const uint16_t *z;
    DocMeta *meta;
    uint32_t *theta;
    unsigned long long *stats;  // [0] += sum nnz, [1] += sum n_d * nnz
    int32_t num_docs;
    int32_t k;
    uint32_t count_bits;
};

struct InitArgs {
    const SlotDesc *slots;
    uint16_t *z;
    uint64_t seed;
    int32_t num_slots;
    int32_t k;
};

}  // namespace culda
