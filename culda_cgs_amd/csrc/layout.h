// layout.h -- HBM layout shared by the host engine and the gfx950 kernels.
//
// Device state is a set of *models* (one per corpus: n_wk, n_k, head
// denominators, hyper-parameters, Philox seed) and *units* (one per
// data_chunk of a corpus: its tokens, z, doc-major index and theta slabs),
// described by two device tables.  A single corpus with C chunks is one
// model with C units; a batch of corpora is several models.  Work lists
// name a unit (and through it a model), so one launch can serve a single
// chunk, all chunks of a corpus, or every corpus of a batch.
//
// Per unit (all units stay resident):
//   tokens[N_u]   u32  unit-local doc id of each token, word-major (the
//                      reference's wordTokens, doc_chunk.h:39, rebased)
//   z[N_u]        u16  topic of each token, word-major (wordTopics)
//   word_idx[V+1] i64  unit-local token offsets of each word (wordIndices)
//   rev_off[D_u+1] u32 doc-major offsets = cumulative doc lengths
//   rev_idx[N_u]  u32  word-major position of each token, doc-major
//                      (docRevIndices / docRevIdx, doc_chunk.cpp:225-255)
//   meta[D_u]     DocMeta: theta slab start, nnz, n_d
//   theta[cap_u + kThetaPad] u16 entries (count << 10) | topic with
//                      1 <= count <= 63, ascending topic within a doc; a
//                      topic with n_dk > 63 takes ceil(n_dk / 63) consecutive
//                      entries (63, ..., 63, remainder); 0 = padding (count
//                      0).  Slab capacity of doc d is roundup16(min(n_d,
//                      K + ceil(n_d / 63))) entries (model_theta.cpp:49-56
//                      sizes rows by min(n_d, K), roundup32 for 32-lane
//                      warps, with (topic, count) as two int arrays);
//                      kThetaPad padding entries at the end
// Per model: n_wk[V*K] i32 (row per word), n_k[K] i32, den[K] f32.
#pragma once
#include <stdint.h>

namespace culda {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kBlock = 16;         // FP32 prefix block = one DPP row
constexpr int kMaxK = 1024;        // Q block totals scanned by one wave
constexpr int kWavesPerGroup = 4;  // rebuild kernels
constexpr int kLLHistory = 1024;   // LL kept per model for this many sweeps
constexpr int kThetaPad = 64;      // theta entries past the last slab (sampler group loads)
constexpr int kThetaCountShift = 10;    // theta entry: count in bits 10..15
constexpr uint32_t kThetaTopicMask = 1023u;  // topic in bits 0..9 (K <= kMaxK)
constexpr uint32_t kThetaMaxCount = 63u;     // largest count of one entry
constexpr int kJointGrid = 1024;     // workgroups (partials) of each joint-LL launch
constexpr int kJointUnitsMax = 63;   // chunks the allocation-free joint LL serves (its static scratch)
constexpr int kMaxDevices = 64;
constexpr uint32_t kXcds = 8;      // accelerator dies of an MI355X; workgroup i runs on die i mod 8
constexpr int64_t kL2BytesPerXcd = (int64_t)4 << 20;  // each die's L2
constexpr int kQueueClaim = 16;    // slot positions a persistent sampler workgroup claims at once
// The sampler's LDS: phi row and Q block prefixes (static), plus eight
// rows of S block prefixes sized by the longest theta row of any doc
// (dynamic).  Corpora are checked against this budget (the HIP
// per-workgroup limit) before anything is allocated.
constexpr int kSamplerLdsBudget = 65536;
constexpr int kSamplerStaticLds = kMaxK * 4 + kWave * 4;

struct DocMeta {
    uint32_t start;  // theta slab start (entries)
    uint32_t nnz;    // theta entries after the last rebuild
    uint32_t nd;     // doc length
    uint32_t pad;
};

struct ModelDev {
    int32_t *nwk;
    int32_t *nk;
    float *den;
    float *zero_phi;  // phi_k of a zero count, fl(beta / den_k)
    double *rden;     // 1 / den_k in FP64: fl(fl(x * rden_k)) = fl(x / den_k) for every FP32 x
    uint64_t seed;
    double k_alpha;  // (double)K * (double)alpha
    float alpha;
    float beta;
    int32_t k;
    int32_t num_words;
};

struct UnitDev {
    const uint32_t *tokens;
    uint16_t *z;
    uint16_t *zp;               // z the current counts were built from (count deltas)
    const int64_t *word_idx;
    const uint32_t *rev_off;
    const uint32_t *rev_idx;
    DocMeta *meta;
    uint16_t *theta;
    int32_t model;
    int32_t num_docs;
    uint16_t *zm;               // z the rare-word rows of n_wk hold the counts of (one-unit models)
};

// Sampler / z-init work item: tokens [begin, end) of one word of one unit.
// gdelta turns a unit-local token index into the chunking-invariant global
// id g = i + gdelta (the token's index in the one-chunk word-major layout).
struct SlotDesc {
    int32_t word;
    int32_t unit;
    int64_t begin;
    int64_t end;
    int64_t gdelta;
    int32_t rare;  // the word's only slot, in a one-unit model, with fewer than K/16 tokens
    int32_t pad;
};

// Phi work item.  split == 0: every token of `word` in units
// [unit_lo, unit_hi) of the model, written as a whole row (or added to it
// when accumulating).  split == 1: tokens [begin, end) of unit unit_lo,
// added with atomics to a row zeroed beforehand (hot words).
struct PhiSlot {
    int32_t word;
    int32_t unit_lo;
    int32_t unit_hi;
    int32_t split;
    int64_t begin;
    int64_t end;
};

// Items [first, last) of a work list, handled by one workgroup; all of them
// belong to `owner` (a model for phi and LL, a unit for theta).
struct Packet {
    int32_t first;
    int32_t last;
    int32_t owner;
    int32_t pad;
};

struct SampleArgs {
    const SlotDesc *slots;
    const int32_t *order;   // optional permutation of slot indices (balance)
    const UnitDev *units;
    const ModelDev *models;
    double *ll_partial;     // one per slot index
    double *nnz_partial;    // one per slot index: theta entries its tokens read (exact in FP64)
    const uint32_t *sweep;  // sweep index t >= 1, read on the device
    uint32_t *queue;        // persistent mode: next list position; else null
    int32_t num_slots;
    int32_t use_zp;         // counts equal counts(zp): rare slots stage phi from zp
    int32_t rs_stride;      // floats per 8-lane row of S block prefixes (dynamic LDS), sample_rs_stride()
    int32_t rs_stride4;     // floats per 4-lane row (set by launch_sample)
};

struct InitArgs {
    const SlotDesc *slots;
    const UnitDev *units;
    const ModelDev *models;
    int32_t num_slots;
    int32_t pad;
};

struct PhiArgs {
    const PhiSlot *slots;
    const Packet *packets;
    const UnitDev *units;
    const ModelDev *models;
    int32_t num_packets;
    int32_t accumulate;  // 0: write rows; 1: add to rows; 2: write rows, counts of zp known, zp = z
};

// Rare-word tile: consecutive words of one unit whose tokens (at most 64,
// every word whole) sit in z[begin, begin + count).  Words without tokens
// may lie between them; bit r of word_mask marks word w0 + r as present,
// bit i of start_mask marks token i as the first of its word.
struct RareTile {
    uint64_t start_mask;
    uint64_t word_mask;
    int64_t begin;
    int32_t w0;
    int32_t unit;
    int32_t count;
    int32_t pad;
};

// Packets of rare tiles (owner = model).  mode 0 (after a sweep): n_k
// gets the tiles' tokens, zp = z; the rows are left alone.  mode 1 (before
// anything reads n_wk): the rows move from the counts of zm to the counts
// of zp, zm = zp.
struct RareArgs {
    const RareTile *tiles;
    const Packet *packets;
    const UnitDev *units;
    const ModelDev *models;
    int32_t num_packets;
    int32_t mode;
};

struct ThetaArgs {
    const Packet *packets;  // owner = unit, [first, last) = its docs
    const UnitDev *units;
    const ModelDev *models;
    int32_t num_packets;
    int32_t pad;
};

// The per-slot LL partials and gathered-entry counts summed per model, both
// in the same fixed order.
struct LLArgs {
    const double *partial;  // per slot index
    const Packet *packets;  // owner = model, [first, last) = slot indices
    const int32_t *model_packets;  // per model: [begin, end) into packets (2 ints)
    double *scratch;        // one per packet
    double *history;        // [kLLHistory][num_models]
    const double *nnz_partial;  // per slot index
    double *nnz_scratch;        // one per packet
    double *nnz_history;        // [kLLHistory][num_models]: theta entries read by that sweep
    const uint32_t *sweep;
    int32_t num_packets;
    int32_t num_models;
};

// Held-out evaluation work item: evaluated tokens [begin, end) of `word`.
struct EvalSlot {
    int32_t word;
    int32_t pad;
    int64_t begin;
    int64_t end;
};

// Document-completion log-likelihood of evaluated tokens (FP64): per token
// log(sum_k (n_dk + alpha) / (n_d + K alpha) * (n_wk + beta) / (n_k + V beta))
// with n_dk, n_d from the fold-in theta of the held-out doc.
struct HeldoutArgs {
    const EvalSlot *slots;
    const uint32_t *tokens;  // held-out doc id of each evaluated token, word-major
    const DocMeta *meta;     // fold-in theta of the held-out docs
    const uint16_t *theta;
    const int32_t *nwk;
    const int32_t *nk;
    double *partial;         // one per slot
    double alpha;
    double beta;
    int32_t num_slots;
    int32_t k;
    int32_t num_words;
    int32_t pad;
};

// ------------------------------------------------------------ learned MH
// MLP proposal q(k | d, w) = softmax(W2^T GELU(W1^T [sketch(d); emb(w)] +
// b1) + b2): a 256-wide signed count sketch of the doc's topic counts, a
// 128-wide word embedding, 512 hidden units (tanh-GELU), K logits.  BF16
// operands on MFMA, FP32 accumulation, FP32 master weights (Adam).
constexpr int kSketch = 256;
constexpr int kEmb = 128;
constexpr int kIn = kSketch + kEmb;  // 384
constexpr int kHidden = 512;
constexpr int kAliasBits = 20;       // proposal weights are integers summing to K << kAliasBits
constexpr int kMaxProposals = 8;     // MH proposals per token and sweep
constexpr int kMHThetaEntries = 2048;  // longest theta row the MH step holds in registers
constexpr int kGemmTile = 64;        // C tile (rows and columns) of the BF16 GEMM
constexpr int kGemmK = 32;           // reduction step of the BF16 GEMM

// C[m x n] = sum_k A(i, k) B(k, j), A and B in BF16 (raw bits), A stored
// [m][k] (lda) and B stored [n][k] (ldb).  m, n multiples of kGemmTile, k
// a multiple of kGemmK, every row start 16-byte aligned.
enum GemmEpilogue : int32_t {
    kEpiStore = 0,     // c = acc (+ bias[j])
    kEpiBiasGelu = 1,  // c = acc + bias[j] (optional), c16 = bf16(gelu(acc + bias[j]))
    kEpiGeluGrad = 2,  // c16 = bf16(acc * gelu'(aux[i][j]))
};

struct GemmArgs {
    const uint16_t *a;
    const uint16_t *b;
    float *c;              // FP32 output [m][n] (ldc) or null
    uint16_t *c16;         // BF16 output [m][n] (ldc16) or null
    const float *bias;     // per column, or null
    const float *aux;      // kEpiGeluGrad: pre-activations [m][n] (ldaux)
    int64_t lda, ldb, ldc, ldc16, ldaux;
    int32_t m, n, k;       // k: reduction depth of one split
    int32_t epilogue;
    int64_t split_stride;  // split z (gridDim.z > 1) covers k range [z k, (z+1) k) and writes c + z * split_stride
};

// Per-token query of the proposal: doc (global id) and word of every token
// of the corpus, in unit order (tokens of unit u at [unit_base_u, ...)).
struct MHTokenInfo {
    const uint32_t *doc;   // global doc id
    const uint32_t *word;
};

// Inputs of the MH step for a run of tokens of one unit.
struct MHArgs {
    const uint16_t *logits;     // [n][K] BF16 proposal logits (rows = tokens start .. start+n-1)
    const uint32_t *word;       // word id per token of the unit
    const int64_t *gdelta;      // per word: global id = unit-local index + gdelta[w]
    UnitDev unit;               // tokens (doc ids), z, meta, theta
    ModelDev model;
    unsigned long long *accepted;  // += accepted proposals
    int32_t *w_out;             // debug: [n][K] integer weights, or null
    uint32_t *f_out;            // debug: [n][K] bucket thresholds, or null
    uint32_t *a_out;            // debug: [n][K] aliases, or null
    int64_t start;              // first token (unit-local index)
    int32_t n;
    int32_t proposals;
    uint32_t sweep;
    int32_t pad;
};

}  // namespace culda
