// engine.h -- device-resident models, units and work lists behind the C ABI.
//
// Plays the roles of the reference's DocChunk device buffers
// (model/doc_chunk.h:43-57), ModelPhiGPU (model/model_phi_gpu.h:17-44) and
// ModelThetaChunk (model/model_theta_chunk.h:20-59), owned by one engine
// instead of by model objects that each call cudaMalloc.  One engine holds
// one corpus (culda_ctx: a model with one unit per data_chunk) or a batch
// of independent corpora (culda_batch: one model per corpus).
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <vector>

#include "corpus.h"
#include "layout.h"

namespace culda {

void hip_check(hipError_t e, const char *what);
int64_t theta_entries_max(int64_t nd, int32_t K);
// The documented size limits a corpus must meet for a model with K topics
// (fails with CULDA_ERR_LIMIT); needs no device.
void check_corpus_limits(const Corpus &c, int32_t K);

// A phi work list with its workgroup packets and the hot-word rows to zero.
struct PhiList {
    PhiSlot *slots = nullptr;
    Packet *packets = nullptr;
    int2 *zero_rows = nullptr;  // (model, word)
    int32_t num_slots = 0, num_packets = 0, num_zero_rows = 0;
};

struct ModelHost {
    int32_t V = 0, K = 0, num_docs = 0;
    float alpha = 0, beta = 0;
    uint64_t seed = 0;
    int32_t unit_begin = 0, unit_end = 0;
    int32_t slot_begin = 0, slot_end = 0;  // contiguous in the slot array
    int64_t num_tokens = 0;
    std::vector<int64_t> doc_len;
    int32_t *nwk = nullptr, *nk = nullptr;
    float *den = nullptr;
    float *zero_phi = nullptr;
    double *rden = nullptr;
    PhiList phi_fused;  // all units of the model in one pass
};

struct UnitHost {
    int32_t model = 0;
    int32_t doc_start = 0, doc_end = 0;
    int64_t num_tokens = 0, theta_cap = 0;
    int64_t max_doc_len = 0;
    int32_t slot_begin = 0, slot_end = 0;
    uint32_t *tokens = nullptr;
    uint16_t *z = nullptr;
    uint16_t *zp = nullptr;  // z of the last count build (see Engine::zp_valid)
    uint16_t *zm = nullptr;  // z the rare-word rows hold the counts of (Engine::rare_stale)
    int64_t *word_idx = nullptr;
    uint32_t *rev_off = nullptr, *rev_idx = nullptr;
    DocMeta *meta = nullptr;
    uint16_t *theta = nullptr;
    PhiList phi_single;  // this unit alone (chunks-in-turn accumulation)
    Packet *theta_packets = nullptr;
    int32_t num_theta_packets = 0;
    std::vector<int64_t> word_idx_h;
    // word w's tokens (in doc order) cross into XCD doc region r (docs
    // [D r / kXcds, D (r + 1) / kXcds) of the unit) at word_idx_h[w] +
    // region_h[w * (kXcds + 1) + r]; empty if some word's docs are unsorted
    std::vector<uint32_t> region_h;
};

// Evaluated tokens of the held-out docs (document completion), kept by the
// engine that holds their observed halves.
struct EvalHost {
    uint32_t *tokens = nullptr;  // held-out doc id per evaluated token, word-major
    EvalSlot *slots = nullptr;
    double *partial = nullptr;
    int32_t num_slots = 0;
    int64_t num_tokens = 0;
};

struct MHState;  // learned-MH proposal (mh_engine.cpp)

struct TuneResult {
    int32_t slot_tokens = 0;
    double best_ms = 0;
    std::vector<std::pair<int32_t, double>> tried;  // (slot tokens, ms per sweep)
    // multi-chunk corpora: workgroups per CU of a chunk's theta rebuild
    // (docs per workgroup = the chunk's docs / (that * CUs)), timed over
    // whole iterations; 0 when not tuned
    int32_t theta_groups = 0;
    std::vector<std::pair<int32_t, double>> theta_tried;  // (groups per CU, ms per two iterations)
};

class Engine {
public:
    culda_params prm{};  // k, device, slot_tokens, profile, fused_phi (alpha/beta/seed per model)
    hipStream_t stream = nullptr;   // sweeps, LL, phi
    hipStream_t stream2 = nullptr;  // theta rebuilds overlapped with later sweeps
    int32_t cu_count = 256;
    std::vector<ModelHost> models;
    std::vector<UnitHost> units;
    std::vector<const Corpus *> pending;  // corpora added, not yet uploaded
    std::vector<float> pending_alpha, pending_beta;
    std::vector<uint64_t> pending_seed;
    bool uploaded = false, phi_ready = false, sampled = false;
    // theta_ok[u]: unit u's theta rows have been built from some z (an
    // unbuilt row has no entries, which the sampler must never read)
    std::vector<char> theta_ok;
    bool theta_ready() const {
        if (!uploaded) return false;
        for (char c : theta_ok)
            if (!c) return false;
        return true;
    }
    void set_theta_ready(bool v) { theta_ok.assign(units.size(), v ? 1 : 0); }
    // n_wk / n_k equal the counts of every unit's zp: set by full count
    // rebuilds (which copy z to zp), kept by count deltas, cleared by
    // anything else that writes the counts
    bool zp_valid = false;
    // the n_wk rows of the rare-tile words hold the counts of zm, not of zp;
    // fresh_rows() brings them up to date before anything reads n_wk
    bool rare_stale = false;
    bool graph_rare_stale = false;  // a replay of the captured iteration leaves them stale
    bool graph_key = false;         // one-chunk graph: captured with the zp-based count update
    // rebuilds touch only the entries of rare words (at least a quarter of
    // the words below K / 16 tokens); otherwise every row is rewritten
    bool sparse_phi = false;

    // device tables and work lists
    ModelDev *d_models = nullptr;
    UnitDev *d_units = nullptr;
    SlotDesc *d_slots = nullptr;
    int32_t *d_order = nullptr;  // slot indices in the launch order over all units
    int32_t *d_order_claims = nullptr;  // the same, interleaved so each kQueueClaim run spans all lengths
    int32_t num_slots = 0;
    int64_t rare_slots = 0, rare_tokens = 0;  // slots staged from zp when the counts allow
    double *d_ll_partial = nullptr;
    Packet *d_ll_packets = nullptr;
    int32_t *d_model_packets = nullptr;
    double *d_ll_scratch = nullptr;
    double *d_ll_hist = nullptr;
    // theta entries read per slot, per packet and per (sweep, model): the
    // sampler's algorithmic bytes, summed like the LL
    double *d_nnz_partial = nullptr, *d_nnz_scratch = nullptr, *d_nnz_hist = nullptr;
    int32_t num_ll_packets = 0;
    PhiList phi_all;  // every model, fused over its units
    // the zp-based update (update_phi when zp_valid): words with tokens that
    // no rare tile covers, and the rare tiles of single-unit models
    PhiList phi_sparse;
    RareTile *d_rare_tiles = nullptr;
    Packet *d_rare_packets = nullptr;
    int32_t num_rare_packets = 0;
    int64_t num_rare_tiles = 0;
    Packet *d_theta_packets_all = nullptr;
    int32_t num_theta_packets_all = 0;
    int32_t *d_nk_all = nullptr;  // M*K
    uint32_t *d_sweep = nullptr;
    uint32_t *d_queue = nullptr;
    int32_t persistent_groups = 0;  // 0 = per-slot grid
    int32_t rs_stride = 0;          // sampler S-prefix row stride (longest theta row of any doc)
    hipGraphExec_t graph = nullptr;
    std::vector<void *> allocations;

    // profiling / stats
    std::vector<hipEvent_t> ev_sampled;
    std::vector<hipEvent_t> ev_theta_unit;  // per-chunk iteration: theta of unit u rebuilt (stream2)
    // The per-chunk iteration returns with the last chunk's theta rebuild
    // still running on stream2 (settle() waits for it).
    bool stream2_busy = false;
    std::vector<hipEvent_t> ev_iter;             // profiling: two per iteration of iterate_units_n
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;  // batched iteration: theta on stream2
    hipEvent_t ev[4] = {};
    double sampler_ms = 0, phi_ms = 0, theta_ms = 0, iter_ms = 0;
    int32_t sampler_launches = 0;
    double last_sampler_bytes = 0;

    ~Engine();
    void init_device(const culda_params &p);
    void *dalloc(size_t bytes);
    void dfree(void *p);
    int32_t add_corpus(const Corpus &c, float alpha, float beta, uint64_t seed);
    void upload();
    void rebuild_slots(int32_t slot_tokens);
    int32_t default_slot_tokens() const;

    void set_sweep(uint32_t t, hipStream_t s);
    void init_z();
    // ll (optional): the LL stage 1 rides in the count-clear launch
    void rebuild_phi_model(int32_t model, int32_t unit, bool accumulate, hipStream_t s, const LLArgs *ll = nullptr);
    void rebuild_phi_all(hipStream_t s, const LLArgs *ll = nullptr);
    // a one-unit model after a sweep: count update and theta rebuild in one
    // launch (after the count clear, which carries ll's stage 1)
    void update_counts_one_unit(hipStream_t s, const LLArgs *ll);
    void copy_zp(int32_t model, hipStream_t s);  // zp (and zm) = z for the model's units (all when model < 0)
    void fresh_rows(hipStream_t s);  // n_wk rows current (rare-tile words moved from zm to zp)
    // counts of every model after a sweep: a rebuild that rewrites only the
    // touched entries of rare words when zp_valid, else a full rebuild
    // (fused, or chunks in turn for a single corpus)
    void update_phi(hipStream_t s, const LLArgs *ll = nullptr);
    // head denominators; ll: also the LL stage 2; advance (one model): the
    // sweep counter then moves on to the next sweep
    void phi_head(hipStream_t s, const LLArgs *ll = nullptr, bool advance = false);
    LLArgs ll_args() const;
    void rebuild_theta_unit(int32_t unit, hipStream_t s);
    // per-chunk theta packets.  A single chunk: about groups_per_cu
    // workgroups per CU, at least 16 docs each.  Several chunks (each
    // rebuild beside a sweep, on a grid of groups_per_cu workgroups per
    // CU): 4-doc packets dealt to the XCDs in contiguous runs
    void build_theta_packets(int32_t groups_per_cu);
    int32_t theta_groups_per_cu = 0;  // 0: each kernel's default
    int32_t theta_beside_grid(int32_t ui) const;
    // whether unit ui's rebuild beside a sweep runs the one-wave kernel
    // (16-bit counts: docs below 65,536 tokens; 32-bit byte offsets; docs
    // of at least kSlimMinMeanDocLen tokens on average)
    bool theta_slim(int32_t ui) const;
    // a batch's models on the XCDs, balanced by tokens (model_rank: the
    // models largest first); set at upload
    void assign_models_to_xcds();
    std::vector<int32_t> model_xcd, model_rank;
    // a batch's sampler launch order (slot indices, given longest first):
    // each model's slots on its XCD (workgroup b runs on XCD b mod kXcds)
    std::vector<int32_t> deal_models_to_xcds(const std::vector<int32_t> &by_len, const std::vector<SlotDesc> &all) const;
    void rebuild_theta_all(hipStream_t s);
    void sample_unit(int32_t unit, hipStream_t s);
    void sample_all(hipStream_t s);
    void ll_reduce(hipStream_t s);
    // sum_d entries(d) and sum_d n_d * entries(d) over every unit's current theta rows
    void theta_stats(int64_t &doc_nnz, int64_t &token_nnz);
    // wait for the theta rebuilds a per-chunk iteration left on stream2
    // (before anything else touches device state)
    void settle();
    void rebuild_all();

    // culda_ctx iteration: per-chunk sweeps, theta of chunk c overlapped
    // with the sweep of chunk c + 1
    void iterate_units(uint32_t sweep, double *ll, double *seconds);
    // `count` such iterations from sweep `first` with no host round trip
    // between them; ll[i] = LL/token of sweep first + i
    void iterate_units_n(uint32_t first, int32_t count, double *ll, double *seconds);
    // set_sweep: write the sweep index first (else the previous iteration's
    // head launch has moved the counter on to it)
    void enqueue_units_iteration(uint32_t sweep, bool set_sweep, hipEvent_t ev_a, hipEvent_t ev_b);
    // culda_batch iteration: one launch per step for every model
    void enqueue_batch_iteration(hipStream_t s, hipEvent_t before = nullptr, hipEvent_t after = nullptr);
    // algorithmic bytes of one sweep over every unit that read `entries`
    // theta entries in all
    double sweep_bytes(double entries) const;
    // theta entries read by sweep `sweep` of every model (from the history)
    double sweep_entries(const std::vector<double> &nnz_hist, uint32_t sweep) const;
    void iterate_batch(uint32_t first, int32_t count, bool use_graph, double *ll, double *seconds);
    double model_ll(int32_t model, uint32_t sweep);  // per token, from the history
    TuneResult autotune(const std::vector<int32_t> &candidates, bool batch);
    void sync();

    // held-out log-likelihood (mh_engine.cpp): this engine holds the
    // observed halves of the held-out docs (one corpus, one chunk) and the
    // evaluated halves; counts come from the training engine
    EvalHost eval;
    void set_eval(const Corpus &ev);
    double heldout_ll(Engine &train, int32_t sweeps);

    // learned-MH proposal (mh_engine.cpp); one corpus (model 0)
    std::unique_ptr<MHState> mh;
    void mh_enable(const culda_mh_params &p);
    int32_t mh_collect();
    double mh_train(int32_t epochs);
    double mh_iterate(uint32_t sweep, int32_t proposals);
    void mh_sketch_current(hipStream_t s);
    void mh_query_logits(const uint32_t *d_docs, const uint32_t *d_words, int64_t n, uint16_t *d_logits, hipStream_t s);
    MHState &mh_state();
};

}  // namespace culda
