/*
 * culda.h -- C ABI of the MI355X (gfx950) collapsed-Gibbs LDA sampler.
 *
 * Drop-in boundary for the CuLDA_CGS hot path: one plain CGS iteration
 * (per-token sampler with fused log-likelihood, phi count rebuild with
 * on-device accumulation over chunks, sparse theta rebuild) plus the
 * data_chunk corpus layout and the ModelPhi / ModelTheta save hooks.
 *
 * Conventions
 *   - Every call returns CULDA_OK (0) or a negative culda_status; the
 *     message of the last failure on the calling thread is returned by
 *     culda_last_error().  The library never calls exit().
 *   - Plain pointers and sizes only.  Host arrays passed in are owned by
 *     the caller and copied; arrays handed out by culda_corpus_* stay
 *     owned by the corpus object.
 *   - A culda_ctx (one corpus) or culda_batch (many corpora) owns all its
 *     device memory and two HIP streams.  A handle is used by one thread
 *     at a time; different handles are independent.
 *
 * Reference interfaces replaced are cited as file:line relative to the
 * CuLDA_CGS source tree (src_culda/, src_format/).
 */
#ifndef CULDA_H_
#define CULDA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CULDA_ABI_VERSION 1

typedef enum culda_status {
    CULDA_OK = 0,
    CULDA_ERR_ARG = -1,      /* invalid argument                                 */
    CULDA_ERR_IO = -2,       /* file missing, unreadable or malformed            */
    CULDA_ERR_FORMAT = -3,   /* data_chunk files inconsistent with each other    */
    CULDA_ERR_DEVICE = -4,   /* HIP runtime failure (no GPU, launch failure ...) */
    CULDA_ERR_MEMORY = -5,   /* host or device allocation failed                 */
    CULDA_ERR_STATE = -6,    /* call out of order (e.g. sample before upload)    */
    CULDA_ERR_LIMIT = -7     /* size beyond a documented limit                   */
} culda_status;

/* Message for the last failing call on this thread ("" if none). */
const char *culda_last_error(void);
/* CULDA_ABI_VERSION of the loaded library. */
int culda_abi_version(void);

/* ------------------------------------------------------------------------
 * Host corpus: the src_format data_chunk layout.
 *
 * For a prefix P and C chunks the files are
 *   P.vocab          text "token id" per word            (format.cpp:97-99)
 *   P.doc.idx        text "docId cumTokens" per doc       (format.cpp:88)
 *   P.chunk          text "chunkId docIdEnd" per chunk    (format.cpp:131)
 *   P.word.idx[c]    text "wordId cumEnd" for all V words, chunk-local
 *   P.word.data[c]   raw int32 global doc ids, word-major, ascending doc
 *                    within a word, repeats allowed     (data_chunk.h:49-67)
 * The [c] suffix is present only when C > 1 (data_chunk.h:37-43).
 * ---------------------------------------------------------------------- */
typedef struct culda_corpus culda_corpus;

typedef struct culda_corpus_info {
    int32_t num_words;    /* V                                             */
    int32_t num_docs;     /* D                                             */
    int32_t num_chunks;   /* C                                             */
    int32_t max_doc_len;
    int64_t num_tokens;   /* N                                             */
} culda_corpus_info;

typedef struct culda_chunk_view {
    int32_t docid_start;          /* first global doc id of the chunk       */
    int32_t docid_end;            /* one past the last                      */
    int64_t num_tokens;           /* tokens of the chunk                    */
    const int64_t *word_indices;  /* [V+1] chunk-local cumulative offsets   */
    const int32_t *word_tokens;   /* [num_tokens] global doc ids, word-major*/
} culda_chunk_view;

/* Load P.doc.idx, P.chunk, P.word.idx[c], P.word.data[c]; num_chunks must
 * match P.chunk.  Replaces Document::loadDocument (model/doc.cpp:62-156)
 * and DocChunk::loadChunk (model/doc_chunk.cpp:106-256), which exit(0)
 * on error. */
int culda_corpus_load(const char *prefix, int32_t num_chunks, culda_corpus **out);

/* Build a corpus from arrays already in the data_chunk layout.
 * doc_cum[D] holds cumulative token counts (the second column of
 * P.doc.idx); chunk_doc_end[C] the second column of P.chunk;
 * word_indices[c] points to V+1 offsets (first 0), word_tokens[c] to the
 * doc ids of chunk c. */
int culda_corpus_from_arrays(int32_t num_words, int32_t num_docs, const int64_t *doc_cum,
                             int32_t num_chunks, const int32_t *chunk_doc_end,
                             const int64_t *const *word_indices,
                             const int32_t *const *word_tokens, culda_corpus **out);

/* Write the corpus as data_chunk files under prefix.  Word tokens in
 * P.vocab are the loaded/formatted ones, or "w<id>" for generated
 * corpora.  Replaces DataChunk::writeChunk (src_format/data_chunk.h:49-67)
 * and the vocab/doc.idx/chunk writers of src_format/format.cpp:88-131. */
int culda_corpus_write(const culda_corpus *corpus, const char *prefix);

/* Text corpus ("title tok tok ..." per line) -> data_chunk files, the
 * behaviour of ./format input output_prefix numChunks
 * (src_format/format.cpp:19-168): word ids in first-seen order, lines
 * without a title skipped, docs split greedily into chunks of
 * ceil(N/C) tokens. */
int culda_format_text(const char *input_path, const char *prefix, int32_t num_chunks);

/* Re-split a corpus into num_chunks chunks with the same greedy rule. */
int culda_corpus_rechunk(const culda_corpus *corpus, int32_t num_chunks, culda_corpus **out);

/* Document-completion split for held-out log-likelihood (no counterpart in
 * the reference, which reports training LL only).  Docs d with
 * d % every == every - 1 are held out: *train gets the other docs
 * (renumbered in order, train_chunks chunks); *observed and *eval get the
 * held-out docs (renumbered in order, one chunk each), whose tokens, taken
 * in ascending word order, alternate between observed (even positions) and
 * evaluated (odd positions).  All three keep the vocabulary size.  Free
 * each with culda_corpus_free. */
int culda_corpus_heldout_split(const culda_corpus *corpus, int32_t every, int32_t train_chunks,
                               culda_corpus **train, culda_corpus **observed, culda_corpus **eval);

/* Parameters of the synthetic LDA generator (seeded, deterministic for a
 * given parameter set, independent of the thread count).
 *   phi_k ~ Dirichlet(topic_word_prior) over V, multiplied by
 *   (rank+1)^-zipf_s and renormalised; theta_d ~ Dirichlet(doc_topic_prior)
 *   over true_topics and z ~ theta_d, drawn together as the Polya urn that
 *   is their marginal (O(1) per token); w ~ phi_z by an alias table; doc
 *   lengths Poisson(len_mean) (len_kind 0, zero lengths redrawn) or
 *   round(exp(Normal(len_mu, len_sigma))) (len_kind 1, at least 1). */
typedef struct culda_gen_params {
    int32_t num_words;
    int32_t num_docs;
    int32_t true_topics;
    int32_t num_chunks;
    int32_t len_kind;        /* 0 = Poisson, 1 = lognormal                 */
    int32_t num_threads;     /* 0 = all hardware threads                   */
    double  len_mean;
    double  len_mu;
    double  len_sigma;
    double  doc_topic_prior;
    double  topic_word_prior;
    double  zipf_s;
    uint64_t seed;
} culda_gen_params;

int culda_corpus_generate(const culda_gen_params *params, culda_corpus **out);

int culda_corpus_info_get(const culda_corpus *corpus, culda_corpus_info *out);
/* doc_cum[D]: cumulative doc lengths (P.doc.idx second column). */
int culda_corpus_doc_cum(const culda_corpus *corpus, const int64_t **doc_cum);
int culda_corpus_chunk(const culda_corpus *corpus, int32_t chunk, culda_chunk_view *out);
void culda_corpus_free(culda_corpus *corpus);

/* ------------------------------------------------------------------------
 * Device model.
 * ---------------------------------------------------------------------- */
typedef struct culda_ctx culda_ctx;

typedef struct culda_params {
    int32_t  k;              /* topics; multiple of 16, 16 <= k <= 1024     */
    float    alpha;          /* doc-topic smoothing (reference -a)          */
    float    beta;           /* topic-word smoothing (reference -b)         */
    uint64_t seed;           /* Philox key; z init and every sweep derive
                                from (seed, global token id, iteration)    */
    int32_t  device;         /* HIP device ordinal                          */
    int32_t  slot_tokens;    /* sampler tokens per wave slot (0 = default:
                                about 32 slots per CU per launch, 128..1024) */
    int32_t  profile;        /* 1 = time each kernel with HIP events        */
    int32_t  fused_phi;      /* 1 = rebuild phi over all chunks in one pass */
} culda_params;

/* Create a context on params->device.  Replaces the Argument plumbing and
 * ModelPhi / ModelTheta constructors (train/lda_train.cu:44-50). */
int culda_create(const culda_params *params, culda_ctx **out);
void culda_destroy(culda_ctx *ctx);

/* Check a corpus against the documented size limits for k topics, without
 * a device (culda_upload and culda_batch_add run the same checks first):
 * k a multiple of 16 in [16, 1024] (CULDA_ERR_ARG); fewer than 2^31 - 1
 * tokens (int32 counts); fewer than 2^32 tokens and 2^31 theta entries per
 * chunk; every doc's theta row (at most min(n_d, k + ceil(n_d / 63))
 * entries) within the sampler's LDS, 30,576 entries (about 1.86M tokens
 * per doc at k = 1024).  Returns CULDA_ERR_LIMIT naming the first limit
 * exceeded.  No counterpart in the reference, which sizes rows from
 * ModelTheta::InitData (model/model_theta.cpp:44-86) and fails inside a
 * kernel when they do not fit. */
int culda_check_corpus(const culda_corpus *corpus, int32_t k);

/* Copy the corpus to HBM and build the per-chunk layouts (word-major doc
 * stream, doc-major token index, theta slabs, wave slots).  Replaces
 * DocChunk::allocGPU/toGPU (model/doc_chunk.cpp:277-356), ModelPhi::InitData
 * (model/model_phi.cpp:51-68) and ModelTheta::InitData
 * (model/model_theta.cpp:44-86).  All chunks stay resident. */
int culda_upload(culda_ctx *ctx, const culda_corpus *corpus);

/* z_i = mulhi64(Philox4x64-10((g_i, 0, 0, 0), (seed, 0))[0], k) for every
 * token; g_i is the token's index in the one-chunk word-major layout.
 * Replaces DocChunk::generateTopics (model/doc_chunk.cpp:259-264). */
int culda_init_z(culda_ctx *ctx);

/* Replace / read z of one chunk (word-major order, chunk-local). */
int culda_set_z(culda_ctx *ctx, int32_t chunk, const uint16_t *z, int64_t n);
int culda_get_z(culda_ctx *ctx, int32_t chunk, uint16_t *z, int64_t n);

/* Rebuild n_wk [V*k] and n_k [k] from z.  chunk = -1 rebuilds from all
 * chunks in one pass; chunk >= 0 adds that chunk's counts, overwriting
 * when accumulate == 0.  Replaces LDAUpdatePhiAPI
 * (kernel/lda_phi_kernel.cu:136-159) and, for chunks in turn,
 * LDAUpdatePhiReduceAPI (kernel/lda_phi_kernel.cu:185-207). */
int culda_rebuild_phi(culda_ctx *ctx, int32_t chunk, int32_t accumulate);

/* Prepare the head denominators den_k = (float)n_k + beta*(float)V used by
 * every sampler launch.  Replaces LDAComputePhiHeadAPI
 * (kernel/lda_phi_kernel.cu:49-72); the head itself is formed on the fly. */
int culda_phi_head(culda_ctx *ctx);

/* Rebuild the sparse doc-topic rows of one chunk (ascending topics).
 * Replaces LDAUpdateThetaAPI (kernel/lda_theta_kernel.cu:89-125) and
 * ModelThetaChunk::UpdateThetaGPU (model/model_theta_chunk.cpp:106-140). */
int culda_rebuild_theta(culda_ctx *ctx, int32_t chunk);

/* One sweep over the tokens of one chunk with frozen counts (sweep index
 * iteration >= 1 keys the Philox stream).  Writes new z and the chunk's
 * FP64 log-likelihood partials.  Replaces the LDAKernelTrain launch
 * (train/SingleChunkSingleGPU.h:75-100, kernel/lda_train_kernel.cu:20-271). */
int culda_sample(culda_ctx *ctx, int32_t chunk, int32_t iteration);

/* Sum of the log-likelihood terms of the last sample() of every chunk,
 * divided by N (FP64, deterministic order).  Replaces LDATrainPerplexity
 * (kernel/lda_train_kernel.cu:323-386). */
int culda_loglik(culda_ctx *ctx, double *ll_per_token);

/* One full iteration: sample every chunk, LL, phi rebuild + head, theta
 * rebuild.  ll = predictive LL/token of the pre-sweep model, seconds = wall
 * time until the LL is known.  The last chunks' theta rebuilds may still be
 * running when it returns: the next culda_iterate orders each chunk's sweep
 * after that chunk's rebuild, and every other call (culda_sampler_bytes and
 * culda_last_sweeps aside) waits for them first.  Replaces one pass of the
 * loop body train/SingleChunkSingleGPU.h:70-127 (and, for C > 1, of
 * train/MultiChunkMultiGPUequal.h:88-169 on one device). */
int culda_iterate(culda_ctx *ctx, int32_t iteration, double *ll, double *seconds);

/* count culda_iterate steps (sweeps first_iteration .. first_iteration +
 * count - 1, count <= 1024) enqueued back to back with no host round trip
 * between them; ll[i] = LL/token of sweep first_iteration + i (ll may be
 * NULL), seconds = wall time until the last LL is known.  culda_last_sweeps
 * then reports the sums over the count iterations.  Same chain as count
 * culda_iterate calls (no counterpart in the reference: its loop reads the
 * LL back every iteration).  A one-chunk corpus replays its iteration as a
 * HIP graph (captured on the first call): no per-sweep timing then, and
 * culda_last_sweeps reports a sampler time of 0. */
int culda_iterate_n(culda_ctx *ctx, int32_t first_iteration, int32_t count, double *ll, double *seconds);

/* Rebuild phi, head and theta from the current z (initial model state,
 * train/SingleChunkSingleGPU.h:24-41). */
int culda_rebuild_all(culda_ctx *ctx);

/* Joint log p(w, z) in FP64 with lgamma over the current counts. */
int culda_loglik_joint(culda_ctx *ctx, double *out);

/* culda_loglik_joint without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (serving
 * loops, code under HIP graph capture, harnesses that forbid hipMalloc).
 * The partial sums go to a static device array of the library (one per
 * device, 64 * 1024 doubles plus the result, held until the process ends)
 * instead of a buffer allocated per call, and the total is formed on the
 * device.  At most 63 chunks (else CULDA_ERR_LIMIT).  Every call on a
 * device writes all the scratch it reads; calls that overlap on one device
 * share that array and race, so serialise them. */
int culda_loglik_joint_noalloc(culda_ctx *ctx, double *out);

int culda_sync(culda_ctx *ctx);

/* Counts to host.  n_wk [V*k] row-major by word, n_k [k].  Replaces
 * ModelPhi::MasterGPUToCPU (model/model_phi.cpp:88-103). */
int culda_get_phi(culda_ctx *ctx, int32_t *n_wk, int32_t *n_k);
/* phi head (n_wk + beta)/(n_k + beta V) in FP32, [V*k]. */
int culda_get_phi_head(culda_ctx *ctx, float *head);

/* Theta rows of one chunk in compact CSR over the chunk's docs:
 * offsets[chunk_docs+1], counts[nnz_total], topics[nnz_total], one
 * (topic, n_dk) pair per nonzero topic, ascending topic (the device's split
 * entries of a topic are merged).  Call with counts == topics == NULL to
 * obtain offsets (and so the total) first.
 * Replaces ModelThetaChunk::toCPU (model/model_theta_chunk.cpp:288-323). */
int culda_get_theta(culda_ctx *ctx, int32_t chunk, int64_t *offsets, int32_t *counts,
                    int16_t *topics);

/* Save hooks (host side, after a download).
 * save_phi: P.phi.bin  = int32 V, int32 k, int32 n_wk[V*k], int32 n_k[k]
 *           P.phi.txt  = the text dump of ModelPhi::savePhi
 *                        (model/model_phi.cpp:161-191) for every word.
 * save_theta: P.theta.bin = int32 D, int32 k, int64 offsets[D+1],
 *           int32 counts[nnz], int16 topics[nnz] (global doc order);
 *           P.theta.txt = "doc | total (topic,count) ..." per doc, the
 *           commented saveDoc of model/model_theta.cpp:207-241. */
int culda_save_phi(culda_ctx *ctx, const char *prefix);
int culda_save_theta(culda_ctx *ctx, const char *prefix);

/* Checkpoint = (z of every chunk, next iteration, seed): all counts derive
 * from z and the random stream is counter-based, so resume is exact. */
int culda_save_checkpoint(culda_ctx *ctx, const char *path, int32_t next_iteration);
int culda_load_checkpoint(culda_ctx *ctx, const char *path, int32_t *next_iteration);

typedef struct culda_stats {
    int64_t num_tokens;
    int64_t num_slots;          /* sampler wave slots over all chunks          */
    int64_t sum_doc_nnz;        /* sum_d entries(d) after the last theta rebuild
                                   (16-bit entries; a topic with n_dk > 63 takes
                                   ceil(n_dk / 63) of them)                      */
    int64_t sum_token_nnz;      /* sum_d n_d * entries(d): theta entries gathered */
    int64_t theta_capacity;     /* theta slab entries over all chunks           */
    double  sampler_ms;         /* last iteration: the chunks' sweeps (profile=1) */
    double  phi_ms;             /* LL reduce + phi rebuild + head denominators    */
    double  theta_ms;           /* 0: the theta rebuilds run on a second stream,
                                   overlapped with the sweeps and the next
                                   iteration                                    */
    double  other_ms;
    int32_t sampler_launches;   /* launches in the last iteration               */
    int32_t pad;
} culda_stats;

int culda_get_stats(culda_ctx *ctx, culda_stats *out);

/* Algorithmic bytes moved by the last sampler launches of the last
 * iteration (doc id + meta + z write + theta entries gathered + count rows
 * staged), summed over chunks. */
int culda_sampler_bytes(culda_ctx *ctx, double *bytes);

/* The last culda_iterate's (culda_iterate_n's: summed over its
 * iterations) sampler time (HIP events; params->profile = 1, else 0),
 * algorithmic bytes (as culda_sampler_bytes) and launch count, without
 * waiting for the theta rebuilds it left running (no counterpart in the
 * reference; for measurement loops). */
int culda_last_sweeps(culda_ctx *ctx, double *sampler_ms, double *sampler_bytes, int32_t *launches);

/* Autotuner result: the sampler's tokens per wave slot (and, for batches,
 * the persistent grid) chosen by timing one sweep per candidate from the
 * same state.  z, counts and every later result are unchanged by tuning:
 * the sampler's output does not depend on the launch shape. */
typedef struct culda_tune {
    int32_t slot_tokens;          /* chosen tokens per wave slot             */
    int32_t persistent_groups;    /* batches: chosen workgroups (0 = one per slot) */
    int32_t num_tried;
    int32_t theta_groups;         /* several chunks: workgroups per CU of a chunk's theta
                                     rebuild beside the next sweep (the launch's grid; one-wave
                                     workgroups, four-wave ones where a doc has 65,536 tokens or
                                     more), timed over whole iterations; 0 = not tuned */
    double  best_ms;              /* sampler time of the chosen candidate    */
    int32_t tried_slot_tokens[32];
    double  tried_ms[32];
    double  theta_tried_ms[8];    /* two iterations' time per theta_groups candidate
                                     1, 2, 3, 4, 6, 8 (0 where not tried)       */
} culda_tune;

/* Time slot sizes 128..4096 for this corpus and keep the fastest; with
 * several chunks, then time two whole iterations for each theta-rebuild
 * workgroup count (1..8 per CU) and keep the fastest.  z, the counts and
 * theta are left as they were.  Needs the initial rebuild
 * (culda_rebuild_all). */
int culda_autotune(culda_ctx *ctx, culda_tune *out);

/* STREAM-style device copy of n_bytes (multiple of 16) between two device
 * buffers owned by the context, timed with HIP events: achievable HBM
 * bandwidth for the roofline.  gbps = 2*n_bytes/time. */
int culda_stream_copy(culda_ctx *ctx, int64_t n_bytes, int32_t repeats, double *gbps);

/* ------------------------------------------------------------------------
 * Held-out log-likelihood by document completion (no counterpart in the
 * reference).  Split with culda_corpus_heldout_split; train one context on
 * *train; create a second context ("fold") with the same k, alpha, beta
 * (its own seed), upload *observed into it and give it *eval here.
 * ---------------------------------------------------------------------- */
int culda_heldout_set_eval(culda_ctx *fold, const culda_corpus *eval);
/* Copies the training context's n_wk / n_k into the fold context (phi
 * fixed), draws the observed tokens' z from the fold context's seed, runs
 * fold_in_sweeps plain sweeps (each followed by a theta rebuild), then
 * returns the mean over evaluated tokens of
 *   log sum_k (n_dk + alpha) / (n_d + K alpha) * (n_wk + beta) / (n_k + V beta)
 * with n_dk, n_d the fold-in counts of the token's doc, in FP64.  The same
 * inputs give the same value on every call. */
int culda_heldout_ll(culda_ctx *train, culda_ctx *fold, int32_t fold_in_sweeps, double *ll_per_token);

/* ------------------------------------------------------------------------
 * Learned Metropolis-Hastings proposal (no counterpart in the reference).
 * An MLP maps [256-wide signed sketch of the doc's topic counts / n_d;
 * 128-wide learned word embedding] through 512 tanh-GELU units to K
 * logits (BF16 operands on the matrix cores, FP32 accumulation, FP32
 * master weights trained with Adam on (doc, word, z) triples collected from
 * the plain sampler's own chain).  A sweep then proposes, for every token,
 * from the integer-weight alias table of its softmax and accepts with the
 * exact MH ratio of the token's collapsed conditional (the plain sampler's
 * frozen-count conditional), so the plain sampler's per-token target is
 * left invariant.  Requires k % 64 == 0; a corpus of several data_chunks
 * moves every chunk's tokens with the same frozen counts (the chunks-in-turn
 * schedule of train/MultiChunkMultiGPUequal.h:88-172), each chunk's docs
 * through its own theta rows.
 * ---------------------------------------------------------------------- */
typedef struct culda_mh_params {
    int32_t batch;          /* training rows per Adam step, multiple of 128 (0: 8192)  */
    int32_t infer_batch;    /* tokens per proposal launch, multiple of 128 (0: 65536) */
    int32_t max_snapshots;  /* training snapshots kept (0: 21)                         */
    int32_t pad;
    float lr, beta1, beta2, eps;  /* Adam (0: 1e-3, 0.9, 0.999, 1e-8)                 */
    uint64_t init_seed;     /* weight initialisation and training order                */
} culda_mh_params;

/* Allocate the MLP, its optimiser state and buffers for the uploaded corpus. */
int culda_mh_enable(culda_ctx *ctx, const culda_mh_params *params);
/* Keep the current state (z and every doc's sketch, from the last rebuild)
 * as one snapshot of training triples; returns the number of snapshots. */
int culda_mh_collect(culda_ctx *ctx, int32_t *num_snapshots);
/* epochs passes over every (snapshot, token) triple in a seeded order;
 * mean cross-entropy of the last epoch. */
int culda_mh_train(culda_ctx *ctx, int32_t epochs, double *mean_loss, double *seconds);
/* One learned-MH iteration: `proposals` MH steps per token from the
 * current state, then the count rebuilds of a plain iteration. */
int culda_mh_iterate(culda_ctx *ctx, int32_t iteration, int32_t proposals, double *accept_rate, double *seconds);
/* Weights (row-major, input x output): w1 384x512, b1 512, w2 512xK, b2 K,
 * emb Vx128; get copies the FP32 masters out, set replaces them. */
int culda_mh_get_weights(culda_ctx *ctx, float *w1, float *b1, float *w2, float *b2, float *emb);
int culda_mh_set_weights(culda_ctx *ctx, const float *w1, const float *b1, const float *w2, const float *b2,
                         const float *emb);
/* Logits [n][K] for (doc, word) queries, doc sketches from the current
 * counts: FP32 values of the BF16 logits the proposal uses (the last
 * product stores BF16). */
int culda_mh_logits(culda_ctx *ctx, int64_t n, const int32_t *docs, const int32_t *words, float *logits);
/* Gradients of the mean cross-entropy over n rows (n <= batch) of
 * (doc, word, target) at the current weights, without updating them;
 * train_step = 1 also applies one Adam step with them. */
int culda_mh_gradients(culda_ctx *ctx, int32_t n, const int32_t *docs, const int32_t *words, const int32_t *targets,
                       int32_t train_step, double *loss, float *dw1, float *db1, float *dw2, float *db2, float *demb);
/* Proposal tables of tokens [first, first + n) (n <= infer_batch): integer
 * weights, alias thresholds and aliases, [n][K] each; z is unchanged.
 * Tokens are counted over the chunks in order (chunk c's tokens follow
 * chunk c-1's, word-major within a chunk); the range lies in one chunk. */
int culda_mh_tables(culda_ctx *ctx, int64_t first, int32_t n, int32_t *weights, uint32_t *thresholds,
                    uint32_t *aliases);
/* Current doc sketches, [D][256] BF16 bit patterns. */
int culda_mh_sketch(culda_ctx *ctx, uint16_t *out);

/* ------------------------------------------------------------------------
 * Batch: many independent corpora sampled together (no counterpart in the
 * reference, which runs one corpus per process).  Each corpus is its own
 * model (counts, alpha, beta, seed); one iteration of the whole batch is a
 * single sampler launch over every corpus's work (a persistent launch with
 * an atomic work queue when set), one LL, one phi, one head and one theta
 * launch, and may be captured once as a HIP graph and replayed.  Every
 * corpus's chain is bit-identical to the same corpus run alone with
 * culda_ctx and the same k, alpha, beta and seed.
 * ---------------------------------------------------------------------- */
typedef struct culda_batch culda_batch;

/* params->k, device, slot_tokens, profile are used; alpha, beta and seed
 * are given per corpus. */
int culda_batch_create(const culda_params *params, culda_batch **out);
void culda_batch_destroy(culda_batch *batch);
/* The corpus must stay alive until culda_batch_upload returns. */
int culda_batch_add(culda_batch *batch, const culda_corpus *corpus, float alpha, float beta, uint64_t seed,
                    int32_t *model);
int culda_batch_upload(culda_batch *batch);
/* z init and the initial rebuild of every model. */
int culda_batch_init(culda_batch *batch);
/* groups > 0: persistent sampler with that many workgroups taking work
 * from an atomic queue; 0: one workgroup per work item. */
int culda_batch_set_persistent(culda_batch *batch, int32_t groups);
/* Iterations first .. first+count-1 (count <= 1024).  ll[i * M + m] =
 * LL/token of model m at iteration first+i; use_graph = 1 captures one
 * iteration as a HIP graph (once) and replays it. */
int culda_batch_iterate(culda_batch *batch, int32_t first_iteration, int32_t count, int32_t use_graph,
                        double *ll, double *seconds);
int culda_batch_autotune(culda_batch *batch, culda_tune *out);
int culda_batch_get_z(culda_batch *batch, int32_t model, int32_t chunk, uint16_t *z, int64_t n);
/* n_wk [V*k] (may be NULL: n_k only) and n_k [k] of one model. */
int culda_batch_get_phi(culda_batch *batch, int32_t model, int32_t *n_wk, int32_t *n_k);
int culda_batch_info(culda_batch *batch, int32_t *num_models, int64_t *num_tokens, int64_t *num_slots);
/* With params->profile: sampler_ms = the batched sampler launch of the last
 * eager (use_graph = 0) iteration, other_ms = mean ms per iteration of the
 * last culda_batch_iterate call; counts as for culda_get_stats. */
int culda_batch_get_stats(culda_batch *batch, culda_stats *out);
/* Algorithmic bytes of that timed sampler launch (as culda_sampler_bytes). */
int culda_batch_sampler_bytes(culda_batch *batch, double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* CULDA_H_ */
