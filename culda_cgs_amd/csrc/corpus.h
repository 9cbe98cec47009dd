// corpus.h -- host-side corpus in the src_format data_chunk layout.
//
// Mirrors the roles of the reference's Document (model/doc.h:25) and
// DocChunk (model/doc_chunk.h:19) host data: per-doc cumulative lengths,
// and per chunk the word-major doc-id stream with chunk-local word offsets.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "common.h"

namespace culda {

struct Chunk {
    int32_t doc_start = 0;             // first global doc id
    int32_t doc_end = 0;               // one past the last
    std::vector<int64_t> word_idx;     // V+1 chunk-local cumulative offsets
    std::vector<int32_t> word_tokens;  // global doc ids, word-major
    int64_t num_tokens() const { return (int64_t)word_tokens.size(); }
};

struct Corpus {
    int32_t num_words = 0;
    std::vector<int64_t> doc_cum;      // D entries: cumulative lengths
    std::vector<Chunk> chunks;
    std::vector<std::string> vocab;    // optional tokens (empty => "w<id>")

    int32_t num_docs() const { return (int32_t)doc_cum.size(); }
    int64_t num_tokens() const { return doc_cum.empty() ? 0 : doc_cum.back(); }
    int64_t doc_len(int32_t d) const { return doc_cum[d] - (d ? doc_cum[d - 1] : 0); }
    int32_t max_doc_len() const;

    // Checks the invariants the loader relies on: offsets monotone,
    // doc ids inside the chunk's range, per-doc token counts equal to the
    // .doc.idx lengths, chunks contiguous and covering all docs.
    void validate() const;
};

Corpus load_corpus(const std::string &prefix, int32_t num_chunks);
void write_corpus(const Corpus &c, const std::string &prefix);
void format_text(const std::string &input, const std::string &prefix, int32_t num_chunks);

// Greedy doc -> chunk split of src_format/format.cpp:107-132: chunk c takes
// docs until its token count reaches ceil(N / C).  Returns C doc-end ids.
std::vector<int32_t> greedy_chunk_ends(const std::vector<int64_t> &doc_len, int32_t num_chunks);

// Document-completion split: docs d with d % every == every - 1 are held
// out.  train = the other docs (renumbered in order, train_chunks chunks);
// observed / eval = the held-out docs (renumbered in order, one chunk),
// whose tokens, taken in ascending word order, alternate observed (even
// positions) and evaluated (odd positions).
void heldout_split(const Corpus &c, int32_t every, int32_t train_chunks, Corpus &train, Corpus &observed,
                   Corpus &eval);

// Re-split (word order and doc order are preserved).
Corpus rechunk(const Corpus &c, int32_t num_chunks);

// Build word-major chunks from per-token (word, doc) pairs given in doc
// order (docs ascending).  Used by the generator, the text formatter and
// the held-out split; threads <= 0 takes host_threads().
void build_chunks_from_doc_major(Corpus &c, const std::vector<int32_t> &tok_word,
                                 const std::vector<int32_t> &chunk_ends, int threads = 0);

// Threads for host-side layout work: OMP_NUM_THREADS when set, else the
// hardware threads; at most 64.
int host_threads();

// f(a, b) over [0, n) cut into `threads` contiguous ranges, one thread each.
template <class F>
void parallel_for(int64_t n, int threads, F &&f) {
    if (threads <= 1 || n < 2) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> ts;
    const int64_t step = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t a = t * step, b = std::min(n, a + step);
        if (a >= b) break;
        ts.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &t : ts) t.join();
}

Corpus generate_corpus(const culda_gen_params &p);

}  // namespace culda
