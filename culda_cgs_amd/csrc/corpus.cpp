// corpus.cpp -- data_chunk reader / writer / text formatter.
//
// Layout and rules follow src_format/format.cpp:19-168 and
// src_format/data_chunk.h:24-67 (writer) and model/doc.cpp:62-156,
// model/doc_chunk.cpp:106-145 (reader).  Unlike the reference, every
// inconsistency is reported as a status instead of exit(0), and the loader
// checks the files against each other.
#include "corpus.h"

#include <algorithm>
#include <cstdlib>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <thread>
#include <unordered_map>

namespace culda {

int32_t Corpus::max_doc_len() const {
    int64_t m = 0;
    for (int32_t d = 0; d < num_docs(); ++d) m = std::max(m, doc_len(d));
    return (int32_t)std::min<int64_t>(m, INT32_MAX);
}

void Corpus::validate() const {
    const int32_t D = num_docs();
    if (num_words < 0) fail(CULDA_ERR_FORMAT, "negative vocabulary size");
    for (int32_t d = 0; d < D; ++d)
        if (doc_len(d) < 0) fail(CULDA_ERR_FORMAT, "doc.idx not cumulative at doc %d", d);
    if (chunks.empty()) fail(CULDA_ERR_FORMAT, "corpus has no chunks");
    int32_t expect_start = 0;
    for (size_t c = 0; c < chunks.size(); ++c) {
        const Chunk &ch = chunks[c];
        if (ch.doc_start != expect_start || ch.doc_end < ch.doc_start || ch.doc_end > D)
            fail(CULDA_ERR_FORMAT, "chunk %zu doc range [%d,%d) is not contiguous in [0,%d)", c,
                 ch.doc_start, ch.doc_end, D);
        expect_start = ch.doc_end;
        if ((int64_t)ch.word_idx.size() != (int64_t)num_words + 1)
            fail(CULDA_ERR_FORMAT, "chunk %zu has %zu word offsets, expected V+1 = %d", c,
                 ch.word_idx.size(), num_words + 1);
        if (ch.word_idx[0] != 0) fail(CULDA_ERR_FORMAT, "chunk %zu word offsets do not start at 0", c);
        for (int32_t w = 0; w < num_words; ++w)
            if (ch.word_idx[w + 1] < ch.word_idx[w])
                fail(CULDA_ERR_FORMAT, "chunk %zu word offsets decrease at word %d", c, w);
        if (ch.word_idx[num_words] != ch.num_tokens())
            fail(CULDA_ERR_FORMAT, "chunk %zu: word.idx ends at %lld but word.data has %lld tokens",
                 c, (long long)ch.word_idx[num_words], (long long)ch.num_tokens());
    }
    if (expect_start != D) fail(CULDA_ERR_FORMAT, "chunks cover docs [0,%d) of %d", expect_start, D);
    // every token's doc inside its chunk, and per-doc token counts equal to
    // the doc lengths: each thread counts the docs of a range of chunks
    // (one chunk split by docs when there are fewer chunks than threads)
    struct Job {
        size_t chunk;
        int32_t da, db;
    };
    std::vector<Job> jobs;
    const int T = host_threads();
    for (size_t c = 0; c < chunks.size(); ++c) {
        const Chunk &ch = chunks[c];
        const int64_t parts = std::max<int64_t>(1, std::min<int64_t>(T, ch.num_tokens() / (1 << 22)));
        for (int64_t q = 0; q < parts; ++q)
            jobs.push_back(Job{c, (int32_t)(ch.doc_start + (int64_t)(ch.doc_end - ch.doc_start) * q / parts),
                               (int32_t)(ch.doc_start + (int64_t)(ch.doc_end - ch.doc_start) * (q + 1) / parts)});
    }
    std::vector<std::string> errs(jobs.size());
    parallel_for((int64_t)jobs.size(), T, [&](int64_t a, int64_t b) {
        std::vector<int64_t> seen;
        for (int64_t j = a; j < b; ++j) {
            const Job &jb = jobs[j];
            const Chunk &ch = chunks[jb.chunk];
            seen.assign(jb.db - jb.da, 0);
            char buf[256];
            for (int32_t d : ch.word_tokens) {
                if (d < ch.doc_start || d >= ch.doc_end) {
                    snprintf(buf, sizeof buf, "chunk %zu holds doc id %d outside its range [%d,%d)", jb.chunk, d,
                             ch.doc_start, ch.doc_end);
                    errs[j] = buf;
                    break;
                }
                if (d >= jb.da && d < jb.db) seen[d - jb.da]++;
            }
            if (!errs[j].empty()) continue;
            for (int32_t d = jb.da; d < jb.db; ++d)
                if (seen[d - jb.da] != doc_len(d)) {
                    snprintf(buf, sizeof buf, "doc %d has %lld tokens in word.data but length %lld in doc.idx", d,
                             (long long)seen[d - jb.da], (long long)doc_len(d));
                    errs[j] = buf;
                    break;
                }
        }
    });
    for (const std::string &e : errs)
        if (!e.empty()) fail(CULDA_ERR_FORMAT, "%s", e.c_str());
}

// ---------------------------------------------------------------- reading
static std::string slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) fail(CULDA_ERR_IO, "file %s open failed", path.c_str());
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Pairs "a b" of integers, whitespace separated; stops at the first token
// that is not an integer (the reference's `while (stream >> a >> b)`).
static std::vector<std::pair<int64_t, int64_t>> read_pairs(const std::string &path) {
    std::string s = slurp(path);
    std::vector<std::pair<int64_t, int64_t>> out;
    const char *p = s.c_str();
    while (true) {
        char *e1;
        errno = 0;
        long long a = strtoll(p, &e1, 10);
        if (e1 == p) break;
        char *e2;
        long long b = strtoll(e1, &e2, 10);
        if (e2 == e1) break;
        out.emplace_back(a, b);
        p = e2;
    }
    return out;
}

static std::string chunk_name(const std::string &prefix, const char *what, int32_t c, int32_t C) {
    std::string n = prefix + what;
    if (C > 1) n += std::to_string(c);
    return n;
}

Corpus load_corpus(const std::string &prefix, int32_t num_chunks) {
    if (num_chunks < 1) fail(CULDA_ERR_ARG, "number of chunks must be >= 1 (got %d)", num_chunks);
    Corpus c;
    // .doc.idx (model/doc.cpp:67-93)
    for (auto &pr : read_pairs(prefix + ".doc.idx")) c.doc_cum.push_back(pr.second);
    // .chunk (model/doc.cpp:97-114)
    auto chunk_lines = read_pairs(prefix + ".chunk");
    if ((int32_t)chunk_lines.size() != num_chunks)
        fail(CULDA_ERR_FORMAT, "numChunks(%d) does not match the chunk file (%zu lines)", num_chunks,
             chunk_lines.size());
    c.chunks.resize(num_chunks);
    int32_t start = 0;
    for (int32_t k = 0; k < num_chunks; ++k) {
        c.chunks[k].doc_start = start;
        c.chunks[k].doc_end = (int32_t)chunk_lines[k].second;
        start = c.chunks[k].doc_end;
    }
    // per-chunk word index and data (model/doc_chunk.cpp:112-145, :215-223)
    for (int32_t k = 0; k < num_chunks; ++k) {
        Chunk &ch = c.chunks[k];
        auto idx = read_pairs(chunk_name(prefix, ".word.idx", k, num_chunks));
        ch.word_idx.resize(idx.size() + 1);
        ch.word_idx[0] = 0;
        for (size_t i = 0; i < idx.size(); ++i) ch.word_idx[i + 1] = idx[i].second;
        const int32_t V = (int32_t)idx.size();
        if (k == 0)
            c.num_words = V;
        else if (V != c.num_words)
            fail(CULDA_ERR_FORMAT, "chunk %d lists %d words, chunk 0 lists %d", k, V, c.num_words);
        std::string path = chunk_name(prefix, ".word.data", k, num_chunks);
        FILE *f = fopen(path.c_str(), "rb");
        if (!f) fail(CULDA_ERR_IO, "file %s open failed", path.c_str());
        fseek(f, 0, SEEK_END);
        long long bytes = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (bytes % 4) {
            fclose(f);
            fail(CULDA_ERR_FORMAT, "%s size %lld is not a multiple of 4", path.c_str(), bytes);
        }
        ch.word_tokens.resize(bytes / 4);
        size_t got = ch.word_tokens.empty() ? 0 : fread(ch.word_tokens.data(), 4, ch.word_tokens.size(), f);
        fclose(f);
        if (got != ch.word_tokens.size()) fail(CULDA_ERR_IO, "short read on %s", path.c_str());
    }
    // vocabulary is optional for training (main.cpp:181 loads it, training never uses it)
    {
        std::ifstream vf(prefix + ".vocab");
        if (vf.is_open()) {
            std::string tok;
            long long id;
            while (vf >> tok >> id) c.vocab.push_back(tok);
            if ((int32_t)c.vocab.size() != c.num_words) c.vocab.clear();
        }
    }
    c.validate();
    return c;
}

// ---------------------------------------------------------------- writing
void write_corpus(const Corpus &c, const std::string &prefix) {
    c.validate();
    const int32_t C = (int32_t)c.chunks.size();
    {
        OutFile f(prefix + ".vocab", "w");
        for (int32_t w = 0; w < c.num_words; ++w) {
            if (c.vocab.empty())
                f.printf("w%d %d\n", w, w);
            else
                f.printf("%s %d\n", c.vocab[w].c_str(), w);
        }
        f.commit();
    }
    {
        OutFile f(prefix + ".doc.idx", "w");
        for (int32_t d = 0; d < c.num_docs(); ++d) f.printf("%d %lld\n", d, (long long)c.doc_cum[d]);
        f.commit();
    }
    {
        OutFile f(prefix + ".chunk", "w");
        for (int32_t k = 0; k < C; ++k) f.printf("%d %d\n", k, c.chunks[k].doc_end);
        f.commit();
    }
    for (int32_t k = 0; k < C; ++k) {
        const Chunk &ch = c.chunks[k];
        {
            OutFile f(chunk_name(prefix, ".word.idx", k, C), "w");
            for (int32_t w = 0; w < c.num_words; ++w) f.printf("%d %lld\n", w, (long long)ch.word_idx[w + 1]);
            f.commit();
        }
        OutFile f(chunk_name(prefix, ".word.data", k, C), "wb");
        f.write(ch.word_tokens.data(), 4, ch.word_tokens.size());
        f.commit();
    }
}

std::vector<int32_t> greedy_chunk_ends(const std::vector<int64_t> &doc_len, int32_t num_chunks) {
    int64_t N = 0;
    for (int64_t l : doc_len) N += l;
    const int64_t per = (N + num_chunks - 1) / num_chunks;
    std::vector<int32_t> ends(num_chunks);
    size_t d = 0;
    for (int32_t c = 0; c < num_chunks; ++c) {
        int64_t acc = 0;
        while (d < doc_len.size()) {
            acc += doc_len[d];
            ++d;
            if (acc >= per) break;
        }
        ends[c] = (int32_t)d;
    }
    return ends;
}

int host_threads() {
    const char *env = std::getenv("OMP_NUM_THREADS");
    int n = env ? std::atoi(env) : 0;
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

// Stable counting sort by word of each chunk's doc-ordered tokens: thread
// t counts a contiguous token range, then writes it at its own per-word
// offsets (after threads 0..t-1 within each word), so every word's docs
// stay ascending whatever the thread count.
void build_chunks_from_doc_major(Corpus &c, const std::vector<int32_t> &tok_word,
                                 const std::vector<int32_t> &chunk_ends, int threads) {
    const int32_t V = c.num_words, C = (int32_t)chunk_ends.size();
    if (threads <= 0) threads = host_threads();
    c.chunks.assign(C, Chunk());
    int32_t d0 = 0;
    for (int32_t k = 0; k < C; ++k) {
        Chunk &ch = c.chunks[k];
        ch.doc_start = d0;
        ch.doc_end = chunk_ends[k];
        d0 = ch.doc_end;
        const int64_t t0 = ch.doc_start ? c.doc_cum[ch.doc_start - 1] : 0;
        const int64_t t1 = ch.doc_end ? c.doc_cum[ch.doc_end - 1] : 0;
        const int64_t n = t1 - t0;
        const int T = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n / 65536));
        const int64_t step = (n + T - 1) / T;
        std::vector<std::vector<int64_t>> cnt(T, std::vector<int64_t>(V, 0));
        parallel_for(T, T, [&](int64_t a, int64_t b) {
            for (int64_t t = a; t < b; ++t)
                for (int64_t i = t0 + t * step; i < std::min(t1, t0 + (t + 1) * step); ++i) cnt[t][tok_word[i]]++;
        });
        ch.word_idx.assign(V + 1, 0);
        for (int32_t w = 0; w < V; ++w) {
            int64_t acc = ch.word_idx[w];
            for (int t = 0; t < T; ++t) {
                const int64_t x = cnt[t][w];
                cnt[t][w] = acc;  // now thread t's write position for word w
                acc += x;
            }
            ch.word_idx[w + 1] = acc;
        }
        ch.word_tokens.resize(n);
        parallel_for(T, T, [&](int64_t a, int64_t b) {
            for (int64_t t = a; t < b; ++t) {
                const int64_t i0 = t0 + t * step, i1 = std::min(t1, t0 + (t + 1) * step);
                if (i0 >= i1) continue;
                // token i belongs to the first doc whose cumulative length exceeds i
                int32_t d = (int32_t)(std::upper_bound(c.doc_cum.begin(), c.doc_cum.end(), i0) - c.doc_cum.begin());
                std::vector<int64_t> &pos = cnt[t];
                for (int64_t i = i0; i < i1; ++i) {
                    while (c.doc_cum[d] <= i) ++d;
                    ch.word_tokens[pos[tok_word[i]]++] = d;
                }
            }
        });
    }
}

// src_format/format.cpp:19-168 semantics.
void format_text(const std::string &input, const std::string &prefix, int32_t num_chunks) {
    if (num_chunks <= 1) num_chunks = 1;
    std::ifstream in(input);
    if (!in.is_open()) fail(CULDA_ERR_IO, "%s open failed", input.c_str());
    Corpus c;
    std::unordered_map<std::string, int32_t> ids;
    std::vector<int32_t> tok_word;
    std::string line;
    int64_t n = 0;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string title, tok;
        if (!(ls >> title)) continue;  // no title: not a document
        while (ls >> tok) {
            auto it = ids.find(tok);
            int32_t id;
            if (it == ids.end()) {
                id = (int32_t)c.vocab.size();
                ids.emplace(tok, id);
                c.vocab.push_back(tok);
            } else {
                id = it->second;
            }
            tok_word.push_back(id);
            ++n;
        }
        c.doc_cum.push_back(n);
    }
    c.num_words = (int32_t)c.vocab.size();
    std::vector<int64_t> lens(c.doc_cum.size());
    for (size_t d = 0; d < lens.size(); ++d) lens[d] = c.doc_len((int32_t)d);
    build_chunks_from_doc_major(c, tok_word, greedy_chunk_ends(lens, num_chunks));
    write_corpus(c, prefix);
}

// Doc-major word ids of every token: the tokens of doc d, in ascending
// word order, at [doc_cum[d-1], doc_cum[d]).  Each thread owns a range of
// docs (about equal token counts) and takes its docs' tokens from one scan
// of the word-major streams.
static std::vector<int32_t> doc_major_words(const Corpus &c) {
    const int32_t D = c.num_docs();
    const int64_t N = c.num_tokens();
    std::vector<int32_t> out(N);
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), N / 65536));
    std::vector<int32_t> dcut(T + 1, D);
    dcut[0] = 0;
    for (int t = 1; t < T; ++t)
        dcut[t] = (int32_t)(std::lower_bound(c.doc_cum.begin(), c.doc_cum.end(), N / T * t) - c.doc_cum.begin());
    parallel_for(T, T, [&](int64_t a, int64_t b) {
        for (int64_t t = a; t < b; ++t) {
            const int32_t da = dcut[t], db = std::max(dcut[t], dcut[t + 1]);
            if (da >= db) continue;
            std::vector<int64_t> cur(db - da);
            for (int32_t d = da; d < db; ++d) cur[d - da] = d ? c.doc_cum[d - 1] : 0;
            for (const Chunk &ch : c.chunks) {
                if (ch.doc_end <= da || ch.doc_start >= db) continue;
                for (int32_t w = 0; w < c.num_words; ++w)
                    for (int64_t i = ch.word_idx[w]; i < ch.word_idx[w + 1]; ++i) {
                        const int32_t d = ch.word_tokens[i];
                        if (d >= da && d < db) out[cur[d - da]++] = w;
                    }
            }
        }
    });
    return out;
}

void heldout_split(const Corpus &c, int32_t every, int32_t train_chunks, Corpus &train, Corpus &observed,
                   Corpus &eval) {
    if (every < 2) fail(CULDA_ERR_ARG, "held-out period must be >= 2 (every %d-th doc)", every);
    if (train_chunks < 1) fail(CULDA_ERR_ARG, "number of chunks must be >= 1");
    const std::vector<int32_t> words = doc_major_words(c);
    const int32_t D = c.num_docs();
    // output position of every doc's tokens: a training doc's in tw, a
    // held-out doc's even positions in ow and odd ones in ew
    std::vector<int64_t> tl, ol, el, dst(D);
    int64_t nt = 0, no = 0, ne = 0;
    for (int32_t d = 0; d < D; ++d) {
        const int64_t L = c.doc_len(d);
        if (d % every == every - 1) {
            dst[d] = no;  // the odd positions' start in ew: edst
            ol.push_back((L + 1) / 2);
            el.push_back(L / 2);
            no += (L + 1) / 2;
            ne += L / 2;
        } else {
            dst[d] = nt;
            tl.push_back(L);
            nt += L;
        }
    }
    std::vector<int32_t> tw(nt), ow(no), ew(ne);
    std::vector<int64_t> edst(D, 0);
    {
        int64_t e = 0;
        for (int32_t d = 0; d < D; ++d)
            if (d % every == every - 1) {
                edst[d] = e;
                e += c.doc_len(d) / 2;
            }
    }
    parallel_for(D, host_threads(), [&](int64_t a, int64_t b) {
        for (int64_t d = a; d < b; ++d) {
            const int64_t s = d ? c.doc_cum[d - 1] : 0, e = c.doc_cum[d];
            if (d % every == every - 1) {
                int64_t o = dst[d], x = edst[d];
                for (int64_t t = s; t < e; ++t) {
                    if ((t - s) % 2 == 0)
                        ow[o++] = words[t];
                    else
                        ew[x++] = words[t];
                }
            } else {
                std::copy(words.begin() + s, words.begin() + e, tw.begin() + dst[d]);
            }
        }
    });
    auto make = [&](Corpus &out, const std::vector<int32_t> &w, const std::vector<int64_t> &lens, int32_t chunks) {
        out = Corpus();
        out.num_words = c.num_words;
        out.vocab = c.vocab;
        out.doc_cum.resize(lens.size());
        int64_t acc = 0;
        for (size_t d = 0; d < lens.size(); ++d) out.doc_cum[d] = acc += lens[d];
        if (lens.empty()) fail(CULDA_ERR_ARG, "held-out split leaves an empty corpus");
        build_chunks_from_doc_major(out, w, greedy_chunk_ends(lens, chunks));
        out.validate();
    };
    make(train, tw, tl, train_chunks);
    make(observed, ow, ol, 1);
    make(eval, ew, el, 1);
}

Corpus rechunk(const Corpus &src, int32_t num_chunks) {
    if (num_chunks < 1) fail(CULDA_ERR_ARG, "number of chunks must be >= 1");
    const int32_t V = src.num_words, D = src.num_docs();
    std::vector<int64_t> lens(D);
    for (int32_t d = 0; d < D; ++d) lens[d] = src.doc_len(d);
    std::vector<int32_t> ends = greedy_chunk_ends(lens, num_chunks);
    std::vector<int32_t> doc_chunk(D, num_chunks - 1);
    {
        int32_t d = 0;
        for (int32_t k = 0; k < num_chunks; ++k)
            for (; d < ends[k]; ++d) doc_chunk[d] = k;
    }
    Corpus out;
    out.num_words = V;
    out.doc_cum = src.doc_cum;
    out.vocab = src.vocab;
    out.chunks.assign(num_chunks, Chunk());
    int32_t d0 = 0;
    for (int32_t k = 0; k < num_chunks; ++k) {
        out.chunks[k].doc_start = d0;
        out.chunks[k].doc_end = ends[k];
        d0 = ends[k];
        out.chunks[k].word_idx.assign(V + 1, 0);
    }
    // one thread per output chunk: it scans the source and keeps its docs'
    // tokens; word-major over the source chunks in doc order keeps docs
    // ascending within every word
    parallel_for(num_chunks, std::min(num_chunks, host_threads()), [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; ++k) {
            Chunk &oc = out.chunks[k];
            for (const Chunk &ch : src.chunks) {
                if (ch.doc_end <= oc.doc_start || ch.doc_start >= oc.doc_end) continue;
                for (int32_t w = 0; w < V; ++w)
                    for (int64_t i = ch.word_idx[w]; i < ch.word_idx[w + 1]; ++i)
                        oc.word_idx[w + 1] += doc_chunk[ch.word_tokens[i]] == k;
            }
            for (int32_t w = 0; w < V; ++w) oc.word_idx[w + 1] += oc.word_idx[w];
            oc.word_tokens.resize(oc.word_idx[V]);
            int64_t o = 0;
            for (int32_t w = 0; w < V; ++w)
                for (const Chunk &ch : src.chunks) {
                    if (ch.doc_end <= oc.doc_start || ch.doc_start >= oc.doc_end) continue;
                    for (int64_t i = ch.word_idx[w]; i < ch.word_idx[w + 1]; ++i) {
                        const int32_t d = ch.word_tokens[i];
                        if (doc_chunk[d] == k) oc.word_tokens[o++] = d;
                    }
                }
        }
    });
    out.validate();
    return out;
}

}  // namespace culda

// ------------------------------------------------------------------ C ABI
using namespace culda;

struct culda_corpus {
    Corpus c;
};

extern "C" int culda_corpus_load(const char *prefix, int32_t num_chunks, culda_corpus **out) {
    return guarded([&] {
        if (!prefix || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        auto *p = new culda_corpus{load_corpus(prefix, num_chunks)};
        *out = p;
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_from_arrays(int32_t num_words, int32_t num_docs, const int64_t *doc_cum,
                                        int32_t num_chunks, const int32_t *chunk_doc_end,
                                        const int64_t *const *word_indices,
                                        const int32_t *const *word_tokens, culda_corpus **out) {
    return guarded([&] {
        if (!out || num_words < 0 || num_docs < 0 || num_chunks < 1 || !chunk_doc_end ||
            !word_indices || !word_tokens || (num_docs > 0 && !doc_cum))
            fail(CULDA_ERR_ARG, "invalid corpus arrays");
        *out = nullptr;
        auto *p = new culda_corpus;
        Corpus &c = p->c;
        c.num_words = num_words;
        c.doc_cum.assign(doc_cum, doc_cum + num_docs);
        c.chunks.resize(num_chunks);
        int32_t start = 0;
        for (int32_t k = 0; k < num_chunks; ++k) {
            Chunk &ch = c.chunks[k];
            ch.doc_start = start;
            ch.doc_end = chunk_doc_end[k];
            start = ch.doc_end;
            if (!word_indices[k]) {
                delete p;
                fail(CULDA_ERR_ARG, "chunk %d has no word offsets", k);
            }
            ch.word_idx.assign(word_indices[k], word_indices[k] + num_words + 1);
            const int64_t n = ch.word_idx[num_words];
            if (n < 0 || (n > 0 && !word_tokens[k])) {
                delete p;
                fail(CULDA_ERR_ARG, "chunk %d has no tokens array", k);
            }
            ch.word_tokens.assign(word_tokens[k], word_tokens[k] + n);
        }
        try {
            c.validate();
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_write(const culda_corpus *corpus, const char *prefix) {
    return guarded([&] {
        if (!corpus || !prefix) fail(CULDA_ERR_ARG, "null argument");
        write_corpus(corpus->c, prefix);
        return CULDA_OK;
    });
}

extern "C" int culda_format_text(const char *input_path, const char *prefix, int32_t num_chunks) {
    return guarded([&] {
        if (!input_path || !prefix) fail(CULDA_ERR_ARG, "null argument");
        format_text(input_path, prefix, num_chunks);
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_rechunk(const culda_corpus *corpus, int32_t num_chunks, culda_corpus **out) {
    return guarded([&] {
        if (!corpus || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        *out = new culda_corpus{rechunk(corpus->c, num_chunks)};
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_heldout_split(const culda_corpus *corpus, int32_t every, int32_t train_chunks,
                                          culda_corpus **train, culda_corpus **observed, culda_corpus **eval) {
    return guarded([&] {
        if (!corpus || !train || !observed || !eval) fail(CULDA_ERR_ARG, "null argument");
        *train = *observed = *eval = nullptr;
        auto t = std::make_unique<culda_corpus>(), o = std::make_unique<culda_corpus>(),
             e = std::make_unique<culda_corpus>();
        heldout_split(corpus->c, every, train_chunks, t->c, o->c, e->c);
        *train = t.release();
        *observed = o.release();
        *eval = e.release();
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_generate(const culda_gen_params *params, culda_corpus **out) {
    return guarded([&] {
        if (!params || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        *out = new culda_corpus{generate_corpus(*params)};
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_info_get(const culda_corpus *corpus, culda_corpus_info *out) {
    return guarded([&] {
        if (!corpus || !out) fail(CULDA_ERR_ARG, "null argument");
        out->num_words = corpus->c.num_words;
        out->num_docs = corpus->c.num_docs();
        out->num_chunks = (int32_t)corpus->c.chunks.size();
        out->max_doc_len = corpus->c.max_doc_len();
        out->num_tokens = corpus->c.num_tokens();
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_doc_cum(const culda_corpus *corpus, const int64_t **doc_cum) {
    return guarded([&] {
        if (!corpus || !doc_cum) fail(CULDA_ERR_ARG, "null argument");
        *doc_cum = corpus->c.doc_cum.data();
        return CULDA_OK;
    });
}

extern "C" int culda_corpus_chunk(const culda_corpus *corpus, int32_t chunk, culda_chunk_view *out) {
    return guarded([&] {
        if (!corpus || !out) fail(CULDA_ERR_ARG, "null argument");
        if (chunk < 0 || chunk >= (int32_t)corpus->c.chunks.size())
            fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
        const Chunk &ch = corpus->c.chunks[chunk];
        out->docid_start = ch.doc_start;
        out->docid_end = ch.doc_end;
        out->num_tokens = ch.num_tokens();
        out->word_indices = ch.word_idx.data();
        out->word_tokens = ch.word_tokens.data();
        return CULDA_OK;
    });
}

extern "C" void culda_corpus_free(culda_corpus *corpus) { delete corpus; }

const culda::Corpus &corpus_ref(const culda_corpus *p) { return p->c; }
