// generator.cpp -- seeded synthetic corpora drawn from the LDA generative
// model (the inputs of every benchmark configuration; the reference ships
// no data).  Each doc and each topic draws from its own Philox stream, so
// the corpus is identical for any thread count.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <thread>

#include <sys/mman.h>

#include "corpus.h"
#include "philox.h"

namespace culda {
namespace {

// Stream (seed, id): counter = (i, id, 7, 0) for i = 0, 1, ...
struct Rng {
    uint64_t seed, id, ctr = 0;
    u64x4 buf{};
    int pos = 4;
    Rng(uint64_t s, uint64_t i) : seed(s), id(i) {}
    uint64_t next() {
        if (pos == 4) {
            buf = philox4x64_10(ctr++, id, 7, 0, seed, 0x5eed);
            pos = 0;
        }
        return buf.v[pos++];
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0,1)
    double normal() {
        double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    // log of a Gamma(a, 1) variate: Marsaglia & Tsang (2000) for a >= 1;
    // for a < 1, Gamma(a) = Gamma(a + 1) * U^(1/a) taken in log space.
    double log_gamma_variate(double a) {
        double boost = 0.0;
        if (a < 1.0) {
            boost = std::log(uniform()) / a;
            a += 1.0;
        }
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        while (true) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return std::log(d * v) + boost;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return std::log(d * v) + boost;
        }
    }
    int64_t poisson(double lam) {
        if (lam > 500.0) return std::max<int64_t>(0, std::llround(lam + std::sqrt(lam) * normal()));
        double p = std::exp(-lam), cdf = p, u = uniform();
        int64_t k = 0;
        while (u > cdf && k < 100000) {
            ++k;
            p *= lam / (double)k;
            cdf += p;
        }
        return k;
    }
};

// Walker alias table of a distribution given by log-weights (Vose's
// construction): bucket j keeps outcome j when 32 random bits fall below
// the low word of cell[j], else gives the outcome in its high word.
void alias_from_logw(const double *lw, int n, uint64_t *cell, std::vector<double> &scaled,
                     std::vector<int32_t> &small, std::vector<int32_t> &large) {
    double m = -INFINITY;
    for (int i = 0; i < n; ++i) m = std::max(m, lw[i]);
    double acc = 0.0;
    scaled.resize(n);
    for (int i = 0; i < n; ++i) acc += scaled[i] = std::exp(lw[i] - m);
    small.clear();
    large.clear();
    for (int i = 0; i < n; ++i) {
        scaled[i] *= (double)n / acc;
        (scaled[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        const int32_t l = small.back(), g = large.back();
        small.pop_back();
        const uint64_t keep = (uint64_t)std::min(4294967295.0, std::floor(scaled[l] * 4294967296.0));
        cell[l] = keep | ((uint64_t)(uint32_t)g << 32);
        scaled[g] = (scaled[g] + scaled[l]) - 1.0;
        if (scaled[g] < 1.0) {
            large.pop_back();
            small.push_back(g);
        }
    }
    // leftovers (rounding) keep themselves
    for (int32_t i : large) cell[i] = 0xffffffffull | ((uint64_t)(uint32_t)i << 32);
    for (int32_t i : small) cell[i] = 0xffffffffull | ((uint64_t)(uint32_t)i << 32);
}

// Uninitialised array in 2 MB-aligned memory advised for transparent huge
// pages (its pages are first touched by the threads that fill it).
template <class T>
class HugeArray {
public:
    explicit HugeArray(size_t n) : n_(n) {
        const size_t align = (size_t)2 << 20, bytes = (std::max<size_t>(n, 1) * sizeof(T) + align - 1) / align * align;
        p_ = static_cast<T *>(std::aligned_alloc(align, bytes));
        if (!p_) throw std::bad_alloc();
        (void)madvise(p_, bytes, MADV_HUGEPAGE);
    }
    ~HugeArray() { release(); }
    HugeArray(const HugeArray &) = delete;
    HugeArray &operator=(const HugeArray &) = delete;
    T &operator[](size_t i) { return p_[i]; }
    const T &operator[](size_t i) const { return p_[i]; }
    void release() {
        std::free(p_);
        p_ = nullptr;
    }

private:
    size_t n_;
    T *p_ = nullptr;
};

inline uint64_t mulhi(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }

}  // namespace

Corpus generate_corpus(const culda_gen_params &p) {
    if (p.num_words < 1 || p.num_docs < 0 || p.true_topics < 1 || p.num_chunks < 1)
        fail(CULDA_ERR_ARG, "generator: need V >= 1, D >= 0, true topics >= 1, chunks >= 1");
    if (p.doc_topic_prior <= 0 || p.topic_word_prior <= 0)
        fail(CULDA_ERR_ARG, "generator: Dirichlet priors must be positive");
    if (p.len_kind != 0 && p.len_kind != 1) fail(CULDA_ERR_ARG, "generator: len_kind must be 0 or 1");
    const int V = p.num_words, D = p.num_docs, KT = p.true_topics;
    const int threads = p.num_threads > 0 ? std::min(p.num_threads, 64) : host_threads();

    // topic-word alias tables: Dirichlet(topic_word_prior) reweighted by (rank+1)^-s
    // the token draws read random cells of this table (KT * V * 8 bytes,
    // over a GB for the large corpora): back it with huge pages where the
    // kernel offers them, so that a lookup is not also a page walk
    HugeArray<uint64_t> phi_cell((size_t)KT * V);
    parallel_for(KT, threads, [&](int64_t a, int64_t b) {
        std::vector<double> lw(V), scaled;
        std::vector<int32_t> small, large;
        for (int64_t k = a; k < b; ++k) {
            Rng r(p.seed, 0x100000000ULL + (uint64_t)k);
            for (int w = 0; w < V; ++w) lw[w] = r.log_gamma_variate(p.topic_word_prior) - p.zipf_s * std::log(w + 1.0);
            alias_from_logw(lw.data(), V, &phi_cell[(size_t)k * V], scaled, small, large);
        }
    });

    // doc lengths
    std::vector<int64_t> len(D);
    parallel_for(D, threads, [&](int64_t a, int64_t b) {
        for (int64_t d = a; d < b; ++d) {
            Rng r(p.seed, 0x200000000ULL + (uint64_t)d);
            int64_t L = 0;
            if (p.len_kind == 0) {
                do L = r.poisson(p.len_mean);
                while (L == 0 && p.len_mean > 0);
            } else {
                L = std::max<int64_t>(1, std::llround(std::exp(p.len_mu + p.len_sigma * r.normal())));
            }
            len[d] = L;
        }
    });
    Corpus c;
    c.num_words = V;
    c.doc_cum.resize(D);
    int64_t n = 0;
    for (int d = 0; d < D; ++d) {
        n += len[d];
        c.doc_cum[d] = n;
    }
    if (n >= (int64_t)INT32_MAX) fail(CULDA_ERR_LIMIT, "generator: %lld tokens exceed 2^31-1", (long long)n);

    // tokens, doc-major.  theta_d ~ Dirichlet(a) over KT topics followed by
    // z ~ theta_d is drawn as the Polya urn it is marginally: token i of a
    // doc takes a new topic, uniform over KT, with probability
    // KT a / (KT a + i), else the topic of one of the i earlier tokens,
    // uniformly -- the Dirichlet-multinomial sequence, O(1) per token.  The
    // word is then one alias draw from phi_z.  One Philox block per token:
    // urn coin, urn index, alias bucket, alias coin.
    std::vector<int32_t> tok_word(n);
    const double kta = (double)KT * p.doc_topic_prior;
    parallel_for(D, threads, [&](int64_t a, int64_t b) {
        constexpr int kRun = 32;  // table cells looked up together (their loads overlap)
        std::vector<int32_t> zd;
        size_t cell[kRun];
        uint32_t flip[kRun];
        for (int64_t d = a; d < b; ++d) {
            Rng r(p.seed, 0x300000000ULL + (uint64_t)d);
            const int64_t t0 = d ? c.doc_cum[d - 1] : 0, nd = c.doc_cum[d] - t0;
            zd.resize(nd);
            for (int64_t i0 = 0; i0 < nd; i0 += kRun) {
                const int m = (int)std::min<int64_t>(kRun, nd - i0);
                for (int j = 0; j < m; ++j) {
                    const int64_t i = i0 + j;
                    const double coin = ((double)(r.next() >> 11) + 0.5) * 0x1.0p-53;
                    const uint64_t pick = r.next(), bucket = r.next();
                    flip[j] = (uint32_t)(r.next() >> 32);
                    const int32_t k = coin * (kta + (double)i) < kta ? (int32_t)mulhi(pick, (uint64_t)KT)
                                                                     : zd[mulhi(pick, (uint64_t)i)];
                    zd[i] = k;
                    cell[j] = (size_t)k * V + mulhi(bucket, (uint64_t)V);
                    __builtin_prefetch(&phi_cell[cell[j]]);
                }
                for (int j = 0; j < m; ++j) {
                    const uint64_t e = phi_cell[cell[j]];
                    tok_word[t0 + i0 + j] = flip[j] < (uint32_t)e ? (int32_t)(cell[j] % (size_t)V) : (int32_t)(e >> 32);
                }
            }
        }
    });
    phi_cell.release();
    build_chunks_from_doc_major(c, tok_word, greedy_chunk_ends(len, p.num_chunks), threads);
    return c;
}

}  // namespace culda
