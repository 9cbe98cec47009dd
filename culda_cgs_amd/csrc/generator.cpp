// generator.cpp -- seeded synthetic corpora drawn from the LDA generative
// model (the inputs of every benchmark configuration; the reference ships
// no data).  Each doc and each topic draws from its own Philox stream, so
// the corpus is identical for any thread count.
#include <algorithm>
#include <cmath>
#include <thread>

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

// Normalised CDF from log-weights.
void cdf_from_logw(const double *lw, int n, double *cdf) {
    double m = -INFINITY;
    for (int i = 0; i < n; ++i) m = std::max(m, lw[i]);
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
        acc += std::exp(lw[i] - m);
        cdf[i] = acc;
    }
    for (int i = 0; i < n; ++i) cdf[i] /= acc;
    cdf[n - 1] = 1.0;
}

int search_cdf(const double *cdf, int n, double u) {
    int i = (int)(std::upper_bound(cdf, cdf + n, u) - cdf);
    return i < n ? i : n - 1;
}

template <class F>
void parallel_for(int64_t n, int threads, F &&f) {
    if (threads <= 1 || n < 2) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    const int64_t step = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        int64_t a = t * step, b = std::min(n, a + step);
        if (a >= b) break;
        ts.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &t : ts) t.join();
}

}  // namespace

Corpus generate_corpus(const culda_gen_params &p) {
    if (p.num_words < 1 || p.num_docs < 0 || p.true_topics < 1 || p.num_chunks < 1)
        fail(CULDA_ERR_ARG, "generator: need V >= 1, D >= 0, true topics >= 1, chunks >= 1");
    if (p.doc_topic_prior <= 0 || p.topic_word_prior <= 0)
        fail(CULDA_ERR_ARG, "generator: Dirichlet priors must be positive");
    if (p.len_kind != 0 && p.len_kind != 1) fail(CULDA_ERR_ARG, "generator: len_kind must be 0 or 1");
    const int V = p.num_words, D = p.num_docs, KT = p.true_topics;
    int threads = p.num_threads > 0 ? p.num_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 64);

    // topic-word CDFs: Dirichlet(topic_word_prior) reweighted by (rank+1)^-s
    std::vector<double> phi_cdf((size_t)KT * V);
    parallel_for(KT, threads, [&](int64_t a, int64_t b) {
        std::vector<double> lw(V);
        for (int64_t k = a; k < b; ++k) {
            Rng r(p.seed, 0x100000000ULL + (uint64_t)k);
            for (int w = 0; w < V; ++w) lw[w] = r.log_gamma_variate(p.topic_word_prior) - p.zipf_s * std::log(w + 1.0);
            cdf_from_logw(lw.data(), V, &phi_cdf[(size_t)k * V]);
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

    // tokens, doc-major
    std::vector<int32_t> tok_word(n);
    parallel_for(D, threads, [&](int64_t a, int64_t b) {
        std::vector<double> lw(KT), th(KT);
        for (int64_t d = a; d < b; ++d) {
            Rng r(p.seed, 0x300000000ULL + (uint64_t)d);
            for (int k = 0; k < KT; ++k) lw[k] = r.log_gamma_variate(p.doc_topic_prior);
            cdf_from_logw(lw.data(), KT, th.data());
            const int64_t t0 = d ? c.doc_cum[d - 1] : 0;
            for (int64_t t = t0; t < c.doc_cum[d]; ++t) {
                int k = search_cdf(th.data(), KT, r.uniform());
                tok_word[t] = search_cdf(&phi_cdf[(size_t)k * V], V, r.uniform());
            }
        }
    });
    phi_cdf.clear();
    phi_cdf.shrink_to_fit();
    build_chunks_from_doc_major(c, tok_word, greedy_chunk_ends(len, p.num_chunks));
    return c;
}

}  // namespace culda
