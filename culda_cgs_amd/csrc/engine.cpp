// engine.cpp -- layouts, work lists and iteration schedules on the device.
//
// Iteration (train/SingleChunkSingleGPU.h:70-127 for one chunk,
// train/MultiChunkMultiGPUequal.h:88-169 for several, here on one device):
// sample every unit with the frozen counts of the previous iteration,
// reduce the FP64 log-likelihood, rebuild n_wk / n_k from z on the device
// (summed over a corpus's chunks), prepare the head denominators, rebuild
// every unit's theta.  One host synchronisation per iteration, or none
// inside a replayed graph.
#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <numeric>
#include <string>

#include "kernels.h"
#include "mh_state.h"

namespace culda {

// Fault injection for tests of the error paths (the reference aborts the
// process on any CUDA error, culda_argument.h:40-48).  CULDA_FAIL_HIP =
// "<text>" or "<text>:<n>" makes the n-th (default: first) checked HIP call
// whose description contains <text> report a failure after the call itself
// succeeded, so every kernel still runs and exits normally; the count
// restarts when the variable's value changes.
static bool injected_failure(const char *what) {
    static std::mutex mu;
    static std::string spec_seen, text;
    static long target = 0, seen = 0;
    const char *env = std::getenv("CULDA_FAIL_HIP");
    if (!env || !*env) return false;
    std::lock_guard<std::mutex> lock(mu);
    if (spec_seen != env) {
        spec_seen = env;
        text = spec_seen;
        target = 1;
        const size_t colon = text.rfind(':');
        if (colon != std::string::npos) {
            target = std::max(1L, std::strtol(text.c_str() + colon + 1, nullptr, 10));
            text.resize(colon);
        }
        seen = 0;
    }
    if (text.empty() || !std::strstr(what, text.c_str())) return false;
    return ++seen == target;
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(CULDA_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    if (injected_failure(what)) fail(CULDA_ERR_DEVICE, "%s: injected failure (CULDA_FAIL_HIP)", what);
}

Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (graph) (void)hipGraphExecDestroy(graph);
    for (void *p : allocations) (void)hipFree(p);
    for (hipEvent_t &e : ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ev_sampled)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ev_theta_unit)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ev_iter)
        if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
}

void Engine::init_device(const culda_params &p) {
    if (p.k < 16 || p.k > kMaxK || p.k % 16)
        fail(CULDA_ERR_ARG, "k = %d: the sampler needs a multiple of 16 in [16, %d]", p.k, kMaxK);
    if (p.slot_tokens < 0) fail(CULDA_ERR_ARG, "slot_tokens must be >= 0");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) fail(CULDA_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (p.device < 0 || p.device >= n) fail(CULDA_ERR_ARG, "device %d out of range (%d devices)", p.device, n);
    hip_check(hipSetDevice(p.device), "hipSetDevice");
    prm = p;
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "hipStreamCreate");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p.device) != hipSuccess) cus = 0;
    cu_count = cus > 0 ? cus : 256;
    // the allocation-free joint LL's static scratch: its address on this device
    hip_check(joint_scratch_init(p.device), "scratch symbol lookup");
}

void *Engine::dalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) fail(CULDA_ERR_MEMORY, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    allocations.push_back(p);
    return p;
}

void Engine::dfree(void *p) {
    if (!p) return;
    auto it = std::find(allocations.begin(), allocations.end(), p);
    if (it != allocations.end()) {
        (void)hipFree(p);
        allocations.erase(it);
    }
}

template <class T>
static T *upload_vec(Engine &e, const std::vector<T> &v) {
    T *d = (T *)e.dalloc(v.size() * sizeof(T));
    if (!v.empty()) hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

// Largest number of theta entries a doc of nd tokens can need: at most one
// per token, and at most one per topic plus one per kThetaMaxCount tokens.
int64_t theta_entries_max(int64_t nd, int32_t K) {
    return std::min<int64_t>(nd, (int64_t)K + (nd + kThetaMaxCount - 1) / kThetaMaxCount);
}

void check_corpus_limits(const Corpus &c, int32_t K) {
    if (K < 16 || K > kMaxK || K % 16) fail(CULDA_ERR_ARG, "k = %d: the sampler needs a multiple of 16 in [16, %d]", K, kMaxK);
    if (c.num_tokens() >= (int64_t)INT32_MAX)
        fail(CULDA_ERR_LIMIT, "the corpus has %lld tokens; counts are int32 (at most %d)", (long long)c.num_tokens(),
             INT32_MAX - 1);
    int64_t longest = 0;
    int32_t longest_doc = -1;
    for (size_t k = 0; k < c.chunks.size(); ++k) {
        const Chunk &ch = c.chunks[k];
        if (ch.num_tokens() >= (int64_t)UINT32_MAX) fail(CULDA_ERR_LIMIT, "chunk %zu has >= 2^32 tokens", k);
        int64_t cap = 0;
        for (int32_t d = ch.doc_start; d < ch.doc_end; ++d) {
            const int64_t e = theta_entries_max(c.doc_len(d), K);
            if (e > longest) {
                longest = e;
                longest_doc = d;
            }
            cap += (e + 15) / 16 * 16;
            // the sampler addresses a chunk's theta with 32-bit byte offsets
            if (cap + kThetaPad >= ((int64_t)1 << 31))
                fail(CULDA_ERR_LIMIT, "chunk %zu's theta capacity exceeds 2^31 entries: use more data_chunks", k);
        }
    }
    // eight S-prefix rows of the longest theta row in the sampler's LDS
    const int64_t lds = (int64_t)kSamplerStaticLds + (int64_t)sample_dyn_lds(sample_rs_stride(longest));
    if (lds > kSamplerLdsBudget)
        fail(CULDA_ERR_LIMIT,
             "doc %d (%lld tokens) may need %lld theta entries at k = %d; the sampler's LDS holds rows of at most "
             "%lld entries",
             longest_doc, (long long)c.doc_len(longest_doc), (long long)longest, K,
             (long long)((kSamplerLdsBudget - kSamplerStaticLds) / 32 / 4 * 4 - 1) * 16);
}

// Hot words are rebuilt in pieces of at most phi_split(tokens) tokens per
// unit, `tokens` being all the engine's tokens (one phi launch covers them).
// Large launches are bound by their total work: at 97M tokens 8192-token
// pieces took 0.367 ms, 4096- or 32768-token pieces 0.396 ms (more pieces,
// more row atomics; or longer chains).  Small ones are bound by the longest
// piece's chain of loads: at 1M tokens 4096-token pieces took 22 us, 8192-
// token pieces 46 us, 2048 21 us and 1024 30 us.
static int64_t phi_split(int64_t tokens) { return tokens <= ((int64_t)1 << 24) ? 4096 : 8192; }

// Packets of at most `per` consecutive items of [first, last), all owned by `owner`.
// Where each word's token list (word-major, unit-local doc ids `doc`)
// crosses into each XCD doc region: (kXcds + 1) offsets per word from the
// word's first token, the first 0 and the last the word's token count.
// Empty if some word's docs are not in ascending order (no region
// affinity for that unit's slots then).
static std::vector<uint32_t> doc_regions(const std::vector<uint32_t> &doc, const std::vector<int64_t> &word_idx,
                                         int32_t V, int32_t D) {
    const int32_t X = (int32_t)kXcds;
    std::vector<uint32_t> out((size_t)V * (X + 1));
    for (int32_t w = 0; w < V; ++w) {
        const auto b = doc.begin() + word_idx[w], e = doc.begin() + word_idx[w + 1];
        if (!std::is_sorted(b, e)) return {};
        for (int32_t r = 0; r <= X; ++r)
            out[(size_t)w * (X + 1) + r] =
                r == X ? (uint32_t)(e - b) : (uint32_t)(std::lower_bound(b, e, (uint32_t)((int64_t)D * r / X)) - b);
    }
    return out;
}

// A unit's slots (indices, longest first) in launch order: position b
// (workgroup b, which runs on XCD b mod kXcds) takes the longer of the next
// slot of region b mod kXcds and the next untagged slot; an XCD whose own
// and the untagged slots are used up takes the longest remaining one.
// Each XCD still takes its slots longest first.
static std::vector<size_t> deal_slots_to_xcds(const std::vector<size_t> &by_len, const std::vector<int8_t> &tag,
                                              const std::vector<SlotDesc> &sl) {
    const int X = (int)kXcds;
    std::vector<std::vector<size_t>> q(X + 1);
    for (size_t i : by_len) q[tag[i] < 0 ? X : tag[i]].push_back(i);
    std::vector<size_t> head(X + 1, 0), out;
    out.reserve(by_len.size());
    auto len = [&](int y) { return sl[q[y][head[y]]].end - sl[q[y][head[y]]].begin; };
    auto left = [&](int y) { return head[y] < q[y].size(); };
    for (size_t b = 0; b < by_len.size(); ++b) {
        const int x = (int)(b % X);
        int pick = left(x) ? x : -1;
        if (left(X) && (pick < 0 || len(X) > len(pick))) pick = X;
        for (int y = 0; pick < 0 && y < X; ++y)
            if (left(y)) pick = y;
        for (int y = 0; pick >= 0 && pick != x && pick != X && y < X; ++y)
            if (left(y) && len(y) > len(pick)) pick = y;
        out.push_back(q[pick][head[pick]++]);
    }
    return out;
}

// Per-XCD lists merged into one launch order: position b takes XCD
// b mod kXcds's next item (workgroup b runs there); an XCD with none left
// takes one of the XCD with the most left.
template <class T>
static std::vector<T> interleave_xcds(const std::vector<std::vector<T>> &q) {
    const int X = (int)q.size();
    size_t n = 0;
    for (const auto &l : q) n += l.size();
    std::vector<size_t> head(X, 0);
    std::vector<T> out;
    out.reserve(n);
    for (size_t b = 0; b < n; ++b) {
        int x = (int)(b % X);
        if (head[x] == q[x].size())
            for (int y = 0; y < X; ++y)
                if (q[y].size() - head[y] > q[x].size() - head[x]) x = y;
        out.push_back(q[x][head[x]++]);
    }
    return out;
}

static void append_packets(std::vector<Packet> &out, int32_t first, int32_t last, int32_t owner, int32_t per) {
    for (int32_t x = first; x < last; x += per) out.push_back(Packet{x, std::min(last, x + per), owner, 0});
}

// Phi list of model m over units [u0, u1): whole words, hot words as pieces
// of at most `split` tokens per unit (rows zeroed first, atomics); items
// in descending token count, tokens[i] being item i's.
static void build_phi_list(const std::vector<UnitHost> &units, int32_t m, int32_t u0, int32_t u1, int32_t V,
                           int64_t split, std::vector<PhiSlot> &slots, std::vector<int64_t> &tokens,
                           std::vector<int2> &zero) {
    std::vector<std::pair<int64_t, PhiSlot>> items;
    items.reserve(V);
    for (int32_t w = 0; w < V; ++w) {
        int64_t tot = 0;
        for (int32_t u = u0; u < u1; ++u) tot += units[u].word_idx_h[w + 1] - units[u].word_idx_h[w];
        if (tot > split) {
            zero.push_back(make_int2(m, w));
            for (int32_t u = u0; u < u1; ++u) {
                const int64_t b = units[u].word_idx_h[w], e = units[u].word_idx_h[w + 1];
                for (int64_t x = b; x < e; x += split)
                    items.emplace_back(std::min(e, x + split) - x, PhiSlot{w, u, u + 1, 1, x, std::min(e, x + split)});
            }
        } else {
            items.emplace_back(tot, PhiSlot{w, u0, u1, 0, 0, 0});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](auto &a, auto &b) { return a.first > b.first; });
    slots.clear();
    tokens.clear();
    for (auto &pr : items) {
        tokens.push_back(pr.first);
        slots.push_back(pr.second);
    }
}

// Deals a list in descending token order into packets of about `per`
// items, in snake order (round r gives packet p the item of rank r P + p,
// or r P + P - 1 - p when r is odd), so every packet gets a like share of
// long and short items and no packet the longest of every round; within a
// packet, the waves of the rebuild take items w, w + W, w + 2W, ... (W
// waves), so every other run of W items is reversed for the same reason.
// Reorders `slots` to match and appends the packets, their ranges offset
// by `base`.
static void deal_packets(std::vector<PhiSlot> &slots, int32_t owner, int32_t per, int32_t base,
                         std::vector<Packet> &packets) {
    const int32_t n = (int32_t)slots.size();
    if (n == 0) return;
    const int32_t P = (n + per - 1) / per, W = kWavesPerGroup;
    std::vector<PhiSlot> out, pk;
    out.reserve(n);
    for (int32_t p = 0; p < P; ++p) {
        pk.clear();
        for (int32_t r = 0; (int64_t)r * P < n; ++r) {
            const int32_t i = r * P + ((r & 1) ? P - 1 - p : p);
            if (i < n) pk.push_back(slots[i]);
        }
        for (size_t j = W; j < pk.size(); j += 2 * (size_t)W)
            std::reverse(pk.begin() + j, pk.begin() + std::min(pk.size(), j + (size_t)W));
        const int32_t first = base + (int32_t)out.size();
        out.insert(out.end(), pk.begin(), pk.end());
        packets.push_back(Packet{first, base + (int32_t)out.size(), owner, 0});
    }
    slots.swap(out);
}

// Rare tiles of a single-unit model: words with 0 < n and n * 16 < K (so
// n < 64, kMaxK being 1024), packed greedily in word order while the tile
// keeps at most 64 tokens and spans at most 64 word ids; a word with tokens
// that is not rare ends the tile (its tokens would break the contiguous
// range).  covered[w] marks the words the tiles take.
static void build_rare_tiles(const UnitHost &u, int32_t ui, int32_t V, int32_t K, std::vector<RareTile> &tiles,
                             std::vector<char> &covered) {
    RareTile cur{};
    bool open = false;
    for (int32_t w = 0; w < V; ++w) {
        const int64_t b = u.word_idx_h[w], n = u.word_idx_h[w + 1] - b;
        if (n == 0) continue;
        if (n * 16 >= K || n > kWave) {
            if (open) tiles.push_back(cur);
            open = false;
            continue;
        }
        if (open && (cur.count + n > kWave || w - cur.w0 >= kWave)) {
            tiles.push_back(cur);
            open = false;
        }
        if (!open) {
            cur = RareTile{0, 0, b, w, ui, 0, 0};
            open = true;
        }
        cur.start_mask |= 1ull << cur.count;
        cur.word_mask |= 1ull << (w - cur.w0);
        cur.count += (int32_t)n;
        covered[w] = 1;
    }
    if (open) tiles.push_back(cur);
}

int32_t Engine::add_corpus(const Corpus &c, float alpha, float beta, uint64_t seed) {
    if (uploaded) fail(CULDA_ERR_STATE, "corpora must be added before the upload");
    if (!(alpha > 0.0f) || !std::isfinite(alpha)) fail(CULDA_ERR_ARG, "alpha must be positive and finite");
    if (!(beta > 0.0f) || !std::isfinite(beta)) fail(CULDA_ERR_ARG, "beta must be positive and finite");
    c.validate();
    check_corpus_limits(c, prm.k);
    pending.push_back(&c);
    pending_alpha.push_back(alpha);
    pending_beta.push_back(beta);
    pending_seed.push_back(seed);
    return (int32_t)pending.size() - 1;
}

void Engine::upload() {
    if (uploaded) fail(CULDA_ERR_STATE, "already uploaded");
    if (pending.empty()) fail(CULDA_ERR_STATE, "no corpus to upload");
    const int32_t K = prm.k;
    for (size_t mi = 0; mi < pending.size(); ++mi) {
        const Corpus &c = *pending[mi];
        ModelHost m;
        m.V = c.num_words;
        m.K = K;
        m.num_docs = c.num_docs();
        m.alpha = pending_alpha[mi];
        m.beta = pending_beta[mi];
        m.seed = pending_seed[mi];
        m.num_tokens = c.num_tokens();
        if (m.num_tokens >= (int64_t)INT32_MAX)
            fail(CULDA_ERR_LIMIT, "corpus %zu has %lld tokens; counts are int32", mi, (long long)m.num_tokens);
        m.doc_len.resize(m.num_docs);
        for (int32_t d = 0; d < m.num_docs; ++d) m.doc_len[d] = c.doc_len(d);
        m.unit_begin = (int32_t)units.size();
        for (size_t k = 0; k < c.chunks.size(); ++k) {
            const Chunk &ch = c.chunks[k];
            UnitHost u;
            u.model = (int32_t)mi;
            u.doc_start = ch.doc_start;
            u.doc_end = ch.doc_end;
            u.num_tokens = ch.num_tokens();
            if (u.num_tokens >= (int64_t)UINT32_MAX) fail(CULDA_ERR_LIMIT, "a chunk has >= 2^32 tokens");
            const int32_t Dc = ch.doc_end - ch.doc_start;
            std::vector<uint32_t> tok(u.num_tokens), rev_off(Dc + 1, 0), rev_idx(u.num_tokens);
            for (int64_t i = 0; i < u.num_tokens; ++i) tok[i] = (uint32_t)(ch.word_tokens[i] - ch.doc_start);
            for (int32_t d = 0; d < Dc; ++d) rev_off[d + 1] = rev_off[d] + (uint32_t)m.doc_len[ch.doc_start + d];
            {
                std::vector<uint32_t> nxt(rev_off.begin(), rev_off.end() - 1);
                for (int64_t i = 0; i < u.num_tokens; ++i) rev_idx[nxt[tok[i]]++] = (uint32_t)i;
            }
            std::vector<DocMeta> meta(Dc);
            int64_t cap = 0;
            for (int32_t d = 0; d < Dc; ++d) {
                const int64_t nd = m.doc_len[ch.doc_start + d];
                u.max_doc_len = std::max(u.max_doc_len, nd);
                meta[d] = DocMeta{(uint32_t)cap, 0u, (uint32_t)nd, 0u};
                cap += (theta_entries_max(nd, K) + 15) / 16 * 16;
                // the sampler addresses a chunk's theta with 32-bit byte offsets
                if (cap + kThetaPad >= ((int64_t)1 << 31))
                    fail(CULDA_ERR_LIMIT, "a chunk's theta capacity exceeds 2^31 entries: use more data_chunks");
            }
            u.theta_cap = cap;
            u.word_idx_h = ch.word_idx;
            u.region_h = doc_regions(tok, ch.word_idx, c.num_words, Dc);
            u.tokens = upload_vec(*this, tok);
            u.z = (uint16_t *)dalloc(std::max<int64_t>(u.num_tokens, 1) * sizeof(uint16_t));
            hip_check(hipMemset(u.z, 0, std::max<int64_t>(u.num_tokens, 1) * sizeof(uint16_t)), "z clear");
            u.zp = (uint16_t *)dalloc(std::max<int64_t>(u.num_tokens, 1) * sizeof(uint16_t));
            u.word_idx = upload_vec(*this, ch.word_idx);
            u.rev_off = upload_vec(*this, rev_off);
            u.rev_idx = upload_vec(*this, rev_idx);
            u.meta = upload_vec(*this, meta);
            // kThetaPad entries of tail padding: the sampler loads whole
            // 32-entry groups and may read up to 16 entries past the last
            // slab (those values are unused)
            u.theta = (uint16_t *)dalloc((cap + kThetaPad) * sizeof(uint16_t));
            hip_check(hipMemset(u.theta, 0, (cap + kThetaPad) * sizeof(uint16_t)), "theta clear");
            units.push_back(std::move(u));
        }
        m.unit_end = (int32_t)units.size();
        m.nwk = (int32_t *)dalloc((size_t)m.V * K * sizeof(int32_t));
        m.den = (float *)dalloc(K * sizeof(float));
        m.zero_phi = (float *)dalloc(K * sizeof(float));
        m.rden = (double *)dalloc(K * sizeof(double));
        models.push_back(std::move(m));
    }
    const int32_t M = (int32_t)models.size(), U = (int32_t)units.size();
    {
        int64_t longest = 0;
        for (const ModelHost &m : models)
            for (int64_t nd : m.doc_len) longest = std::max(longest, theta_entries_max(nd, K));
        rs_stride = sample_rs_stride(longest);
    }
    {
        int64_t words = 0, rare = 0;
        for (const ModelHost &m : models)
            for (int32_t w = 0; w < m.V; ++w) {
                int64_t n = 0;
                for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui)
                    n += units[ui].word_idx_h[w + 1] - units[ui].word_idx_h[w];
                ++words;
                rare += n * 16 < K;
            }
        sparse_phi = rare * 4 >= words;
    }
    d_nk_all = (int32_t *)dalloc((size_t)M * K * sizeof(int32_t));
    for (int32_t mi = 0; mi < M; ++mi) models[mi].nk = d_nk_all + (size_t)mi * K;
    // device tables
    std::vector<ModelDev> md(M);
    for (int32_t mi = 0; mi < M; ++mi) {
        const ModelHost &m = models[mi];
        md[mi] = ModelDev{m.nwk, m.nk, m.den, m.zero_phi, m.rden, m.seed, (double)K * (double)m.alpha, m.alpha, m.beta, K,
                          m.V};
    }
    d_models = upload_vec(*this, md);
    // zm for the units whose rare words are updated in tiles (one-unit models)
    if (sparse_phi)
        for (ModelHost &m : models)
            if (m.unit_end - m.unit_begin == 1) {
                UnitHost &u = units[m.unit_begin];
                u.zm = (uint16_t *)dalloc(std::max<int64_t>(u.num_tokens, 1) * sizeof(uint16_t));
            }
    std::vector<UnitDev> ud(U);
    for (int32_t ui = 0; ui < U; ++ui) {
        const UnitHost &u = units[ui];
        ud[ui] = UnitDev{u.tokens, u.z, u.zp, u.word_idx, u.rev_off, u.rev_idx, u.meta, u.theta, u.model,
                         u.doc_end - u.doc_start, u.zm};
    }
    d_units = upload_vec(*this, ud);
    // phi work lists: fused per model (and all models concatenated), and per unit
    {
        int64_t all_tokens = 0;
        for (const ModelHost &m : models) all_tokens += m.num_tokens;
        const int64_t split = phi_split(all_tokens);
        // whole-word lists in packets of one resident round of the rebuild
        // launch (a second round of a few workgroups would double its time)
        const int32_t per_cu = phi_blocks_per_cu();
        const size_t phi_groups = (size_t)cu_count * (size_t)(per_cu > 0 ? per_cu : 8);
        std::vector<PhiSlot> all_slots, sparse_slots;
        std::vector<Packet> all_packets;
        std::vector<int2> all_zero;
        std::vector<RareTile> tiles;
        std::vector<int64_t> model_tiles(M + 1, 0);
        std::vector<std::vector<PhiSlot>> model_sparse(M);
        for (int32_t mi = 0; mi < M; ++mi) {
            ModelHost &m = models[mi];
            std::vector<PhiSlot> sl;
            std::vector<int64_t> tok;
            std::vector<int2> zr;
            build_phi_list(units, mi, m.unit_begin, m.unit_end, m.V, split, sl, tok, zr);
            // the zp-based update skips words without tokens and the words
            // of the rare tiles
            std::vector<char> covered(m.V, 0);
            if (m.unit_end - m.unit_begin == 1)
                build_rare_tiles(units[m.unit_begin], m.unit_begin, m.V, K, tiles, covered);
            model_tiles[mi + 1] = (int64_t)tiles.size();
            for (size_t i = 0; i < sl.size(); ++i)
                if (sl[i].split || (tok[i] > 0 && !covered[sl[i].word])) model_sparse[mi].push_back(sl[i]);
            const int32_t per = std::max<int32_t>(4, (int32_t)((sl.size() + phi_groups - 1) / phi_groups));
            std::vector<Packet> pk;
            deal_packets(sl, mi, per, 0, pk);
            m.phi_fused.slots = upload_vec(*this, sl);
            m.phi_fused.packets = upload_vec(*this, pk);
            m.phi_fused.zero_rows = upload_vec(*this, zr);
            m.phi_fused.num_slots = (int32_t)sl.size();
            m.phi_fused.num_packets = (int32_t)pk.size();
            m.phi_fused.num_zero_rows = (int32_t)zr.size();
            const int32_t off = (int32_t)all_slots.size();
            for (Packet p : pk) all_packets.push_back(Packet{p.first + off, p.last + off, p.owner, 0});
            all_slots.insert(all_slots.end(), sl.begin(), sl.end());
            all_zero.insert(all_zero.end(), zr.begin(), zr.end());
            for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
                std::vector<PhiSlot> s1;
                std::vector<int64_t> t1;
                std::vector<int2> z1;
                build_phi_list(units, mi, ui, ui + 1, m.V, split, s1, t1, z1);
                const int32_t per1 =
                    std::max<int32_t>(4, (int32_t)((s1.size() + phi_groups - 1) / phi_groups));
                std::vector<Packet> p1;
                deal_packets(s1, mi, per1, 0, p1);
                UnitHost &u = units[ui];
                u.phi_single.slots = upload_vec(*this, s1);
                u.phi_single.packets = upload_vec(*this, p1);
                u.phi_single.zero_rows = upload_vec(*this, z1);
                u.phi_single.num_slots = (int32_t)s1.size();
                u.phi_single.num_packets = (int32_t)p1.size();
                u.phi_single.num_zero_rows = (int32_t)z1.size();
            }
        }
        phi_all.slots = upload_vec(*this, all_slots);
        phi_all.packets = upload_vec(*this, all_packets);
        phi_all.zero_rows = upload_vec(*this, all_zero);
        phi_all.num_slots = (int32_t)all_slots.size();
        phi_all.num_packets = (int32_t)all_packets.size();
        phi_all.num_zero_rows = (int32_t)all_zero.size();
        // packet sizes from the totals over every model, so that a batch of
        // many small corpora still gets a few hundred items per workgroup
        size_t sparse_total = 0;
        for (const auto &v : model_sparse) sparse_total += v.size();
        const size_t groups = (size_t)cu_count * 8;
        const int32_t per_s = std::max<int32_t>(4, (int32_t)((sparse_total + groups - 1) / groups));
        const int32_t per_t = std::max<int32_t>(4, (int32_t)((tiles.size() + groups - 1) / groups));
        std::vector<Packet> sparse_packets, tile_packets;
        for (int32_t mi = 0; mi < M; ++mi) {
            deal_packets(model_sparse[mi], mi, per_s, (int32_t)sparse_slots.size(), sparse_packets);
            sparse_slots.insert(sparse_slots.end(), model_sparse[mi].begin(), model_sparse[mi].end());
            append_packets(tile_packets, (int32_t)model_tiles[mi], (int32_t)model_tiles[mi + 1], mi, per_t);
        }
        if (tiles.size() >= (size_t)INT32_MAX) fail(CULDA_ERR_LIMIT, "too many rare-word tiles");
        phi_sparse.slots = upload_vec(*this, sparse_slots);
        phi_sparse.packets = upload_vec(*this, sparse_packets);
        phi_sparse.zero_rows = phi_all.zero_rows;
        phi_sparse.num_slots = (int32_t)sparse_slots.size();
        phi_sparse.num_packets = (int32_t)sparse_packets.size();
        phi_sparse.num_zero_rows = phi_all.num_zero_rows;
        d_rare_tiles = upload_vec(*this, tiles);
        d_rare_packets = upload_vec(*this, tile_packets);
        num_rare_packets = (int32_t)tile_packets.size();
        num_rare_tiles = (int64_t)tiles.size();
    }
    // theta packets.  Per unit (a chunk's rebuild, which runs beside the
    // next chunk's sweep): about two workgroups per CU by default, so that
    // it keeps pace with the sweeps without taking the CUs from them (the
    // autotuner may pick another count).  All units at once (batches, full
    // rebuilds): eight workgroups per CU over the total.
    {
        int64_t total_docs = 0;
        for (const UnitHost &u : units) total_docs += u.doc_end - u.doc_start;
        const int64_t per_all = std::max<int64_t>(16, (total_docs + (int64_t)cu_count * 8 - 1) / ((int64_t)cu_count * 8));
        // a batch's models each on the XCD that samples them (their z, just
        // written there, may still be in its L2)
        if (M > 1) assign_models_to_xcds();
        std::vector<std::vector<Packet>> q(M > 1 ? kXcds : 1);
        for (int32_t ui = 0; ui < U; ++ui) {
            const int64_t Du = units[ui].doc_end - units[ui].doc_start;
            append_packets(q[M > 1 ? model_xcd[units[ui].model] : 0], 0, (int32_t)Du, ui, (int32_t)per_all);
        }
        const std::vector<Packet> all = M > 1 ? interleave_xcds(q) : q[0];
        d_theta_packets_all = upload_vec(*this, all);
        num_theta_packets_all = (int32_t)all.size();
        build_theta_packets(theta_groups_per_cu);
    }
    d_sweep = (uint32_t *)dalloc(16);
    d_queue = (uint32_t *)dalloc(16);
    hip_check(hipMemset(d_sweep, 0, 16), "sweep clear");
    d_ll_hist = (double *)dalloc(sizeof(double) * kLLHistory * M);
    hip_check(hipMemset(d_ll_hist, 0, sizeof(double) * kLLHistory * M), "ll clear");
    d_nnz_hist = (double *)dalloc(sizeof(double) * kLLHistory * M);
    hip_check(hipMemset(d_nnz_hist, 0, sizeof(double) * kLLHistory * M), "history clear");
    ev_sampled.assign(U, nullptr);
    for (hipEvent_t &e : ev_sampled) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    ev_theta_unit.assign(U, nullptr);
    for (hipEvent_t &e : ev_theta_unit) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    uploaded = true;
    rebuild_slots(prm.slot_tokens);
    hip_check(hipDeviceSynchronize(), "upload");
    pending.clear();
    set_theta_ready(false);
    phi_ready = sampled = false;
    zp_valid = false;
}

// Default tokens per sampler slot: about 32 slots per CU for the tokens of
// one launch (a unit; every unit of a batch), as a power of two in
// [128, 1024] -- a small corpus needs short slots to fill the device, a
// large one gains nothing below 1024.
int32_t Engine::default_slot_tokens() const {
    int64_t per_launch = 0, total = 0;
    for (const UnitHost &u : units) {
        per_launch = std::max(per_launch, u.num_tokens);
        total += u.num_tokens;
    }
    if (models.size() > 1) per_launch = total;
    const int64_t want = per_launch / ((int64_t)cu_count * 32);
    int32_t st = 128;
    while (st < 1024 && (int64_t)st * 2 <= want) st *= 2;
    return st;
}

// Sampler slots: each unit's word ranges cut into pieces of at most
// slot_tokens (and, where the unit's theta rows outgrow the L2s, at its
// XCD doc regions), longest first within the unit (dealt to the XCDs by
// region there); units in order, so a model's slots are contiguous (its LL
// partials too).  d_order is the launch order over all units: for a batch,
// each model's slots on one XCD.
void Engine::rebuild_slots(int32_t slot_tokens_in) {
    if (graph) {
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    dfree(d_slots);
    dfree(d_order);
    dfree(d_order_claims);
    dfree(d_ll_partial);
    dfree(d_ll_packets);
    dfree(d_model_packets);
    dfree(d_ll_scratch);
    dfree(d_nnz_partial);
    dfree(d_nnz_scratch);
    const int64_t slot_tokens = slot_tokens_in > 0 ? slot_tokens_in : default_slot_tokens();
    std::vector<SlotDesc> all;
    std::vector<Packet> llp;
    std::vector<int32_t> mp(2 * models.size());
    for (size_t mi = 0; mi < models.size(); ++mi) {
        ModelHost &m = models[mi];
        m.slot_begin = (int32_t)all.size();
        // g = sum_{w'<w} cnt(w') + sum_{c'<c} cnt_c'(w) + local
        std::vector<int64_t> pre(m.V + 1, 0);
        for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui)
            for (int32_t w = 0; w < m.V; ++w) pre[w + 1] += units[ui].word_idx_h[w + 1] - units[ui].word_idx_h[w];
        for (int32_t w = 0; w < m.V; ++w) pre[w + 1] += pre[w];
        std::vector<int64_t> run(pre.begin(), pre.end() - 1);
        for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
            UnitHost &u = units[ui];
            // XCD affinity where the unit's theta rows outgrow the XCDs' L2s
            // together (smaller ones stay cached anyway, and the extra slots
            // would only cost): doc regions of at least 64 docs, docs sorted
            // in every word
            const bool xcd = !u.region_h.empty() && u.doc_end - u.doc_start >= 64 * (int32_t)kXcds &&
                             u.theta_cap * (int64_t)sizeof(uint16_t) > (int64_t)kXcds * kL2BytesPerXcd;
            std::vector<SlotDesc> sl;
            std::vector<int8_t> tag;  // the XCD region holding all of a slot's docs, or -1
            for (int32_t w = 0; w < m.V; ++w) {
                const int64_t b = u.word_idx_h[w], e = u.word_idx_h[w + 1], n = e - b;
                if (n > 0) {
                    // A rare word (fewer than K / 16 tokens, so at most 63) of
                    // a one-unit model keeps all its tokens in one slot
                    // whatever the slot size: the sampler may stage its row
                    // from the tokens' zp topics, one token per lane.
                    const int32_t rare = m.unit_end - m.unit_begin == 1 && n * 16 < m.K;
                    auto add = [&](int64_t x0, int64_t x1, int8_t t) {
                        const int64_t nx = x1 - x0;
                        const int64_t pieces = rare ? 1 : (nx + slot_tokens - 1) / slot_tokens;
                        const int64_t step = (nx + pieces - 1) / pieces;
                        for (int64_t x = x0; x < x1; x += step) {
                            sl.push_back(SlotDesc{w, ui, x, std::min(x1, x + step), run[w] - b, rare, 0});
                            tag.push_back(t);
                        }
                    };
                    const uint32_t *rb = xcd ? &u.region_h[(size_t)w * (kXcds + 1)] : nullptr;
                    if (rb && !rare && n > slot_tokens) {
                        // a word of several slots is cut at the region
                        // boundaries too: each piece's docs (whose theta rows
                        // and metadata the sampler gathers) lie in one region
                        for (uint32_t r = 0; r < kXcds; ++r)
                            if (rb[r + 1] > rb[r]) add(b + rb[r], b + rb[r + 1], (int8_t)r);
                    } else {
                        int8_t t = -1;
                        for (uint32_t r = 0; rb && r < kXcds; ++r)
                            if (rb[r] == 0 && rb[r + 1] == (uint32_t)n) t = (int8_t)r;
                        add(b, e, t);
                    }
                }
                run[w] += n;
            }
            std::vector<size_t> by_len(sl.size());
            std::iota(by_len.begin(), by_len.end(), 0);
            std::stable_sort(by_len.begin(), by_len.end(),
                             [&](size_t a, size_t b) { return sl[a].end - sl[a].begin > sl[b].end - sl[b].begin; });
            u.slot_begin = (int32_t)all.size();
            for (size_t i : xcd ? deal_slots_to_xcds(by_len, tag, sl) : by_len) all.push_back(sl[i]);
            u.slot_end = (int32_t)all.size();
        }
        m.slot_end = (int32_t)all.size();
        mp[2 * mi] = (int32_t)llp.size();
        append_packets(llp, m.slot_begin, m.slot_end, (int32_t)mi, 2048);
        mp[2 * mi + 1] = (int32_t)llp.size();
    }
    if (all.size() >= (size_t)INT32_MAX) fail(CULDA_ERR_LIMIT, "too many sampler slots");
    rare_slots = rare_tokens = 0;
    for (const SlotDesc &sd : all)
        if (sd.rare) {
            ++rare_slots;
            rare_tokens += sd.end - sd.begin;
        }
    num_slots = (int32_t)all.size();
    std::vector<int32_t> by_len(all.size());
    std::iota(by_len.begin(), by_len.end(), 0);
    std::stable_sort(by_len.begin(), by_len.end(), [&](int32_t a, int32_t b) {
        return all[a].end - all[a].begin > all[b].end - all[b].begin;
    });
    // the one-launch order of a batch: every model on one XCD
    const std::vector<int32_t> order = models.size() > 1 ? deal_models_to_xcds(by_len, all) : by_len;
    // persistent claims: run j of kQueueClaim positions holds the slots of
    // rank j, j + R, j + 2R, ... (R runs), so every claim mixes long and
    // short slots and the longest are still taken first
    std::vector<int32_t> claims;
    claims.reserve(by_len.size());
    {
        const size_t R = (by_len.size() + kQueueClaim - 1) / kQueueClaim;
        for (size_t j = 0; j < R; ++j)
            for (size_t i = j; i < by_len.size(); i += R) claims.push_back(by_len[i]);
    }
    d_slots = upload_vec(*this, all);
    d_order = upload_vec(*this, order);
    d_order_claims = upload_vec(*this, claims);
    d_ll_partial = (double *)dalloc(sizeof(double) * std::max<size_t>(all.size(), 1));
    hip_check(hipMemset(d_ll_partial, 0, sizeof(double) * std::max<size_t>(all.size(), 1)), "ll clear");
    d_ll_packets = upload_vec(*this, llp);
    d_model_packets = upload_vec(*this, mp);
    d_ll_scratch = (double *)dalloc(sizeof(double) * std::max<size_t>(llp.size(), 1));
    d_nnz_partial = (double *)dalloc(sizeof(double) * std::max<size_t>(all.size(), 1));
    hip_check(hipMemset(d_nnz_partial, 0, sizeof(double) * std::max<size_t>(all.size(), 1)), "partials clear");
    d_nnz_scratch = (double *)dalloc(sizeof(double) * std::max<size_t>(llp.size(), 1));
    num_ll_packets = (int32_t)llp.size();
    prm.slot_tokens = (int32_t)slot_tokens;
    // the copies and the memset above run on the null stream, which does not
    // order against the engine's non-blocking streams
    hip_check(hipDeviceSynchronize(), "slot upload");
}

void Engine::assign_models_to_xcds() {
    // largest first onto the least-loaded XCD (tokens)
    std::vector<int32_t> ms(models.size());
    std::iota(ms.begin(), ms.end(), 0);
    std::stable_sort(ms.begin(), ms.end(),
                     [&](int32_t a, int32_t b) { return models[a].num_tokens > models[b].num_tokens; });
    std::vector<int64_t> load(kXcds, 0);
    model_xcd.assign(models.size(), 0);
    model_rank.assign(models.size(), 0);
    for (size_t i = 0; i < ms.size(); ++i) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        model_xcd[ms[i]] = x;
        model_rank[ms[i]] = (int32_t)i;
        load[x] += models[ms[i]].num_tokens;
    }
}

std::vector<int32_t> Engine::deal_models_to_xcds(const std::vector<int32_t> &by_len,
                                                const std::vector<SlotDesc> &all) const {
    // each XCD's slots model by model (largest model first), each model's
    // longest first: the XCD's L2 then mostly holds one model's theta rows
    std::vector<int32_t> sorted(by_len);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int32_t a, int32_t b) {
        return model_rank[units[all[a].unit].model] < model_rank[units[all[b].unit].model];
    });
    std::vector<std::vector<int32_t>> q(kXcds);
    for (int32_t i : sorted) q[model_xcd[units[all[i].unit].model]].push_back(i);
    return interleave_xcds(q);
}

void Engine::set_sweep(uint32_t t, hipStream_t s) {
    hip_check(launch_fill_u32(d_sweep, 1, t, s), "sweep set");
}

void Engine::init_z() {
    if (!uploaded) fail(CULDA_ERR_STATE, "init_z before upload");
    InitArgs a{d_slots, d_units, d_models, num_slots, 0};
    hip_check(launch_init_z(a, stream), "init_z launch");
    set_theta_ready(false);
    phi_ready = false;
}

// unit < 0: model's chunks in one pass (overwrite).  unit >= 0: that unit
// alone, added to the counts when accumulate (chunks in turn).
void Engine::rebuild_phi_model(int32_t mi, int32_t unit, bool accumulate, hipStream_t s, const LLArgs *ll) {
    ModelHost &m = models[mi];
    const PhiList &L = unit < 0 ? m.phi_fused : units[unit].phi_single;
    if (unit < 0) accumulate = false;
    if (!accumulate) {
        hip_check(launch_phi_clear((uint32_t *)m.nk, m.K, d_models, L.zero_rows, L.num_zero_rows, s, ll), "phi clear");
    } else if (ll) {
        hip_check(launch_phi_clear(nullptr, 0, d_models, nullptr, 0, s, ll), "phi clear");
    }
    PhiArgs a{L.slots, L.packets, d_units, d_models, L.num_packets, accumulate ? 1 : 0};
    hip_check(launch_phi_rebuild(a, s), "phi rebuild launch");
    phi_ready = false;
    // rows written from z: zp and zm follow (a one-unit model's unit list
    // is the whole model), so a later pass over the rare tiles keeps them
    const bool whole = unit < 0 || m.unit_end - m.unit_begin == 1;
    if (whole && sparse_phi) copy_zp(mi, s);
    if (whole && models.size() == 1) rare_stale = false;
    if (unit < 0)
        zp_valid = sparse_phi && models.size() == 1;
    else
        zp_valid = false;  // one chunk's counts: complete only after the caller's loop
}

void Engine::copy_zp(int32_t model, hipStream_t s) {
    for (const UnitHost &u : units)
        if (model < 0 || u.model == model) {
            hip_check(launch_copy_u16(u.z, u.zp, u.num_tokens, s), "zp copy");
            if (u.zm) hip_check(launch_copy_u16(u.z, u.zm, u.num_tokens, s), "zm copy");
        }
}

void Engine::fresh_rows(hipStream_t s) {
    if (!rare_stale) return;
    RareArgs r{d_rare_tiles, d_rare_packets, d_units, d_models, num_rare_packets, 1};
    hip_check(launch_phi_rare(r, s), "rare rows launch");
    rare_stale = false;
}

void Engine::update_phi(hipStream_t s, const LLArgs *ll) {
    if (zp_valid && sparse_phi) {
        hip_check(launch_phi_clear((uint32_t *)d_nk_all, (int64_t)models.size() * prm.k, d_models, phi_all.zero_rows,
                                   phi_all.num_zero_rows, s, ll),
                  "phi clear");
        PhiArgs a{phi_sparse.slots, phi_sparse.packets, d_units, d_models, phi_sparse.num_packets, 2};
        RareArgs r{d_rare_tiles, d_rare_packets, d_units, d_models, num_rare_packets, 0};
        hip_check(launch_phi_update(a, r, s), "phi update launch");
        rare_stale = num_rare_packets > 0;
        phi_ready = false;
    } else if (prm.fused_phi || units.size() == models.size()) {
        rebuild_phi_all(s, ll);
    } else {
        for (size_t mi = 0; mi < models.size(); ++mi)
            for (int32_t ui = models[mi].unit_begin; ui < models[mi].unit_end; ++ui)
                rebuild_phi_model((int32_t)mi, ui, ui > models[mi].unit_begin, s, mi == 0 && ui == 0 ? ll : nullptr);
        if (sparse_phi) copy_zp(-1, s);
        zp_valid = sparse_phi;
        rare_stale = false;
    }
}

void Engine::update_counts_one_unit(hipStream_t s, const LLArgs *ll) {
    const ThetaArgs t{units[0].theta_packets, d_units, d_models, units[0].num_theta_packets, 0};
    const RareArgs none{d_rare_tiles, d_rare_packets, d_units, d_models, 0, 0};
    if (zp_valid && sparse_phi) {
        hip_check(launch_phi_clear((uint32_t *)d_nk_all, prm.k, d_models, phi_all.zero_rows, phi_all.num_zero_rows, s,
                                   ll),
                  "phi clear");
        const PhiArgs a{phi_sparse.slots, phi_sparse.packets, d_units, d_models, phi_sparse.num_packets, 2};
        const RareArgs r{d_rare_tiles, d_rare_packets, d_units, d_models, num_rare_packets, 0};
        hip_check(launch_count_update(a, r, t, s), "count update launch");
        rare_stale = num_rare_packets > 0;
    } else {
        hip_check(launch_phi_clear((uint32_t *)d_nk_all, prm.k, d_models, phi_all.zero_rows, phi_all.num_zero_rows, s,
                                   ll),
                  "phi clear");
        const PhiArgs a{phi_all.slots, phi_all.packets, d_units, d_models, phi_all.num_packets, 0};
        hip_check(launch_count_update(a, none, t, s), "count update launch");
        if (sparse_phi) copy_zp(-1, s);
        zp_valid = sparse_phi;
        rare_stale = false;
    }
    phi_ready = false;
    theta_ok[0] = 1;
}

void Engine::rebuild_phi_all(hipStream_t s, const LLArgs *ll) {
    hip_check(launch_phi_clear((uint32_t *)d_nk_all, (int64_t)models.size() * prm.k, d_models, phi_all.zero_rows,
                               phi_all.num_zero_rows, s, ll),
              "phi clear");
    PhiArgs a{phi_all.slots, phi_all.packets, d_units, d_models, phi_all.num_packets, 0};
    hip_check(launch_phi_rebuild(a, s), "phi rebuild launch");
    phi_ready = false;
    if (sparse_phi) copy_zp(-1, s);
    zp_valid = sparse_phi;
    rare_stale = false;
}

void Engine::phi_head(hipStream_t s, const LLArgs *ll, bool advance) {
    hip_check(launch_den(d_models, (int32_t)models.size(), s, ll, advance ? d_sweep : nullptr), "den launch");
    phi_ready = true;
}

LLArgs Engine::ll_args() const {
    return LLArgs{d_ll_partial,    d_ll_packets,  d_model_packets, d_ll_scratch,   d_ll_hist,
                  d_nnz_partial,   d_nnz_scratch, d_nnz_hist,      d_sweep,        num_ll_packets,
                  (int32_t)models.size()};
}

// Packets of `per` consecutive docs of [0, D) in the order a grid of `grid`
// workgroups takes them (workgroup b takes packets b, b + grid, ... and runs
// on XCD b mod kXcds): each XCD walks its own contiguous share of the docs
// with its workgroups side by side, so the docs in flight on one XCD are
// neighbours, and the z lines of the words they share (a word's tokens lie
// in doc order in the word-major z) are gathered from that XCD's L2 instead
// of the fabric.  Empty packets pad the XCDs' lists to whole rounds.
static void append_xcd_packets(std::vector<Packet> &out, int32_t D, int32_t owner, int32_t per, int32_t grid) {
    const int32_t X = (int32_t)kXcds;
    if (grid < X) {
        append_packets(out, 0, D, owner, per);
        return;
    }
    std::vector<std::vector<Packet>> lists(X);
    for (int32_t x = 0; x < X; ++x)
        append_packets(lists[x], (int32_t)((int64_t)D * x / X), (int32_t)((int64_t)D * (x + 1) / X), owner, per);
    const int64_t groups_per_xcd = grid / X;  // workgroups b < X * groups_per_xcd; the rest take only padding
    int64_t rounds = 0;
    for (const auto &l : lists) rounds = std::max<int64_t>(rounds, ((int64_t)l.size() + groups_per_xcd - 1) / groups_per_xcd);
    for (int64_t r = 0; r < rounds; ++r)
        for (int32_t b = 0; b < grid; ++b) {
            const int64_t i = r * groups_per_xcd + b / X;
            const auto &l = lists[b % X];
            out.push_back(b < X * groups_per_xcd && i < (int64_t)l.size() ? l[i] : Packet{0, 0, owner, 0});
        }
}

void Engine::build_theta_packets(int32_t groups_per_cu) {
    theta_groups_per_cu = groups_per_cu;  // read by theta_beside_grid
    const int64_t groups = (int64_t)cu_count * (groups_per_cu > 0 ? groups_per_cu : 2);
    for (int32_t ui = 0; ui < (int32_t)units.size(); ++ui) {
        UnitHost &u = units[ui];
        const int64_t Du = u.doc_end - u.doc_start;
        const int64_t per_u = std::max<int64_t>(16, (Du + groups - 1) / groups);
        std::vector<Packet> pk;
        if (units.size() > 1)
            append_xcd_packets(pk, (int32_t)Du, ui, kWavesPerGroup, theta_beside_grid(ui));
        else
            append_packets(pk, 0, (int32_t)Du, ui, (int32_t)per_u);
        dfree(u.theta_packets);
        u.theta_packets = upload_vec(*this, pk);
        u.num_theta_packets = (int32_t)pk.size();
    }
}

// workgroups of a chunk's theta rebuild beside a sweep: a bounded share of
// the CUs, theta_groups_per_cu per CU (by default 4 of the one-wave kernel,
// 2 of the four-wave one)
int32_t Engine::theta_beside_grid(int32_t ui) const {
    const int32_t g = theta_groups_per_cu > 0 ? theta_groups_per_cu : (theta_slim(ui) ? 4 : 2);
    return g * cu_count;
}

// The one-wave kernel's waves take their packets' docs in turn, and every
// doc has a fixed cost besides its tokens: a chunk of short docs needs more
// waves than fit beside the sweep, so it keeps the four-wave kernel (16
// chunks of 90-token docs, 664M tokens: 20.5 G tokens/s with it, 20.2 with
// the one-wave kernel at 4 per CU; 4 chunks of 323-token docs, 97M tokens:
// 13.85 with the four-wave kernel, 13.93 with the one-wave one).
static constexpr int64_t kSlimMinMeanDocLen = 192;

bool Engine::theta_slim(int32_t ui) const {
    const UnitHost &u = units[ui];
    return units.size() > 1 && u.max_doc_len < 65536 && u.num_tokens < ((int64_t)1 << 30) &&
           u.num_tokens >= kSlimMinMeanDocLen * (int64_t)(u.doc_end - u.doc_start);
}

void Engine::rebuild_theta_unit(int32_t ui, hipStream_t s) {
    ThetaArgs a{units[ui].theta_packets, d_units, d_models, units[ui].num_theta_packets, 0};
    if (theta_slim(ui)) {
        hip_check(launch_theta_rebuild_slim(a, theta_beside_grid(ui), s), "theta rebuild launch");
        theta_ok[ui] = 1;
        return;
    }
    // a single chunk's rebuild has nothing beside it: the next sweep waits;
    // a chunk's rebuild beside the next chunk's sweep holds a bounded grid
    // (theta_beside_grid), so that the sweep keeps most of the CUs (the
    // sweeps are the critical path)
    const bool alone = units.size() == 1;
    hip_check(launch_theta_rebuild(a, alone, alone ? 0 : theta_beside_grid(ui), s), "theta rebuild launch");
    theta_ok[ui] = 1;
}

void Engine::rebuild_theta_all(hipStream_t s) {
    ThetaArgs a{d_theta_packets_all, d_units, d_models, num_theta_packets_all, 0};
    hip_check(launch_theta_rebuild(a, false, 0, s), "theta rebuild launch");
    set_theta_ready(true);
}

void Engine::sample_unit(int32_t ui, hipStream_t s) {
    UnitHost &u = units[ui];
    SampleArgs a{d_slots + u.slot_begin, nullptr, d_units, d_models, d_ll_partial + u.slot_begin,
                 d_nnz_partial + u.slot_begin, d_sweep, nullptr,
                 u.slot_end - u.slot_begin, (zp_valid && sparse_phi) ? 1 : 0, rs_stride, 0};
    hip_check(launch_sample(a, 0, s), "sampler launch");
}

void Engine::sample_all(hipStream_t s) {
    SampleArgs a{d_slots, d_order, d_units, d_models, d_ll_partial, d_nnz_partial, d_sweep, nullptr, num_slots,
                 (zp_valid && sparse_phi) ? 1 : 0, rs_stride, 0};
    int32_t grid = 0;
    if (persistent_groups > 0) {
        hip_check(launch_fill_u32(d_queue, 1, 0, s), "queue reset");
        a.queue = d_queue;
        a.order = d_order_claims;
        grid = std::min<int32_t>(num_slots, persistent_groups);
    }
    hip_check(launch_sample(a, grid, s), "sampler launch");
}

void Engine::ll_reduce(hipStream_t s) {
    hip_check(launch_ll_reduce(ll_args(), s), "ll reduce launch");
}

void Engine::theta_stats(int64_t &doc_nnz, int64_t &token_nnz) {
    doc_nnz = token_nnz = 0;
    if (!uploaded) return;
    sync();
    unsigned long long *d = (unsigned long long *)dalloc(2 * sizeof(unsigned long long));
    hip_check(launch_fill_u32((uint32_t *)d, 4, 0, stream), "stats clear");
    for (const UnitHost &u : units)
        hip_check(launch_meta_stats(u.meta, u.doc_end - u.doc_start, d, stream), "theta statistics");
    unsigned long long h[2] = {0, 0};
    hip_check(hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, stream), "stats copy");
    hip_check(hipStreamSynchronize(stream), "stats sync");
    dfree(d);
    doc_nnz = (int64_t)h[0];
    token_nnz = (int64_t)h[1];
}

void Engine::settle() {
    if (stream2_busy) {
        hip_check(hipStreamSynchronize(stream2), "sync");
        stream2_busy = false;
    }
}

void Engine::rebuild_all() {
    if (!uploaded) fail(CULDA_ERR_STATE, "rebuild before upload");
    hip_check(hipStreamSynchronize(stream2), "sync");
    if (prm.fused_phi) {
        rebuild_phi_all(stream);
    } else {
        for (size_t mi = 0; mi < models.size(); ++mi)
            for (int32_t ui = models[mi].unit_begin; ui < models[mi].unit_end; ++ui)
                rebuild_phi_model((int32_t)mi, ui, ui > models[mi].unit_begin, stream);
        if (sparse_phi) copy_zp(-1, stream);
        zp_valid = sparse_phi;
        rare_stale = false;
    }
    phi_head(stream);
    rebuild_theta_all(stream);
    sync();
}

double Engine::model_ll(int32_t mi, uint32_t sweep) {
    double h = 0;
    hip_check(hipMemcpy(&h, d_ll_hist + (size_t)(sweep % kLLHistory) * models.size() + mi, sizeof(double),
                        hipMemcpyDeviceToHost),
              "ll copy");
    return h / (double)models[mi].num_tokens;
}

// One per-chunk iteration enqueued on the two streams.
// stream: sweeps of all chunks, LL, phi + head denominators.
// stream2: theta of chunk c as soon as chunk c's sweep is done, so the
// (gather-bound) theta rebuilds overlap the next chunks' sweeps.  Theta of
// chunk c is read only by chunk c's sweep: the last chunk's rebuild also
// overlaps the next iteration, whose sweep of that chunk waits for
// ev_theta_unit[c].  A theta rebuild reads z of its own chunk, which no
// other launch of the iteration writes.  The rebuilds' statistics go to
// the pinned stats_dst; ev_a / ev_b (profiling) bracket the sweeps.
void Engine::enqueue_units_iteration(uint32_t sweep, bool set, hipEvent_t ev_a, hipEvent_t ev_b) {
    const int32_t U = (int32_t)units.size();
    if (set) set_sweep(sweep, stream);
    if (U == 1) {
        // one chunk: sweep, then the count update and the theta rebuild in
        // one launch, all on one stream (the next sweep needs both)
        if (stream2_busy) hip_check(hipStreamWaitEvent(stream, ev_theta_unit[0], 0), "wait");
        stream2_busy = false;
        if (ev_a) hip_check(hipEventRecord(ev_a, stream), "event");
        sample_unit(0, stream);
        if (ev_b) hip_check(hipEventRecord(ev_b, stream), "event");
        const LLArgs ll = ll_args();
        update_counts_one_unit(stream, &ll);
        phi_head(stream, &ll, true);
        return;
    }
    const bool after_iteration = stream2_busy;
    for (int32_t u = 0; u < U; ++u) {
        if (after_iteration) hip_check(hipStreamWaitEvent(stream, ev_theta_unit[u], 0), "wait");
        // the sweeps' time starts once the first one may run
        if (u == 0 && ev_a) hip_check(hipEventRecord(ev_a, stream), "event");
        sample_unit(u, stream);
        if (u + 1 == U) break;  // the last chunk's rebuild follows the count update (below)
        hip_check(hipEventRecord(ev_sampled[u], stream), "event");
        hip_check(hipStreamWaitEvent(stream2, ev_sampled[u], 0), "wait");
        rebuild_theta_unit(u, stream2);
        hip_check(hipEventRecord(ev_theta_unit[u], stream2), "event");
    }
    if (ev_b) hip_check(hipEventRecord(ev_b, stream), "event");
    // the LL reduction rides in the count update's clear launch (stage 1)
    // and in the head-denominator launch (stage 2), which then moves the
    // sweep counter on to the next sweep
    const LLArgs ll = ll_args();
    update_phi(stream, &ll);
    phi_head(stream, &ll, true);
    // the last chunk's theta rebuild is needed only by that chunk's next
    // sweep: started after the count update (on the critical path, and a
    // gather-bound rebuild beside it slows it), it runs beside the next
    // iteration's first sweeps instead
    hip_check(hipEventRecord(ev_sampled[U - 1], stream), "event");
    hip_check(hipStreamWaitEvent(stream2, ev_sampled[U - 1], 0), "wait");
    rebuild_theta_unit(U - 1, stream2);
    hip_check(hipEventRecord(ev_theta_unit[U - 1], stream2), "event");
    // all rebuilds but the last one before the next iteration's sweeps:
    // they run beside the sweeps and may fall behind (many chunks, short
    // sweeps); here they catch up while nothing else competes, so their
    // lag cannot carry from one iteration to the next
    if (U >= 2) hip_check(hipStreamWaitEvent(stream, ev_theta_unit[U - 2], 0), "wait");
    stream2_busy = true;
}

void Engine::iterate_units(uint32_t sweep, double *ll, double *seconds) {
    if (!uploaded || !theta_ready() || !phi_ready) fail(CULDA_ERR_STATE, "iterate before the initial rebuild");
    if (models.size() != 1) fail(CULDA_ERR_STATE, "per-chunk iteration serves a single corpus");
    const int32_t U = (int32_t)units.size();
    const bool prof = prm.profile != 0;
    if (prof && !ev[0])
        for (hipEvent_t &e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    auto t0 = std::chrono::steady_clock::now();
    enqueue_units_iteration(sweep, true, prof ? ev[0] : nullptr, prof ? ev[1] : nullptr);
    if (prof) hip_check(hipEventRecord(ev[2], stream), "event");
    double h[2] = {0, 0};  // LL sum, theta entries read
    hip_check(hipMemcpyAsync(&h[0], d_ll_hist + (size_t)(sweep % kLLHistory), sizeof(double), hipMemcpyDeviceToHost,
                             stream),
              "ll copy");
    hip_check(hipMemcpyAsync(&h[1], d_nnz_hist + (size_t)(sweep % kLLHistory), sizeof(double), hipMemcpyDeviceToHost,
                             stream),
              "ll copy");
    hip_check(hipStreamSynchronize(stream), "iteration sync");
    auto t1 = std::chrono::steady_clock::now();
    const double h_ll = h[0];
    const double bytes = sweep_bytes(h[1]);
    if (prof) {
        float a = 0, b = 0;
        hip_check(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
        hip_check(hipEventElapsedTime(&b, ev[1], ev[2]), "elapsed");
        sampler_ms = a;
        phi_ms = b;
        theta_ms = 0;  // the rebuilds overlap the sweeps and the next iteration
    }
    sampler_launches = U;
    last_sampler_bytes = bytes;
    sampled = true;
    if (ll) *ll = h_ll / (double)models[0].num_tokens;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

void Engine::iterate_units_n(uint32_t first, int32_t count, double *ll, double *seconds) {
    if (!uploaded || !theta_ready() || !phi_ready) fail(CULDA_ERR_STATE, "iterate before the initial rebuild");
    if (models.size() != 1) fail(CULDA_ERR_STATE, "per-chunk iteration serves a single corpus");
    if (count < 1 || count > kLLHistory) fail(CULDA_ERR_LIMIT, "count must be in [1, %d]", kLLHistory);
    const int32_t U = (int32_t)units.size();
    const bool prof = prm.profile != 0;
    if (prof)
        while ((int32_t)ev_iter.size() < 2 * count) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreate(&e), "hipEventCreate");
            ev_iter.push_back(e);
        }
    // One chunk: the iteration (sweep, count clear, count update + theta,
    // head) is captured once as a graph and replayed; it reads the sweep
    // index on the device and advances it.  The capture depends on whether
    // the counts follow zp (the sparse update), so a change of that state
    // captures again.  A replayed graph has no per-sweep timing events: the
    // call then reports a sampler time of 0 (culda_iterate times one).
    const bool use_graph = U == 1;
    const bool key = zp_valid && sparse_phi;
    if (use_graph && (!graph || graph_key != key)) {
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        settle();
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue_units_iteration(first, false, nullptr, nullptr);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &g), "end capture");
        hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "graph instantiate");
        graph_key = key;
        graph_rare_stale = rare_stale;
    }
    auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
        set_sweep(first, stream);
        for (int32_t i = 0; i < count; ++i) hip_check(hipGraphLaunch(graph, stream), "graph launch");
        rare_stale = graph_rare_stale;
    } else {
        for (int32_t i = 0; i < count; ++i)
            enqueue_units_iteration(first + (uint32_t)i, i == 0, prof ? ev_iter[2 * i] : nullptr,
                                    prof ? ev_iter[2 * i + 1] : nullptr);
    }
    std::vector<double> hist(kLLHistory), nnz(kLLHistory);
    hip_check(hipMemcpyAsync(hist.data(), d_ll_hist, kLLHistory * sizeof(double), hipMemcpyDeviceToHost, stream),
              "ll copy");
    hip_check(hipMemcpyAsync(nnz.data(), d_nnz_hist, kLLHistory * sizeof(double), hipMemcpyDeviceToHost, stream),
              "ll copy");
    hip_check(hipStreamSynchronize(stream), "iteration sync");
    auto t1 = std::chrono::steady_clock::now();
    double bytes = 0, ms = 0;
    for (int32_t i = 0; i < count; ++i) {
        bytes += sweep_bytes(sweep_entries(nnz, first + (uint32_t)i));
        if (prof && !use_graph) {
            float a = 0;
            hip_check(hipEventElapsedTime(&a, ev_iter[2 * i], ev_iter[2 * i + 1]), "elapsed");
            ms += a;
        }
    }
    if (prof) {
        sampler_ms = ms;
        phi_ms = 0;
        theta_ms = 0;
    }
    sampler_launches = U * count;
    last_sampler_bytes = bytes;
    sampled = true;
    for (int32_t i = 0; i < count; ++i)
        if (ll) ll[i] = hist[(first + (uint32_t)i) % kLLHistory] / (double)models[0].num_tokens;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

// One iteration of every model: a single sampler launch over all slots
// (persistent when persistent_groups > 0), LL, phi, head, theta, then the
// sweep counter advances -- everything reads the sweep index on the device,
// so the sequence can be captured once and replayed.  No host memset and
// no host synchronisation inside.
// One batched iteration on stream s.  The theta rebuild and the count
// update both only read the new z, so theta runs on stream2 beside the LL,
// phi and head-denominator launches; s waits for it before the sweep
// counter advances.  Captured into a graph, this is a fork and a join.
void Engine::enqueue_batch_iteration(hipStream_t s, hipEvent_t before, hipEvent_t after) {
    if (!ev_fork) {
        hip_check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
    }
    if (before) hip_check(hipEventRecord(before, s), "event");
    sample_all(s);
    if (after) hip_check(hipEventRecord(after, s), "event");
    hip_check(hipEventRecord(ev_fork, s), "event");
    hip_check(hipStreamWaitEvent(stream2, ev_fork, 0), "wait");
    rebuild_theta_all(stream2);
    hip_check(hipEventRecord(ev_join, stream2), "event");
    // LL stage 1 in the count update's clear launch, stage 2 in the head
    // launch, which also advances the sweep counter for a single model
    const LLArgs ll = ll_args();
    const bool one = models.size() == 1;
    update_phi(s, &ll);
    phi_head(s, &ll, one);
    hip_check(hipStreamWaitEvent(s, ev_join, 0), "wait");
    if (!one) hip_check(launch_add_u32(d_sweep, 1, s), "sweep advance");
}

void Engine::iterate_batch(uint32_t first, int32_t count, bool use_graph, double *ll, double *seconds) {
    if (!uploaded || !theta_ready() || !phi_ready) fail(CULDA_ERR_STATE, "iterate before the initial rebuild");
    if (first < 1 || count < 1) fail(CULDA_ERR_ARG, "first iteration must be >= 1 and count >= 1");
    if (count > kLLHistory) fail(CULDA_ERR_LIMIT, "at most %d iterations per call", kLLHistory);
    hip_check(hipStreamSynchronize(stream2), "sync");
    const bool prof = prm.profile != 0;
    if (prof && !ev[0])
        for (hipEvent_t &e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    if (use_graph && !graph) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue_batch_iteration(stream);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &g), "end capture");
        hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "graph instantiate");
        // the captured count update leaves the rare-tile rows stale after
        // every replay, which no host code sees: remember it here
        graph_rare_stale = rare_stale;
    }
    set_sweep(first, stream);
    auto t0 = std::chrono::steady_clock::now();
    if (prof) hip_check(hipEventRecord(ev[0], stream), "event");
    // the last eager iteration's sampler launch is timed on its own
    const bool time_sampler = prof && !use_graph;
    for (int32_t i = 0; i < count; ++i) {
        if (use_graph)
            hip_check(hipGraphLaunch(graph, stream), "graph launch");
        else if (time_sampler && i == count - 1)
            enqueue_batch_iteration(stream, ev[2], ev[3]);
        else
            enqueue_batch_iteration(stream);
    }
    if (use_graph && graph_rare_stale) rare_stale = true;
    if (prof) hip_check(hipEventRecord(ev[1], stream), "event");
    hip_check(hipStreamSynchronize(stream), "iteration sync");
    auto t1 = std::chrono::steady_clock::now();
    if (prof) {
        float a = 0, b = 0;
        hip_check(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
        iter_ms = a / count;
        if (time_sampler) hip_check(hipEventElapsedTime(&b, ev[2], ev[3]), "elapsed");
        sampler_ms = b;
        sampler_launches = time_sampler ? 1 : 0;
    }
    sampled = true;
    const size_t M = models.size();
    if (time_sampler) {
        std::vector<double> nnz((size_t)kLLHistory * M);
        hip_check(hipMemcpy(nnz.data(), d_nnz_hist, nnz.size() * 8, hipMemcpyDeviceToHost), "history copy");
        last_sampler_bytes = sweep_bytes(sweep_entries(nnz, first + (uint32_t)count - 1));
    }
    if (ll) {
        std::vector<double> hist((size_t)kLLHistory * M);
        hip_check(hipMemcpy(hist.data(), d_ll_hist, hist.size() * 8, hipMemcpyDeviceToHost), "ll copy");
        for (int32_t i = 0; i < count; ++i)
            for (size_t mi = 0; mi < M; ++mi)
                ll[(size_t)i * M + mi] = hist[(size_t)((first + i) % kLLHistory) * M + mi] / (double)models[mi].num_tokens;
    }
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

// Per token: doc id, doc meta, z write, the doc's theta nonzeros; per slot
// the word's n_wk row, or for a rare slot staged from zp (when the counts
// allow) zp, the count and the denominator of each token's topic.
double Engine::sweep_entries(const std::vector<double> &nnz_hist, uint32_t sweep) const {
    double e = 0;
    for (size_t mi = 0; mi < models.size(); ++mi) e += nnz_hist[(size_t)(sweep % kLLHistory) * models.size() + mi];
    return e;
}

double Engine::sweep_bytes(double nnz) const {
    int64_t n = 0;
    for (auto &m : models) n += m.num_tokens;
    const bool zp_stage = zp_valid && sparse_phi;
    const double rows = (double)(num_slots - (zp_stage ? rare_slots : 0)) * 4.0 * (double)prm.k;
    const double rare = zp_stage ? 6.0 * (double)rare_tokens : 0.0;  // zp topic + den_k
    return (double)n * (4.0 + 16.0 + 2.0) + 2.0 * (double)nnz + rows + rare;
}

// Times one sweep per candidate slot size from the same state (z is saved
// and restored, counts rebuilt), then keeps the fastest.  Results never
// depend on the slot size, so tuning leaves the chain unchanged.
TuneResult Engine::autotune(const std::vector<int32_t> &candidates, bool batch) {
    if (!uploaded || !theta_ready() || !phi_ready) fail(CULDA_ERR_STATE, "autotune needs the initial rebuild");
    hip_check(hipStreamSynchronize(stream2), "sync");
    std::vector<uint16_t *> saved(units.size());
    for (size_t u = 0; u < units.size(); ++u) {
        saved[u] = (uint16_t *)dalloc(std::max<int64_t>(units[u].num_tokens, 1) * 2);
        hip_check(hipMemcpyAsync(saved[u], units[u].z, units[u].num_tokens * 2, hipMemcpyDeviceToDevice, stream), "z save");
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    TuneResult r;
    r.best_ms = 1e30;
    for (int32_t cand : candidates) {
        rebuild_slots(cand);
        set_sweep(1, stream);
        hip_check(hipEventRecord(e0, stream), "event");
        if (batch) {
            sample_all(stream);
        } else {
            for (size_t u = 0; u < units.size(); ++u) sample_unit((int32_t)u, stream);
        }
        hip_check(hipEventRecord(e1, stream), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        r.tried.emplace_back(prm.slot_tokens, ms);
        if (ms < r.best_ms) {
            r.best_ms = ms;
            r.slot_tokens = prm.slot_tokens;
        }
        for (size_t u = 0; u < units.size(); ++u)
            hip_check(hipMemcpyAsync(units[u].z, saved[u], units[u].num_tokens * 2, hipMemcpyDeviceToDevice, stream),
                      "z restore");
    }
    rebuild_slots(r.slot_tokens);
    hip_check(hipStreamSynchronize(stream), "sync");
    if (!batch && models.size() == 1 && units.size() >= 2) {
        // a chunk's theta rebuild runs beside the next chunk's sweep, so its
        // workgroup count is timed over whole iterations (two, back to back,
        // each candidate from the same z and freshly rebuilt counts), not alone
        const int32_t U = (int32_t)units.size();
        double best = 1e30;
        for (int32_t g : {1, 2, 3, 4, 6, 8}) {
            build_theta_packets(g);
            for (size_t u = 0; u < units.size(); ++u)
                hip_check(hipMemcpyAsync(units[u].z, saved[u], units[u].num_tokens * 2, hipMemcpyDeviceToDevice, stream),
                          "z restore");
            rebuild_all();
            stream2_busy = false;
            hip_check(hipEventRecord(e0, stream), "event");
            enqueue_units_iteration(1, true, nullptr, nullptr);
            enqueue_units_iteration(2, true, nullptr, nullptr);
            hip_check(hipStreamWaitEvent(stream, ev_theta_unit[U - 1], 0), "wait");
            hip_check(hipEventRecord(e1, stream), "event");
            hip_check(hipEventSynchronize(e1), "event sync");
            stream2_busy = false;
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            r.theta_tried.emplace_back(g, ms);
            if (ms < best) {
                best = ms;
                r.theta_groups = g;
            }
        }
        build_theta_packets(r.theta_groups);
        for (size_t u = 0; u < units.size(); ++u)
            hip_check(hipMemcpyAsync(units[u].z, saved[u], units[u].num_tokens * 2, hipMemcpyDeviceToDevice, stream),
                      "z restore");
        rebuild_all();
    }
    for (uint16_t *p : saved) dfree(p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

void Engine::sync() {
    hip_check(hipStreamSynchronize(stream), "sync");
    hip_check(hipStreamSynchronize(stream2), "sync");
}

}  // namespace culda
