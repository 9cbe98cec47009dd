// context.cpp -- device state, layouts and the iteration driver behind the
// C ABI.
//
// Iteration schedule (train/SingleChunkSingleGPU.h:70-127 for one chunk,
// train/MultiChunkMultiGPUequal.h:88-169 for several, here on one device):
//   sample every chunk with the frozen counts of the previous iteration,
//   reduce the FP64 log-likelihood, rebuild n_wk / n_k over all chunks (on
//   device, no host round trip), prepare the head denominators, rebuild
//   every chunk's theta.  One host synchronisation per iteration.
#include "context.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "kernels.h"

namespace culda {

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(CULDA_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

Context::~Context() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : allocations) (void)hipFree(p);
    if (stream2) (void)hipStreamSynchronize(stream2);
    for (hipEvent_t &e : ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : ev_sampled)
        if (e) (void)hipEventDestroy(e);
    if (ev_theta_done) (void)hipEventDestroy(ev_theta_done);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
}

void *Context::dalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) fail(CULDA_ERR_MEMORY, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    allocations.push_back(p);
    return p;
}

template <class T>
static T *upload_vec(Context &ctx, const std::vector<T> &v) {
    T *d = (T *)ctx.dalloc(v.size() * sizeof(T));
    if (!v.empty()) hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

static uint32_t ceil_log2(uint32_t k) {
    uint32_t b = 0;
    while ((1u << b) < k) ++b;
    return b ? b : 1;
}

static constexpr int64_t kPhiSplit = 8192;  // hot words are rebuilt in pieces

// Phi work list over chunks [c0, c1): whole words first-class, hot words as
// per-chunk pieces of at most kPhiSplit tokens (rows pre-zeroed, atomics).
static void build_phi_slots(const std::vector<const std::vector<int64_t> *> &widx, int32_t V,
                            std::vector<PhiSlot> &slots, std::vector<int32_t> &split_words) {
    std::vector<std::pair<int64_t, int32_t>> whole;
    whole.reserve(V);
    for (int32_t w = 0; w < V; ++w) {
        int64_t tot = 0;
        for (auto *wi : widx) tot += (*wi)[w + 1] - (*wi)[w];
        if (tot > kPhiSplit) {
            split_words.push_back(w);
            for (size_t c = 0; c < widx.size(); ++c) {
                const int64_t b = (*widx[c])[w], e = (*widx[c])[w + 1];
                for (int64_t x = b; x < e; x += kPhiSplit)
                    slots.push_back(PhiSlot{w, (int32_t)c, x, std::min(e, x + kPhiSplit)});
            }
        } else {
            whole.emplace_back(tot, w);
        }
    }
    std::stable_sort(whole.begin(), whole.end(), [](auto &a, auto &b) { return a.first > b.first; });
    for (auto &pr : whole) slots.push_back(PhiSlot{pr.second, -1, 0, 0});
}

void Context::upload(const Corpus &c) {
    c.validate();
    const int32_t K = prm.k;
    num_words = c.num_words;
    num_docs = c.num_docs();
    num_tokens = c.num_tokens();
    if (num_tokens >= (int64_t)INT32_MAX) fail(CULDA_ERR_LIMIT, "corpus has %lld tokens; counts are int32", (long long)num_tokens);
    if ((int64_t)num_words * K >= ((int64_t)1 << 40)) fail(CULDA_ERR_LIMIT, "V*K too large");
    count_bits = 30 - ceil_log2((uint32_t)K);
    const int64_t max_count = ((int64_t)1 << count_bits) - 1;
    if (c.max_doc_len() > max_count)
        fail(CULDA_ERR_LIMIT, "doc length %d exceeds %lld, the largest count a packed theta entry holds at K=%d",
             c.max_doc_len(), (long long)max_count, K);
    doc_len.resize(num_docs);
    for (int32_t d = 0; d < num_docs; ++d) doc_len[d] = c.doc_len(d);

    const int32_t C = (int32_t)c.chunks.size();
    const int32_t V = num_words;
    chunks.assign(C, ChunkState());
    // chunking-invariant global ids: g = sum_{w'<w} cnt(w') + sum_{c'<c} cnt_c'(w) + local
    std::vector<int64_t> pre(V + 1, 0);
    for (const Chunk &ch : c.chunks)
        for (int32_t w = 0; w < V; ++w) pre[w + 1] += ch.word_idx[w + 1] - ch.word_idx[w];
    for (int32_t w = 0; w < V; ++w) pre[w + 1] += pre[w];
    std::vector<int64_t> run(pre.begin(), pre.end() - 1);

    const int64_t slot_tokens = prm.slot_tokens > 0 ? prm.slot_tokens : 1024;
    std::vector<int64_t> slot_count(C);
    int64_t total_slots = 0;
    std::vector<std::vector<SlotDesc>> all_slots(C);
    for (int32_t k = 0; k < C; ++k) {
        const Chunk &ch = c.chunks[k];
        ChunkState &cs = chunks[k];
        cs.doc_start = ch.doc_start;
        cs.doc_end = ch.doc_end;
        cs.num_tokens = ch.num_tokens();
        if (cs.num_tokens >= (int64_t)UINT32_MAX) fail(CULDA_ERR_LIMIT, "chunk %d has >= 2^32 tokens", k);
        const int32_t Dc = ch.doc_end - ch.doc_start;
        // sampler slots (longest first)
        std::vector<SlotDesc> &sl = all_slots[k];
        for (int32_t w = 0; w < V; ++w) {
            const int64_t b = ch.word_idx[w], e = ch.word_idx[w + 1], n = e - b;
            if (n > 0) {
                const int64_t pieces = (n + slot_tokens - 1) / slot_tokens;
                const int64_t step = (n + pieces - 1) / pieces;
                for (int64_t x = b; x < e; x += step)
                    sl.push_back(SlotDesc{w, 0, x, std::min(e, x + step), run[w] - b});
            }
            run[w] += n;
        }
        std::stable_sort(sl.begin(), sl.end(),
                         [](const SlotDesc &a, const SlotDesc &b) { return a.end - a.begin > b.end - b.begin; });
        cs.num_slots = (int32_t)sl.size();
        cs.ll_offset = total_slots;
        total_slots += cs.num_slots;
        // doc-major index and theta slabs
        std::vector<uint32_t> tok(cs.num_tokens), rev_off(Dc + 1, 0), rev_idx(cs.num_tokens);
        for (int64_t i = 0; i < cs.num_tokens; ++i) tok[i] = (uint32_t)(ch.word_tokens[i] - ch.doc_start);
        for (int32_t d = 0; d < Dc; ++d) rev_off[d + 1] = rev_off[d] + (uint32_t)doc_len[ch.doc_start + d];
        {
            std::vector<uint32_t> nxt(rev_off.begin(), rev_off.end() - 1);
            for (int64_t i = 0; i < cs.num_tokens; ++i) rev_idx[nxt[tok[i]]++] = (uint32_t)i;
        }
        std::vector<DocMeta> meta(Dc);
        int64_t cap = 0;
        cs.meta_start_h.resize(Dc);
        for (int32_t d = 0; d < Dc; ++d) {
            const int64_t nd = doc_len[ch.doc_start + d];
            const int64_t slab = (std::min<int64_t>(nd, K) + 15) / 16 * 16;
            meta[d] = DocMeta{(uint32_t)cap, 0u, (uint32_t)nd, 0u};
            cs.meta_start_h[d] = (uint32_t)cap;
            cap += slab;
            if (cap + 64 >= (int64_t)UINT32_MAX) fail(CULDA_ERR_LIMIT, "chunk %d theta capacity exceeds 2^32 entries", k);
        }
        cs.theta_cap = cap;
        cs.word_idx_h = ch.word_idx;
        // device copies
        cs.tokens = upload_vec(*this, tok);
        cs.z = (uint16_t *)dalloc(std::max<int64_t>(cs.num_tokens, 1) * sizeof(uint16_t));
        hip_check(hipMemset(cs.z, 0, std::max<int64_t>(cs.num_tokens, 1) * sizeof(uint16_t)), "z clear");
        cs.word_idx = upload_vec(*this, ch.word_idx);
        cs.rev_off = upload_vec(*this, rev_off);
        cs.rev_idx = upload_vec(*this, rev_idx);
        cs.meta = upload_vec(*this, meta);
        // 64 entries of tail padding: the sampler loads whole 64-entry
        // groups and may read past the last slab (those values are unused)
        cs.theta = (uint32_t *)dalloc((cap + 64) * sizeof(uint32_t));
        hip_check(hipMemset(cs.theta, 0, (cap + 64) * sizeof(uint32_t)), "theta clear");
        cs.slots = upload_vec(*this, sl);
        // phi work list of this chunk alone
        std::vector<PhiSlot> ps;
        std::vector<int32_t> sw;
        build_phi_slots({&cs.word_idx_h}, V, ps, sw);
        for (PhiSlot &p : ps)
            if (p.chunk >= 0) p.chunk = k;
        cs.phi_slots = upload_vec(*this, ps);
        cs.num_phi_slots = (int32_t)ps.size();
        cs.split_words = upload_vec(*this, sw);
        cs.num_split_words = (int32_t)sw.size();
    }
    // fused phi work list over all chunks
    {
        std::vector<const std::vector<int64_t> *> widx;
        for (auto &cs : chunks) widx.push_back(&cs.word_idx_h);
        std::vector<PhiSlot> ps;
        std::vector<int32_t> sw;
        build_phi_slots(widx, V, ps, sw);
        phi_slots_all = upload_vec(*this, ps);
        num_phi_slots_all = (int32_t)ps.size();
        split_words_all = upload_vec(*this, sw);
        num_split_words_all = (int32_t)sw.size();
    }
    std::vector<ChunkPtrs> cp(C);
    for (int32_t k = 0; k < C; ++k) cp[k] = ChunkPtrs{chunks[k].z, chunks[k].word_idx};
    chunk_ptrs = upload_vec(*this, cp);
    nwk = (int32_t *)dalloc((size_t)V * K * sizeof(int32_t));
    nk = (int32_t *)dalloc(K * sizeof(int32_t));
    den = (float *)dalloc(K * sizeof(float));
    ll_partials = total_slots;
    ll_partial = (double *)dalloc(std::max<int64_t>(total_slots, 1) * sizeof(double));
    hip_check(hipMemset(ll_partial, 0, std::max<int64_t>(total_slots, 1) * sizeof(double)), "hipMemset");
    ll_out = (double *)dalloc(sizeof(double));
    ll_scratch = (double *)dalloc(kLLGroups * sizeof(double));
    stats = (unsigned long long *)dalloc(2 * sizeof(unsigned long long));
    hip_check(hipMemset(stats, 0, 2 * sizeof(unsigned long long)), "hipMemset");
    ev_sampled.resize(C, nullptr);
    for (hipEvent_t &e : ev_sampled) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    if (!ev_theta_done) hip_check(hipEventCreateWithFlags(&ev_theta_done, hipEventDisableTiming), "event");
    hip_check(hipDeviceSynchronize(), "upload");
    uploaded = true;
    theta_ready = phi_ready = sampled = false;
}

void Context::init_z() {
    if (!uploaded) fail(CULDA_ERR_STATE, "init_z before upload");
    for (auto &cs : chunks) {
        InitArgs a{cs.slots, cs.z, prm.seed, cs.num_slots, prm.k};
        hip_check(launch_init_z(a, stream), "init_z launch");
    }
    theta_ready = phi_ready = false;
}

void Context::rebuild_phi(int32_t chunk, bool accumulate) {
    if (!uploaded) fail(CULDA_ERR_STATE, "rebuild_phi before upload");
    const int32_t C = (int32_t)chunks.size();
    if (chunk < -1 || chunk >= C) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
    const int32_t grid = std::max(1, cu_count) * 8;
    PhiArgs a{};
    a.chunks = chunk_ptrs;
    a.nwk = nwk;
    a.nk = nk;
    a.k = prm.k;
    if (chunk < 0) {
        accumulate = false;
        hip_check(launch_fill_u32((uint32_t *)nk, prm.k, 0, stream), "fill n_k");
        hip_check(launch_zero_rows(nwk, split_words_all, num_split_words_all, prm.k, stream), "zero rows");
        a.slots = phi_slots_all;
        a.num_slots = num_phi_slots_all;
        a.chunk_begin = 0;
        a.chunk_end = C;
    } else {
        const ChunkState &cs = chunks[chunk];
        if (!accumulate) {
            hip_check(launch_fill_u32((uint32_t *)nk, prm.k, 0, stream), "fill n_k");
            hip_check(launch_zero_rows(nwk, cs.split_words, cs.num_split_words, prm.k, stream), "zero rows");
        }
        a.slots = cs.phi_slots;
        a.num_slots = cs.num_phi_slots;
        a.chunk_begin = chunk;
        a.chunk_end = chunk + 1;
    }
    a.accumulate = accumulate ? 1 : 0;
    hip_check(launch_phi_rebuild(a, grid, stream), "phi rebuild launch");
    phi_ready = false;
}

void Context::phi_head() {
    if (!uploaded) fail(CULDA_ERR_STATE, "phi_head before upload");
    hip_check(launch_den(nk, den, prm.k, prm.beta, num_words, stream), "den launch");
    phi_ready = true;
}

void Context::rebuild_theta(int32_t chunk, hipStream_t s) {
    if (!uploaded) fail(CULDA_ERR_STATE, "rebuild_theta before upload");
    if (chunk < 0 || chunk >= (int32_t)chunks.size()) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
    ChunkState &cs = chunks[chunk];
    ThetaArgs a{cs.rev_off, cs.rev_idx, cs.z, cs.meta, cs.theta, stats, cs.doc_end - cs.doc_start, prm.k, count_bits};
    hip_check(launch_theta_rebuild(a, std::max(1, cu_count) * 8, s ? s : stream), "theta rebuild launch");
    theta_ready = true;
}

void Context::sample(int32_t chunk, int32_t sweep) {
    if (!uploaded) fail(CULDA_ERR_STATE, "sample before upload");
    if (!theta_ready || !phi_ready) fail(CULDA_ERR_STATE, "sample before the phi/theta rebuild");
    if (chunk < 0 || chunk >= (int32_t)chunks.size()) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
    if (sweep < 1) fail(CULDA_ERR_ARG, "sweep index must be >= 1 (0 keys the z initialisation)");
    ChunkState &cs = chunks[chunk];
    SampleArgs a{};
    a.slots = cs.slots;
    a.tokens = cs.tokens;
    a.z = cs.z;
    a.meta = cs.meta;
    a.theta = cs.theta;
    a.nwk = nwk;
    a.den = den;
    a.ll_partial = ll_partial + cs.ll_offset;
    a.seed = prm.seed;
    a.k_alpha = (double)prm.k * (double)prm.alpha;
    a.alpha = prm.alpha;
    a.beta = prm.beta;
    a.num_slots = cs.num_slots;
    a.k = prm.k;
    a.count_bits = count_bits;
    a.sweep = (uint32_t)sweep;
    hip_check(launch_sample(a, stream), "sampler launch");
    sampled = true;
    last_sweep = sweep;
}

double Context::loglik() {
    if (!sampled) fail(CULDA_ERR_STATE, "loglik before any sweep");
    hip_check(launch_ll_reduce(ll_partial, ll_partials, ll_scratch, ll_out, stream), "ll reduce launch");
    double h = 0;
    hip_check(hipMemcpyAsync(&h, ll_out, sizeof(double), hipMemcpyDeviceToHost, stream), "ll copy");
    hip_check(hipStreamSynchronize(stream), "ll sync");
    return h / (double)num_tokens;
}

void Context::read_stats() {
    unsigned long long h[2] = {0, 0};
    hip_check(hipMemcpyAsync(h, stats, sizeof h, hipMemcpyDeviceToHost, stream), "stats copy");
    hip_check(hipStreamSynchronize(stream), "stats sync");
    sum_doc_nnz = (int64_t)h[0];
    sum_token_nnz = (int64_t)h[1];
}

void Context::rebuild_all() {
    if (!uploaded) fail(CULDA_ERR_STATE, "rebuild before upload");
    if (prm.fused_phi || chunks.size() == 1) {
        rebuild_phi(-1, false);
    } else {
        for (int32_t k = 0; k < (int32_t)chunks.size(); ++k) rebuild_phi(k, k > 0);
    }
    phi_head();
    hip_check(launch_fill_u32((uint32_t *)stats, 4, 0, stream), "stats clear");
    for (int32_t k = 0; k < (int32_t)chunks.size(); ++k) rebuild_theta(k);
    read_stats();
}

void Context::iterate(int32_t sweep, double *ll, double *seconds) {
    if (!uploaded || !theta_ready || !phi_ready) fail(CULDA_ERR_STATE, "iterate before the initial rebuild");
    const int32_t C = (int32_t)chunks.size();
    const bool prof = prm.profile != 0;
    if (prof && !ev[0])
        for (hipEvent_t &e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    const int64_t token_nnz_used = sum_token_nnz;  // theta the sweep reads
    int64_t slots_total = 0;
    for (auto &cs : chunks) slots_total += cs.num_slots;
    auto t0 = std::chrono::steady_clock::now();
    // stream: sweeps of all chunks, LL, phi + head denominators.
    // stream2: theta of chunk c as soon as chunk c's sweep is done, so the
    // (gather-bound) theta rebuilds overlap the next chunks' sweeps.  Theta
    // of chunk c is read only by chunk c's sweep, which has finished.
    if (prof) hip_check(hipEventRecord(ev[0], stream), "event");
    hip_check(hipStreamWaitEvent(stream2, ev_theta_done, 0), "wait");  // previous stats read
    hip_check(launch_fill_u32((uint32_t *)stats, 4, 0, stream2), "stats clear");
    for (int32_t k = 0; k < C; ++k) {
        sample(k, sweep);
        hip_check(hipEventRecord(ev_sampled[k], stream), "event");
        hip_check(hipStreamWaitEvent(stream2, ev_sampled[k], 0), "wait");
        rebuild_theta(k, stream2);
    }
    if (prof) hip_check(hipEventRecord(ev[1], stream), "event");
    hip_check(launch_ll_reduce(ll_partial, ll_partials, ll_scratch, ll_out, stream), "ll reduce launch");
    if (prm.fused_phi || C == 1) {
        rebuild_phi(-1, false);
    } else {
        for (int32_t k = 0; k < C; ++k) rebuild_phi(k, k > 0);
    }
    phi_head();
    if (prof) hip_check(hipEventRecord(ev[2], stream), "event");
    hip_check(hipEventRecord(ev_theta_done, stream2), "event");
    hip_check(hipStreamWaitEvent(stream, ev_theta_done, 0), "wait");
    if (prof) hip_check(hipEventRecord(ev[3], stream), "event");
    double h_ll = 0;
    unsigned long long h_st[2] = {0, 0};
    hip_check(hipMemcpyAsync(&h_ll, ll_out, sizeof(double), hipMemcpyDeviceToHost, stream), "ll copy");
    hip_check(hipMemcpyAsync(h_st, stats, sizeof h_st, hipMemcpyDeviceToHost, stream), "stats copy");
    hip_check(hipStreamSynchronize(stream), "iteration sync");
    auto t1 = std::chrono::steady_clock::now();
    sum_doc_nnz = (int64_t)h_st[0];
    sum_token_nnz = (int64_t)h_st[1];
    if (prof) {
        float a = 0, b = 0, c = 0;
        hip_check(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
        hip_check(hipEventElapsedTime(&b, ev[1], ev[2]), "elapsed");
        hip_check(hipEventElapsedTime(&c, ev[2], ev[3]), "elapsed");
        sampler_ms = a;
        phi_ms = b;
        theta_ms = c;
    }
    sampler_launches = C;
    last_iter_token_nnz = token_nnz_used;
    last_sampler_bytes = (double)num_tokens * (4.0 + 16.0 + 2.0) + 4.0 * (double)token_nnz_used +
                         4.0 * (double)prm.k * (double)slots_total;
    if (ll) *ll = h_ll / (double)num_tokens;
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
}

void Context::sync() { hip_check(hipStreamSynchronize(stream), "sync"); }

}  // namespace culda
