// mh_engine.cpp -- held-out evaluation (document completion) on the device.
//
// No counterpart in the reference, which reports only the training LL
// (kernel/lda_train_kernel.cu:256-258).  Held-out docs are split into
// observed and evaluated tokens (corpus.cpp heldout_split); the observed
// halves are folded in with phi fixed at the training counts, then every
// evaluated token scores log sum_k theta_dk phi_wk in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "kernels.h"
#include "mh_state.h"

namespace culda {

void Engine::set_eval(const Corpus &ev) {
    if (!uploaded || models.size() != 1 || units.size() != 1)
        fail(CULDA_ERR_STATE, "the held-out engine needs its observed corpus (one chunk) uploaded first");
    const ModelHost &m = models[0];
    if (ev.num_words != m.V || ev.num_docs() != m.num_docs || ev.chunks.size() != 1)
        fail(CULDA_ERR_ARG, "evaluated corpus must match the observed one (V %d, %d docs, one chunk)", m.V, m.num_docs);
    ev.validate();
    dfree(eval.tokens);
    dfree(eval.slots);
    dfree(eval.partial);
    const Chunk &ch = ev.chunks[0];
    std::vector<uint32_t> tok(ch.word_tokens.begin(), ch.word_tokens.end());
    std::vector<EvalSlot> slots;
    for (int32_t w = 0; w < ev.num_words; ++w)
        for (int64_t b = ch.word_idx[w]; b < ch.word_idx[w + 1]; b += 1024)
            slots.push_back(EvalSlot{w, 0, b, std::min<int64_t>(ch.word_idx[w + 1], b + 1024)});
    eval.tokens = (uint32_t *)dalloc(std::max<size_t>(tok.size(), 1) * 4);
    if (!tok.empty()) hip_check(hipMemcpy(eval.tokens, tok.data(), tok.size() * 4, hipMemcpyHostToDevice), "eval upload");
    eval.slots = (EvalSlot *)dalloc(std::max<size_t>(slots.size(), 1) * sizeof(EvalSlot));
    if (!slots.empty())
        hip_check(hipMemcpy(eval.slots, slots.data(), slots.size() * sizeof(EvalSlot), hipMemcpyHostToDevice),
                  "eval upload");
    eval.partial = (double *)dalloc(std::max<size_t>(slots.size(), 1) * sizeof(double));
    eval.num_slots = (int32_t)slots.size();
    eval.num_tokens = (int64_t)tok.size();
}

// Fold-in: n_wk / n_k copied from the training engine (phi fixed), z drawn
// from this engine's seed, `sweeps` sweeps of the plain sampler each
// followed by a theta rebuild (no phi rebuild); then the evaluation.
double Engine::heldout_ll(Engine &train, int32_t sweeps) {
    if (!eval.num_tokens) fail(CULDA_ERR_STATE, "no evaluated tokens (set them first)");
    if (sweeps < 0) fail(CULDA_ERR_ARG, "fold-in sweeps must be >= 0");
    if (train.models.size() != 1 || !train.uploaded) fail(CULDA_ERR_STATE, "training context not ready");
    const ModelHost &tm = train.models[0];
    ModelHost &m = models[0];
    if (tm.V != m.V || tm.K != m.K) fail(CULDA_ERR_ARG, "held-out and training contexts differ in V or K");
    if (!train.phi_ready) fail(CULDA_ERR_STATE, "training counts not built");
    train.fresh_rows(train.stream);
    train.sync();
    hip_check(hipStreamSynchronize(stream2), "sync");
    hip_check(hipMemcpyAsync(m.nwk, tm.nwk, (size_t)m.V * m.K * 4, hipMemcpyDeviceToDevice, stream), "phi copy");
    hip_check(hipMemcpyAsync(m.nk, tm.nk, (size_t)m.K * 4, hipMemcpyDeviceToDevice, stream), "phi copy");
    zp_valid = false;  // these counts are the training context's
    rare_stale = false;
    phi_head(stream);
    init_z();
    rebuild_theta_all(stream);
    for (int32_t t = 1; t <= sweeps; ++t) {
        set_sweep((uint32_t)t, stream);
        for (size_t u = 0; u < units.size(); ++u) sample_unit((int32_t)u, stream);
        rebuild_theta_all(stream);
    }
    const UnitHost &u = units[0];
    HeldoutArgs a{eval.slots, eval.tokens, u.meta, u.theta, m.nwk, m.nk, eval.partial, (double)m.alpha,
                  (double)m.beta, eval.num_slots, m.K, m.V, 0};
    hip_check(launch_heldout(a, stream), "held-out launch");
    std::vector<double> part(eval.num_slots);
    hip_check(hipMemcpyAsync(part.data(), eval.partial, part.size() * 8, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "held-out sync");
    sync();
    double sum = 0;
    for (double v : part) sum += v;
    phi_ready = true;
    set_theta_ready(true);
    return sum / (double)eval.num_tokens;
}

}  // namespace culda

// ====================================================== learned MH proposal
namespace culda {


MHState &Engine::mh_state() {
    if (!mh) fail(CULDA_ERR_STATE, "learned MH not enabled (culda_mh_enable)");
    return *mh;
}

static uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void fill_uniform(std::vector<float> &v, float bound, uint64_t &s) {
    for (float &x : v) x = (float)(((double)(splitmix64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * bound);
}

static int64_t round128(int64_t n) { return (n + 127) / 128 * 128; }

// Split-K factor of an m x n x k weight-gradient product: about 512
// workgroups (128 x 128 tiles when the shape allows, else 64 x 64), each
// split a multiple of kGemmK and at least 4 stages deep.
int32_t mh_splits(int32_t m, int32_t n, int32_t k) {
    const int32_t tile = (m % 128 == 0 && n % 128 == 0) ? 128 : kGemmTile;
    const int32_t tiles = (m / tile) * (n / tile);
    int32_t s = 1;
    while (s * 2 * tiles <= 512 && k % (s * 2 * kGemmK) == 0 && k / (s * 2) >= 4 * kGemmK) s *= 2;
    return s;
}

void Engine::mh_enable(const culda_mh_params &p_in) {
    if (!uploaded || models.size() != 1) fail(CULDA_ERR_STATE, "learned MH needs one uploaded corpus");
    const ModelHost &m = models[0];
    if (m.K % 64) fail(CULDA_ERR_ARG, "learned MH needs k %% 64 == 0 (k = %d)", m.K);
    for (int64_t nd : m.doc_len)
        if (theta_entries_max(nd, m.K) > kMHThetaEntries)
            fail(CULDA_ERR_LIMIT, "learned MH holds theta rows of at most %d entries; a doc of %lld tokens may need more",
                 kMHThetaEntries, (long long)nd);
    culda_mh_params p = p_in;
    if (!p.batch) p.batch = 8192;
    if (!p.infer_batch) p.infer_batch = 65536;
    if (!p.max_snapshots) p.max_snapshots = 21;
    if (!(p.lr > 0)) p.lr = 1e-3f;
    if (!(p.beta1 > 0)) p.beta1 = 0.9f;
    if (!(p.beta2 > 0)) p.beta2 = 0.999f;
    if (!(p.eps > 0)) p.eps = 1e-8f;
    if (p.batch < 128 || p.batch % 128 || p.infer_batch < 128 || p.infer_batch % 128 || p.max_snapshots < 1)
        fail(CULDA_ERR_ARG, "batch and infer_batch must be positive multiples of 128, max_snapshots >= 1");
    if ((int64_t)p.infer_batch / 64 > 65535) fail(CULDA_ERR_ARG, "infer_batch too large");
    mh.reset(new MHState);
    MHState &s = *mh;
    s.prm = p;
    s.K = m.K;
    s.V = m.V;
    s.D = m.num_docs;
    s.N = m.num_tokens;
    if ((uint64_t)s.N * (uint64_t)p.max_snapshots >= (1ULL << 32))
        fail(CULDA_ERR_LIMIT, "training triples (tokens x snapshots) must stay below 2^32");
    const int64_t K = s.K, V = s.V;
    auto falloc = [&](int64_t n) { return (float *)dalloc(std::max<int64_t>(n, 1) * 4); };
    auto hzero = [&](void *ptr, int64_t bytes) {
        if (bytes > 0) hip_check(hipMemset(ptr, 0, bytes), "clear");
    };
    // weights (torch Linear-style uniform init, embeddings U(-0.1, 0.1))
    uint64_t seed = p.init_seed ^ 0x6d68u;
    std::vector<float> hw1((size_t)kIn * kHidden), hb1(kHidden), hw2((size_t)kHidden * K), hb2(K),
        he((size_t)V * kEmb);
    fill_uniform(hw1, 1.0f / std::sqrt((float)kIn), seed);
    fill_uniform(hb1, 1.0f / std::sqrt((float)kIn), seed);
    fill_uniform(hw2, 1.0f / std::sqrt((float)kHidden), seed);
    fill_uniform(hb2, 1.0f / std::sqrt((float)kHidden), seed);
    fill_uniform(he, 0.1f, seed);
    s.w1 = falloc(hw1.size());
    s.b1 = falloc(hb1.size());
    s.w2 = falloc(hw2.size());
    s.b2 = falloc(hb2.size());
    s.emb = falloc(he.size());
    float **moments[] = {&s.m_w1, &s.v_w1, &s.m_b1, &s.v_b1, &s.m_w2, &s.v_w2, &s.m_b2, &s.v_b2, &s.m_emb, &s.v_emb};
    const int64_t sizes[] = {kIn * kHidden, kHidden, kHidden * K, K, V * kEmb};
    for (int i = 0; i < 10; ++i) {
        *moments[i] = falloc(sizes[i / 2]);
        hzero(*moments[i], sizes[i / 2] * 4);
    }
    s.w1h = (uint16_t *)dalloc(hw1.size() * 2);
    s.w2h = (uint16_t *)dalloc(hw2.size() * 2);
    s.w1t = (uint16_t *)dalloc(hw1.size() * 2);
    s.w2t = (uint16_t *)dalloc(hw2.size() * 2);
    s.embh = (uint16_t *)dalloc(he.size() * 2);
    s.g_w1 = falloc(hw1.size());
    s.g_b1 = falloc(kHidden);
    s.g_w2 = falloc(hw2.size());
    s.g_b2 = falloc(K);
    s.g_emb = (long long *)dalloc(he.size() * 8);
    hzero(s.g_emb, he.size() * 8);
    // per-token doc / word (global token order = units in order)
    s.tok_doc = (uint32_t *)dalloc(std::max<int64_t>(s.N, 1) * 4);
    s.tok_word = (uint32_t *)dalloc(std::max<int64_t>(s.N, 1) * 4);
    {
        std::vector<int64_t> pre(V + 1, 0);
        for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui)
            for (int32_t w = 0; w < V; ++w) pre[w + 1] += units[ui].word_idx_h[w + 1] - units[ui].word_idx_h[w];
        for (int32_t w = 0; w < V; ++w) pre[w + 1] += pre[w];
        std::vector<int64_t> run(pre.begin(), pre.end() - 1);
        int64_t base = 0;
        for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
            const UnitHost &u = units[ui];
            std::vector<uint32_t> loc(u.num_tokens), doc(u.num_tokens), word(u.num_tokens);
            if (u.num_tokens)
                hip_check(hipMemcpy(loc.data(), u.tokens, u.num_tokens * 4, hipMemcpyDeviceToHost), "tokens copy");
            std::vector<int64_t> gd(V);
            for (int32_t w = 0; w < V; ++w) {
                gd[w] = run[w] - u.word_idx_h[w];
                for (int64_t i = u.word_idx_h[w]; i < u.word_idx_h[w + 1]; ++i) {
                    word[i] = (uint32_t)w;
                    doc[i] = loc[i] + (uint32_t)u.doc_start;
                }
                run[w] += u.word_idx_h[w + 1] - u.word_idx_h[w];
            }
            if (u.num_tokens) {
                hip_check(hipMemcpy(s.tok_doc + base, doc.data(), u.num_tokens * 4, hipMemcpyHostToDevice), "copy");
                hip_check(hipMemcpy(s.tok_word + base, word.data(), u.num_tokens * 4, hipMemcpyHostToDevice), "copy");
            }
            int64_t *dg = (int64_t *)dalloc(V * 8);
            hip_check(hipMemcpy(dg, gd.data(), V * 8, hipMemcpyHostToDevice), "copy");
            s.gdelta.push_back(dg);
            s.unit_base.push_back(base);
            base += u.num_tokens;
        }
    }
    s.zs = (uint16_t *)dalloc(std::max<int64_t>(s.N, 1) * 2 * p.max_snapshots);
    s.sks = (uint16_t *)dalloc((int64_t)std::max(s.D, 1) * kSketch * 2 * p.max_snapshots);
    s.sk_cur = (uint16_t *)dalloc((int64_t)std::max(s.D, 1) * kSketch * 2);
    const int64_t B = p.batch;
    s.X = (uint16_t *)dalloc(B * kIn * 2);
    s.H = (uint16_t *)dalloc(B * kHidden * 2);
    s.G = (uint16_t *)dalloc(B * K * 2);
    s.dP = (uint16_t *)dalloc(B * kHidden * 2);
    s.Xt = (uint16_t *)dalloc(B * kIn * 2);
    s.Ht = (uint16_t *)dalloc(B * kHidden * 2);
    s.Gt = (uint16_t *)dalloc(B * K * 2);
    s.dPt = (uint16_t *)dalloc(B * kHidden * 2);
    s.P = falloc(B * kHidden);
    s.L = falloc(B * K);
    s.dXe = falloc(B * kEmb);
    s.loss = falloc(B);
    s.tgt = (int32_t *)dalloc(B * 4);
    s.wrow = (int32_t *)dalloc(B * 4);
    s.q_doc = (uint32_t *)dalloc(std::max<int64_t>(B, p.infer_batch) * 4);
    s.q_word = (uint32_t *)dalloc(std::max<int64_t>(B, p.infer_batch) * 4);
    const int64_t IB = p.infer_batch;
    s.Xi = (uint16_t *)dalloc(IB * kIn * 2);
    s.Hi = (uint16_t *)dalloc(IB * kHidden * 2);
    s.Li = (uint16_t *)dalloc(IB * K * 2);
    s.accepted = (unsigned long long *)dalloc(8);
    {
        const int64_t w2 = (int64_t)mh_splits(kHidden, (int32_t)K, (int32_t)B) * kHidden * K;
        const int64_t w1 = (int64_t)mh_splits(kIn, kHidden, (int32_t)B) * kIn * kHidden;
        const int64_t cs = (int64_t)colsum_parts((int32_t)B) * std::max<int64_t>(K, kHidden);
        s.part = falloc(std::max({w2, w1, cs}));
    }
    hip_check(hipDeviceSynchronize(), "mh buffers");
    // upload the initial weights (and their BF16 copies)
    hip_check(hipMemcpy(s.w1, hw1.data(), hw1.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(s.b1, hb1.data(), hb1.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(s.w2, hw2.data(), hw2.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(s.b2, hb2.data(), hb2.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(s.emb, he.data(), he.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(launch_to_bf16(s.w1, s.w1h, s.w1t, kIn, hw1.size(), stream), "bf16");
    hip_check(launch_to_bf16(s.w2, s.w2h, s.w2t, kHidden, hw2.size(), stream), "bf16");
    hip_check(launch_to_bf16(s.emb, s.embh, nullptr, 1, he.size(), stream), "bf16");
    hip_check(hipStreamSynchronize(stream), "mh init");
}

void Engine::mh_sketch_current(hipStream_t st) {
    MHState &s = mh_state();
    const ModelHost &m = models[0];
    for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
        const UnitHost &u = units[ui];
        hip_check(launch_sketch(u.meta, u.theta, u.doc_end - u.doc_start, s.sk_cur + (int64_t)u.doc_start * kSketch,
                                st),
                  "sketch");
    }
}

int32_t Engine::mh_collect() {
    MHState &s = mh_state();
    if (!theta_ready()) fail(CULDA_ERR_STATE, "collect needs rebuilt counts");
    if (s.nsnap >= s.prm.max_snapshots) fail(CULDA_ERR_LIMIT, "all %d snapshots taken", s.prm.max_snapshots);
    hip_check(hipStreamSynchronize(stream2), "sync");
    const ModelHost &m = models[0];
    for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
        const UnitHost &u = units[ui];
        if (u.num_tokens)
            hip_check(hipMemcpyAsync(s.zs + (int64_t)s.nsnap * s.N + s.unit_base[ui - m.unit_begin], u.z,
                                     u.num_tokens * 2, hipMemcpyDeviceToDevice, stream),
                      "z snapshot");
        hip_check(launch_sketch(u.meta, u.theta, u.doc_end - u.doc_start,
                                s.sks + ((int64_t)s.nsnap * s.D + u.doc_start) * kSketch, stream),
                  "sketch");
    }
    hip_check(hipStreamSynchronize(stream), "collect");
    return ++s.nsnap;
}

// Forward + backward of the rows already gathered into s.X / s.tgt /
// s.wrow (rows_pad = batch rows), gradients into s.g_*; per-row loss sum
// into *loss_fx (device, 2^-24 units).  Then optionally one Adam step.
void mh_batch_step(Engine &e, MHState &s, int32_t rows, bool update, unsigned long long *loss_fx, hipStream_t st) {
    const int32_t B = s.prm.batch, K = s.K;
    auto gemm = [&](const GemmArgs &g, const char *what) { hip_check(launch_gemm_bf16(g, st), what); };
    // every product has both operands k-contiguous (A [m][k], B [n][k]):
    // weights through their transposed copies, weight gradients through
    // feature-major copies of the activations
    GemmArgs f1{s.X, s.w1t, s.P, s.H, s.b1, nullptr, kIn, kIn, kHidden, kHidden, 0, B, kHidden, kIn, kEpiBiasGelu};
    gemm(f1, "gemm x.w1");
    GemmArgs f2{s.H, s.w2t, s.L, nullptr, s.b2, nullptr, kHidden, kHidden, K, 0, 0, B, K, kHidden, kEpiStore};
    gemm(f2, "gemm h.w2");
    hip_check(launch_softmax_ce(s.L, K, s.tgt, 1.0f / (float)std::max(rows, 1), s.G, s.loss, loss_fx, B, st),
              "softmax");
    hip_check(launch_transpose_bf16(s.H, B, kHidden, kHidden, s.Ht, B, st), "transpose");
    hip_check(launch_transpose_bf16(s.G, B, K, K, s.Gt, B, st), "transpose");
    hip_check(launch_transpose_bf16(s.X, B, kIn, kIn, s.Xt, B, st), "transpose");
    GemmArgs gw2{s.Ht, s.Gt, s.g_w2, nullptr, nullptr, nullptr, B, B, K, 0, 0, kHidden, K, B, kEpiStore};
    hip_check(launch_gemm_bf16_splitk(gw2, mh_splits(kHidden, K, B), s.part, st), "gemm h^T.g");
    hip_check(launch_colsum_bf16(s.G, K, B, K, s.part, s.g_b2, st), "colsum");
    GemmArgs dp{s.G, s.w2h, nullptr, s.dP, nullptr, s.P, K, K, 0, kHidden, kHidden, B, kHidden, K, kEpiGeluGrad};
    gemm(dp, "gemm g.w2^T");
    hip_check(launch_colsum_bf16(s.dP, kHidden, B, kHidden, s.part, s.g_b1, st), "colsum");
    hip_check(launch_transpose_bf16(s.dP, B, kHidden, kHidden, s.dPt, B, st), "transpose");
    GemmArgs gw1{s.Xt, s.dPt, s.g_w1, nullptr, nullptr, nullptr, B, B, kHidden, 0, 0, kIn, kHidden, B, kEpiStore};
    hip_check(launch_gemm_bf16_splitk(gw1, mh_splits(kIn, kHidden, B), s.part, st), "gemm x^T.dp");
    GemmArgs dx{s.dP, s.w1h + (int64_t)kSketch * kHidden, s.dXe, nullptr, nullptr, nullptr, kHidden, kHidden, kEmb,
                0, 0, B, kEmb, kHidden, kEpiStore};
    gemm(dx, "gemm dp.w1e^T");
    hip_check(launch_emb_grad(s.dXe, s.wrow, B, s.g_emb, st), "emb grad");
    if (update) mh_apply_adam(e, s, st);
}

// One Adam step on every weight from the gradients in s.g_* (the
// embedding accumulator is cleared by the update).
void mh_apply_adam(Engine &e, MHState &s, hipStream_t st) {
    const int32_t K = s.K;
    ++s.adam_t;
    const double bc1 = 1.0 - std::pow((double)s.prm.beta1, (double)s.adam_t);
    const double bc2 = 1.0 - std::pow((double)s.prm.beta2, (double)s.adam_t);
    const float step = (float)(s.prm.lr / bc1), sq = (float)std::sqrt(bc2);
    const float b1 = s.prm.beta1, b2 = s.prm.beta2, eps = s.prm.eps;
    hip_check(launch_adam(s.w1, s.m_w1, s.v_w1, s.g_w1, nullptr, s.w1h, s.w1t, kIn, (int64_t)kIn * kHidden, step, b1,
                          b2, eps, sq, st),
              "adam");
    hip_check(launch_adam(s.b1, s.m_b1, s.v_b1, s.g_b1, nullptr, nullptr, nullptr, 1, kHidden, step, b1, b2, eps, sq, st),
              "adam");
    hip_check(launch_adam(s.w2, s.m_w2, s.v_w2, s.g_w2, nullptr, s.w2h, s.w2t, kHidden, (int64_t)kHidden * K, step, b1,
                          b2, eps, sq, st),
              "adam");
    hip_check(launch_adam(s.b2, s.m_b2, s.v_b2, s.g_b2, nullptr, nullptr, nullptr, 1, K, step, b1, b2, eps, sq, st),
              "adam");
    hip_check(launch_adam(s.emb, s.m_emb, s.v_emb, nullptr, s.g_emb, s.embh, nullptr, 1, (int64_t)s.V * kEmb, step, b1,
                          b2, eps, sq, st),
              "adam");
    (void)e;
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

double Engine::mh_train(int32_t epochs) {
    MHState &s = mh_state();
    if (epochs < 1) fail(CULDA_ERR_ARG, "epochs must be >= 1");
    if (!s.nsnap) fail(CULDA_ERR_STATE, "no training snapshots (culda_mh_collect)");
    const uint64_t T = (uint64_t)s.nsnap * (uint64_t)s.N;
    const int64_t B = s.prm.batch;
    const int64_t steps = (int64_t)((T + B - 1) / B);
    unsigned long long *loss_fx = (unsigned long long *)dalloc(8);
    uint64_t rs = s.prm.init_seed ^ 0x747261696eULL;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        uint64_t mul = T > 1 ? splitmix64(rs) % (T - 1) + 1 : 1;
        while (gcd64(mul, T) != 1) mul = mul % (T - 1) + 1;
        const uint64_t add = T > 1 ? splitmix64(rs) % T : 0;
        hip_check(launch_fill_u32((uint32_t *)loss_fx, 2, 0, stream), "clear");  // loss of the last epoch
        for (int64_t st = 0; st < steps; ++st) {
            const int64_t p0 = st * B;
            const int32_t rows = (int32_t)std::min<int64_t>(B, (int64_t)T - p0);
            hip_check(launch_gather_train(s.sks, s.zs, s.tok_doc, s.tok_word, s.embh, mul, add, T, (uint64_t)s.N, s.D, p0,
                                          rows, (int32_t)B, s.X, s.tgt, s.wrow, stream),
                      "gather");
            mh_batch_step(*this, s, rows, true, loss_fx, stream);
        }
    }
    unsigned long long fx = 0;
    hip_check(hipMemcpyAsync(&fx, loss_fx, 8, hipMemcpyDeviceToHost, stream), "loss copy");
    hip_check(hipStreamSynchronize(stream), "train");
    dfree(loss_fx);
    return (double)(long long)fx / 16777216.0 / (double)T;
}

// BF16 logits [n][K] (rows padded to 128 inside) for explicit (global doc,
// word) queries against s.sk_cur: the proposal needs no more precision,
// and the product writes (and the MH step reads) half the bytes of FP32.
void Engine::mh_query_logits(const uint32_t *d_docs, const uint32_t *d_words, int64_t n, uint16_t *d_logits,
                             hipStream_t st) {
    MHState &s = mh_state();
    const int64_t np = round128(n);
    hip_check(launch_gather_query(s.sk_cur, d_docs, d_words, s.embh, n, np, s.Xi, st), "gather");
    GemmArgs f1{s.Xi, s.w1t, nullptr, s.Hi, s.b1, nullptr, kIn, kIn, 0, kHidden, 0, (int32_t)np, kHidden, kIn,
                kEpiBiasGelu};
    hip_check(launch_gemm_bf16(f1, st), "gemm x.w1");
    GemmArgs f2{s.Hi, s.w2t, nullptr, d_logits, s.b2, nullptr, kHidden, kHidden, 0, s.K, 0, (int32_t)np, s.K, kHidden,
                kEpiStore};
    hip_check(launch_gemm_bf16(f2, st), "gemm h.w2");
}

double Engine::mh_iterate(uint32_t sweep, int32_t proposals) {
    MHState &s = mh_state();
    if (!theta_ready() || !phi_ready) fail(CULDA_ERR_STATE, "MH iteration needs rebuilt counts");
    if (proposals < 0 || proposals > kMaxProposals) fail(CULDA_ERR_ARG, "proposals must be in [0, %d]", kMaxProposals);
    hip_check(hipStreamSynchronize(stream2), "sync");
    const ModelHost &m = models[0];
    hip_check(launch_fill_u32((uint32_t *)s.accepted, 2, 0, stream), "clear");
    fresh_rows(stream);  // the MH step reads n_wk
    mh_sketch_current(stream);
    const int64_t IB = s.prm.infer_batch;
    for (int32_t ui = m.unit_begin; ui < m.unit_end; ++ui) {
        const UnitHost &u = units[ui];
        const int64_t base = s.unit_base[ui - m.unit_begin];
        for (int64_t start = 0; start < u.num_tokens; start += IB) {
            const int64_t n = std::min(IB, u.num_tokens - start);
            mh_query_logits(s.tok_doc + base + start, s.tok_word + base + start, n, s.Li, stream);
            MHArgs a{};
            a.logits = s.Li;
            a.word = s.tok_word + base;
            a.gdelta = s.gdelta[ui - m.unit_begin];
            a.unit = UnitDev{u.tokens, u.z, u.zp, u.word_idx, u.rev_off, u.rev_idx, u.meta, u.theta, u.model,
                             u.doc_end - u.doc_start};
            a.model = ModelDev{m.nwk, m.nk, m.den, m.zero_phi, m.rden, m.seed, (double)m.K * (double)m.alpha, m.alpha, m.beta, m.K,
                               m.V};
            a.accepted = s.accepted;
            a.start = start;
            a.n = (int32_t)n;
            a.proposals = proposals;
            a.sweep = sweep;
            hip_check(launch_mh(a, stream), "mh launch");
        }
    }
    // the count updates of a plain iteration
    update_phi(stream);
    phi_head(stream);
    rebuild_theta_all(stream);
    unsigned long long acc = 0;
    hip_check(hipMemcpyAsync(&acc, s.accepted, 8, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "mh iteration");
    sync();
    sampled = true;
    const double tried = (double)s.N * (double)proposals;
    return tried > 0 ? (double)acc / tried : 0.0;
}

}  // namespace culda
