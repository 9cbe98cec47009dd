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
    train.sync();
    hip_check(hipStreamSynchronize(stream2), "sync");
    hip_check(hipMemcpyAsync(m.nwk, tm.nwk, (size_t)m.V * m.K * 4, hipMemcpyDeviceToDevice, stream), "phi copy");
    hip_check(hipMemcpyAsync(m.nk, tm.nk, (size_t)m.K * 4, hipMemcpyDeviceToDevice, stream), "phi copy");
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
                  (double)m.beta, eval.num_slots, m.K, m.V, m.count_bits};
    hip_check(launch_heldout(a, stream), "held-out launch");
    std::vector<double> part(eval.num_slots);
    hip_check(hipMemcpyAsync(part.data(), eval.partial, part.size() * 8, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "held-out sync");
    read_stats();
    double sum = 0;
    for (double v : part) sum += v;
    phi_ready = true;
    theta_ready = true;
    return sum / (double)eval.num_tokens;
}

}  // namespace culda
