// mh_state.h -- device state of the learned-MH proposal of one corpus.
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "common.h"

namespace culda {

struct MHState {
    culda_mh_params prm{};
    int32_t K = 0, V = 0, D = 0;
    int64_t N = 0;
    std::vector<int64_t> unit_base;   // first global token of each unit
    // weights: FP32 masters, Adam moments, BF16 copies, gradients
    float *w1 = nullptr, *b1 = nullptr, *w2 = nullptr, *b2 = nullptr, *emb = nullptr;
    float *m_w1, *v_w1, *m_b1, *v_b1, *m_w2, *v_w2, *m_b2, *v_b2, *m_emb, *v_emb;
    uint16_t *w1h = nullptr, *w2h = nullptr, *embh = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr;  // BF16 transposes: [512][384], [K][512]
    float *g_w1, *g_b1, *g_w2, *g_b2;
    long long *g_emb = nullptr;  // fixed point, 2^-32
    int64_t adam_t = 0;
    // per token (global order) and per unit
    uint32_t *tok_doc = nullptr, *tok_word = nullptr;
    std::vector<int64_t *> gdelta;  // per unit, V entries
    // training snapshots
    uint16_t *zs = nullptr, *sks = nullptr;
    int32_t nsnap = 0;
    uint16_t *sk_cur = nullptr;  // D x 256
    // training buffers (batch rows)
    uint16_t *X = nullptr, *H = nullptr, *G = nullptr, *dP = nullptr;
    uint16_t *Xt = nullptr, *Ht = nullptr, *Gt = nullptr, *dPt = nullptr;  // feature-major copies
    float *P = nullptr, *L = nullptr, *dXe = nullptr, *loss = nullptr;
    int32_t *tgt = nullptr, *wrow = nullptr;
    uint32_t *q_doc = nullptr, *q_word = nullptr;  // explicit query rows
    // inference buffers
    uint16_t *Xi = nullptr, *Hi = nullptr;
    uint16_t *Li = nullptr;  // BF16 logits [infer_batch][K], as the proposal uses them
    unsigned long long *accepted = nullptr;
    float *part = nullptr;  // split-K and column-sum partials
};

class Engine;
// Forward + backward of the rows gathered into X / tgt / wrow (gradients
// into g_*, row losses added to *loss_fx in 2^-24 units), then one Adam
// step if update.
void mh_batch_step(Engine &e, MHState &s, int32_t rows, bool update, unsigned long long *loss_fx, hipStream_t st);
void mh_apply_adam(Engine &e, MHState &s, hipStream_t st);
int32_t mh_splits(int32_t m, int32_t n, int32_t k);

}  // namespace culda
