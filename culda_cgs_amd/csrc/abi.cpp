// abi.cpp -- extern "C" entry points over culda::Engine: the single-corpus
// context (culda_ctx), the multi-corpus batch (culda_batch), the host side
// of the ModelPhi / ModelTheta save hooks and checkpoints.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "mh_state.h"
#include "kernels.h"

using namespace culda;

struct culda_ctx {
    Engine e;
    float alpha = 0, beta = 0;
    uint64_t seed = 0;
    uint32_t last_sweep = 0;
};

struct culda_batch {
    Engine e;
};

const culda::Corpus &corpus_ref(const culda_corpus *p);

namespace {

Engine &eng(culda_ctx *p) {
    if (!p) fail(CULDA_ERR_ARG, "null context");
    hip_check(hipSetDevice(p->e.prm.device), "hipSetDevice");
    return p->e;
}

Engine &eng(culda_batch *p) {
    if (!p) fail(CULDA_ERR_ARG, "null batch");
    hip_check(hipSetDevice(p->e.prm.device), "hipSetDevice");
    return p->e;
}

void check_unit(Engine &e, int32_t unit) {
    if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
    if (unit < 0 || unit >= (int32_t)e.units.size()) fail(CULDA_ERR_ARG, "chunk %d out of range", unit);
}

void check_model(Engine &e, int32_t m) {
    if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
    if (m < 0 || m >= (int32_t)e.models.size()) fail(CULDA_ERR_ARG, "model %d out of range", m);
}

// Host mirror of one model's counts (ModelPhi::hostPhiTopicWordShort /
// hostPhiTopic, model/model_phi.h:32-33).
struct HostPhi {
    std::vector<int32_t> nwk, nk;
};

HostPhi download_phi(Engine &e, int32_t m) {
    check_model(e, m);
    const ModelHost &mh = e.models[m];
    HostPhi h;
    h.nwk.resize((size_t)mh.V * mh.K);
    h.nk.resize(mh.K);
    e.fresh_rows(e.stream);
    e.sync();
    if (!h.nwk.empty()) hip_check(hipMemcpy(h.nwk.data(), mh.nwk, h.nwk.size() * 4, hipMemcpyDeviceToHost), "phi copy");
    hip_check(hipMemcpy(h.nk.data(), mh.nk, h.nk.size() * 4, hipMemcpyDeviceToHost), "n_k copy");
    return h;
}

// Compact CSR of one unit's theta rows.
// Theta rows as (topic, n_dk) pairs: the device's entries of one topic
// (several when n_dk > kThetaMaxCount) are merged.
void download_theta(Engine &e, int32_t unit, std::vector<int64_t> &off, std::vector<int32_t> &cnt,
                    std::vector<int16_t> &top) {
    check_unit(e, unit);
    UnitHost &u = e.units[unit];
    const int32_t Dc = u.doc_end - u.doc_start;
    std::vector<DocMeta> meta(Dc);
    std::vector<uint16_t> th(u.theta_cap);
    e.sync();
    if (Dc) hip_check(hipMemcpy(meta.data(), u.meta, Dc * sizeof(DocMeta), hipMemcpyDeviceToHost), "meta copy");
    if (u.theta_cap) hip_check(hipMemcpy(th.data(), u.theta, u.theta_cap * 2, hipMemcpyDeviceToHost), "theta copy");
    off.assign(Dc + 1, 0);
    cnt.clear();
    top.clear();
    for (int32_t d = 0; d < Dc; ++d) {
        for (uint32_t j = 0; j < meta[d].nnz; ++j) {
            const uint32_t v = th[meta[d].start + j];
            const int16_t k = (int16_t)(v & kThetaTopicMask);
            const int32_t c = (int32_t)(v >> kThetaCountShift);
            if (j > 0 && top.back() == k)
                cnt.back() += c;
            else {
                top.push_back(k);
                cnt.push_back(c);
            }
        }
        off[d + 1] = (int64_t)top.size();
    }
}

FILE *open_or_fail(const std::string &path, const char *mode) {
    FILE *f = fopen(path.c_str(), mode);
    if (!f) fail(CULDA_ERR_IO, "%s open failed", path.c_str());
    return f;
}

void copy_z(Engine &e, int32_t unit, uint16_t *z, int64_t n, bool to_device) {
    check_unit(e, unit);
    UnitHost &u = e.units[unit];
    if (n != u.num_tokens)
        fail(CULDA_ERR_ARG, "chunk %d has %lld tokens, got %lld", unit, (long long)u.num_tokens, (long long)n);
    if (n && !z) fail(CULDA_ERR_ARG, "null z");
    e.sync();
    if (!n) return;
    if (to_device) {
        const uint32_t K = (uint32_t)e.prm.k;
        for (int64_t i = 0; i < n; ++i)
            if (z[i] >= K) fail(CULDA_ERR_ARG, "z[%lld] = %u is not below k", (long long)i, z[i]);
        hip_check(hipMemcpy(u.z, z, n * 2, hipMemcpyHostToDevice), "z copy");
        e.theta_ok[unit] = 0;  // this chunk's rows no longer match its z
        e.phi_ready = false;
    } else {
        hip_check(hipMemcpy(z, u.z, n * 2, hipMemcpyDeviceToHost), "z copy");
    }
}

void fill_tune(const TuneResult &r, culda_tune *out) {
    std::memset(out, 0, sizeof *out);
    out->slot_tokens = r.slot_tokens;
    out->theta_groups = r.theta_groups;
    for (size_t i = 0; i < r.theta_tried.size() && i < 8; ++i) out->theta_tried_ms[i] = r.theta_tried[i].second;
    out->best_ms = r.best_ms;
    out->num_tried = (int32_t)std::min<size_t>(r.tried.size(), 32);
    for (int32_t i = 0; i < out->num_tried; ++i) {
        out->tried_slot_tokens[i] = r.tried[i].first;
        out->tried_ms[i] = r.tried[i].second;
    }
}

const std::vector<int32_t> kSlotCandidates = {128, 256, 512, 1024, 2048, 4096};

}  // namespace

// Every call first waits for the theta rebuilds a culda_iterate may have
// left running on the second stream (CTX_CALL_LIVE: calls that read only
// host-side values, or order their launches after those rebuilds).
#define CTX_CALL(...)         \
    return guarded([&] {      \
        Engine &e = eng(ctx); \
        e.settle();           \
        __VA_ARGS__;          \
        return CULDA_OK;      \
    })
#define CTX_CALL_LIVE(...)    \
    return guarded([&] {      \
        Engine &e = eng(ctx); \
        (void)e;              \
        __VA_ARGS__;          \
        return CULDA_OK;      \
    })

// ====================================================== single corpus

extern "C" int culda_create(const culda_params *params, culda_ctx **out) {
    return guarded([&] {
        if (!params || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        const culda_params &p = *params;
        if (!(p.alpha > 0.0f) || !std::isfinite(p.alpha)) fail(CULDA_ERR_ARG, "alpha must be positive and finite");
        if (!(p.beta > 0.0f) || !std::isfinite(p.beta)) fail(CULDA_ERR_ARG, "beta must be positive and finite");
        auto *h = new culda_ctx;
        try {
            h->e.init_device(p);
        } catch (...) {
            delete h;
            throw;
        }
        h->alpha = p.alpha;
        h->beta = p.beta;
        h->seed = p.seed;
        *out = h;
        return CULDA_OK;
    });
}

extern "C" void culda_destroy(culda_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->e.prm.device);
    delete ctx;
}

extern "C" int culda_check_corpus(const culda_corpus *corpus, int32_t k) {
    return guarded([&] {
        if (!corpus) fail(CULDA_ERR_ARG, "null corpus");
        check_corpus_limits(corpus_ref(corpus), k);
        return CULDA_OK;
    });
}

extern "C" int culda_upload(culda_ctx *ctx, const culda_corpus *corpus) {
    CTX_CALL({
        if (!corpus) fail(CULDA_ERR_ARG, "null corpus");
        if (e.uploaded) fail(CULDA_ERR_STATE, "a corpus is already uploaded to this context");
        e.add_corpus(corpus_ref(corpus), ctx->alpha, ctx->beta, ctx->seed);
        e.upload();
    });
}

extern "C" int culda_init_z(culda_ctx *ctx) { CTX_CALL(e.init_z()); }

extern "C" int culda_set_z(culda_ctx *ctx, int32_t chunk, const uint16_t *z, int64_t n) {
    CTX_CALL(copy_z(e, chunk, const_cast<uint16_t *>(z), n, true));
}

extern "C" int culda_get_z(culda_ctx *ctx, int32_t chunk, uint16_t *z, int64_t n) {
    CTX_CALL(copy_z(e, chunk, z, n, false));
}

extern "C" int culda_rebuild_phi(culda_ctx *ctx, int32_t chunk, int32_t accumulate) {
    CTX_CALL({
        if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        if (chunk < -1 || chunk >= (int32_t)e.units.size()) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
        hip_check(hipStreamSynchronize(e.stream2), "sync");
        e.rebuild_phi_model(0, chunk, accumulate != 0, e.stream);
    });
}

extern "C" int culda_phi_head(culda_ctx *ctx) {
    CTX_CALL({
        if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        e.phi_head(e.stream);
    });
}

extern "C" int culda_rebuild_theta(culda_ctx *ctx, int32_t chunk) {
    CTX_CALL({
        check_unit(e, chunk);
        e.rebuild_theta_unit(chunk, e.stream);
    });
}

extern "C" int culda_sample(culda_ctx *ctx, int32_t chunk, int32_t iteration) {
    CTX_CALL({
        check_unit(e, chunk);
        if (!e.phi_ready) fail(CULDA_ERR_STATE, "sample before the phi rebuild and head (culda_rebuild_phi, culda_phi_head)");
        if (!e.theta_ok[chunk]) fail(CULDA_ERR_STATE, "sample of chunk %d before its theta rebuild (culda_rebuild_theta)", chunk);
        if (iteration < 1) fail(CULDA_ERR_ARG, "sweep index must be >= 1 (0 keys the z initialisation)");
        hip_check(hipStreamSynchronize(e.stream2), "sync");
        e.set_sweep((uint32_t)iteration, e.stream);
        e.sample_unit(chunk, e.stream);
        e.sampled = true;
        ctx->last_sweep = (uint32_t)iteration;
    });
}

extern "C" int culda_loglik(culda_ctx *ctx, double *ll_per_token) {
    CTX_CALL({
        if (!ll_per_token) fail(CULDA_ERR_ARG, "null output");
        if (!e.sampled) fail(CULDA_ERR_STATE, "loglik before any sweep");
        e.set_sweep(ctx->last_sweep, e.stream);
        e.ll_reduce(e.stream);
        hip_check(hipStreamSynchronize(e.stream), "sync");
        *ll_per_token = e.model_ll(0, ctx->last_sweep);
    });
}

extern "C" int culda_iterate(culda_ctx *ctx, int32_t iteration, double *ll, double *seconds) {
    CTX_CALL_LIVE({
        if (iteration < 1) fail(CULDA_ERR_ARG, "iteration must be >= 1");
        e.iterate_units((uint32_t)iteration, ll, seconds);
        ctx->last_sweep = (uint32_t)iteration;
    });
}

extern "C" int culda_iterate_n(culda_ctx *ctx, int32_t first_iteration, int32_t count, double *ll, double *seconds) {
    CTX_CALL_LIVE({
        if (first_iteration < 1) fail(CULDA_ERR_ARG, "iteration must be >= 1");
        e.iterate_units_n((uint32_t)first_iteration, count, ll, seconds);
        ctx->last_sweep = (uint32_t)(first_iteration + count - 1);
    });
}

extern "C" int culda_rebuild_all(culda_ctx *ctx) { CTX_CALL(e.rebuild_all()); }

extern "C" int culda_sync(culda_ctx *ctx) { CTX_CALL(e.sync()); }

extern "C" int culda_autotune(culda_ctx *ctx, culda_tune *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        fill_tune(e.autotune(kSlotCandidates, false), out);
    });
}

extern "C" int culda_loglik_joint(culda_ctx *ctx, double *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        if (!e.theta_ready()) fail(CULDA_ERR_STATE, "joint log-likelihood needs rebuilt counts");
        e.fresh_rows(e.stream);
        e.sync();
        const ModelHost &m = e.models[0];
        const int32_t K = m.K, V = m.V;
        const double alpha = m.alpha, beta = m.beta;
        const int32_t G = 1024;
        const size_t U = e.units.size();
        double *part = (double *)e.dalloc(G * sizeof(double) * (1 + U));
        hip_check(launch_joint_ll_words(m.nwk, V, K, std::lgamma(beta), m.beta, part, G, e.stream), "joint");
        for (size_t k = 0; k < U; ++k) {
            UnitHost &u = e.units[k];
            hip_check(launch_joint_ll_docs(u.meta, u.theta, u.doc_end - u.doc_start, m.alpha,
                                           part + G * (1 + k), G, e.stream),
                      "joint");
        }
        std::vector<double> h(G * (1 + U));
        hip_check(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, e.stream), "copy");
        std::vector<int32_t> nk(K);
        hip_check(hipMemcpyAsync(nk.data(), m.nk, K * 4, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
        e.dfree(part);
        double s = 0;
        for (double v : h) s += v;
        const double vb = (double)V * beta, ka = (double)K * alpha;
        for (int32_t k = 0; k < K; ++k) s += std::lgamma(vb) - std::lgamma((double)nk[k] + vb);
        for (int64_t nd : m.doc_len) s += std::lgamma(ka) - std::lgamma((double)nd + ka);
        *out = s;
    });
}

// The joint LL with no allocation: the partials go to the library's static
// device scratch (its address looked up at culda_create), the final sum is
// formed on the device and 8 bytes come back.  Calls that overlap on one
// device share that scratch and race.
extern "C" int culda_loglik_joint_noalloc(culda_ctx *ctx, double *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        if (!e.theta_ready()) fail(CULDA_ERR_STATE, "joint log-likelihood needs rebuilt counts");
        const int32_t U = (int32_t)e.units.size();
        if (U > kJointUnitsMax)
            fail(CULDA_ERR_LIMIT, "the allocation-free joint log-likelihood serves at most %d chunks (this corpus has %d)",
                 kJointUnitsMax, U);
        double *scr = joint_scratch_ptr(e.prm.device);
        if (!scr) fail(CULDA_ERR_DEVICE, "static scratch not set up on device %d", e.prm.device);
        const ModelHost &m = e.models[0];
        const int32_t K = m.K, V = m.V;
        e.fresh_rows(e.stream);
        hip_check(launch_joint_ll_words(m.nwk, V, K, std::lgamma((double)m.beta), m.beta, scr, kJointGrid, e.stream),
                  "joint");
        for (int32_t k = 0; k < U; ++k) {
            const UnitHost &u = e.units[k];
            hip_check(launch_joint_ll_docs(u.meta, u.theta, u.doc_end - u.doc_start, m.alpha,
                                           scr + (size_t)kJointGrid * (1 + k), kJointGrid, e.stream),
                      "joint");
        }
        double *res = scr + (size_t)kJointGrid * (1 + kJointUnitsMax);
        hip_check(launch_joint_final(scr, kJointGrid * (1 + U), m.nk, K, (double)V * (double)m.beta, e.d_units,
                                     m.unit_begin, m.unit_end, (double)K * (double)m.alpha, res, e.stream),
                  "joint");
        hip_check(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
    });
}

extern "C" int culda_get_phi(culda_ctx *ctx, int32_t *n_wk, int32_t *n_k) {
    CTX_CALL({
        if (!n_wk || !n_k) fail(CULDA_ERR_ARG, "null output");
        HostPhi h = download_phi(e, 0);
        std::memcpy(n_wk, h.nwk.data(), h.nwk.size() * 4);
        std::memcpy(n_k, h.nk.data(), h.nk.size() * 4);
    });
}

extern "C" int culda_get_phi_head(culda_ctx *ctx, float *head) {
    CTX_CALL({
        if (!head) fail(CULDA_ERR_ARG, "null output");
        if (!e.uploaded || !e.phi_ready) fail(CULDA_ERR_STATE, "phi head not prepared");
        const ModelHost &m = e.models[0];
        const size_t n = (size_t)m.V * m.K;
        float *d = (float *)e.dalloc(n * 4);
        e.fresh_rows(e.stream);
        hip_check(launch_phi_head(m.nwk, m.den, d, m.V, m.K, m.beta, e.stream), "head");
        hip_check(hipStreamSynchronize(e.stream), "sync");
        hip_check(hipMemcpy(head, d, n * 4, hipMemcpyDeviceToHost), "head copy");
        e.dfree(d);
    });
}

extern "C" int culda_get_theta(culda_ctx *ctx, int32_t chunk, int64_t *offsets, int32_t *counts, int16_t *topics) {
    CTX_CALL({
        if (!offsets) fail(CULDA_ERR_ARG, "null offsets");
        std::vector<int64_t> off;
        std::vector<int32_t> cnt;
        std::vector<int16_t> top;
        download_theta(e, chunk, off, cnt, top);
        std::memcpy(offsets, off.data(), off.size() * 8);
        if (counts) std::memcpy(counts, cnt.data(), cnt.size() * 4);
        if (topics) std::memcpy(topics, top.data(), top.size() * 2);
    });
}

// ModelPhi::savePhi (model/model_phi.cpp:161-191) wrote a text dump of the
// first 1000 words only; this writes every word, plus a binary file.
extern "C" int culda_save_phi(culda_ctx *ctx, const char *prefix) {
    CTX_CALL({
        if (!prefix) fail(CULDA_ERR_ARG, "null prefix");
        HostPhi h = download_phi(e, 0);
        const int32_t V = e.models[0].V, K = e.models[0].K;
        {
            OutFile f(std::string(prefix) + ".phi.bin", "wb");
            f.write(&V, 4, 1);
            f.write(&K, 4, 1);
            f.write(h.nwk.data(), 4, h.nwk.size());
            f.write(h.nk.data(), 4, h.nk.size());
            f.commit();
        }
        OutFile f(std::string(prefix) + ".phi.txt", "w");
        for (int32_t w = 0; w < V; ++w) {
            f.printf("Word id:%d, ", w);
            for (int32_t k = 0; k < K; ++k) {
                if (k % 32 == 0) f.printf("\n%2d:", k / 32);
                const int32_t v = h.nwk[(size_t)w * K + k];
                if (v == 0)
                    f.printf("_,");
                else
                    f.printf("%d,", v);
            }
            f.printf("\n");
        }
        f.commit();
    });
}

// ModelTheta::saveTheta / saveDoc exist only as commented code in the
// reference (model/model_theta.cpp:161-242); this writes CSR theta in
// global doc order (binary) and the saveDoc text layout.
extern "C" int culda_save_theta(culda_ctx *ctx, const char *prefix) {
    CTX_CALL({
        if (!prefix) fail(CULDA_ERR_ARG, "null prefix");
        if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        const int32_t D = e.models[0].num_docs, K = e.models[0].K;
        std::vector<int64_t> goff(D + 1, 0);
        std::vector<int32_t> gcnt;
        std::vector<int16_t> gtop;
        for (size_t k = 0; k < e.units.size(); ++k) {
            std::vector<int64_t> off;
            std::vector<int32_t> cnt;
            std::vector<int16_t> top;
            download_theta(e, (int32_t)k, off, cnt, top);
            const int32_t ds = e.units[k].doc_start;
            for (size_t d = 0; d + 1 < off.size(); ++d) goff[ds + d + 1] = off[d + 1] - off[d];
            gcnt.insert(gcnt.end(), cnt.begin(), cnt.end());
            gtop.insert(gtop.end(), top.begin(), top.end());
        }
        for (int32_t d = 0; d < D; ++d) goff[d + 1] += goff[d];
        {
            OutFile f(std::string(prefix) + ".theta.bin", "wb");
            f.write(&D, 4, 1);
            f.write(&K, 4, 1);
            f.write(goff.data(), 8, goff.size());
            f.write(gcnt.data(), 4, gcnt.size());
            f.write(gtop.data(), 2, gtop.size());
            f.commit();
        }
        OutFile f(std::string(prefix) + ".theta.txt", "w");
        f.printf("Doc, total, top topics\n");
        for (int32_t d = 0; d < D; ++d) {
            std::vector<std::pair<int32_t, int32_t>> row;
            int64_t tot = 0;
            for (int64_t j = goff[d]; j < goff[d + 1]; ++j) {
                row.emplace_back(gcnt[j], gtop[j]);
                tot += gcnt[j];
            }
            std::stable_sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first > b.first; });
            f.printf("%4d| %lld ", d, (long long)tot);
            for (auto &pr : row) f.printf("(%d,%d), ", pr.second, pr.first);
            f.printf("\n");
        }
        f.commit();
    });
}

static const char kCkptMagic[8] = {'C', 'U', 'L', 'D', 'A', 'C', 'K', '1'};

extern "C" int culda_save_checkpoint(culda_ctx *ctx, const char *path, int32_t next_iteration) {
    CTX_CALL({
        if (!path) fail(CULDA_ERR_ARG, "null path");
        if (!e.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        e.sync();
        // written to path.tmp and renamed: an earlier checkpoint survives a failed save
        OutFile f(path, "wb");
        const int32_t C = (int32_t)e.units.size(), K = e.prm.k;
        f.write(kCkptMagic, 1, 8);
        f.write(&C, 4, 1);
        f.write(&K, 4, 1);
        f.write(&ctx->seed, 8, 1);
        f.write(&next_iteration, 4, 1);
        for (auto &u : e.units) {
            std::vector<uint16_t> z(u.num_tokens);
            if (u.num_tokens) hip_check(hipMemcpy(z.data(), u.z, u.num_tokens * 2, hipMemcpyDeviceToHost), "z");
            f.write(&u.num_tokens, 8, 1);
            f.write(z.data(), 2, z.size());
        }
        f.commit();
    });
}

extern "C" int culda_load_checkpoint(culda_ctx *ctx, const char *path, int32_t *next_iteration) {
    CTX_CALL({
        if (!path || !next_iteration) fail(CULDA_ERR_ARG, "null argument");
        if (!e.uploaded) fail(CULDA_ERR_STATE, "upload the corpus before loading a checkpoint");
        FILE *f = open_or_fail(path, "rb");
        char magic[8];
        int32_t C = 0, K = 0, it = 0;
        uint64_t seed = 0;
        bool ok = fread(magic, 1, 8, f) == 8 && !std::memcmp(magic, kCkptMagic, 8) && fread(&C, 4, 1, f) == 1 &&
                  fread(&K, 4, 1, f) == 1 && fread(&seed, 8, 1, f) == 1 && fread(&it, 4, 1, f) == 1;
        if (!ok || C != (int32_t)e.units.size() || K != e.prm.k || seed != ctx->seed) {
            fclose(f);
            fail(CULDA_ERR_FORMAT, "%s is not a checkpoint of this corpus / k / seed", path);
        }
        std::vector<std::vector<uint16_t>> zs(C);
        for (int32_t k = 0; k < C; ++k) {
            int64_t n = 0;
            if (fread(&n, 8, 1, f) != 1 || n != e.units[k].num_tokens) {
                fclose(f);
                fail(CULDA_ERR_FORMAT, "checkpoint chunk %d size mismatch", k);
            }
            zs[k].resize(n);
            if (n && fread(zs[k].data(), 2, n, f) != (size_t)n) {
                fclose(f);
                fail(CULDA_ERR_FORMAT, "truncated checkpoint");
            }
            for (uint16_t v : zs[k])
                if (v >= (uint32_t)K) {
                    fclose(f);
                    fail(CULDA_ERR_FORMAT, "checkpoint topic out of range");
                }
        }
        fclose(f);
        e.sync();
        for (int32_t k = 0; k < C; ++k)
            if (!zs[k].empty())
                hip_check(hipMemcpy(e.units[k].z, zs[k].data(), zs[k].size() * 2, hipMemcpyHostToDevice), "z");
        e.set_theta_ready(false);
        e.phi_ready = false;
        *next_iteration = it;
    });
}

extern "C" int culda_get_stats(culda_ctx *ctx, culda_stats *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        std::memset(out, 0, sizeof *out);
        for (auto &m : e.models) out->num_tokens += m.num_tokens;
        out->num_slots = e.num_slots;
        for (auto &u : e.units) out->theta_capacity += u.theta_cap;
        e.theta_stats(out->sum_doc_nnz, out->sum_token_nnz);
        out->sampler_ms = e.sampler_ms;
        out->phi_ms = e.phi_ms;
        out->theta_ms = e.theta_ms;
        out->sampler_launches = e.sampler_launches;
    });
}

extern "C" int culda_sampler_bytes(culda_ctx *ctx, double *bytes) {
    CTX_CALL_LIVE({
        if (!bytes) fail(CULDA_ERR_ARG, "null output");
        *bytes = e.last_sampler_bytes;
    });
}

extern "C" int culda_last_sweeps(culda_ctx *ctx, double *sampler_ms, double *sampler_bytes, int32_t *launches) {
    CTX_CALL_LIVE({
        if (!sampler_ms || !sampler_bytes || !launches) fail(CULDA_ERR_ARG, "null output");
        *sampler_ms = e.sampler_ms;
        *sampler_bytes = e.last_sampler_bytes;
        *launches = e.sampler_launches;
    });
}

extern "C" int culda_stream_copy(culda_ctx *ctx, int64_t n_bytes, int32_t repeats, double *gbps) {
    CTX_CALL({
        if (!gbps || n_bytes <= 0 || n_bytes % 16 || repeats < 1) fail(CULDA_ERR_ARG, "bad stream copy arguments");
        float4 *a = (float4 *)e.dalloc(n_bytes), *b = (float4 *)e.dalloc(n_bytes);
        hip_check(launch_fill_u32((uint32_t *)a, n_bytes / 4, 0x3f800000u, e.stream), "fill");
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(launch_copy(a, b, n_bytes / 16, e.stream), "copy warmup");
        hip_check(hipEventRecord(e0, e.stream), "event");
        for (int32_t r = 0; r < repeats; ++r) hip_check(launch_copy(a, b, n_bytes / 16, e.stream), "copy");
        hip_check(hipEventRecord(e1, e.stream), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * (double)n_bytes * repeats / (ms * 1e-3) / 1e9;
        e.dfree(a);
        e.dfree(b);
    });
}

// ====================================================== batch of corpora

#define BATCH_CALL(...)         \
    return guarded([&] {        \
        Engine &e = eng(batch); \
        (void)e;                \
        __VA_ARGS__;            \
        return CULDA_OK;        \
    })

extern "C" int culda_batch_create(const culda_params *params, culda_batch **out) {
    return guarded([&] {
        if (!params || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        auto *h = new culda_batch;
        try {
            h->e.init_device(*params);
        } catch (...) {
            delete h;
            throw;
        }
        h->e.prm.fused_phi = 1;
        *out = h;
        return CULDA_OK;
    });
}

extern "C" void culda_batch_destroy(culda_batch *batch) {
    if (!batch) return;
    (void)hipSetDevice(batch->e.prm.device);
    delete batch;
}

extern "C" int culda_batch_add(culda_batch *batch, const culda_corpus *corpus, float alpha, float beta, uint64_t seed,
                               int32_t *model) {
    BATCH_CALL({
        if (!corpus) fail(CULDA_ERR_ARG, "null corpus");
        const int32_t m = e.add_corpus(corpus_ref(corpus), alpha, beta, seed);
        if (model) *model = m;
    });
}

extern "C" int culda_batch_upload(culda_batch *batch) { BATCH_CALL(e.upload()); }

extern "C" int culda_batch_init(culda_batch *batch) {
    BATCH_CALL({
        e.init_z();
        e.rebuild_all();
    });
}

extern "C" int culda_batch_set_persistent(culda_batch *batch, int32_t groups) {
    BATCH_CALL({
        if (groups < 0) fail(CULDA_ERR_ARG, "groups must be >= 0");
        if (e.graph) {
            (void)hipGraphExecDestroy(e.graph);
            e.graph = nullptr;
        }
        e.persistent_groups = groups;
    });
}

extern "C" int culda_batch_iterate(culda_batch *batch, int32_t first_iteration, int32_t count, int32_t use_graph,
                                   double *ll, double *seconds) {
    BATCH_CALL(e.iterate_batch((uint32_t)std::max(first_iteration, 0), count, use_graph != 0, ll, seconds));
}

// ====================================================== learned MH

extern "C" int culda_mh_enable(culda_ctx *ctx, const culda_mh_params *params) {
    CTX_CALL({
        if (!params) fail(CULDA_ERR_ARG, "null params");
        e.mh_enable(*params);
    });
}

extern "C" int culda_mh_collect(culda_ctx *ctx, int32_t *num_snapshots) {
    CTX_CALL({
        const int32_t n = e.mh_collect();
        if (num_snapshots) *num_snapshots = n;
    });
}

extern "C" int culda_mh_train(culda_ctx *ctx, int32_t epochs, double *mean_loss, double *seconds) {
    CTX_CALL({
        auto t0 = std::chrono::steady_clock::now();
        const double l = e.mh_train(epochs);
        if (mean_loss) *mean_loss = l;
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    });
}

extern "C" int culda_mh_iterate(culda_ctx *ctx, int32_t iteration, int32_t proposals, double *accept_rate,
                                double *seconds) {
    CTX_CALL({
        if (iteration < 1) fail(CULDA_ERR_ARG, "iteration must be >= 1");
        auto t0 = std::chrono::steady_clock::now();
        const double r = e.mh_iterate((uint32_t)iteration, proposals);
        if (accept_rate) *accept_rate = r;
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ctx->last_sweep = (uint32_t)iteration;
    });
}

namespace {
struct MHSizes {
    size_t w1, b1, w2, b2, emb;
};
MHSizes mh_sizes(const MHState &s) {
    return MHSizes{(size_t)kIn * kHidden, (size_t)kHidden, (size_t)kHidden * s.K, (size_t)s.K, (size_t)s.V * kEmb};
}
}  // namespace

extern "C" int culda_mh_get_weights(culda_ctx *ctx, float *w1, float *b1, float *w2, float *b2, float *emb) {
    CTX_CALL({
        MHState &s = e.mh_state();
        const MHSizes z = mh_sizes(s);
        e.sync();
        const std::pair<float *, std::pair<const float *, size_t>> parts[] = {
            {w1, {s.w1, z.w1}}, {b1, {s.b1, z.b1}}, {w2, {s.w2, z.w2}}, {b2, {s.b2, z.b2}}, {emb, {s.emb, z.emb}}};
        for (auto &p : parts)
            if (p.first) hip_check(hipMemcpy(p.first, p.second.first, p.second.second * 4, hipMemcpyDeviceToHost), "copy");
    });
}

extern "C" int culda_mh_set_weights(culda_ctx *ctx, const float *w1, const float *b1, const float *w2, const float *b2,
                                    const float *emb) {
    CTX_CALL({
        MHState &s = e.mh_state();
        const MHSizes z = mh_sizes(s);
        if (!w1 || !b1 || !w2 || !b2 || !emb) fail(CULDA_ERR_ARG, "null weights");
        e.sync();
        hip_check(hipMemcpy(s.w1, w1, z.w1 * 4, hipMemcpyHostToDevice), "copy");
        hip_check(hipMemcpy(s.b1, b1, z.b1 * 4, hipMemcpyHostToDevice), "copy");
        hip_check(hipMemcpy(s.w2, w2, z.w2 * 4, hipMemcpyHostToDevice), "copy");
        hip_check(hipMemcpy(s.b2, b2, z.b2 * 4, hipMemcpyHostToDevice), "copy");
        hip_check(hipMemcpy(s.emb, emb, z.emb * 4, hipMemcpyHostToDevice), "copy");
        hip_check(launch_to_bf16(s.w1, s.w1h, s.w1t, kIn, z.w1, e.stream), "bf16");
        hip_check(launch_to_bf16(s.w2, s.w2h, s.w2t, kHidden, z.w2, e.stream), "bf16");
        hip_check(launch_to_bf16(s.emb, s.embh, nullptr, 1, z.emb, e.stream), "bf16");
        hip_check(hipStreamSynchronize(e.stream), "sync");
    });
}

static void upload_queries(Engine &e, MHState &s, int64_t n, const int32_t *docs, const int32_t *words) {
    if (!docs || !words) fail(CULDA_ERR_ARG, "null queries");
    for (int64_t i = 0; i < n; ++i) {
        if (docs[i] < 0 || docs[i] >= s.D) fail(CULDA_ERR_ARG, "query %lld: doc %d out of range", (long long)i, docs[i]);
        if (words[i] < 0 || words[i] >= s.V)
            fail(CULDA_ERR_ARG, "query %lld: word %d out of range", (long long)i, words[i]);
    }
    hip_check(hipMemcpy(s.q_doc, docs, n * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(s.q_word, words, n * 4, hipMemcpyHostToDevice), "copy");
    (void)e;
}

extern "C" int culda_mh_logits(culda_ctx *ctx, int64_t n, const int32_t *docs, const int32_t *words, float *logits) {
    CTX_CALL({
        MHState &s = e.mh_state();
        if (!logits || n < 0 || n > s.prm.infer_batch) fail(CULDA_ERR_ARG, "n must be in [0, infer_batch]");
        if (!e.theta_ready()) fail(CULDA_ERR_STATE, "needs rebuilt counts");
        e.sync();
        upload_queries(e, s, n, docs, words);
        e.mh_sketch_current(e.stream);
        e.mh_query_logits(s.q_doc, s.q_word, n, s.Li, e.stream);
        std::vector<uint16_t> h((size_t)n * s.K);
        if (n) hip_check(hipMemcpyAsync(h.data(), s.Li, h.size() * 2, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
        for (size_t i = 0; i < h.size(); ++i) {
            const uint32_t u = (uint32_t)h[i] << 16;  // BF16 -> FP32, exact
            std::memcpy(logits + i, &u, 4);
        }
    });
}

extern "C" int culda_mh_gradients(culda_ctx *ctx, int32_t n, const int32_t *docs, const int32_t *words,
                                  const int32_t *targets, int32_t train_step, double *loss, float *dw1, float *db1,
                                  float *dw2, float *db2, float *demb) {
    CTX_CALL({
        MHState &s = e.mh_state();
        if (n < 1 || n > s.prm.batch) fail(CULDA_ERR_ARG, "n must be in [1, batch]");
        if (!targets) fail(CULDA_ERR_ARG, "null targets");
        if (!e.theta_ready()) fail(CULDA_ERR_STATE, "needs rebuilt counts");
        for (int32_t i = 0; i < n; ++i)
            if (targets[i] < 0 || targets[i] >= s.K) fail(CULDA_ERR_ARG, "target %d out of range", targets[i]);
        e.sync();
        upload_queries(e, s, n, docs, words);
        e.mh_sketch_current(e.stream);
        const int32_t B = s.prm.batch;
        hip_check(launch_gather_query(s.sk_cur, s.q_doc, s.q_word, s.embh, n, B, s.X, e.stream), "gather");
        std::vector<int32_t> tg(B, -1), wr(B, -1);
        for (int32_t i = 0; i < n; ++i) {
            tg[i] = targets[i];
            wr[i] = words[i];
        }
        hip_check(hipMemcpy(s.tgt, tg.data(), B * 4, hipMemcpyHostToDevice), "copy");
        hip_check(hipMemcpy(s.wrow, wr.data(), B * 4, hipMemcpyHostToDevice), "copy");
        unsigned long long *lfx = (unsigned long long *)e.dalloc(16);
        hip_check(launch_fill_u32((uint32_t *)lfx, 2, 0, e.stream), "clear");
        mh_batch_step(e, s, n, false, lfx, e.stream);
        const MHSizes z = mh_sizes(s);
        unsigned long long hl = 0;
        hip_check(hipMemcpyAsync(&hl, lfx, 8, hipMemcpyDeviceToHost, e.stream), "copy");
        const std::pair<float *, std::pair<const float *, size_t>> parts[] = {
            {dw1, {s.g_w1, z.w1}}, {db1, {s.g_b1, z.b1}}, {dw2, {s.g_w2, z.w2}}, {db2, {s.g_b2, z.b2}}};
        for (auto &p : parts)
            if (p.first)
                hip_check(hipMemcpyAsync(p.first, p.second.first, p.second.second * 4, hipMemcpyDeviceToHost, e.stream),
                          "copy");
        std::vector<long long> ge(z.emb);
        hip_check(hipMemcpyAsync(ge.data(), s.g_emb, z.emb * 8, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
        e.dfree(lfx);
        if (demb)
            for (size_t i = 0; i < z.emb; ++i) demb[i] = (float)((double)ge[i] * (1.0 / 4294967296.0));
        if (loss) *loss = (double)(long long)hl / 16777216.0 / (double)n;
        if (train_step) {
            mh_apply_adam(e, s, e.stream);  // consumes (and clears) the embedding accumulator
        } else {
            hip_check(launch_fill_u32((uint32_t *)s.g_emb, (int64_t)z.emb * 2, 0, e.stream), "clear");
        }
        hip_check(hipStreamSynchronize(e.stream), "sync");
    });
}

extern "C" int culda_mh_tables(culda_ctx *ctx, int64_t first, int32_t n, int32_t *weights, uint32_t *thresholds,
                               uint32_t *aliases) {
    CTX_CALL({
        MHState &s = e.mh_state();
        if (!weights || !thresholds || !aliases) fail(CULDA_ERR_ARG, "null output");
        // first counts tokens over the chunks in order; the range stays in one chunk
        int32_t ui = 0;
        while (ui + 1 < (int32_t)e.units.size() && first >= s.unit_base[ui + 1]) ++ui;
        const UnitHost &u = e.units[ui];
        const int64_t base = s.unit_base[ui];
        if (n < 0 || n > s.prm.infer_batch || first < 0 || first + n > base + u.num_tokens)
            fail(CULDA_ERR_ARG, "token range out of bounds (it must lie in one chunk)");
        if (!e.theta_ready() || !e.phi_ready) fail(CULDA_ERR_STATE, "needs rebuilt counts");
        e.sync();
        const size_t cells = (size_t)n * s.K;
        int32_t *dw = (int32_t *)e.dalloc(std::max<size_t>(cells, 1) * 4);
        uint32_t *df = (uint32_t *)e.dalloc(std::max<size_t>(cells, 1) * 4);
        uint32_t *da = (uint32_t *)e.dalloc(std::max<size_t>(cells, 1) * 4);
        e.fresh_rows(e.stream);
        e.mh_sketch_current(e.stream);
        e.mh_query_logits(s.tok_doc + first, s.tok_word + first, n, s.Li, e.stream);
        const ModelHost &m = e.models[0];
        MHArgs a{};
        a.logits = s.Li;
        a.word = s.tok_word + base;
        a.gdelta = s.gdelta[ui];
        a.unit = UnitDev{u.tokens, u.z, u.zp, u.word_idx, u.rev_off, u.rev_idx, u.meta, u.theta, u.model, u.doc_end - u.doc_start};
        a.model = ModelDev{m.nwk, m.nk, m.den, m.zero_phi, m.rden, m.seed, (double)m.K * (double)m.alpha, m.alpha, m.beta, m.K, m.V};
        a.accepted = s.accepted;
        a.w_out = dw;
        a.f_out = df;
        a.a_out = da;
        a.start = first - base;
        a.n = n;
        a.proposals = 0;
        a.sweep = 0;
        hip_check(launch_mh(a, e.stream), "mh tables");
        hip_check(hipMemcpyAsync(weights, dw, cells * 4, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipMemcpyAsync(thresholds, df, cells * 4, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipMemcpyAsync(aliases, da, cells * 4, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
        e.dfree(dw);
        e.dfree(df);
        e.dfree(da);
    });
}

extern "C" int culda_mh_sketch(culda_ctx *ctx, uint16_t *out) {
    CTX_CALL({
        MHState &s = e.mh_state();
        if (!out) fail(CULDA_ERR_ARG, "null output");
        if (!e.theta_ready()) fail(CULDA_ERR_STATE, "needs rebuilt counts");
        e.sync();
        e.mh_sketch_current(e.stream);
        hip_check(hipMemcpyAsync(out, s.sk_cur, (size_t)s.D * kSketch * 2, hipMemcpyDeviceToHost, e.stream), "copy");
        hip_check(hipStreamSynchronize(e.stream), "sync");
    });
}

// ====================================================== held-out LL

extern "C" int culda_heldout_set_eval(culda_ctx *fold, const culda_corpus *eval) {
    return guarded([&] {
        if (!eval) fail(CULDA_ERR_ARG, "null corpus");
        eng(fold).settle();
        eng(fold).set_eval(corpus_ref(eval));
        return CULDA_OK;
    });
}

extern "C" int culda_heldout_ll(culda_ctx *train, culda_ctx *fold, int32_t fold_in_sweeps, double *ll_per_token) {
    return guarded([&] {
        if (!ll_per_token) fail(CULDA_ERR_ARG, "null output");
        if (train == fold) fail(CULDA_ERR_ARG, "training and held-out contexts must differ");
        Engine &t = eng(train);
        Engine &f = eng(fold);
        if (t.prm.device != f.prm.device) fail(CULDA_ERR_ARG, "contexts on different devices");
        t.settle();
        f.settle();
        *ll_per_token = f.heldout_ll(t, fold_in_sweeps);
        return CULDA_OK;
    });
}

// Tries each slot size with one workgroup per slot and with a persistent
// grid of all (and half) the sampler workgroups the device holds at once;
// keeps the fastest pair.
extern "C" int culda_batch_autotune(culda_batch *batch, culda_tune *out) {
    BATCH_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        TuneResult best;
        best.best_ms = 1e30;
        int32_t best_groups = 0;
        std::vector<std::pair<int32_t, double>> all;
        const int32_t per_cu = std::max(1, sample_blocks_per_cu(e.rs_stride));
        const int32_t resident = per_cu * e.cu_count;
        for (int32_t groups : {0, resident, std::max(1, resident / 2)}) {
            e.persistent_groups = groups;
            TuneResult r = e.autotune(kSlotCandidates, true);
            for (auto &t : r.tried) all.push_back(t);
            if (r.best_ms < best.best_ms) {
                best = r;
                best_groups = groups;
            }
        }
        e.persistent_groups = best_groups;
        e.rebuild_slots(best.slot_tokens);
        best.tried = all;
        fill_tune(best, out);
        out->persistent_groups = best_groups;
    });
}

extern "C" int culda_batch_get_z(culda_batch *batch, int32_t model, int32_t chunk, uint16_t *z, int64_t n) {
    BATCH_CALL({
        check_model(e, model);
        const ModelHost &m = e.models[model];
        if (chunk < 0 || chunk >= m.unit_end - m.unit_begin) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
        copy_z(e, m.unit_begin + chunk, z, n, false);
    });
}

extern "C" int culda_batch_get_phi(culda_batch *batch, int32_t model, int32_t *n_wk, int32_t *n_k) {
    BATCH_CALL({
        if (!n_k) fail(CULDA_ERR_ARG, "null output");
        if (!n_wk) {  // n_k only
            check_model(e, model);
            e.sync();
            hip_check(hipMemcpy(n_k, e.models[model].nk, (size_t)e.models[model].K * 4, hipMemcpyDeviceToHost),
                      "n_k copy");
            return CULDA_OK;
        }
        HostPhi h = download_phi(e, model);
        std::memcpy(n_wk, h.nwk.data(), h.nwk.size() * 4);
        std::memcpy(n_k, h.nk.data(), h.nk.size() * 4);
    });
}

extern "C" int culda_batch_get_stats(culda_batch *batch, culda_stats *out) {
    BATCH_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        std::memset(out, 0, sizeof *out);
        for (auto &m : e.models) out->num_tokens += m.num_tokens;
        out->num_slots = e.num_slots;
        for (auto &u : e.units) out->theta_capacity += u.theta_cap;
        e.theta_stats(out->sum_doc_nnz, out->sum_token_nnz);
        out->sampler_ms = e.sampler_ms;
        out->other_ms = e.iter_ms;
        out->sampler_launches = e.sampler_launches;
    });
}

extern "C" int culda_batch_sampler_bytes(culda_batch *batch, double *bytes) {
    BATCH_CALL({
        if (!bytes) fail(CULDA_ERR_ARG, "null output");
        *bytes = e.last_sampler_bytes;
    });
}

extern "C" int culda_batch_info(culda_batch *batch, int32_t *num_models, int64_t *num_tokens, int64_t *num_slots) {
    BATCH_CALL({
        int64_t n = 0;
        for (auto &m : e.models) n += m.num_tokens;
        if (num_models) *num_models = (int32_t)(e.uploaded ? e.models.size() : e.pending.size());
        if (num_tokens) *num_tokens = n;
        if (num_slots) *num_slots = e.num_slots;
    });
}
