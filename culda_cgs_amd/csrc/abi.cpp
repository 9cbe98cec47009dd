// abi.cpp -- extern "C" entry points over culda::Context, plus the host
// side of the ModelPhi / ModelTheta save hooks and checkpoints.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.h"
#include "kernels.h"

using namespace culda;

struct culda_ctx {
    Context c;
};

const culda::Corpus &corpus_ref(const culda_corpus *p);

namespace {

Context &ctx_of(culda_ctx *p) {
    if (!p) fail(CULDA_ERR_ARG, "null context");
    return p->c;
}

void check_chunk(Context &c, int32_t chunk) {
    if (!c.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
    if (chunk < 0 || chunk >= (int32_t)c.chunks.size()) fail(CULDA_ERR_ARG, "chunk %d out of range", chunk);
}

// Host mirror of the phi counts (ModelPhi::hostPhiTopicWordShort / hostPhiTopic,
// model/model_phi.h:32-33).
struct HostPhi {
    std::vector<int32_t> nwk, nk;
};

HostPhi download_phi(Context &c) {
    if (!c.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
    HostPhi h;
    h.nwk.resize((size_t)c.num_words * c.prm.k);
    h.nk.resize(c.prm.k);
    hip_check(hipStreamSynchronize(c.stream), "sync");
    if (!h.nwk.empty())
        hip_check(hipMemcpy(h.nwk.data(), c.nwk, h.nwk.size() * 4, hipMemcpyDeviceToHost), "phi copy");
    hip_check(hipMemcpy(h.nk.data(), c.nk, h.nk.size() * 4, hipMemcpyDeviceToHost), "n_k copy");
    return h;
}

// Compact CSR of one chunk's theta rows.
void download_theta(Context &c, int32_t chunk, std::vector<int64_t> &off, std::vector<int32_t> &cnt,
                    std::vector<int16_t> &top) {
    check_chunk(c, chunk);
    ChunkState &cs = c.chunks[chunk];
    const int32_t Dc = cs.doc_end - cs.doc_start;
    std::vector<DocMeta> meta(Dc);
    std::vector<uint32_t> th(cs.theta_cap);
    hip_check(hipStreamSynchronize(c.stream), "sync");
    if (Dc) hip_check(hipMemcpy(meta.data(), cs.meta, Dc * sizeof(DocMeta), hipMemcpyDeviceToHost), "meta copy");
    if (cs.theta_cap)
        hip_check(hipMemcpy(th.data(), cs.theta, cs.theta_cap * 4, hipMemcpyDeviceToHost), "theta copy");
    off.assign(Dc + 1, 0);
    for (int32_t d = 0; d < Dc; ++d) off[d + 1] = off[d] + meta[d].nnz;
    cnt.resize(off[Dc]);
    top.resize(off[Dc]);
    const uint32_t cmask = (1u << c.count_bits) - 1u;
    for (int32_t d = 0; d < Dc; ++d)
        for (uint32_t j = 0; j < meta[d].nnz; ++j) {
            const uint32_t e = th[meta[d].start + j];
            cnt[off[d] + j] = (int32_t)(e & cmask);
            top[off[d] + j] = (int16_t)(e >> (c.count_bits + 2));
        }
}

FILE *open_or_fail(const std::string &path, const char *mode) {
    FILE *f = fopen(path.c_str(), mode);
    if (!f) fail(CULDA_ERR_IO, "%s open failed", path.c_str());
    return f;
}

}  // namespace

extern "C" int culda_create(const culda_params *params, culda_ctx **out) {
    return guarded([&] {
        if (!params || !out) fail(CULDA_ERR_ARG, "null argument");
        *out = nullptr;
        const culda_params &p = *params;
        if (p.k < 16 || p.k > kMaxK || p.k % 16)
            fail(CULDA_ERR_ARG, "k = %d: the sampler needs a multiple of 16 in [16, %d]", p.k, kMaxK);
        if (!(p.alpha > 0.0f) || !std::isfinite(p.alpha)) fail(CULDA_ERR_ARG, "alpha must be positive and finite");
        if (!(p.beta > 0.0f) || !std::isfinite(p.beta)) fail(CULDA_ERR_ARG, "beta must be positive and finite");
        if (p.slot_tokens < 0) fail(CULDA_ERR_ARG, "slot_tokens must be >= 0");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) fail(CULDA_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
        if (p.device < 0 || p.device >= n) fail(CULDA_ERR_ARG, "device %d out of range (%d devices)", p.device, n);
        hip_check(hipSetDevice(p.device), "hipSetDevice");
        auto *h = new culda_ctx;
        h->c.prm = p;
        e = hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->c.stream2, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete h;
            fail(CULDA_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p.device) != hipSuccess) cus = 0;
        h->c.cu_count = cus > 0 ? cus : 256;
        *out = h;
        return CULDA_OK;
    });
}

extern "C" void culda_destroy(culda_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.prm.device);
    delete ctx;
}

#define CTX_CALL(...)                                 \
    return guarded([&] {                              \
        Context &c = ctx_of(ctx);                     \
        hip_check(hipSetDevice(c.prm.device), "dev"); \
        __VA_ARGS__;                                  \
        return CULDA_OK;                              \
    })

extern "C" int culda_upload(culda_ctx *ctx, const culda_corpus *corpus) {
    CTX_CALL({
        if (!corpus) fail(CULDA_ERR_ARG, "null corpus");
        if (c.uploaded) fail(CULDA_ERR_STATE, "a corpus is already uploaded to this context");
        c.upload(corpus_ref(corpus));
    });
}

extern "C" int culda_init_z(culda_ctx *ctx) { CTX_CALL(c.init_z()); }

extern "C" int culda_set_z(culda_ctx *ctx, int32_t chunk, const uint16_t *z, int64_t n) {
    CTX_CALL({
        check_chunk(c, chunk);
        ChunkState &cs = c.chunks[chunk];
        if (n != cs.num_tokens) fail(CULDA_ERR_ARG, "chunk %d has %lld tokens, got %lld", chunk,
                                     (long long)cs.num_tokens, (long long)n);
        if (n && !z) fail(CULDA_ERR_ARG, "null z");
        for (int64_t i = 0; i < n; ++i)
            if (z[i] >= (uint32_t)c.prm.k) fail(CULDA_ERR_ARG, "z[%lld] = %u is not below k", (long long)i, z[i]);
        hip_check(hipStreamSynchronize(c.stream), "sync");
        if (n) hip_check(hipMemcpy(cs.z, z, n * 2, hipMemcpyHostToDevice), "z copy");
        c.theta_ready = c.phi_ready = false;
    });
}

extern "C" int culda_get_z(culda_ctx *ctx, int32_t chunk, uint16_t *z, int64_t n) {
    CTX_CALL({
        check_chunk(c, chunk);
        ChunkState &cs = c.chunks[chunk];
        if (n != cs.num_tokens) fail(CULDA_ERR_ARG, "chunk %d has %lld tokens, got %lld", chunk,
                                     (long long)cs.num_tokens, (long long)n);
        if (n && !z) fail(CULDA_ERR_ARG, "null z");
        hip_check(hipStreamSynchronize(c.stream), "sync");
        if (n) hip_check(hipMemcpy(z, cs.z, n * 2, hipMemcpyDeviceToHost), "z copy");
    });
}

extern "C" int culda_rebuild_phi(culda_ctx *ctx, int32_t chunk, int32_t accumulate) {
    CTX_CALL(c.rebuild_phi(chunk, accumulate != 0));
}

extern "C" int culda_phi_head(culda_ctx *ctx) { CTX_CALL(c.phi_head()); }

extern "C" int culda_rebuild_theta(culda_ctx *ctx, int32_t chunk) {
    CTX_CALL({
        check_chunk(c, chunk);
        c.rebuild_theta(chunk);
    });
}

extern "C" int culda_sample(culda_ctx *ctx, int32_t chunk, int32_t iteration) {
    CTX_CALL(c.sample(chunk, iteration));
}

extern "C" int culda_loglik(culda_ctx *ctx, double *ll_per_token) {
    CTX_CALL({
        if (!ll_per_token) fail(CULDA_ERR_ARG, "null output");
        *ll_per_token = c.loglik();
    });
}

extern "C" int culda_iterate(culda_ctx *ctx, int32_t iteration, double *ll, double *seconds) {
    CTX_CALL({
        if (iteration < 1) fail(CULDA_ERR_ARG, "iteration must be >= 1");
        c.iterate(iteration, ll, seconds);
    });
}

extern "C" int culda_rebuild_all(culda_ctx *ctx) { CTX_CALL(c.rebuild_all()); }

extern "C" int culda_sync(culda_ctx *ctx) { CTX_CALL(c.sync()); }

extern "C" int culda_loglik_joint(culda_ctx *ctx, double *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        if (!c.theta_ready) fail(CULDA_ERR_STATE, "joint log-likelihood needs rebuilt counts");
        const int32_t K = c.prm.k, V = c.num_words;
        const double alpha = c.prm.alpha, beta = c.prm.beta;
        const int32_t G = 1024;
        double *part = (double *)c.dalloc(G * sizeof(double) * (1 + c.chunks.size()));
        hip_check(launch_joint_ll_words(c.nwk, V, K, std::lgamma(beta), c.prm.beta, part, G, c.stream), "joint");
        for (size_t k = 0; k < c.chunks.size(); ++k) {
            ChunkState &cs = c.chunks[k];
            hip_check(launch_joint_ll_docs(cs.meta, cs.theta, cs.doc_end - cs.doc_start, c.count_bits, c.prm.alpha,
                                           part + G * (1 + k), G, c.stream),
                      "joint");
        }
        std::vector<double> h(G * (1 + c.chunks.size()));
        hip_check(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c.stream), "copy");
        std::vector<int32_t> nk(K);
        hip_check(hipMemcpyAsync(nk.data(), c.nk, K * 4, hipMemcpyDeviceToHost, c.stream), "copy");
        hip_check(hipStreamSynchronize(c.stream), "sync");
        double s = 0;
        for (double v : h) s += v;
        const double vb = (double)V * beta, ka = (double)K * alpha;
        for (int32_t k = 0; k < K; ++k) s += std::lgamma(vb) - std::lgamma((double)nk[k] + vb);
        for (int64_t nd : c.doc_len) s += std::lgamma(ka) - std::lgamma((double)nd + ka);
        *out = s;
        // release the scratch partials
        (void)hipFree(part);
        c.allocations.erase(std::find(c.allocations.begin(), c.allocations.end(), (void *)part));
    });
}

extern "C" int culda_get_phi(culda_ctx *ctx, int32_t *n_wk, int32_t *n_k) {
    CTX_CALL({
        if (!n_wk || !n_k) fail(CULDA_ERR_ARG, "null output");
        HostPhi h = download_phi(c);
        std::memcpy(n_wk, h.nwk.data(), h.nwk.size() * 4);
        std::memcpy(n_k, h.nk.data(), h.nk.size() * 4);
    });
}

extern "C" int culda_get_phi_head(culda_ctx *ctx, float *head) {
    CTX_CALL({
        if (!head) fail(CULDA_ERR_ARG, "null output");
        if (!c.phi_ready) fail(CULDA_ERR_STATE, "phi head not prepared");
        const size_t n = (size_t)c.num_words * c.prm.k;
        float *d = (float *)c.dalloc(n * 4);
        hip_check(launch_phi_head(c.nwk, c.den, d, c.num_words, c.prm.k, c.prm.beta, c.stream), "head");
        hip_check(hipStreamSynchronize(c.stream), "sync");
        hip_check(hipMemcpy(head, d, n * 4, hipMemcpyDeviceToHost), "head copy");
        (void)hipFree(d);
        c.allocations.erase(std::find(c.allocations.begin(), c.allocations.end(), (void *)d));
    });
}

extern "C" int culda_get_theta(culda_ctx *ctx, int32_t chunk, int64_t *offsets, int32_t *counts, int16_t *topics) {
    CTX_CALL({
        if (!offsets) fail(CULDA_ERR_ARG, "null offsets");
        std::vector<int64_t> off;
        std::vector<int32_t> cnt;
        std::vector<int16_t> top;
        download_theta(c, chunk, off, cnt, top);
        std::memcpy(offsets, off.data(), off.size() * 8);
        if (counts) std::memcpy(counts, cnt.data(), cnt.size() * 4);
        if (topics) std::memcpy(topics, top.data(), top.size() * 2);
    });
}

// ModelPhi::savePhi (model/model_phi.cpp:161-191) wrote a text dump of the
// first 1000 words only (and nothing else); this writes every word, plus a
// binary file a program can load.
extern "C" int culda_save_phi(culda_ctx *ctx, const char *prefix) {
    CTX_CALL({
        if (!prefix) fail(CULDA_ERR_ARG, "null prefix");
        HostPhi h = download_phi(c);
        const int32_t V = c.num_words, K = c.prm.k;
        FILE *f = open_or_fail(std::string(prefix) + ".phi.bin", "wb");
        fwrite(&V, 4, 1, f);
        fwrite(&K, 4, 1, f);
        if (!h.nwk.empty()) fwrite(h.nwk.data(), 4, h.nwk.size(), f);
        fwrite(h.nk.data(), 4, h.nk.size(), f);
        fclose(f);
        f = open_or_fail(std::string(prefix) + ".phi.txt", "w");
        for (int32_t w = 0; w < V; ++w) {
            fprintf(f, "Word id:%d, ", w);
            for (int32_t k = 0; k < K; ++k) {
                if (k % 32 == 0) fprintf(f, "\n%2d:", k / 32);
                const int32_t v = h.nwk[(size_t)w * K + k];
                if (v == 0)
                    fprintf(f, "_,");
                else
                    fprintf(f, "%d,", v);
            }
            fprintf(f, "\n");
        }
        fclose(f);
    });
}

// ModelTheta::saveTheta / saveDoc exist only as commented code in the
// reference (model/model_theta.cpp:161-242); this writes CSR theta in
// global doc order (binary) and the saveDoc text layout.
extern "C" int culda_save_theta(culda_ctx *ctx, const char *prefix) {
    CTX_CALL({
        if (!prefix) fail(CULDA_ERR_ARG, "null prefix");
        if (!c.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        const int32_t D = c.num_docs, K = c.prm.k;
        std::vector<int64_t> goff(D + 1, 0);
        std::vector<int32_t> gcnt;
        std::vector<int16_t> gtop;
        for (size_t k = 0; k < c.chunks.size(); ++k) {
            std::vector<int64_t> off;
            std::vector<int32_t> cnt;
            std::vector<int16_t> top;
            download_theta(c, (int32_t)k, off, cnt, top);
            const int32_t ds = c.chunks[k].doc_start;
            for (size_t d = 0; d + 1 < off.size(); ++d) goff[ds + d + 1] = off[d + 1] - off[d];
            gcnt.insert(gcnt.end(), cnt.begin(), cnt.end());
            gtop.insert(gtop.end(), top.begin(), top.end());
        }
        for (int32_t d = 0; d < D; ++d) goff[d + 1] += goff[d];
        FILE *f = open_or_fail(std::string(prefix) + ".theta.bin", "wb");
        fwrite(&D, 4, 1, f);
        fwrite(&K, 4, 1, f);
        fwrite(goff.data(), 8, goff.size(), f);
        if (!gcnt.empty()) fwrite(gcnt.data(), 4, gcnt.size(), f);
        if (!gtop.empty()) fwrite(gtop.data(), 2, gtop.size(), f);
        fclose(f);
        f = open_or_fail(std::string(prefix) + ".theta.txt", "w");
        fprintf(f, "Doc, total, top topics\n");
        for (int32_t d = 0; d < D; ++d) {
            std::vector<std::pair<int32_t, int32_t>> row;
            int64_t tot = 0;
            for (int64_t j = goff[d]; j < goff[d + 1]; ++j) {
                row.emplace_back(gcnt[j], gtop[j]);
                tot += gcnt[j];
            }
            std::stable_sort(row.begin(), row.end(), [](auto &a, auto &b) { return a.first > b.first; });
            fprintf(f, "%4d| %lld ", d, (long long)tot);
            for (auto &pr : row) fprintf(f, "(%d,%d), ", pr.second, pr.first);
            fprintf(f, "\n");
        }
        fclose(f);
    });
}

static const char kCkptMagic[8] = {'C', 'U', 'L', 'D', 'A', 'C', 'K', '1'};

extern "C" int culda_save_checkpoint(culda_ctx *ctx, const char *path, int32_t next_iteration) {
    CTX_CALL({
        if (!path) fail(CULDA_ERR_ARG, "null path");
        if (!c.uploaded) fail(CULDA_ERR_STATE, "no corpus uploaded");
        hip_check(hipStreamSynchronize(c.stream), "sync");
        FILE *f = open_or_fail(path, "wb");
        const int32_t C = (int32_t)c.chunks.size(), K = c.prm.k;
        fwrite(kCkptMagic, 1, 8, f);
        fwrite(&C, 4, 1, f);
        fwrite(&K, 4, 1, f);
        fwrite(&c.prm.seed, 8, 1, f);
        fwrite(&next_iteration, 4, 1, f);
        for (auto &cs : c.chunks) {
            std::vector<uint16_t> z(cs.num_tokens);
            if (cs.num_tokens) hip_check(hipMemcpy(z.data(), cs.z, cs.num_tokens * 2, hipMemcpyDeviceToHost), "z");
            fwrite(&cs.num_tokens, 8, 1, f);
            if (!z.empty()) fwrite(z.data(), 2, z.size(), f);
        }
        fclose(f);
    });
}

extern "C" int culda_load_checkpoint(culda_ctx *ctx, const char *path, int32_t *next_iteration) {
    CTX_CALL({
        if (!path || !next_iteration) fail(CULDA_ERR_ARG, "null argument");
        if (!c.uploaded) fail(CULDA_ERR_STATE, "upload the corpus before loading a checkpoint");
        FILE *f = open_or_fail(path, "rb");
        char magic[8];
        int32_t C = 0, K = 0, it = 0;
        uint64_t seed = 0;
        bool ok = fread(magic, 1, 8, f) == 8 && !std::memcmp(magic, kCkptMagic, 8) && fread(&C, 4, 1, f) == 1 &&
                  fread(&K, 4, 1, f) == 1 && fread(&seed, 8, 1, f) == 1 && fread(&it, 4, 1, f) == 1;
        if (!ok || C != (int32_t)c.chunks.size() || K != c.prm.k || seed != c.prm.seed) {
            fclose(f);
            fail(CULDA_ERR_FORMAT, "%s is not a checkpoint of this corpus / k / seed", path);
        }
        std::vector<std::vector<uint16_t>> zs(C);
        for (int32_t k = 0; k < C; ++k) {
            int64_t n = 0;
            if (fread(&n, 8, 1, f) != 1 || n != c.chunks[k].num_tokens) {
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
        hip_check(hipStreamSynchronize(c.stream), "sync");
        for (int32_t k = 0; k < C; ++k)
            if (!zs[k].empty())
                hip_check(hipMemcpy(c.chunks[k].z, zs[k].data(), zs[k].size() * 2, hipMemcpyHostToDevice), "z");
        c.theta_ready = c.phi_ready = false;
        *next_iteration = it;
    });
}

extern "C" int culda_get_stats(culda_ctx *ctx, culda_stats *out) {
    CTX_CALL({
        if (!out) fail(CULDA_ERR_ARG, "null output");
        std::memset(out, 0, sizeof *out);
        out->num_tokens = c.num_tokens;
        for (auto &cs : c.chunks) {
            out->num_slots += cs.num_slots;
            out->theta_capacity += cs.theta_cap;
        }
        out->sum_doc_nnz = c.sum_doc_nnz;
        out->sum_token_nnz = c.sum_token_nnz;
        out->sampler_ms = c.sampler_ms;
        out->phi_ms = c.phi_ms;
        out->theta_ms = c.theta_ms;
        out->other_ms = c.other_ms;
        out->sampler_launches = c.sampler_launches;
    });
}

extern "C" int culda_sampler_bytes(culda_ctx *ctx, double *bytes) {
    CTX_CALL({
        if (!bytes) fail(CULDA_ERR_ARG, "null output");
        *bytes = c.last_sampler_bytes;
    });
}

extern "C" int culda_stream_copy(culda_ctx *ctx, int64_t n_bytes, int32_t repeats, double *gbps) {
    CTX_CALL({
        if (!gbps || n_bytes <= 0 || n_bytes % 16 || repeats < 1) fail(CULDA_ERR_ARG, "bad stream copy arguments");
        float4 *a = (float4 *)c.dalloc(n_bytes), *b = (float4 *)c.dalloc(n_bytes);
        hip_check(launch_fill_u32((uint32_t *)a, n_bytes / 4, 0x3f800000u, c.stream), "fill");
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(launch_copy(a, b, n_bytes / 16, c.stream), "copy warmup");
        hip_check(hipEventRecord(e0, c.stream), "event");
        for (int32_t r = 0; r < repeats; ++r) hip_check(launch_copy(a, b, n_bytes / 16, c.stream), "copy");
        hip_check(hipEventRecord(e1, c.stream), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * (double)n_bytes * repeats / (ms * 1e-3) / 1e9;
        (void)hipFree(a);
        (void)hipFree(b);
        c.allocations.erase(std::find(c.allocations.begin(), c.allocations.end(), (void *)a));
        c.allocations.erase(std::find(c.allocations.begin(), c.allocations.end(), (void *)b));
    });
}
