/*
 * culda_oracle.c -- CPU oracle for the plain collapsed-Gibbs LDA iteration.
 *
 * TEST INFRASTRUCTURE.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load this file's library; the product
 * (libculda.so) never links or calls it.
 *
 * It restates, in plain C99 on the host, the update schedule of CuLDA_CGS
 * (src_culda/) with the build's pinned decisions:
 *   - frozen counts for a whole sweep, own assignment included
 *     (kernel/lda_train_kernel.cu:104-259, train/SingleChunkSingleGPU.h:75-116);
 *   - two buckets S (theta nonzeros) and Q (alpha * phi head)
 *     (kernel/lda_train_kernel.cu:68-137, :145-253); S runs over the
 *     doc's theta *entries*: ascending topics, a topic with n_dk > 63
 *     split into ceil(n_dk / 63) entries (63, ..., 63, remainder), the
 *     layout of the device's 16-bit theta entries;
 *   - counts rebuilt from z after the sweep
 *     (kernel/lda_phi_kernel.cu:76-117, kernel/lda_theta_kernel.cu:5-87),
 *     summed over chunks (kernel/lda_phi_kernel.cu:162-183);
 *   - Philox4x64-10 keyed by (seed, global token id, sweep, purpose)
 *     instead of the reference's clock-seeded curand state (a sweep draw
 *     takes word g & 3 of the block of counter g >> 2);
 *   - FP32 head (n_wk + beta) / (n_k + beta V) (kernel/lda_phi_kernel.cu:26);
 *   - FP32 bucket arithmetic in a fixed order: blocks of 16 products
 *     whose pair sums fuse the even product into the odd one (fmaf, one
 *     rounding, pairs16) and whose totals are the balanced trees of the 8
 *     pair sums (block16), in-block prefixes from the pair sums scanned
 *     Kogge-Stone (pfx16), S block totals carried sequentially,
 *     Q block totals scanned Kogge-Stone over 64 lanes; t = u * (S + Q)
 *     picks the S bucket when t < S (target t) and otherwise the Q bucket
 *     (target t - S) -- the reference's u < S/(S+Q) test and
 *     (u - S/(S+Q)) * (S+Q) target without the division; two-level search
 *     (first block, then first element) with last-element fallbacks;
 *   - FP64 log-likelihood term log((S+Q) / (n_d + K alpha))
 *     (kernel/lda_train_kernel.cu:258, accumulated in FP64 instead of
 *     float partials truncated to int, :101, :289-290).
 * Mode 1 ("exact") evaluates the same distribution in FP64 with a plain
 * linear inverse CDF and flags draws whose target lies within 1e-6 (relative
 * to S+Q) of a bucket boundary; the FP32 ("faithful") mode is what the
 * device reproduces bit for bit.
 *
 * Build: gcc -O2 -mfma -fopenmp -ffp-contract=off (x86-64 SSE arithmetic, so
 * every float operation is rounded to FP32; fmaf is the hardware's single-
 * rounding FMA, and nothing is contracted into one implicitly).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ Philox */
static inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

/* Philox4x64-10 (Salmon et al., SC'11); numpy.random.Philox uses the same
 * constants and round function. */
void oc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t p0 = 0xD2E7470EE14C6C93ULL * c0, h0 = mulhi64(0xD2E7470EE14C6C93ULL, c0);
        uint64_t p1 = 0xCA5A826395121157ULL * c2, h1 = mulhi64(0xCA5A826395121157ULL, c2);
        uint64_t n0 = h1 ^ c1 ^ k0, n1 = p1, n2 = h0 ^ c3 ^ k1, n3 = p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static inline uint64_t token_word(uint64_t seed, uint64_t gid, uint64_t sweep, uint64_t purpose) {
    uint64_t c[4] = {gid, sweep, purpose, 0}, k[2] = {seed, 0}, o[4];
    oc_philox4x64_10(c, k, o);
    return o[0];
}

/* Sweep draw of token g: word g & 3 of Philox((g >> 2, sweep, 1, 0), (seed, 0)),
 * one block per four consecutive global token ids. */
static inline uint64_t sweep_word(uint64_t seed, uint64_t gid, uint64_t sweep) {
    uint64_t c[4] = {gid >> 2, sweep, 1, 0}, k[2] = {seed, 0}, o[4];
    oc_philox4x64_10(c, k, o);
    return o[gid & 3];
}

/* ------------------------------------------------------------------ corpus */
typedef struct oc_corpus {
    int32_t V, D, C, K;
    const int64_t *doc_cum;        /* [D] cumulative doc lengths            */
    const int32_t *chunk_doc_end;  /* [C]                                   */
    const int64_t *const *word_idx;/* C x [V+1] chunk-local word offsets    */
    const int32_t *const *tokens;  /* C x [N_c] global doc id per token     */
    uint16_t *const *z;            /* C x [N_c] topics                      */
} oc_corpus;

static int64_t doc_len(const oc_corpus *c, int32_t d) { return c->doc_cum[d] - (d ? c->doc_cum[d - 1] : 0); }

/* global token id of every token: g = sum_{w'<w} cnt(w') + sum_{c'<c}
 * cnt_c'(w) + position in the chunk's word range.  With each word's docs
 * ascending (src_format/format.cpp:86, :146-156) this is the token's index
 * in the one-chunk word-major layout. */
static int64_t *word_base(const oc_corpus *c) {
    int64_t *pre = (int64_t *)calloc((size_t)c->V + 1, sizeof(int64_t));
    for (int32_t k = 0; k < c->C; ++k)
        for (int32_t w = 0; w < c->V; ++w) pre[w + 1] += c->word_idx[k][w + 1] - c->word_idx[k][w];
    for (int32_t w = 0; w < c->V; ++w) pre[w + 1] += pre[w];
    return pre;
}

/* z_i = floor(x * K / 2^64), x = Philox((g, 0, 0, 0), (seed, 0))[0].
 * (The reference draws rand() % K after srand(time), doc_chunk.cpp:259-264.) */
void oc_init_z(const oc_corpus *c, uint64_t seed) {
    int64_t *pre = word_base(c);
    int64_t *run = (int64_t *)malloc(sizeof(int64_t) * (size_t)(c->V + 1));
    memcpy(run, pre, sizeof(int64_t) * (size_t)(c->V + 1));
    for (int32_t k = 0; k < c->C; ++k) {
        for (int32_t w = 0; w < c->V; ++w) {
            const int64_t b = c->word_idx[k][w], e = c->word_idx[k][w + 1];
            for (int64_t i = b; i < e; ++i)
                c->z[k][i] = (uint16_t)mulhi64(token_word(seed, (uint64_t)(run[w] + i - b), 0, 0), (uint64_t)c->K);
            run[w] += e - b;
        }
    }
    free(run);
    free(pre);
}

/* ------------------------------------------------------------------ counts */
/* n_wk, n_k summed over chunks (kernel/lda_phi_kernel.cu:94-116, :162-183).
 * Rows are independent (one word each); n_k is summed per thread and then
 * added, integer sums, so the result is the same for any thread count. */
void oc_rebuild_phi(const oc_corpus *c, int32_t *nwk, int32_t *nk) {
    const int32_t K = c->K;
    memset(nk, 0, sizeof(int32_t) * K);
#pragma omp parallel
    {
        int32_t *lk = (int32_t *)calloc((size_t)K, sizeof(int32_t));
#pragma omp for schedule(dynamic, 64)
        for (int32_t w = 0; w < c->V; ++w) {
            int32_t *row = nwk + (size_t)w * K;
            memset(row, 0, sizeof(int32_t) * K);
            for (int32_t k = 0; k < c->C; ++k)
                for (int64_t i = c->word_idx[k][w]; i < c->word_idx[k][w + 1]; ++i) row[c->z[k][i]]++;
            for (int32_t k = 0; k < K; ++k) lk[k] += row[k];
        }
#pragma omp critical
        for (int32_t k = 0; k < K; ++k) nk[k] += lk[k];
        free(lk);
    }
}

#define OC_THETA_MAX_COUNT 63

/* Most 16-entry blocks any doc's theta row has (from th_off). */
static int64_t max_blocks(const oc_corpus *c, const int64_t *th_off) {
    int64_t m = 0;
    for (int32_t d = 0; d < c->D; ++d) {
        const int64_t nb = (th_off[d + 1] - th_off[d] + 15) / 16;
        if (nb > m) m = nb;
    }
    return m;
}

/* Entries of one topic with count c: ceil(c / 63). */
static inline int64_t theta_entries(int32_t c) { return (c + OC_THETA_MAX_COUNT - 1) / OC_THETA_MAX_COUNT; }

/* Theta rows, global doc order, ascending topics (kernel/lda_theta_kernel.cu:44-84),
 * as entries: a topic with count c > 63 takes ceil(c / 63) consecutive
 * entries, 63 each and the remainder last.  th_off has D+1 entries; th_cnt /
 * th_top have room for sum_d min(n_d, K + ceil(n_d / 63)) entries.  Returns
 * sum_d n_d * entries(d). */
int64_t oc_rebuild_theta(const oc_corpus *c, int64_t *th_off, int32_t *th_cnt, int16_t *th_top) {
    const int32_t K = c->K, D = c->D;
    /* doc-major list of token topics */
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    pos[0] = 0;
    for (int32_t d = 0; d < D; ++d) pos[d + 1] = c->doc_cum[d];
    const int64_t N = D ? c->doc_cum[D - 1] : 0;
    uint16_t *dz = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(N + 1));
    int64_t *nxt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(D + 1));
    memcpy(nxt, pos, sizeof(int64_t) * (size_t)(D + 1));
    for (int32_t k = 0; k < c->C; ++k) {
        const int64_t n = c->word_idx[k][c->V];
        for (int64_t i = 0; i < n; ++i) dz[nxt[c->tokens[k][i]]++] = c->z[k][i];
    }
    free(nxt);
    int64_t *nnz = (int64_t *)calloc((size_t)D + 1, sizeof(int64_t));
    int64_t total = 0;
#pragma omp parallel
    {
        int32_t *h = (int32_t *)calloc((size_t)K, sizeof(int32_t));
#pragma omp for schedule(dynamic, 256)
        for (int32_t d = 0; d < D; ++d) {
            int64_t m = 0;
            for (int64_t t = pos[d]; t < pos[d + 1]; ++t) h[dz[t]]++;
            for (int32_t k = 0; k < K; ++k) m += theta_entries(h[k]);
            nnz[d] = m;
            for (int64_t t = pos[d]; t < pos[d + 1]; ++t) h[dz[t]] = 0;
        }
        free(h);
    }
    th_off[0] = 0;
    for (int32_t d = 0; d < D; ++d) th_off[d + 1] = th_off[d] + nnz[d];
#pragma omp parallel
    {
        int32_t *h = (int32_t *)calloc((size_t)K, sizeof(int32_t));
#pragma omp for schedule(dynamic, 256) reduction(+ : total)
        for (int32_t d = 0; d < D; ++d) {
            for (int64_t t = pos[d]; t < pos[d + 1]; ++t) h[dz[t]]++;
            int64_t o = th_off[d];
            for (int32_t k = 0; k < K; ++k) {
                for (int32_t c = h[k]; c > 0; c -= OC_THETA_MAX_COUNT) {
                    th_cnt[o] = c > OC_THETA_MAX_COUNT ? OC_THETA_MAX_COUNT : c;
                    th_top[o] = (int16_t)k;
                    ++o;
                }
                h[k] = 0;
            }
            total += (pos[d + 1] - pos[d]) * nnz[d];
        }
        free(h);
    }
    free(nnz);
    free(dz);
    free(pos);
    return total;
}

/* FP32 head, (kernel/lda_phi_kernel.cu:26); the reference then rounds to FP16 (:29). */
void oc_phi_head(int32_t V, int32_t K, float beta, const int32_t *nwk, const int32_t *nk, float *head) {
    const float bv = beta * (float)V;
    for (int64_t w = 0; w < V; ++w)
        for (int32_t k = 0; k < K; ++k) {
            const float den = (float)nk[k] + bv;
            head[w * K + k] = ((float)nwk[w * K + k] + beta) / den;
        }
}

/* ------------------------------------------------------------------ sweep */
/* Pair sums of 16 products x_i * y_i: q_m = fma(x_2m, y_2m, fl(x_2m+1 *
 * y_2m+1)) -- the even product fused into the odd one, one rounding. */
static void pairs16(const float *x, const float *y, float *q) {
    for (int m = 0; m < 8; ++m) q[m] = fmaf(x[2 * m], y[2 * m], x[2 * m + 1] * y[2 * m + 1]);
}

/* Inclusive Kogge-Stone scan of 8 pair sums, offsets 1, 2, 4. */
static void ks8(float *P) {
    for (int d = 1; d < 8; d <<= 1) {
        float z[8];
        for (int m = 0; m < 8; ++m) z[m] = m >= d ? P[m] + P[m - d] : P[m];
        memcpy(P, z, sizeof z);
    }
}

/* Block total of 16 products: the balanced tree of the 8 pair sums
 * ((q0 + q1) + (q2 + q3)) + ((q4 + q5) + (q6 + q7)), the last element of
 * their Kogge-Stone scan. */
static float block16(const float *x, const float *y) {
    float q[8];
    pairs16(x, y, q);
    ks8(q);
    return q[7];
}

/* In-block prefixes of 16 products into c: the pair sums' scan P_m,
 * c_2m+1 = P_m and c_2m = fma(x_2m, y_2m, P_m-1) (P_-1 = 0).  c_15 is the
 * block total. */
static void pfx16(const float *x, const float *y, float *c) {
    float P[8];
    pairs16(x, y, P);
    ks8(P);
    for (int m = 0; m < 8; ++m) {
        c[2 * m] = fmaf(x[2 * m], y[2 * m], m ? P[m - 1] : 0.0f);
        c[2 * m + 1] = P[m];
    }
}

/* Inclusive Kogge-Stone scan of 64 values, offsets 1 .. 32. */
static void ks64(float *x) {
    for (int d = 1; d < 64; d <<= 1) {
        float y[64];
        for (int m = 0; m < 64; ++m) y[m] = m >= d ? x[m] + x[m - d] : x[m];
        memcpy(x, y, sizeof y);
    }
}

typedef struct oc_sweep_out {
    double ll_sum;        /* sum of FP64 LL terms (faithful mode)           */
    double ll_sum_exact;  /* same terms with S, Q in FP64                   */
    int64_t boundary;     /* draws within 1e-6 (relative) of a boundary     */
    int64_t mismatch;     /* faithful topic != exact topic                  */
    int64_t s_bucket;     /* draws that took the S bucket                   */
    int64_t tokens;       /* tokens swept                                   */
} oc_sweep_out;

/* One sweep with frozen counts (nwk, nk, theta rows) over the tokens of
 * words [w_lo, w_hi) (every word: 0, V).  Writes the faithful FP32 topics
 * into c->z (the sweep never reads z).  With want_exact, every draw is also
 * evaluated in FP64 by a plain linear inverse CDF: exact_z (optional, C
 * arrays) receives those topics, flags (optional) 1 for boundary draws, and
 * out counts them; without it only the FP32 path runs (the CPU baseline). */
void oc_sweep_range(const oc_corpus *c, float alpha, float beta, uint64_t seed, uint32_t sweep, const int32_t *nwk,
                    const int32_t *nk, const int64_t *th_off, const int32_t *th_cnt, const int16_t *th_top,
                    int32_t w_lo, int32_t w_hi, int want_exact, uint16_t *const *exact_z, uint8_t *const *flags,
                    oc_sweep_out *out) {
    const int32_t K = c->K, V = c->V, nbq = K / 16;
    const float bv = beta * (float)V;
    const double k_alpha = (double)K * (double)alpha;
    if (w_lo < 0) w_lo = 0;
    if (w_hi > V) w_hi = V;
    float *den = (float *)malloc(sizeof(float) * K);
    double *den_d = (double *)malloc(sizeof(double) * K);
    for (int32_t k = 0; k < K; ++k) {
        den[k] = (float)nk[k] + bv;
        den_d[k] = (double)nk[k] + (double)beta * (double)V;
    }
    int64_t *pre = word_base(c);
    /* per chunk running offsets: base of word w in chunk k = pre[w] + sum_{k'<k} cnt */
    int64_t *cbase = (int64_t *)malloc(sizeof(int64_t) * (size_t)c->C * V);
    for (int32_t w = 0; w < V; ++w) {
        int64_t r = pre[w];
        for (int32_t k = 0; k < c->C; ++k) {
            cbase[(size_t)k * V + w] = r;
            r += c->word_idx[k][w + 1] - c->word_idx[k][w];
        }
    }
    double ll = 0, lle = 0;
    int64_t nbound = 0, nmis = 0, ns = 0, ntot = 0;
    const int64_t nbmax = max_blocks(c, th_off);
#pragma omp parallel reduction(+ : ll, lle, nbound, nmis, ns, ntot)
    {
        float *ph = (float *)malloc(sizeof(float) * K);
        double *phd = (double *)malloc(sizeof(double) * K);
        float rq[64];
        float *rs = (float *)malloc(sizeof(float) * (size_t)(nbmax + 1)); /* S prefix after each block */
#pragma omp for schedule(dynamic, 16)
        for (int32_t w = w_lo; w < w_hi; ++w) {
            int64_t ntok = 0;
            for (int32_t k = 0; k < c->C; ++k) ntok += c->word_idx[k][w + 1] - c->word_idx[k][w];
            if (!ntok) continue;
            ntot += ntok;
            /* head row and Q table */
            double Qd = 0;
            for (int32_t k = 0; k < K; ++k) ph[k] = ((float)nwk[(size_t)w * K + k] + beta) / den[k];
            if (want_exact)
                for (int32_t k = 0; k < K; ++k) {
                    phd[k] = ((double)nwk[(size_t)w * K + k] + (double)beta) / den_d[k];
                    Qd += (double)alpha * phd[k];
                }
            float av[16];
            for (int m = 0; m < 16; ++m) av[m] = alpha;
            for (int b = 0; b < 64; ++b) rq[b] = b < nbq ? block16(av, ph + b * 16) : 0.0f;
            ks64(rq);
            const float Q = rq[nbq - 1];
            for (int32_t ck = 0; ck < c->C; ++ck) {
                const int64_t b0 = c->word_idx[ck][w], e0 = c->word_idx[ck][w + 1];
                for (int64_t i = b0; i < e0; ++i) {
                    const int32_t d = c->tokens[ck][i];
                    const int64_t o = th_off[d];
                    const int64_t nnz = th_off[d + 1] - o;
                    const uint64_t g = (uint64_t)(cbase[(size_t)ck * V + w] + (i - b0));
                    const float u = (float)((uint32_t)(sweep_word(seed, g, sweep) >> 40) + 1u) * 5.9604644775390625e-8f;
                    /* S: blocks of 16 */
                    const int64_t nb = (nnz + 15) / 16;
                    float R = 0.0f;
                    for (int64_t b = 0; b < nb; ++b) {
                        float x[16], y[16];
                        for (int m = 0; m < 16; ++m) {
                            const int64_t j = b * 16 + m;
                            x[m] = j < nnz ? (float)th_cnt[o + j] : 0.0f; /* padding: exact zero products */
                            y[m] = j < nnz ? ph[th_top[o + j]] : 0.0f;
                        }
                        R = R + block16(x, y);
                        rs[b + 1] = R;
                    }
                    const float S = R, SQ = S + Q, tu = u * SQ;
                    int32_t topic;
                    if (tu < S) {
                        ns++;
                        const float t = tu;
                        int64_t bsel = -1;
                        for (int64_t b = 0; b < nb; ++b)
                            if (t < rs[b + 1]) {
                                bsel = b;
                                break;
                            }
                        if (bsel < 0) {
                            topic = th_top[o + nnz - 1];
                        } else {
                            float x[16], y[16], p[16];
                            for (int m = 0; m < 16; ++m) {
                                const int64_t j = bsel * 16 + m;
                                x[m] = j < nnz ? (float)th_cnt[o + j] : 0.0f;
                                y[m] = j < nnz ? ph[th_top[o + j]] : 0.0f;
                            }
                            pfx16(x, y, p);
                            const int64_t nvalid = nnz - bsel * 16 < 16 ? nnz - bsel * 16 : 16;
                            int64_t msel = nvalid - 1;
                            for (int64_t m = 0; m < nvalid; ++m) {
                                const float C = bsel == 0 ? p[m] : rs[bsel] + p[m];
                                if (t < C) {
                                    msel = m;
                                    break;
                                }
                            }
                            topic = th_top[o + bsel * 16 + msel];
                        }
                    } else {
                        const float t = tu - S;
                        int bsel = -1;
                        for (int b = 0; b < nbq; ++b)
                            if (t < rq[b]) {
                                bsel = b;
                                break;
                            }
                        if (bsel < 0) {
                            topic = K - 1;
                        } else {
                            float v[16];
                            pfx16(av, ph + bsel * 16, v);
                            int msel = 15;
                            for (int m = 0; m < 16; ++m) {
                                const float C = bsel == 0 ? v[m] : rq[bsel - 1] + v[m];
                                if (t < C) {
                                    msel = m;
                                    break;
                                }
                            }
                            topic = bsel * 16 + msel;
                        }
                    }
                    c->z[ck][i] = (uint16_t)topic;
                    ll += log((double)SQ / ((double)doc_len(c, d) + k_alpha));
                    if (!want_exact) continue;

                    /* exact FP64 inverse CDF over the concatenated S then Q masses */
                    double Sd = 0;
                    for (int64_t j = 0; j < nnz; ++j) Sd += (double)th_cnt[o + j] * phd[th_top[o + j]];
                    const double SQd = Sd + Qd, ud = (double)u;
                    lle += log(SQd / ((double)doc_len(c, d) + k_alpha));
                    const double tol = 1e-6 * SQd;
                    const double target = ud * SQd;
                    int32_t ex = K - 1;
                    int bound = fabs(target - Sd) <= tol;
                    if (target < Sd) {
                        double acc = 0;
                        ex = th_top[o + nnz - 1];
                        for (int64_t j = 0; j < nnz; ++j) {
                            const double lo = acc;
                            acc += (double)th_cnt[o + j] * phd[th_top[o + j]];
                            if (target < acc) {
                                ex = th_top[o + j];
                                bound |= fabs(target - lo) <= tol || fabs(target - acc) <= tol;
                                break;
                            }
                        }
                    } else {
                        double acc = Sd;
                        for (int32_t k = 0; k < K; ++k) {
                            const double lo = acc;
                            acc += (double)alpha * phd[k];
                            if (target < acc) {
                                ex = k;
                                bound |= fabs(target - lo) <= tol || fabs(target - acc) <= tol;
                                break;
                            }
                        }
                    }
                    if (exact_z) exact_z[ck][i] = (uint16_t)ex;
                    if (flags) flags[ck][i] = (uint8_t)bound;
                    nbound += bound;
                    nmis += ex != topic;
                }
            }
        }
        free(ph);
        free(rs);
        free(phd);
    }
    free(cbase);
    free(pre);
    free(den);
    free(den_d);
    if (out) {
        out->ll_sum = ll;
        out->ll_sum_exact = want_exact ? lle : 0.0;
        out->boundary = nbound;
        out->mismatch = nmis;
        out->s_bucket = ns;
        out->tokens = ntot;
    }
}

/* Every word, FP32 choice plus the FP64 shadow. */
void oc_sweep(const oc_corpus *c, float alpha, float beta, uint64_t seed, uint32_t sweep, const int32_t *nwk,
              const int32_t *nk, const int64_t *th_off, const int32_t *th_cnt, const int16_t *th_top,
              uint16_t *const *exact_z, uint8_t *const *flags, oc_sweep_out *out) {
    oc_sweep_range(c, alpha, beta, seed, sweep, nwk, nk, th_off, th_cnt, th_top, 0, c->V, 1, exact_z, flags, out);
}

/* Every word, the FP32 path only (timing): the same code as oc_sweep with
 * the FP64 shadow switched off. */
void oc_sweep_fast(const oc_corpus *c, float alpha, float beta, uint64_t seed, uint32_t sweep, const int32_t *nwk,
                   const int32_t *nk, const int64_t *th_off, const int32_t *th_cnt, const int16_t *th_top,
                   double *ll_sum) {
    oc_sweep_out o;
    oc_sweep_range(c, alpha, beta, seed, sweep, nwk, nk, th_off, th_cnt, th_top, 0, c->V, 0, NULL, NULL, &o);
    *ll_sum = o.ll_sum;
}

/* Joint log p(w, z) in FP64 (Griffiths & Steyvers 2004, eq. 2-3 collapsed
 * form); the reference declares LDALikelihood but never defines it
 * (train/lda_train.h:16). */
double oc_loglik_joint(const oc_corpus *c, float alpha, float beta, const int32_t *nwk, const int32_t *nk,
                       const int64_t *th_off, const int32_t *th_cnt, const int16_t *th_top) {
    const int32_t K = c->K, V = c->V, D = c->D;
    const double a = alpha, b = beta, vb = (double)V * b, ka = (double)K * a;
    double s = 0;
    for (int32_t k = 0; k < K; ++k) s += lgamma(vb) - lgamma((double)nk[k] + vb);
    for (int64_t i = 0; i < (int64_t)V * K; ++i)
        if (nwk[i]) s += lgamma((double)nwk[i] + b) - lgamma(b);
    for (int32_t d = 0; d < D; ++d) {
        s += lgamma(ka) - lgamma((double)doc_len(c, d) + ka);
        /* n_dk = the sum over the topic's consecutive entries */
        for (int64_t j = th_off[d]; j < th_off[d + 1];) {
            int64_t n = 0, i = j;
            for (; i < th_off[d + 1] && th_top[i] == th_top[j]; ++i) n += th_cnt[i];
            s += lgamma((double)n + a) - lgamma(a);
            j = i;
        }
    }
    return s;
}

int oc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
