"""Pins the CPU oracle against independent checks: numpy's Philox golden
vectors, numpy histograms, closed forms, the exact FP64 conditional, a
chi-square test and scipy's gammaln.  No GPU."""
import json
from pathlib import Path

import numpy as np
import pytest
from scipy import special, stats

import oracle
from helpers import chunks_from_docs, random_docs

GOLD = json.loads((Path(__file__).parent / "golden" / "philox4x64_10.json").read_text())
M64 = (1 << 64) - 1


def test_philox_known_answers():
    for case in GOLD["cases"]:
        ctr = [int(x) for x in case["counter"]]
        key = [int(x) for x in case["key"]]
        got = oracle.philox4x64_10(ctr, key)
        assert [int(x) for x in got] == [int(x) for x in case["block"]], case
    # consecutive blocks of one numpy stream: counters 1, 2, 3
    key = [int(x) for x in GOLD["stream_key"]]
    stream = [int(x) for x in GOLD["stream"]]
    for b in range(3):
        assert [int(x) for x in oracle.philox4x64_10([b + 1, 0, 0, 0], key)] == stream[4 * b:4 * b + 4]


def _small(seed=0, V=60, D=30, mean=25, chunks=1, K=32, empty=2):
    rng = np.random.default_rng(seed)
    docs = random_docs(rng, V, D, mean, empty_docs=empty)
    arrays = chunks_from_docs(V, docs, chunks)
    return docs, arrays, oracle.OracleModel(V, K, *arrays)


def test_init_z_matches_numpy_philox():
    docs, (doc_cum, ends, wi, wt), om = _small(chunks=3, K=48)
    seed = 77
    om.init_z(seed)
    V = 60
    # global id: word-major over all chunks, chunk order within a word
    gid = 0
    for w in range(V):
        for c in range(3):
            for i in range(wi[c][w], wi[c][w + 1]):
                x = int(oracle.philox4x64_10([gid, 0, 0, 0], [seed, 0])[0])
                assert om.z[c][i] == (x * 48) >> 64
                gid += 1
    # numpy's own stream for one token id (counter convention: numpy adds 1)
    from numpy.random import Philox
    x = int(Philox(counter=[4, 0, 0, 0], key=[seed, 0]).random_raw(1)[0])
    assert x == int(oracle.philox4x64_10([5, 0, 0, 0], [seed, 0])[0])


@pytest.mark.parametrize("chunks", [1, 2, 4])
def test_rebuild_matches_numpy_histograms(chunks):
    docs, (doc_cum, ends, wi, wt), om = _small(seed=1, chunks=chunks)
    rng = np.random.default_rng(9)
    K, V = 32, 60
    zs = [rng.integers(0, K, size=t.size).astype(np.uint16) for t in wt]
    om.set_z(zs)
    om.rebuild()
    # n_wk / n_k from scratch with numpy
    nwk = np.zeros((V, K), np.int64)
    ndk = np.zeros((len(docs), K), np.int64)
    for c in range(chunks):
        words = np.repeat(np.arange(V), np.diff(wi[c]))
        np.add.at(nwk, (words, zs[c]), 1)
        np.add.at(ndk, (wt[c], zs[c]), 1)
    assert np.array_equal(om.nwk.reshape(V, K), nwk)
    assert np.array_equal(om.nk, nwk.sum(0))
    lens = np.array([len(d) for d in docs])
    tok_nnz = 0
    for d in range(len(docs)):
        cnt, top = om.theta_rows(d)
        nz = np.flatnonzero(ndk[d])
        assert np.array_equal(top, nz)                   # ascending topics
        assert np.array_equal(cnt, ndk[d][nz])
        assert cnt.sum() == lens[d]
        tok_nnz += lens[d] * int(((ndk[d][nz] + 62) // 63).sum())
    assert om.token_nnz == tok_nnz
    # validPhi invariants (model/model_phi.cpp:194-280)
    assert np.array_equal(nwk.sum(1), np.concatenate([np.diff(w) for w in wi]).reshape(chunks, V).sum(0))


def test_phi_head_closed_form():
    docs, arrays, om = _small(seed=2, K=64)
    rng = np.random.default_rng(3)
    om.set_z([rng.integers(0, 64, size=t.size).astype(np.uint16) for t in arrays[3]])
    om.rebuild()
    beta = np.float32(0.01)
    h = om.phi_head(float(beta))
    nwk = om.nwk.reshape(60, 64)
    # numpy float32: each operation rounded to FP32
    ref = (nwk.astype(np.float32) + beta) / (om.nk.astype(np.float32) + beta * np.float32(60))
    assert np.array_equal(h, ref)
    exact = (nwk + 0.01) / (om.nk + 0.01 * 60)
    assert np.allclose(h, exact, rtol=2e-7, atol=0)
    # each row of the head over K sums like sum_k (n_wk+b)/(n_k+bV)
    assert np.all(h > 0)


def _conditional(om, d, w, alpha, beta):
    """Exact conditional of the frozen-count sweep, FP64: the concatenation
    [S entries (ascending topic) ..., Q entries over all topics]."""
    K, V = om.K, om.V
    nwk = om.nwk.reshape(V, K).astype(np.float64)
    phi = (nwk[w] + beta) / (om.nk + beta * V)
    cnt, top = om.theta_rows(d)
    s = cnt * phi[top]
    q = alpha * phi
    return np.concatenate([s, q]), np.concatenate([top, np.arange(K)])


def test_sweep_matches_exact_inverse_cdf():
    """Faithful FP32 choice == independent FP64 inverse CDF away from boundaries."""
    docs, (doc_cum, ends, wi, wt), om = _small(seed=4, V=80, D=40, mean=60, K=64)
    alpha, beta, seed, sweep = float(np.float32(0.3)), float(np.float32(0.05)), 5, 2
    rng = np.random.default_rng(4)
    om.set_z([rng.integers(0, 64, size=wt[0].size).astype(np.uint16)])
    om.rebuild()
    gid_words = np.repeat(np.arange(80), np.diff(wi[0]))
    om_before = oracle.OracleModel(80, 64, doc_cum, ends, wi, wt, z=om.z)
    om_before.rebuild()
    out, ex, fl = om.sweep(alpha, beta, seed, sweep, want_exact=True)
    checked = 0
    for i in range(wt[0].size):
        d, w = int(wt[0][i]), int(gid_words[i])
        x = int(oracle.philox4x64_10([i >> 2, sweep, 1, 0], [seed, 0])[i & 3])
        u = np.float32(((x >> 40) + 1) * 2.0**-24)
        mass, topics = _conditional(om_before, d, w, alpha, beta)
        cum = np.cumsum(mass)
        t = float(u) * cum[-1]
        j = int(np.searchsorted(cum, t, side="right"))
        j = min(j, len(cum) - 1)
        margin = min(abs(t - cum[j]), abs(t - (cum[j - 1] if j else 0.0)))
        if margin > 1e-5 * cum[-1]:
            assert om.z[0][i] == topics[j], (i, d, w)
            assert ex[0][i] == topics[j]
            checked += 1
    assert checked > 0.95 * wt[0].size
    assert out.mismatch <= out.boundary


def test_sweep_distribution_chi_square():
    """Many tokens of one (doc, word) share one conditional: their new topics
    must follow p_k ∝ (n_dk + alpha) * phi_wk."""
    K, V = 16, 3
    n_rep = 20000
    docs = [np.array([0] * n_rep + [1] * 50 + [2] * 30)]
    arrays = chunks_from_docs(V, docs, 1)
    om = oracle.OracleModel(V, K, *arrays)
    rng = np.random.default_rng(0)
    z0 = rng.integers(0, 4, size=arrays[3][0].size).astype(np.uint16) * 3  # topics 0, 3, 6, 9
    om.set_z([z0])
    om.rebuild()
    alpha, beta = 200.0, 0.5
    mass, topics = _conditional(om, 0, 0, alpha, beta)
    p = np.zeros(K)
    np.add.at(p, topics, mass)
    p /= p.sum()
    om.sweep(alpha, beta, 123, 1)
    got = np.bincount(om.z[0][:n_rep], minlength=K)
    chi2, pval = stats.chisquare(got, p * n_rep)
    assert pval > 1e-4, (got, p * n_rep)


def test_special_cases():
    K, V = 32, 4
    # one topic in the doc and alpha -> 0: the S bucket takes every draw
    docs = [np.array([0, 1, 2, 3] * 10)]
    arrays = chunks_from_docs(V, docs, 1)
    om = oracle.OracleModel(V, K, *arrays)
    om.set_z([np.full(40, 7, np.uint16)])
    om.rebuild()
    out, _, _ = om.sweep(1e-30, 0.01, 1, 1)
    assert np.all(om.z[0] == 7) and out.s_bucket == 40
    # huge beta: phi is flat, p_k ∝ n_dk + alpha; S/Q split frequency = n_d / (n_d + K alpha)
    docs = [np.arange(V).repeat(500)]
    arrays = chunks_from_docs(V, docs, 1)
    om = oracle.OracleModel(V, K, *arrays)
    om.set_z([np.zeros(2000, np.uint16)])
    om.rebuild()
    alpha = 2000.0 / K  # S and Q equally heavy
    out, _, _ = om.sweep(alpha, 1e6, 3, 1)
    assert abs(out.s_bucket / 2000 - 0.5) < 0.05
    q_draws = om.z[0][om.z[0] != 0]
    assert abs(q_draws.size / 2000 - 0.5 * (K - 1) / K) < 0.05


def test_loglik_closed_forms():
    docs, (doc_cum, ends, wi, wt), om = _small(seed=6, V=70, D=35, mean=40, K=32)
    rng = np.random.default_rng(6)
    om.set_z([rng.integers(0, 32, size=wt[0].size).astype(np.uint16)])
    om.rebuild()
    alpha, beta = 0.25, float(np.float32(0.02))  # the oracle takes FP32 parameters
    before = oracle.OracleModel(70, 32, doc_cum, ends, wi, wt, z=om.z)
    before.rebuild()
    out, _, _ = om.sweep(alpha, beta, 1, 1)
    words = np.repeat(np.arange(70), np.diff(wi[0]))
    lens = np.diff(np.concatenate([[0], doc_cum]))
    ll = 0.0
    for i in range(wt[0].size):
        mass, _ = _conditional(before, int(wt[0][i]), int(words[i]), alpha, beta)
        ll += np.log(mass.sum() / (lens[wt[0][i]] + 32 * alpha))
    assert abs(out.ll_sum_exact - ll) <= 1e-10 * abs(ll)
    assert abs(out.ll_sum - ll) <= 1e-5 * abs(ll)
    # joint log p(w, z) against scipy.special.gammaln
    V, K = 70, 32
    nwk = before.nwk.reshape(V, K)
    ndk = np.zeros((len(docs), K))
    np.add.at(ndk, (wt[0], before.z[0]), 1)
    g = special.gammaln
    ref = (K * (g(V * beta) - V * g(beta)) + (g(nwk + beta)).sum() - g(before.nk + V * beta).sum()
           + len(docs) * (g(K * alpha) - K * g(alpha)) + g(ndk + alpha).sum() - g(lens + K * alpha).sum())
    got = before.loglik_joint(alpha, beta)
    assert abs(got - ref) <= 1e-9 * abs(ref)


def test_theta_entries_split_counts_above_63():
    """A topic with n_dk > 63 takes ceil(n_dk / 63) entries (63, ..., 63,
    remainder): the entry list S is summed over.  Merged rows, the joint
    log-likelihood (gammaln of the merged n_dk) and a sweep's exact FP64
    log-likelihood are those of the merged counts."""
    K, V = 32, 5
    counts = {3: 300, 4: 63, 9: 64, 20: 126, 31: 1}
    z = np.concatenate([np.full(c, k, np.uint16) for k, c in counts.items()])
    docs = [np.arange(z.size) % V, np.array([0, 1, 2])]
    arrays = chunks_from_docs(V, docs, 1)
    om = oracle.OracleModel(V, K, *arrays)
    # z is word-major: give doc 0's tokens the planted topics in that order
    words = np.repeat(np.arange(V), np.diff(arrays[2][0]))
    zw = np.zeros(arrays[3][0].size, np.uint16)
    d0 = np.flatnonzero(arrays[3][0] == 0)
    zw[d0] = np.sort(z)  # any assignment of the planted multiset to doc 0's tokens
    zw[arrays[3][0] == 1] = [5, 5, 6]
    om.set_z([zw])
    om.rebuild()
    cnt, top = om.theta_entries(0)
    assert list(top) == [3] * 5 + [4] + [9] * 2 + [20] * 2 + [31]
    assert list(cnt) == [63, 63, 63, 63, 48, 63, 63, 1, 63, 63, 1]
    mc, mt = om.theta_rows(0)
    assert list(mt) == sorted(counts) and list(mc) == [counts[k] for k in sorted(counts)]
    cnt1, top1 = om.theta_entries(1)
    assert list(top1) == [5, 6] and list(cnt1) == [2, 1]
    assert om.token_nnz == z.size * 11 + 3 * 2
    # joint log p(w, z): gammaln of the merged counts
    alpha, beta = 0.5, float(np.float32(0.1))
    nwk = om.nwk.reshape(V, K)
    ndk = np.zeros((2, K))
    ndk[0, list(counts)] = list(counts.values())
    ndk[1, [5, 6]] = [2, 1]
    lens = np.array([z.size, 3])
    g = special.gammaln
    ref = (K * (g(V * beta) - V * g(beta)) + (g(nwk + beta)).sum() - g(om.nk + V * beta).sum()
           + 2 * (g(K * alpha) - K * g(alpha)) + g(ndk + alpha).sum() - g(lens + K * alpha).sum())
    assert abs(om.loglik_joint(alpha, beta) - ref) <= 1e-9 * abs(ref)
    # the sweep's exact LL uses the merged conditional
    before = oracle.OracleModel(V, K, *arrays, z=om.z)
    before.rebuild()
    out, _, _ = om.sweep(alpha, beta, 1, 1)
    ll = 0.0
    for i in range(zw.size):
        mass, _ = _conditional(before, int(arrays[3][0][i]), int(words[i]), alpha, beta)
        ll += np.log(mass.sum() / (lens[arrays[3][0][i]] + K * alpha))
    assert abs(out.ll_sum_exact - ll) <= 1e-10 * abs(ll)
    assert abs(out.ll_sum - ll) <= 1e-5 * abs(ll)


def _random_config(seed):
    """The seeded configurations of the GPU parity tests: K in multiples of
    16 up to 1024, 1-3 chunks, shuffled word ids, empty docs, one long doc
    (counts above 63 split theta entries) on odd seeds."""
    from helpers import random_docs as rd
    rng = np.random.default_rng(1000 + seed)
    K = int(rng.choice([16, 48, 96, 256, 640, 1024]))
    V = int(rng.integers(40, 2500))
    D = int(rng.integers(5, 120))
    chunks = int(rng.integers(1, 4))
    perm = rng.permutation(V)
    docs = [perm[d] for d in rd(rng, V, D, float(rng.uniform(5, 80)), empty_docs=int(rng.integers(0, 3)),
                                zipf=float(rng.uniform(0.6, 1.3)))]
    if seed % 2:
        docs.append(perm[rng.integers(0, min(V, 30), int(rng.integers(2000, 9000)))])
    arrays = chunks_from_docs(V, docs, min(chunks, len(docs)))
    alpha, beta = float(rng.uniform(0.01, 0.5)), float(rng.uniform(0.005, 0.1))
    return V, K, arrays, alpha, beta, int(rng.integers(1, 1 << 30))


@pytest.mark.parametrize("seed", range(8))
def test_sweep_fast_equals_pinned_sweep(seed):
    """The timing sweep (no FP64 shadow) and the word-range sweep give the
    pinned sweep's z and LL, three sweeps in a row from a common state."""
    V, K, arrays, alpha, beta, s_seed = _random_config(seed)
    a = oracle.OracleModel(V, K, *arrays)
    a.init_z(s_seed)
    b = oracle.OracleModel(V, K, *arrays, z=a.z)
    c = oracle.OracleModel(V, K, *arrays, z=a.z)
    for m in (a, b, c):
        m.rebuild()
    for it in range(1, 4):
        out, _, _ = a.sweep(alpha, beta, s_seed, it)
        ll_fast = b.sweep_fast(alpha, beta, s_seed, it)
        # c: the words in three ranges
        cuts = [0, V // 3, (2 * V) // 3, V]
        ll_r, ntok = 0.0, 0
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            o = c.sweep_words(alpha, beta, s_seed, it, lo, hi)
            ll_r += o.ll_sum
            ntok += o.tokens
        assert ntok == a.num_tokens
        for ch in range(a.C):
            assert np.array_equal(a.z[ch], b.z[ch]), f"K={K} sweep {it} chunk {ch}: fast sweep differs"
            assert np.array_equal(a.z[ch], c.z[ch]), f"K={K} sweep {it} chunk {ch}: range sweep differs"
        assert abs(ll_fast - out.ll_sum) <= 1e-12 * abs(out.ll_sum)
        assert abs(ll_r - out.ll_sum) <= 1e-12 * abs(out.ll_sum)
        for m in (a, b, c):
            m.rebuild()
        assert np.array_equal(a.nwk, b.nwk) and np.array_equal(a.th_top, b.th_top)


def test_sweep_fast_equals_pinned_sweep_on_the_small_corpus():
    """Two sweeps of the 1M-token LDA corpus (V=10k Zipf(1.07), D=10k
    Poisson(100), K=1024): the timing sweep equals the pinned sweep."""
    import culda_cgs_amd as cg
    corpus = cg.Corpus.generate(10_000, 10_000, 100, 1, len_mean=100.0, doc_topic_prior=0.1,
                                topic_word_prior=0.05, zipf_s=1.07, seed=1, num_threads=8)
    info = corpus.info()
    _, end, wi, wt = corpus.chunk(0)
    K, alpha, beta, seed = 1024, 50.0 / 1024, 0.01, 1
    a = oracle.OracleModel(info.num_words, K, corpus.doc_cum(), [end], [wi], [wt])
    a.init_z(seed)
    b = oracle.OracleModel(info.num_words, K, corpus.doc_cum(), [end], [wi], [wt], z=a.z)
    a.rebuild()
    b.rebuild()
    for it in (1, 2):
        out, _, _ = a.sweep(alpha, beta, seed, it)
        ll = b.sweep_fast(alpha, beta, seed, it)
        assert np.array_equal(a.z[0], b.z[0]), f"sweep {it}"
        assert abs(ll - out.ll_sum) <= 1e-12 * abs(out.ll_sum)
        a.rebuild()
        b.rebuild()


def test_phi_rebuild_independent_of_thread_count():
    V, K, arrays, alpha, beta, s_seed = _random_config(3)
    ref = None
    for threads in (1, 3, 8):
        oracle.set_threads(threads)
        m = oracle.OracleModel(V, K, *arrays)
        m.init_z(s_seed)
        m.rebuild()
        got = (m.nwk.copy(), m.nk.copy())
        if ref is None:
            ref = got
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    oracle.set_threads(oracle.max_threads())
