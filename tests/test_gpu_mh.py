"""Learned-MH proposal on the GPU, through the C ABI.

* doc sketches bit-exact against the oracle's (from the GPU's theta);
* MLP logits and gradients against torch float32 autograd on the same BF16
  roundings (logits 1e-2 absolute; gradients 3e-2 relative in norm);
* one Adam step against torch.optim.Adam applied to the GPU's gradients;
* proposal weights (sum K * 2^20, >= 1, within 2 units of the softmax),
  alias tables bit-exact against the oracle's sequential sweep;
* a full MH iteration: every token's new z bit-exact against the oracle's
  MH steps given the same weights, counts and Philox stream;
* training is deterministic and lowers the cross-entropy.

K = 64 runs the 64 x 64-tile GEMM, K = 128 and K = 1024 the 128 x 128-tile
one, a full 65,536-query batch at K = 1024 the ring kernel; K = 1024 also
fills all 16 weight and theta registers per lane of the MH step.  A corpus
of three data_chunks checks the per-chunk token ids.
"""
import numpy as np
import pytest

import culda_cgs_amd as cg
from oracle import mh

pytestmark = pytest.mark.gpu

ALPHA, BETA, SEED = 0.1, 0.05, 5
KS = [64, 128, 1024]
M = 1 << mh.ALIAS_BITS


def _corpus(chunks=1):
    return cg.Corpus.generate(300, 200, 8, chunks, len_mean=40.0, doc_topic_prior=0.2, topic_word_prior=0.05,
                              zipf_s=1.0, seed=4, num_threads=4)


def _sampler(corpus, iters=3, K=64, **mhkw):
    s = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=SEED)
    s.upload(corpus)
    s.init_z()
    s.rebuild_all()
    for it in range(1, iters + 1):
        s.iterate(it)
    s.mh_enable(**mhkw)
    return s


def _token_arrays(corpus):
    _, _, wi, wt = corpus.chunk(0)
    words = np.repeat(np.arange(len(wi) - 1), np.diff(wi)).astype(np.int32)
    return wt.astype(np.int32), words


def _theta_dense(s, D, K):
    off, cnt, top = s.get_theta(0)
    th = np.zeros((D, K), np.int64)
    for d in range(D):
        th[d, top[off[d]:off[d + 1]]] = cnt[off[d]:off[d + 1]]
    return off, cnt, top, th


def _bf(x):
    return mh.bf16_to_f32(mh.bf16_bits(x))


def test_sketch_matches_oracle():
    K = 64
    c = _corpus()
    s = _sampler(c, K=K)
    D = c.info().num_docs
    sk = s.mh_sketch()
    off, cnt, top = s.get_theta(0)
    lens = np.diff(np.concatenate([[0], c.doc_cum()]))
    for d in range(D):
        ref = mh.sketch(top[off[d]:off[d + 1]], cnt[off[d]:off[d + 1]], lens[d])
        assert np.array_equal(sk[d], ref), f"doc {d}"
    s.close()


def _torch_forward(s, w, docs, words):
    import torch
    sk = torch.tensor(mh.bf16_to_f32(s.mh_sketch()))
    E = torch.tensor(w["emb"], requires_grad=True)
    W1 = torch.tensor(w["w1"], requires_grad=True)
    b1 = torch.tensor(w["b1"], requires_grad=True)
    W2 = torch.tensor(w["w2"], requires_grad=True)
    b2 = torch.tensor(w["b2"], requires_grad=True)

    def bf(t):  # BF16 rounding with a straight-through gradient
        return t + (torch.tensor(_bf(t.detach().numpy())) - t).detach()

    d = torch.tensor(docs, dtype=torch.long)
    x = torch.cat([sk[d], bf(E[torch.tensor(words, dtype=torch.long)])], dim=1)
    pre = x @ bf(W1) + b1
    h = bf(torch.nn.functional.gelu(pre, approximate="tanh"))
    logits = h @ bf(W2) + b2
    return logits, (W1, b1, W2, b2, E)


def _random_weights(s, rng, scale=1.0):
    w = s.mh_get_weights()
    return {k: (v * scale + rng.normal(0, 0.02, v.shape)).astype(np.float32) for k, v in w.items()}


@pytest.mark.parametrize("K", KS)
def test_logits_match_torch(K):
    c = _corpus()
    s = _sampler(c, K=K)
    rng = np.random.default_rng(1)
    w = _random_weights(s, rng, 3.0)
    s.mh_set_weights(w)
    D, V = c.info().num_docs, c.info().num_words
    n = 300
    docs = rng.integers(0, D, n).astype(np.int32)
    words = rng.integers(0, V, n).astype(np.int32)
    got = s.mh_logits(docs, words)
    ref, _ = _torch_forward(s, w, docs, words)
    ref = ref.detach().numpy()
    assert got.shape == (n, K)
    err = np.abs(got - ref).max()
    assert err <= 1e-2 * max(1.0, np.abs(ref).max()), err
    s.close()


def test_logits_full_batch_match_torch():
    """65,536 queries at K = 1024: both products (65,536 x 384 x 512 and
    65,536 x 512 x 1024) take the ring kernel with the staged BF16 epilogue.
    Sampled rows against torch; the first 12,800 rows bit-exact against a
    12,800-query call, whose products are too small for the ring kernel and
    take the 128-tile kernel (same MFMA k order and epilogue arithmetic)."""
    K = 1024
    c = _corpus()
    s = _sampler(c, K=K)
    rng = np.random.default_rng(2)
    w = _random_weights(s, rng, 3.0)
    s.mh_set_weights(w)
    D, V = c.info().num_docs, c.info().num_words
    n, n_small = 65536, 12800
    docs = rng.integers(0, D, n).astype(np.int32)
    words = rng.integers(0, V, n).astype(np.int32)
    got = s.mh_logits(docs, words)
    assert got.shape == (n, K)
    sel = np.sort(rng.choice(n, 512, replace=False))
    ref, _ = _torch_forward(s, w, docs[sel], words[sel])
    ref = ref.detach().numpy()
    err = np.abs(got[sel] - ref).max()
    assert err <= 1e-2 * max(1.0, np.abs(ref).max()), err
    small = s.mh_logits(docs[:n_small], words[:n_small])
    assert np.array_equal(small, got[:n_small])
    s.close()


@pytest.mark.parametrize("K", KS)
def test_gradients_match_torch_and_adam_step(K):
    import torch
    c = _corpus()
    s = _sampler(c, K=K, batch=256)
    rng = np.random.default_rng(2)
    w = _random_weights(s, rng, 2.0)
    s.mh_set_weights(w)
    D, V = c.info().num_docs, c.info().num_words
    n = 200
    docs = rng.integers(0, D, n).astype(np.int32)
    words = rng.integers(0, V, n).astype(np.int32)
    tg = rng.integers(0, K, n).astype(np.int32)
    loss, g = s.mh_gradients(docs, words, tg)
    logits, params = _torch_forward(s, w, docs, words)
    ref_loss = torch.nn.functional.cross_entropy(logits, torch.tensor(tg, dtype=torch.long))
    ref_loss.backward()
    assert abs(loss - ref_loss.item()) <= 1e-3 * abs(ref_loss.item()) + 1e-4
    for name, p in zip(("w1", "b1", "w2", "b2", "emb"), params):
        ref = p.grad.numpy()
        rel = np.linalg.norm(g[name] - ref) / max(np.linalg.norm(ref), 1e-12)
        assert rel < 3e-2, f"{name}: relative error {rel}"
    # weights unchanged by a gradient-only call
    w_now = s.mh_get_weights()
    assert all(np.array_equal(w_now[k], w[k]) for k in w)
    # one Adam step: compare with torch.optim.Adam on the GPU's own gradients
    loss2, g2 = s.mh_gradients(docs, words, tg, train_step=True)
    for k in g:
        assert np.array_equal(g[k], g2[k]), f"gradient {k} not reproducible"
    after = s.mh_get_weights()
    tp = {k: torch.tensor(w[k], requires_grad=True) for k in w}
    opt = torch.optim.Adam(list(tp.values()), lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for k in tp:
        tp[k].grad = torch.tensor(g[k])
    opt.step()
    for k in tp:
        ref = tp[k].detach().numpy()
        assert np.allclose(after[k], ref, rtol=1e-5, atol=1e-7), k
    s.close()


def _tables_ok(W, F, A, logits, K):
    assert np.all(W >= 1)
    assert np.all(W.astype(np.int64).sum(axis=1) == K * M)
    ex = np.exp(logits.astype(np.float64) - logits.max(axis=1, keepdims=True))
    soft = np.floor(ex / ex.sum(axis=1, keepdims=True) * K * M)
    diff = np.abs(W - np.maximum(soft, 1))
    diff[np.arange(len(W)), np.argmax(W, axis=1)] = 0  # takes the rounding remainder
    assert diff.max() <= np.maximum(2, 1e-5 * W).max()
    step = max(1, len(W) // 400)  # the Python sweep at K = 1024 costs ~5 ms per row
    for i in range(0, len(W), step):
        f, a = mh.alias_sweep(W[i])
        assert np.array_equal(F[i], f) and np.array_equal(A[i], a), f"alias table of row {i}"


@pytest.mark.parametrize("K", KS)
def test_tables_and_mh_iteration_match_oracle(K):
    import oracle
    c = _corpus()
    s = _sampler(c, iters=4, K=K)
    rng = np.random.default_rng(3)
    s.mh_set_weights(_random_weights(s, rng, 4.0))
    info = c.info()
    D, V, N = info.num_docs, info.num_words, info.num_tokens
    docs, words = _token_arrays(c)
    z0 = s.get_z(0)
    nwk, nk = s.get_phi()
    _, _, _, th = _theta_dense(s, D, K)
    W, F, A = s.mh_tables(0, N)
    _tables_ok(W, F, A, s.mh_logits(docs, words), K)
    # a sub-range gives the same rows
    W2, F2, A2 = s.mh_tables(100, 50)
    assert np.array_equal(W2, W[100:150]) and np.array_equal(A2, A[100:150])
    it, props = 5, 2
    rate, _ = s.mh_iterate(it, props)
    z1 = s.get_z(0)
    acc = 0
    ref = np.empty(N, np.int64)
    for i in range(N):
        d, w = int(docs[i]), int(words[i])
        p_of = lambda k: mh.target(th[d, k], nwk[w, k], nk[k], V, ALPHA, BETA)
        ref[i] = mh.mh_token(W[i], F[i], A[i], p_of, z0[i], SEED, it, i, props)
    bad = np.flatnonzero(z1 != ref)
    assert bad.size == 0, f"{bad.size} of {N} tokens differ (first {bad[:5]})"
    # acceptance count of the oracle replay (an accepted move can return to z0)
    assert 0.0 < rate <= 1.0
    # counts were rebuilt from the new z
    om = oracle.OracleModel(V, K, c.doc_cum(), [D], [c.chunk(0)[2]], [c.chunk(0)[3]], z=[z1])
    om.rebuild()
    nwk1, nk1 = s.get_phi()
    assert np.array_equal(nwk1.ravel(), om.nwk) and np.array_equal(nk1, om.nk)
    s.close()


@pytest.mark.parametrize("K", KS)
def test_training_deterministic_and_learns(K):
    c = _corpus()
    runs = []
    for _ in range(2):
        s = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=SEED)
        s.upload(c)
        s.init_z()
        s.rebuild_all()
        s.mh_enable(batch=512, max_snapshots=4, init_seed=9)
        for it in range(1, 9):
            s.iterate(it)
            if it >= 5:
                s.mh_collect()
        l1, _ = s.mh_train(1)
        l2, _ = s.mh_train(2)
        runs.append((l1, l2, s.mh_get_weights()))
        s.close()
    (a1, a2, wa), (b1, b2, wb) = runs
    assert a1 == b1 and a2 == b2
    assert all(np.array_equal(wa[k], wb[k]) for k in wa), "training is not deterministic"
    assert a2 < a1 < np.log(K), (a1, a2)


def test_mh_errors():
    K = 64
    c = _corpus()
    s = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=SEED)
    s.upload(c)
    s.init_z()
    s.rebuild_all()
    with pytest.raises(cg.CuldaError):
        s.mh_iterate(1)  # not enabled
    s.mh_enable(max_snapshots=1)
    with pytest.raises(cg.CuldaError):
        s.mh_train(1)  # no snapshot
    s.mh_collect()
    with pytest.raises(cg.CuldaError):
        s.mh_collect()  # over max_snapshots
    with pytest.raises(cg.CuldaError):
        s.mh_logits([c.info().num_docs], [0])  # doc out of range
    s.close()
    t = cg.Sampler(k=48, alpha=ALPHA, beta=BETA, seed=SEED)
    t.upload(c)
    with pytest.raises(cg.CuldaError):
        t.mh_enable()  # K not a multiple of 64
    t.close()


def test_mh_refuses_rows_longer_than_it_holds():
    # the MH step holds a doc's theta row in registers (2048 16-bit
    # entries); a 200,000-token doc at K = 1024 may need 1024 + 3175
    from helpers import chunks_from_docs, product_corpus
    rng = np.random.default_rng(3)
    V = 50
    docs = [rng.integers(0, V, 200_000), rng.integers(0, V, 30)]
    c = product_corpus(V, *chunks_from_docs(V, docs, 1))
    s = cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=SEED)
    s.upload(c)
    s.init_z()
    s.rebuild_all()
    with pytest.raises(cg.CuldaError):
        s.mh_enable()
    s.close()


def test_mh_iteration_over_chunks_matches_oracle():
    """Three data_chunks: every chunk's tokens move with the same frozen
    counts and their own chunk's theta rows; the Philox stream is keyed by
    the global token id (chunk-invariant), so each token's new z equals the
    oracle's replay."""
    import oracle
    K, C = 128, 3
    c = _corpus(C)
    s = _sampler(c, iters=3, K=K)
    rng = np.random.default_rng(7)
    s.mh_set_weights(_random_weights(s, rng, 4.0))
    info = c.info()
    D, V = info.num_docs, info.num_words
    assert info.num_chunks == C
    chunks = [c.chunk(k) for k in range(C)]
    # global id of each token: word-major over the one-chunk layout
    cnt = np.array([np.diff(ch[2]) for ch in chunks])  # [C][V]
    pre = np.concatenate([[0], np.cumsum(cnt.sum(0))])
    nwk, nk = s.get_phi()
    th = np.zeros((D, K), np.int64)
    for k in range(C):
        off, ct, tp = s.get_theta(k)
        ds = chunks[k][0]
        for d in range(chunks[k][1] - ds):
            th[ds + d, tp[off[d]:off[d + 1]]] = ct[off[d]:off[d + 1]]
    z0 = [s.get_z(k) for k in range(C)]
    tables, base = [], 0
    for k in range(C):
        n = chunks[k][3].size
        tables.append(s.mh_tables(base, n))
        base += n
    with pytest.raises(cg.CuldaError):
        s.mh_tables(chunks[0][3].size - 5, 10)  # a range across two chunks
    it, props = 4, 2
    s.mh_iterate(it, props)
    for k in range(C):
        W, F, A = tables[k]
        wi, wt = chunks[k][2], chunks[k][3]
        words = np.repeat(np.arange(V), np.diff(wi))
        z1 = s.get_z(k)
        for i in range(wt.size):
            d, w = int(wt[i]), int(words[i])
            g = int(pre[w] + cnt[:k, w].sum() + (i - wi[w]))
            p_of = lambda t: mh.target(th[d, t], nwk[w, t], nk[t], V, ALPHA, BETA)
            ref = mh.mh_token(W[i], F[i], A[i], p_of, z0[k][i], SEED, it, g, props)
            assert z1[i] == ref, f"chunk {k} token {i}"
    s.close()
