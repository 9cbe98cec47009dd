"""The benchmark configurations at their full sizes on the GPU.

* the ~100M-token corpus (4 chunks): one sweep from the GPU's state after
  two iterations equals the oracle's bit for bit (z of every chunk, LL);
* the batched workload (64 corpora, ~42M tokens, the benchmark's seeds):
  two graph-replayed iterations; count invariants for every corpus, the
  smallest and the largest corpus bit-exact against the oracle;
* the large corpus (V=140k, D=8.2M, ~738M tokens) in 16 data_chunks: count
  invariants after two iterations, and the same chain as 4 chunks.
Progress goes to stderr (run with -s to see it).
"""
import sys
import time

import numpy as np
import pytest

import culda_cgs_amd as cg

pytestmark = pytest.mark.gpu

K = 1024
ALPHA = 50.0 / K
BETA = 0.01
THREADS = 16


def _say(msg):
    print(f"[configs] {msg}", file=sys.stderr, flush=True)


def _oracle_threads():
    import oracle
    oracle.set_threads(THREADS)
    return oracle


@pytest.mark.timeout(600)
def test_c2_full_size_sweep_bit_exact_vs_oracle():
    oracle = _oracle_threads()
    corpus = cg.Corpus.generate(100_000, 300_000, 500, 4, len_kind=1, len_mu=5.6, len_sigma=0.6,
                                doc_topic_prior=0.1, topic_word_prior=0.01, zipf_s=1.0, seed=2,
                                num_threads=THREADS)
    info = corpus.info()
    C = info.num_chunks
    parts = [corpus.chunk(c) for c in range(C)]
    with cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=1) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        s.iterate(1)
        s.iterate(2)
        om = oracle.OracleModel(info.num_words, K, corpus.doc_cum(), [p[1] for p in parts], [p[2] for p in parts],
                                [p[3] for p in parts], z=[s.get_z(c) for c in range(C)])
        om.rebuild()
        ll, _ = s.iterate(3)
        t0 = time.time()
        ref = om.sweep_fast(ALPHA, BETA, 1, 3) / om.num_tokens
        _say(f"c2 oracle sweep {time.time() - t0:.1f}s")
        for c in range(C):
            bad = np.count_nonzero(s.get_z(c) != om.z[c])
            assert bad == 0, f"chunk {c}: {bad} tokens differ"
        assert abs(ll - ref) <= 1e-5 * abs(ref)
        om.rebuild()
        nwk, nk = s.get_phi()
        assert np.array_equal(nk, om.nk) and np.array_equal(nwk.ravel(), om.nwk)


def _c3_corpora():
    rng = np.random.default_rng(3)
    out = []
    for i in range(64):
        V = int(rng.integers(5000, 50001))
        D = int(rng.integers(1000, 5001))
        mean = float(rng.uniform(30.0, 400.0))
        out.append(cg.Corpus.generate(V, D, 100, 1, len_kind=0, len_mean=mean, doc_topic_prior=0.1,
                                      topic_word_prior=0.01, zipf_s=1.0, seed=1000 + i, num_threads=THREADS))
    return out


@pytest.mark.timeout(600)
def test_c3_batch_of_64_corpora():
    oracle = _oracle_threads()
    t0 = time.time()
    corpora = _c3_corpora()
    sizes = [c.info().num_tokens for c in corpora]
    assert 3.5e7 < sum(sizes) < 5e7
    _say(f"c3 corpora {time.time() - t0:.1f}s, {sum(sizes)} tokens")
    seed0 = 1
    with cg.Batch(k=K) as b:
        for i, c in enumerate(corpora):
            b.add(c, ALPHA, BETA, seed0 + i)
        b.upload()
        b.init()
        b.set_persistent(2048)  # the persistent sampler with its atomic work queue
        lls, _ = b.iterate(1, count=2, use_graph=True)
        lls = np.asarray(lls).reshape(2, 64)
        assert np.all(np.isfinite(lls))
        # count invariants of every corpus (validPhi, model_phi.cpp:194-280)
        for m, c in enumerate(corpora):
            z = b.get_z(m, 0)
            nk = b.get_nk(m)
            assert int(nk.sum()) == sizes[m], m
            assert np.array_equal(nk, np.bincount(z, minlength=K)), m
        # the smallest and the largest corpus against the oracle's two sweeps
        for m in (int(np.argmin(sizes)), int(np.argmax(sizes))):
            c = corpora[m]
            info = c.info()
            _, end, wi, wt = c.chunk(0)
            om = oracle.OracleModel(info.num_words, K, c.doc_cum(), [end], [wi], [wt])
            om.init_z(seed0 + m)
            om.rebuild()
            for it in (1, 2):
                out, _, _ = om.sweep(ALPHA, BETA, seed0 + m, it)
                om.rebuild()
                assert abs(lls[it - 1, m] - out.ll_sum / om.num_tokens) <= 1e-5 * abs(lls[it - 1, m])
            bad = np.count_nonzero(b.get_z(m, 0) != om.z[0])
            assert bad == 0, f"corpus {m} ({sizes[m]} tokens): {bad} tokens differ"
            nwk, nk = b.get_phi(m)
            assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)


def _chain(corpus, chunks, iters):
    with cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=1) as s:
        t0 = time.time()
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        _say(f"c5 {chunks} chunks: upload + init {time.time() - t0:.1f}s")
        lls = [s.iterate(it)[0] for it in range(1, iters + 1)]
        nwk, nk = s.get_phi()
        z = [s.get_z(c) for c in range(chunks)]
        return lls, nwk, nk, z


@pytest.mark.timeout(900)
def test_c5_full_size_16_chunks():
    t0 = time.time()
    c16 = cg.Corpus.generate(140_000, 8_200_000, 1000, 16, len_kind=0, len_mean=90.0, doc_topic_prior=0.05,
                             topic_word_prior=0.01, zipf_s=1.05, seed=5, num_threads=THREADS)
    info = c16.info()
    N, V = info.num_tokens, info.num_words
    assert 7.2e8 < N < 7.6e8 and info.num_chunks == 16
    _say(f"c5 corpus {time.time() - t0:.1f}s, {N} tokens")
    lls16, nwk, nk, z16 = _chain(c16, 16, 2)
    assert all(np.isfinite(lls16))
    assert int(nk.sum(dtype=np.int64)) == N
    assert np.array_equal(nwk.sum(axis=0, dtype=np.int64), nk.astype(np.int64))
    word_cnt = np.zeros(V, np.int64)
    for c in range(16):
        wi = c16.chunk(c)[2]
        word_cnt += np.diff(wi)
        # per chunk: z within range, the chunk's histogram inside the global one
        assert z16[c].max() < K
    assert np.array_equal(nwk.sum(axis=1, dtype=np.int64), word_cnt)
    nk16 = nk.copy()
    del nwk
    c4 = c16.rechunk(4)
    _say(f"c5 rechunked to 4 ({time.time() - t0:.1f}s)")
    lls4, _, nk4, z4 = _chain(c4, 4, 2)
    assert np.array_equal(nk4, nk16)
    assert all(abs(a - b) <= 1e-12 * abs(a) for a, b in zip(lls4, lls16))
    # word-major z interleaved per word over the chunks: equal for 16 and 4 chunks
    def merged(corpus, zs, chunks):
        wis = [corpus.chunk(c)[2] for c in range(chunks)]
        tot = sum(np.diff(w) for w in wis)
        base = np.concatenate([[0], np.cumsum(tot)])[:-1]
        out = np.empty(N, np.uint16)
        pos = base.copy()
        for wi, z in zip(wis, zs):
            cnt = np.diff(wi)
            out[np.repeat(pos - wi[:-1], cnt) + np.arange(z.size)] = z
            pos += cnt
        return out
    assert np.array_equal(merged(c16, z16, 16), merged(c4, z4, 4))
    _say(f"c5 done {time.time() - t0:.1f}s")
