"""The per-chunk drop-in entry points, driven through the C ABI alone in the
reference's own multi-chunk schedule, against the CPU oracle; state errors
of that surface; HIP failure propagation (forced from the host); rare-word
slots whatever the slot size.

Reference schedule (train/MultiChunkMultiGPUequal.h:95-160): every chunk
sampled with the previous iteration's counts (LDAKernelTrain), the phi
counts rebuilt per chunk and reduced (LDAUpdatePhiAPI, LDAUpdatePhiReduceAPI),
the log-likelihood (LDATrainPerplexity), every chunk's theta rebuilt
(LDAUpdateThetaAPI), the head (LDAComputePhiHeadAPI).
"""
import os

import numpy as np
import pytest

import culda_cgs_amd as cg
from helpers import chunks_from_docs, oracle_model, product_corpus, random_docs

pytestmark = pytest.mark.gpu

LL_RTOL = 1e-5


def _corpus(seed, V=400, D=160, mean_len=60, chunks=4, long_doc=False):
    rng = np.random.default_rng(seed)
    docs = random_docs(rng, V, D, mean_len, empty_docs=2)
    if long_doc:
        docs.append(rng.integers(0, 20, 3000))  # counts above 63: split theta entries
    arrays = chunks_from_docs(V, docs, chunks)
    return V, arrays


@pytest.mark.parametrize("K", [64, 1024])
def test_per_chunk_schedule_through_the_abi_matches_oracle(K):
    V, arrays = _corpus(11, long_doc=True)
    alpha, beta, seed = 0.1, 0.02, 21
    C = len(arrays[1])
    assert C == 4
    corpus = product_corpus(V, *arrays)
    om = oracle_model(V, K, *arrays)
    with cg.Sampler(k=K, alpha=alpha, beta=beta, seed=seed, fused_phi=False) as s:
        s.upload(corpus)
        s.init_z()
        om.init_z(seed)
        # initial model: phi chunk by chunk (overwrite, then +=), head, theta per chunk
        for c in range(C):
            s.rebuild_phi(c, accumulate=c > 0)
        s.phi_head()
        for c in range(C):
            s.rebuild_theta(c)
        om.rebuild()
        for it in range(1, 4):
            for c in range(C):
                s.sample(c, it)
            ll = s.loglik()
            out, _, _ = om.sweep(alpha, beta, seed, it)
            for c in range(C):
                s.rebuild_phi(c, accumulate=c > 0)
            for c in range(C):
                s.rebuild_theta(c)
            s.phi_head()
            om.rebuild()
            for c in range(C):
                bad = np.flatnonzero(s.get_z(c) != om.z[c])
                assert bad.size == 0, f"K={K} sweep {it} chunk {c}: {bad.size} tokens differ"
            ref = out.ll_sum / om.num_tokens
            assert abs(ll - ref) <= LL_RTOL * abs(ref), (it, ll, ref)
            nwk, nk = s.get_phi()
            assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
            assert np.array_equal(s.get_phi_head(), om.phi_head(beta)), f"sweep {it}: head differs"
            start = 0
            for c, e in enumerate(arrays[1]):
                off, cnt, top = s.get_theta(c)
                for d in range(start, e):
                    oc, ot = om.theta_rows(d)
                    lo, hi = off[d - start], off[d - start + 1]
                    assert np.array_equal(cnt[lo:hi], oc) and np.array_equal(top[lo:hi], ot), (it, c, d)
                start = e
        # the per-chunk drop-ins and the one-call iteration continue the same chain
        ll_a, _ = s.iterate(4)
        out, _, _ = om.sweep(alpha, beta, seed, 4)
        om.rebuild()
        for c in range(C):
            assert np.array_equal(s.get_z(c), om.z[c])
        assert abs(ll_a - out.ll_sum / om.num_tokens) <= LL_RTOL * abs(ll_a)


def test_state_errors_of_the_per_chunk_surface():
    V, arrays = _corpus(12, chunks=2)
    corpus = product_corpus(V, *arrays)
    with cg.Sampler(k=64, alpha=0.1, beta=0.02, seed=3) as s:
        with pytest.raises(cg.CuldaError) as e:
            s.sample(0, 1)  # before upload
        assert e.value.status == -6
        s.upload(corpus)
        s.init_z()
        s.rebuild_phi(-1)
        s.phi_head()
        s.rebuild_theta(0)
        with pytest.raises(cg.CuldaError) as e:
            s.sample(1, 1)  # chunk 1's theta was never built
        assert e.value.status == -6 and "theta" in str(e.value)
        s.sample(0, 1)
        s.rebuild_theta(1)
        s.sample(1, 1)
        with pytest.raises(cg.CuldaError) as e:
            s.sample(0, 0)  # sweep 0 keys the z initialisation
        assert e.value.status == -1
        with pytest.raises(cg.CuldaError) as e:
            s.sample(2, 1)
        assert e.value.status == -1
        # replacing a chunk's z invalidates that chunk's theta and the counts
        s.set_z(1, s.get_z(1))
        with pytest.raises(cg.CuldaError) as e:
            s.sample(0, 2)  # phi no longer matches z
        assert e.value.status == -6
        s.rebuild_phi(-1)
        s.phi_head()
        s.sample(0, 2)
        with pytest.raises(cg.CuldaError) as e:
            s.sample(1, 2)
        assert e.value.status == -6
        s.rebuild_theta(1)
        s.sample(1, 2)
        assert np.isfinite(s.loglik())


def test_injected_hip_failure_is_reported_not_fatal(monkeypatch):
    """CULDA_FAIL_HIP makes a named HIP call report failure after it
    succeeded (every kernel still runs normally): the call returns
    CULDA_ERR_DEVICE with the call's name, the process lives on, the
    context is destroyed cleanly and a new one produces the same chain as an
    undisturbed run."""
    V, arrays = _corpus(13, chunks=2)
    corpus = product_corpus(V, *arrays)

    def run(fail_at=None):
        s = cg.Sampler(k=64, alpha=0.1, beta=0.02, seed=5)
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        if fail_at:
            monkeypatch.setenv("CULDA_FAIL_HIP", fail_at)
            with pytest.raises(cg.CuldaError) as e:
                s.iterate(1)
            monkeypatch.delenv("CULDA_FAIL_HIP")
            s.close()
            return e.value
        s.iterate(1)
        s.iterate(2)
        z = [s.get_z(c) for c in range(2)]
        s.close()
        return z

    ref = run()
    for spec in ("theta rebuild launch", "sampler launch:2", "phi clear", "iteration sync"):
        err = run(spec)
        assert err.status == -4, spec
        assert "injected failure" in str(err) and spec.split(":")[0] in str(err), str(err)
    assert "CULDA_FAIL_HIP" not in os.environ
    again = run()
    assert all(np.array_equal(a, b) for a, b in zip(ref, again))


@pytest.mark.parametrize("slot_tokens", [16, 32])
def test_rare_words_with_slots_shorter_than_their_tokens(slot_tokens):
    """Slots shorter than K/16 tokens: a rare word (fewer than K/16 tokens)
    still keeps all its tokens in one slot, so the sampler's staging of its
    row from the tokens' previous topics stays exact; z and counts equal the
    oracle's over several iterations at K = 1024."""
    import oracle
    rng = np.random.default_rng(31)
    V, K = 3000, 1024
    docs = random_docs(rng, V, 150, 70, zipf=0.7)
    arrays = chunks_from_docs(V, docs, 1)
    counts = np.diff(arrays[2][0])
    assert ((counts > 0) & (counts * 16 < K) & (counts > slot_tokens)).sum() >= 10
    alpha, beta, seed = 0.05, 0.01, 8
    corpus = product_corpus(V, *arrays)
    om = oracle.OracleModel(V, K, *arrays)
    with cg.Sampler(k=K, alpha=alpha, beta=beta, seed=seed, slot_tokens=slot_tokens) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        om.init_z(seed)
        om.rebuild()
        for it in range(1, 6):
            ll, _ = s.iterate(it)
            out, _, _ = om.sweep(alpha, beta, seed, it)
            om.rebuild()
            bad = np.flatnonzero(s.get_z(0) != om.z[0])
            assert bad.size == 0, f"sweep {it}: {bad.size} tokens differ"
            assert abs(ll - out.ll_sum / om.num_tokens) <= LL_RTOL * abs(ll)
        nwk, nk = s.get_phi()
        assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
