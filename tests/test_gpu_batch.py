"""Batched multi-corpus sampling, graph replay and the autotuner on the GPU.

* every corpus of a batch follows the chain it follows alone (bit-exact z
  and counts, identical LL), and the first one matches the oracle;
* graph-replayed and persistent (atomic work queue) iterations equal the
  eager per-slot launches after every replay;
* autotuning (which samples from the current state and restores it) leaves
  the chain unchanged, single corpus and batch.
"""
import numpy as np
import pytest

import culda_cgs_amd as cg
from helpers import chunks_from_docs, oracle_model, product_corpus, random_docs

pytestmark = pytest.mark.gpu

K = 64
# (V, D, mean doc length, chunks, alpha, beta, seed, empty docs)
SPECS = [
    (300, 120, 50, 1, 0.1, 0.01, 11, 2),
    (800, 40, 200, 2, 0.5, 0.1, 12, 0),
    (150, 300, 12, 3, 50.0 / K, 0.01, 13, 5),
    (1200, 90, 80, 1, 0.05, 0.02, 14, 0),
]


def _arrays(i, spec):
    V, D, mean_len, chunks, _, _, _, empty = spec
    rng = np.random.default_rng(100 + i)
    docs = random_docs(rng, V, D, mean_len, empty_docs=empty)
    return chunks_from_docs(V, docs, chunks)


@pytest.fixture(scope="module")
def corpora():
    out = []
    for i, sp in enumerate(SPECS):
        arr = _arrays(i, sp)
        out.append((sp, arr, product_corpus(sp[0], *arr)))
    return out


def _batch(corpora, **kw):
    b = cg.Batch(k=K, **kw)
    for sp, _, c in corpora:
        b.add(c, sp[4], sp[5], sp[6])
    b.upload()
    b.init()
    return b


def _state(b, m, chunks):
    return [b.get_z(m, c) for c in range(chunks)], b.get_phi(m)


def _assert_same(state_a, state_b, what):
    za, (nwa, nka) = state_a
    zb, (nwb, nkb) = state_b
    for c, (x, y) in enumerate(zip(za, zb)):
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: z differs in chunk {c} at {bad.size} tokens"
    assert np.array_equal(nwa, nwb), f"{what}: n_wk differs"
    assert np.array_equal(nka, nkb), f"{what}: n_k differs"


def test_batch_equals_each_corpus_alone(corpora):
    iters = 4
    b = _batch(corpora)
    lls = np.stack([b.iterate(it)[0][0] for it in range(1, iters + 1)])
    assert b.info()[0] == len(SPECS)
    for m, (sp, arr, c) in enumerate(corpora):
        s = cg.Sampler(k=K, alpha=sp[4], beta=sp[5], seed=sp[6])
        s.upload(c)
        s.init_z()
        s.rebuild_all()
        for it in range(1, iters + 1):
            ll, _ = s.iterate(it)
            assert ll == lls[it - 1, m], f"model {m} iteration {it}: LL {ll} vs batch {lls[it - 1, m]}"
        alone = ([s.get_z(ch) for ch in range(sp[3])], s.get_phi())
        _assert_same(_state(b, m, sp[3]), alone, f"model {m}")
        s.close()
    b.close()


def test_batch_first_model_matches_oracle(corpora):
    sp, arr, _ = corpora[0]
    V, alpha, beta, seed = sp[0], sp[4], sp[5], sp[6]
    b = _batch(corpora)
    om = oracle_model(V, K, *arr)
    om.init_z(seed)
    om.rebuild()
    for it in range(1, 4):
        ll, _ = b.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        z = b.get_z(0, 0)
        assert np.array_equal(z, om.z[0]), f"sweep {it}: z differs from the oracle"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll[0, 0] - ref) <= 1e-5 * abs(ref)
    nwk, nk = b.get_phi(0)
    assert np.array_equal(nwk.ravel(), om.nwk)
    assert np.array_equal(nk, om.nk)
    b.close()


@pytest.mark.parametrize("groups", [0, 7, 512])
def test_graph_and_persistent_equal_eager(corpora, groups):
    eager = _batch(corpora)
    other = _batch(corpora)
    other.set_persistent(groups)
    for it in range(1, 5):
        ll_e, _ = eager.iterate(it)
        ll_g, _ = other.iterate(it, use_graph=True)
        assert np.array_equal(ll_e, ll_g), f"groups={groups} iteration {it}: LL differs"
        for m, (sp, _, _) in enumerate(corpora):
            _assert_same(_state(other, m, sp[3]), _state(eager, m, sp[3]), f"groups={groups} it={it} model {m}")
    # several replays in one call, LL of each iteration returned
    ll_e = np.concatenate([eager.iterate(it)[0] for it in range(5, 9)])
    ll_g, _ = other.iterate(5, count=4, use_graph=True)
    assert np.array_equal(ll_e, ll_g)
    for m, (sp, _, _) in enumerate(corpora):
        _assert_same(_state(other, m, sp[3]), _state(eager, m, sp[3]), f"groups={groups} after 4 replays, model {m}")
    eager.close()
    other.close()


def test_autotune_leaves_chain_unchanged(corpora):
    sp, arr, c = corpora[2]
    plain = cg.Sampler(k=K, alpha=sp[4], beta=sp[5], seed=sp[6])
    tuned = cg.Sampler(k=K, alpha=sp[4], beta=sp[5], seed=sp[6])
    for s in (plain, tuned):
        s.upload(c)
        s.init_z()
        s.rebuild_all()
    t = tuned.autotune()
    assert t.num_tried == 6 and t.slot_tokens in [x for x, _ in t.tried()]
    assert sp[3] > 1 and t.theta_groups in (1, 2, 3, 4, 6, 8)  # theta rebuild grid, tuned
    assert len(t.theta_tried()) == 6 and all(ms > 0 for _, ms in t.theta_tried())
    assert all(ms > 0 for _, ms in t.tried())
    # the slot partition changes how the FP64 LL partials group: same value
    # up to rounding; z and the counts are identical
    for it in range(1, 4):
        assert np.isclose(plain.iterate(it)[0], tuned.iterate(it)[0], rtol=1e-12, atol=0)
    for ch in range(sp[3]):
        assert np.array_equal(plain.get_z(ch), tuned.get_z(ch))
    plain.close()
    tuned.close()

    eager = _batch(corpora)
    tb = _batch(corpora)
    t = tb.autotune()
    assert t.num_tried == 18 and t.persistent_groups >= 0 and t.theta_groups == 0
    for it in range(1, 4):
        assert np.allclose(eager.iterate(it)[0], tb.iterate(it, use_graph=True)[0], rtol=1e-12, atol=0)
    for m, (sp, _, _) in enumerate(corpora):
        _assert_same(_state(tb, m, sp[3]), _state(eager, m, sp[3]), f"tuned batch model {m}")
    eager.close()
    tb.close()


def test_batch_errors(corpora):
    b = cg.Batch(k=K)
    with pytest.raises(cg.CuldaError):
        b.upload()  # nothing added
    b.add(corpora[0][2], 0.1, 0.01, 1)
    with pytest.raises(cg.CuldaError):
        b.add(corpora[1][2], -1.0, 0.01, 1)
    b.upload()
    with pytest.raises(cg.CuldaError):
        b.iterate(1)  # before the initial rebuild
    b.init()
    with pytest.raises(cg.CuldaError):
        b.set_persistent(-1)
    with pytest.raises(cg.CuldaError):
        b.iterate(1, count=2000)
    b.close()
