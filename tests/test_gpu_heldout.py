"""Held-out log-likelihood (document completion) on the GPU.

* fold-in sweeps with phi fixed at the training counts: z bit-exact against
  the oracle's sweep given the same n_wk / n_k;
* the evaluated LL against a dense FP64 brute force
  mean log(theta_hat @ phi_hat^T) over the evaluated (doc, word) pairs;
* evaluating does not disturb the training chain, and is repeatable.
"""
import numpy as np
import pytest

import culda_cgs_amd as cg
from helpers import chunks_from_docs, random_docs

pytestmark = pytest.mark.gpu

K = 64
ALPHA, BETA = 0.2, 0.05
SEED, FOLD_SEED = 3, 99


def _split(every=5, chunks=2):
    rng = np.random.default_rng(21)
    V = 250
    docs = random_docs(rng, V, 140, 40, empty_docs=3)
    full = cg.Corpus.from_arrays(V, *chunks_from_docs(V, docs, 1))
    return V, full.heldout_split(every, chunks)


def _arrays(corpus):
    info = corpus.info()
    parts = [corpus.chunk(c) for c in range(info.num_chunks)]
    return corpus.doc_cum(), [p[1] for p in parts], [p[2] for p in parts], [p[3] for p in parts]


def _train(train, iters):
    s = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=SEED)
    s.upload(train)
    s.init_z()
    s.rebuild_all()
    for it in range(1, iters + 1):
        s.iterate(it)
    return s


def _brute_force(V, nwk, nk, theta_dense, nd, ev):
    # alpha and beta as the library holds them (FP32), widened exactly
    alpha, beta = float(np.float32(ALPHA)), float(np.float32(BETA))
    phi = (nwk.astype(np.float64) + beta) / (nk.astype(np.float64) + V * beta)
    th = (theta_dense + alpha) / (nd[:, None] + K * alpha)
    _, _, wi, wt = ev.chunk(0)
    words = np.repeat(np.arange(V), np.diff(wi))
    p = np.einsum("ik,ik->i", th[wt], phi[words])
    return float(np.mean(np.log(p)))


@pytest.mark.parametrize("sweeps", [0, 1, 4])
def test_heldout_ll_matches_oracle_and_brute_force(sweeps):
    import oracle
    V, (train, obs, ev) = _split()
    s = _train(train, 3)
    f = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=FOLD_SEED)
    f.upload(obs)
    f.set_eval(ev)
    ll = f.heldout_ll(s, sweeps)
    nwk, nk = s.get_phi()
    # fold-in replayed by the oracle with the same fixed counts
    doc_cum, ends, wi, wt = _arrays(obs)
    om = oracle.OracleModel(V, K, doc_cum, ends, wi, wt)
    om.nwk[:] = nwk.ravel()
    om.nk[:] = nk
    om.init_z(FOLD_SEED)
    om.rebuild_theta()
    for t in range(1, sweeps + 1):
        om.sweep(ALPHA, BETA, FOLD_SEED, t)
        om.rebuild_theta()
    assert np.array_equal(f.get_z(0), om.z[0]), "fold-in z differs from the oracle"
    H = len(doc_cum)
    theta = np.zeros((H, K))
    off, cnt, top = f.get_theta(0)
    for d in range(H):
        theta[d, top[off[d]:off[d + 1]]] = cnt[off[d]:off[d + 1]]
    for d in range(H):
        oc, ot = om.theta_rows(d)
        assert np.array_equal(theta[d, ot], oc) and theta[d].sum() == oc.sum()
    nd = np.diff(np.concatenate([[0], doc_cum])).astype(np.float64)
    ref = _brute_force(V, nwk, nk, theta, nd, ev)
    assert abs(ll - ref) <= 1e-10 * abs(ref), (ll, ref)
    assert f.heldout_ll(s, sweeps) == ll, "same inputs must give the same value"
    f.close()
    s.close()


def test_heldout_does_not_disturb_training():
    V, (train, obs, ev) = _split(every=4)
    a = _train(train, 2)
    b = _train(train, 2)
    f = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=FOLD_SEED)
    f.upload(obs)
    f.set_eval(ev)
    lls = []
    for it in range(3, 7):
        lls.append(f.heldout_ll(a, 3))
        la, _ = a.iterate(it)
        lb, _ = b.iterate(it)
        assert la == lb
    for c in range(train.info().num_chunks):
        assert np.array_equal(a.get_z(c), b.get_z(c))
    assert all(np.isfinite(lls))
    f.close()
    a.close()
    b.close()


def test_heldout_errors():
    V, (train, obs, ev) = _split()
    s = _train(train, 1)
    f = cg.Sampler(k=K, alpha=ALPHA, beta=BETA, seed=FOLD_SEED)
    f.upload(obs)
    with pytest.raises(cg.CuldaError):
        f.heldout_ll(s, 1)  # no evaluated tokens yet
    with pytest.raises(cg.CuldaError):
        f.set_eval(train)  # doc count differs
    f.set_eval(ev)
    with pytest.raises(cg.CuldaError):
        f.heldout_ll(f, 1)
    g = cg.Sampler(k=2 * K, alpha=ALPHA, beta=BETA, seed=FOLD_SEED)
    g.upload(obs)
    g.set_eval(ev)
    with pytest.raises(cg.CuldaError):
        g.heldout_ll(s, 1)  # K differs
    f.close()
    g.close()
    s.close()
