"""The learned-MH oracle (oracle/mh.py) pinned by mathematics, on the CPU.

* the alias table keeps every integer weight's mass exactly;
* the transition kernel of one MH step, enumerated bucket by bucket from the
  alias table, leaves the target distribution invariant and satisfies
  detailed balance (tiny K, FP64);
* alias draws driven by the Philox stream follow W / (K M) (chi-square);
* repeated MH steps of ``mh_token`` reach the target (chi-square), so the
  acceptance rule as coded is the MH one;
* sketch special cases: one topic gives +-1 in its bucket, a colliding pair
  of opposite signs cancels.
"""
import numpy as np
import pytest
from scipy import stats

from oracle import mh

M = 1 << mh.ALIAS_BITS


def _weights(rng, K, kind):
    if kind == "uniform":
        W = np.full(K, M, np.int64)
    elif kind == "skewed":
        p = rng.dirichlet(np.full(K, 0.05))
        W = np.maximum(1, np.floor(p * K * M)).astype(np.int64)
    else:
        p = rng.dirichlet(np.ones(K))
        W = np.maximum(1, np.floor(p * K * M)).astype(np.int64)
    W[np.argmax(W)] += K * M - W.sum()
    return W


@pytest.mark.parametrize("K", [1, 2, 5, 16, 64, 1024])
@pytest.mark.parametrize("kind", ["uniform", "skewed", "flat"])
def test_alias_mass_is_exact(K, kind):
    rng = np.random.default_rng(K * 7 + len(kind))
    W = _weights(rng, K, kind)
    F, A = mh.alias_sweep(W)
    assert np.all((F >= 0) & (F <= M)) and np.all((A >= 0) & (A < K))
    assert np.array_equal(mh.alias_mass(F, A), W)


def test_alias_mass_with_exact_heavies():
    # weights equal to M (zero excess) among lights and heavies
    W = np.array([M, M // 2, M, 3 * M // 2, M], np.int64)
    F, A = mh.alias_sweep(W)
    assert np.array_equal(mh.alias_mass(F, A), W)


def _enumerated_kernel(W, F, A, p):
    """T[k, k'] of one step: bucket j (1/K) gives j with F_j / M and A_j
    otherwise; the move is accepted with min(1, p(k') W(k) / (p(k) W(k')))."""
    K = len(W)
    q = np.zeros(K)
    for j in range(K):
        q[j] += F[j] / M / K
        q[A[j]] += (M - F[j]) / M / K
    T = np.zeros((K, K))
    for k in range(K):
        for kp in range(K):
            if kp != k:
                T[k, kp] = q[kp] * min(1.0, (p[kp] * W[k]) / (p[k] * W[kp]))
        T[k, k] = 1.0 - T[k].sum()
    return T


@pytest.mark.parametrize("seed", range(5))
def test_mh_kernel_invariance_by_enumeration(seed):
    rng = np.random.default_rng(seed)
    K = int(rng.integers(2, 9))
    W = _weights(rng, K, "skewed" if seed % 2 else "flat")
    F, A = mh.alias_sweep(W)
    p = rng.gamma(0.5, size=K) + 1e-3
    T = _enumerated_kernel(W, F, A, p)
    pi = p / p.sum()
    assert np.all(T >= -1e-15)
    assert np.allclose(T.sum(axis=1), 1.0, atol=1e-14)
    assert np.max(np.abs(pi @ T - pi)) < 1e-14
    D = pi[:, None] * T
    assert np.max(np.abs(D - D.T)) < 1e-14
    # the generic closed form agrees with the enumeration
    assert np.allclose(T, mh.mh_transition(W, p), atol=1e-14)


def test_mh_kernel_with_a_wrong_ratio_is_not_invariant():
    # the invariance check has power: dropping the proposal correction breaks it
    rng = np.random.default_rng(11)
    K = 6
    W = _weights(rng, K, "skewed")
    p = rng.gamma(0.5, size=K) + 1e-3
    q = W / W.sum()
    T = np.zeros((K, K))
    for k in range(K):
        for kp in range(K):
            if kp != k:
                T[k, kp] = q[kp] * min(1.0, p[kp] / p[k])
        T[k, k] = 1.0 - T[k].sum()
    pi = p / p.sum()
    assert np.max(np.abs(pi @ T - pi)) > 1e-3


def test_alias_draws_chi_square():
    import oracle
    rng = np.random.default_rng(5)
    K = 16
    W = _weights(rng, K, "flat")
    F, A = mh.alias_sweep(W)
    n = 40000
    counts = np.zeros(K, np.int64)
    for i in range(n):
        x = oracle.philox4x64_10([i, 7, mh.PURPOSE_MH, 0], [3, 0])
        counts[mh.alias_draw(F, A, x[0], x[1], K)] += 1
    exp = W / W.sum() * n
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, K - 1) > 1e-4


def test_mh_token_reaches_target():
    rng = np.random.default_rng(9)
    K = 8
    W = _weights(rng, K, "flat")
    F, A = mh.alias_sweep(W)
    p = rng.gamma(1.0, size=K) + 0.05
    n = 6000
    counts = np.zeros(K, np.int64)
    for g in range(n):
        counts[mh.mh_token(W, F, A, lambda k: p[k], 0, 17, 1, g, 30)] += 1
    exp = p / p.sum() * n
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, K - 1) > 1e-4


def test_sketch_special_cases():
    # one topic: +-1.0 in its bucket, zero elsewhere
    for k in [0, 5, 1023]:
        s = mh.bf16_to_f32(mh.sketch([k], [37], 37))
        b = int(mh.sketch_bucket(k))
        assert s[b] == (-1.0 if mh.sketch_negative(k) else 1.0)
        assert np.count_nonzero(s) == 1
    # a pair in one bucket with opposite signs cancels
    buckets = mh.sketch_bucket(np.arange(1024)).astype(np.int64)
    neg = mh.sketch_negative(np.arange(1024))
    pair = None
    for b in range(256):
        ks = np.flatnonzero(buckets == b)
        pos, ng = ks[~neg[ks]], ks[neg[ks]]
        if pos.size and ng.size:
            pair = (int(pos[0]), int(ng[0]))
            break
    assert pair is not None
    s = mh.bf16_to_f32(mh.sketch(list(pair), [4, 4], 8))
    assert np.count_nonzero(s) == 0
    # every bucket is used by the 1024 topics and none is crowded
    occ = np.bincount(buckets, minlength=256)
    assert occ.min() >= 1 and occ.max() <= 12
    # empty doc
    assert np.count_nonzero(mh.sketch([], [], 0)) == 0


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38], np.float32)
    b = mh.bf16_bits(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins
    assert list(b[:2]) == [0x3F80, 0x3F80]
    assert b[2] == 0x3F81  # 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert mh.bf16_to_f32(b)[3] == -2.5


def test_target_matches_the_pinned_phi_head():
    # target(k) = (n_dk + a)(n_wk + b) / (n_k + V b) must be proportional to
    # (n_dk + a) * phi_hat_wk, with phi_hat from the C oracle's FP32 head
    from helpers import chunks_from_docs, oracle_model, random_docs
    rng = np.random.default_rng(4)
    V, K = 40, 16
    docs = random_docs(rng, V, 30, 25)
    om = oracle_model(V, K, *chunks_from_docs(V, docs, 1))
    om.init_z(3)
    om.rebuild()
    alpha, beta = 0.1, 0.05
    head = om.phi_head(beta).astype(np.float64)
    nwk = om.nwk.reshape(V, K)
    for d in range(5):
        cnt, top = om.theta_rows(d)
        ndk = np.zeros(K, np.int64)
        ndk[top] = cnt
        for w in range(0, V, 7):
            t = np.array([mh.target(ndk[k], nwk[w, k], om.nk[k], V, alpha, beta) for k in range(K)])
            ref = (ndk + np.float32(alpha)) * head[w]
            assert np.allclose(t / t.sum(), ref / ref.sum(), rtol=2e-6, atol=0)
