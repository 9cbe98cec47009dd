"""HIP path vs the CPU oracle, called through the C ABI (libculda.so).

Bit-exact for z, n_wk, n_k and theta; LL/token within 1e-5 relative of the
oracle's FP64 sum.
"""
import numpy as np
import pytest

from helpers import chunks_from_docs, oracle_model, product_corpus, random_docs

pytestmark = pytest.mark.gpu

LL_RTOL = 1e-5


def _setup(seed, V=300, D=120, mean_len=50, chunks=1, K=64, empty_docs=0):
    rng = np.random.default_rng(seed)
    docs = random_docs(rng, V, D, mean_len, empty_docs=empty_docs)
    arrays = chunks_from_docs(V, docs, chunks)
    return V, K, arrays


def _pair(V, K, arrays, alpha, beta, seed, **kw):
    import culda_cgs_amd as cg
    corpus = product_corpus(V, *arrays)
    s = cg.Sampler(k=K, alpha=alpha, beta=beta, seed=seed, **kw)
    s.upload(corpus)
    om = oracle_model(V, K, *arrays)
    return s, om


def _theta_equal(s, om, arrays):
    ends = arrays[1]
    start = 0
    for c, e in enumerate(ends):
        off, cnt, top = s.get_theta(c)
        for d in range(start, e):
            oc, ot = om.theta_rows(d)
            lo, hi = off[d - start], off[d - start + 1]
            assert np.array_equal(cnt[lo:hi], oc), f"theta counts differ at doc {d}"
            assert np.array_equal(top[lo:hi], ot), f"theta topics differ at doc {d}"
        start = e


@pytest.mark.parametrize("chunks", [1, 3])
def test_init_z_and_rebuild(chunks):
    V, K, arrays = _setup(1, chunks=chunks, empty_docs=3)
    s, om = _pair(V, K, arrays, 0.1, 0.01, 11)
    s.init_z()
    om.init_z(11)
    for c in range(chunks):
        assert np.array_equal(s.get_z(c), om.z[c]), f"init z differs in chunk {c}"
    s.rebuild_all()
    om.rebuild()
    nwk, nk = s.get_phi()
    assert np.array_equal(nwk.ravel(), om.nwk)
    assert np.array_equal(nk, om.nk)
    _theta_equal(s, om, arrays)
    head = s.get_phi_head()
    assert np.array_equal(head, om.phi_head(0.01)), "FP32 head differs"
    s.close()


@pytest.mark.parametrize("K", [16, 48, 64, 256, 1024])
def test_sweeps_bit_exact(K):
    V, _, arrays = _setup(2, V=400, D=150, mean_len=70, K=K)
    alpha, beta, seed = 50.0 / K, 0.01, 5
    s, om = _pair(V, K, arrays, alpha, beta, seed)
    s.init_z()
    s.rebuild_all()
    om.init_z(seed)
    om.rebuild()
    for it in range(1, 6):
        ll, _ = s.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        z = s.get_z(0)
        bad = np.flatnonzero(z != om.z[0])
        assert bad.size == 0, f"K={K} sweep {it}: {bad.size} tokens differ (first {bad[:5]})"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll - ref) <= LL_RTOL * abs(ref)
    nwk, nk = s.get_phi()
    assert np.array_equal(nwk.ravel(), om.nwk)
    _theta_equal(s, om, arrays)
    s.close()


@pytest.mark.parametrize("fused", [True, False])
def test_multichunk_matches_oracle_and_one_chunk(fused):
    V, K, arrays4 = _setup(3, V=500, D=200, mean_len=60, chunks=4, K=128)
    _, _, arrays1 = _setup(3, V=500, D=200, mean_len=60, chunks=1, K=128)
    alpha, beta, seed = 0.2, 0.02, 9
    s4, om = _pair(V, K, arrays4, alpha, beta, seed, fused_phi=fused)
    s1, _ = _pair(V, K, arrays1, alpha, beta, seed)
    for s in (s4, s1):
        s.init_z()
        s.rebuild_all()
    om.init_z(seed)
    om.rebuild()
    for it in range(1, 4):
        l4, _ = s4.iterate(it)
        l1, _ = s1.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        for c in range(4):
            assert np.array_equal(s4.get_z(c), om.z[c])
        ref = out.ll_sum / om.num_tokens
        assert abs(l4 - ref) <= LL_RTOL * abs(ref)
        assert abs(l1 - ref) <= LL_RTOL * abs(ref)
    # 4 chunks == 1 chunk, bit for bit
    n4, k4 = s4.get_phi()
    n1, k1 = s1.get_phi()
    assert np.array_equal(n4, n1) and np.array_equal(k4, k1)
    s4.close()
    s1.close()


def test_single_sweep_from_common_state_and_boundaries():
    """One sweep from an arbitrary (non-initial) z; report boundary draws."""
    V, K, arrays = _setup(4, V=1000, D=400, mean_len=120, K=256)
    alpha, beta, seed = 0.1, 0.05, 21
    s, om = _pair(V, K, arrays, alpha, beta, seed)
    rng = np.random.default_rng(4)
    z0 = (rng.integers(0, 16, size=arrays[3][0].size) * 7 % K).astype(np.uint16)
    s.set_z(0, z0)
    s.rebuild_all()
    om.set_z([z0])
    om.rebuild()
    s.sample(0, 3)
    ll = s.loglik()
    out, ex, fl = om.sweep(alpha, beta, seed, 3, want_exact=True)
    z = s.get_z(0)
    assert np.array_equal(z, om.z[0])
    # faithful FP32 vs exact FP64 disagree only at flagged boundary draws
    mis = np.flatnonzero(om.z[0] != ex[0])
    assert np.all(fl[0][mis] == 1)
    ref = out.ll_sum / om.num_tokens
    assert abs(ll - ref) <= LL_RTOL * abs(ref)
    s.close()


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("K", [64, 256])
def test_counts_after_each_iteration_equal_a_full_rebuild(fused, K):
    """After every iteration n_wk / n_k equal the oracle's rebuild from the
    device's own z, also after z is replaced (set_z + rebuild_all) mid-run.
    At K = 256 most words have fewer than K / 16 tokens, so the rebuild
    touches only their previous and new topics' entries; at K = 64 it
    rewrites every row."""
    V, _, arrays = _setup(6, V=350, D=140, mean_len=60, chunks=3, K=K)
    s, om = _pair(V, K, arrays, 0.1, 0.02, 9, fused_phi=fused)
    s.init_z()
    s.rebuild_all()
    for it in range(1, 7):
        s.iterate(it)
        om.set_z([s.get_z(c) for c in range(3)])
        om.rebuild()
        nwk, nk = s.get_phi()
        assert np.array_equal(nwk.ravel(), om.nwk), f"iteration {it}: n_wk differs from a rebuild"
        assert np.array_equal(nk, om.nk), f"iteration {it}: n_k differs from a rebuild"
        if it == 3:
            rng = np.random.default_rng(1)
            for c in range(3):
                s.set_z(c, rng.integers(0, K, s.get_z(c).size).astype(np.uint16))
            s.rebuild_all()
    s.close()


@pytest.mark.parametrize("V,D,K", [(350, 140, 256), (350, 140, 1024), (3000, 100, 1024), (3000, 100, 256)])
def test_rare_tiles_equal_a_full_rebuild(V, D, K):
    """One chunk: the rare words' counts are updated in tiles of up to 64
    tokens of consecutive words.  Word ids are shuffled, so frequent words
    end tiles between rare ones; V = 350 fills tiles up to 64 tokens;
    V = 3000 leaves many words without tokens, so tiles also end at 64 word
    ids.  After every iteration (and after set_z + rebuild_all) n_wk / n_k
    equal the oracle's rebuild."""
    rng = np.random.default_rng(8)
    perm = rng.permutation(V)
    docs = [perm[d] for d in random_docs(rng, V, D, 60)]
    arrays = chunks_from_docs(V, docs, 1)
    s, om = _pair(V, K, arrays, 0.1, 0.02, 4)
    s.init_z()
    s.rebuild_all()
    for it in range(1, 6):
        s.iterate(it)
        om.set_z([s.get_z(0)])
        om.rebuild()
        nwk, nk = s.get_phi()
        assert np.array_equal(nwk.ravel(), om.nwk), f"iteration {it}: n_wk differs from a rebuild"
        assert np.array_equal(nk, om.nk), f"iteration {it}: n_k differs from a rebuild"
        if it == 2:
            s.set_z(0, np.random.default_rng(2).integers(0, K, s.get_z(0).size).astype(np.uint16))
            s.rebuild_all()
    s.close()


def test_long_doc_with_more_than_1024_theta_entries():
    """A 100,000-token doc at K = 1024: every topic holds ~98 of its tokens,
    so each takes two entries and the row has ~2048 entries (128 blocks of
    16, more than the 64 of any row with one entry per topic).  z, n_wk,
    n_k and theta equal the oracle's over three sweeps; the LL within
    1e-5."""
    V, K = 300, 1024
    rng = np.random.default_rng(12)
    docs = random_docs(rng, V, 8, 40) + [rng.integers(0, V, 100_000)]
    arrays = chunks_from_docs(V, docs, 1)
    alpha, beta, seed = 50.0 / K, 0.01, 3
    s, om = _pair(V, K, arrays, alpha, beta, seed)
    s.init_z()
    s.rebuild_all()
    om.init_z(seed)
    om.rebuild()
    assert int(np.diff(om.th_off).max()) > 1024 + 512
    for it in range(1, 4):
        ll, _ = s.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        bad = np.flatnonzero(s.get_z(0) != om.z[0])
        assert bad.size == 0, f"sweep {it}: {bad.size} tokens differ (first {bad[:5]})"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll - ref) <= LL_RTOL * abs(ref)
    nwk, nk = s.get_phi()
    assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
    _theta_equal(s, om, arrays)
    s.close()


def test_theta_counts_near_and_above_16_bits():
    """Topic counts at and past 16 bits in the per-chunk theta rebuilds.
    Two chunks hold docs of 70,000 and 65,535 tokens (chunk 0) and 65,536
    tokens (chunk 1), each with every token on one topic, so their counts
    are 70,000, 65,535 (the largest 16-bit value) and 65,536, split into
    entries of at most 63.  Theta equals the oracle's after the full
    rebuild and after each of two sweeps (the per-chunk rebuilds); z
    bit-exact, LL within 1e-5."""
    V, K = 40, 32
    rng = np.random.default_rng(21)
    small = random_docs(rng, V, 30, 30)
    long_a, long_b, long_c = (rng.integers(0, V, n) for n in (70_000, 65_535, 65_536))
    docs = small[:15] + [long_a] + small[15:] + [long_b, long_c]
    ia, ib, ic = 15, 31, 32
    arrays = chunks_from_docs(V, docs, 2)
    ends, tokens = arrays[1], arrays[3]
    assert list(ends) == [32, 33], ends
    alpha, beta, seed = 50.0 / K, 0.01, 5
    s, om = _pair(V, K, arrays, alpha, beta, seed)
    s.init_z()
    zs = []
    for c in range(2):
        z = s.get_z(c).copy()
        for d, t in ((ia, 5), (ib, 6), (ic, 7)):
            z[tokens[c] == d] = t
        zs.append(z)
        s.set_z(c, z)
    om.set_z(zs)
    s.rebuild_all()
    om.rebuild()
    _theta_equal(s, om, arrays)
    for it in range(1, 3):
        ll, _ = s.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        for c in range(2):
            bad = np.flatnonzero(s.get_z(c) != om.z[c])
            assert bad.size == 0, f"sweep {it}, chunk {c}: {bad.size} tokens differ (first {bad[:5]})"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll - ref) <= LL_RTOL * abs(ref)
    _theta_equal(s, om, arrays)
    s.close()


def test_theta_rebuild_kernels_beside_a_sweep_at_the_16_bit_bound():
    """The two rebuilds that run beside a sweep.  Chunk 0's longest doc has
    65,535 tokens, all on one topic: its rebuild takes the one-wave kernel
    with 16-bit counts, at the largest count they hold.  Chunk 1 holds a
    65,536-token doc, so its rebuild takes the four-wave kernel with 32-bit
    counts.  Theta equals the oracle's after the full rebuild and after
    each of two sweeps; z bit-exact, LL within 1e-5."""
    V, K = 40, 32
    rng = np.random.default_rng(22)
    small = random_docs(rng, V, 30, 30)
    long_b, long_c = rng.integers(0, V, 65_535), rng.integers(0, V, 65_536)
    docs = small[:15] + [long_b] + small[15:] + [long_c]
    ib, ic = 15, 31
    arrays = chunks_from_docs(V, docs, 2)
    ends, tokens = arrays[1], arrays[3]
    assert list(ends) == [16, 32], ends
    alpha, beta, seed = 50.0 / K, 0.01, 9
    s, om = _pair(V, K, arrays, alpha, beta, seed)
    s.init_z()
    zs = []
    for c in range(2):
        z = s.get_z(c).copy()
        for d, t in ((ib, 3), (ic, 11)):
            z[tokens[c] == d] = t
        zs.append(z)
        s.set_z(c, z)
    om.set_z(zs)
    s.rebuild_all()
    om.rebuild()
    _theta_equal(s, om, arrays)
    # the per-chunk rebuilds (the kernels that run beside a sweep) on the
    # same state: counts of exactly 65,535 and 65,536
    for c in range(2):
        s.rebuild_theta(c)
    s.sync()
    _theta_equal(s, om, arrays)
    for it in range(1, 3):
        ll, _ = s.iterate(it)
        out, _, _ = om.sweep(alpha, beta, seed, it)
        om.rebuild()
        for c in range(2):
            bad = np.flatnonzero(s.get_z(c) != om.z[c])
            assert bad.size == 0, f"sweep {it}, chunk {c}: {bad.size} tokens differ (first {bad[:5]})"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll - ref) <= LL_RTOL * abs(ref)
        _theta_equal(s, om, arrays)
    s.close()


def test_row_readers_after_lazy_rare_updates():
    """One chunk, K = 1024, mostly rare words: after each sweep only n_k and
    zp follow the new z, the rare words' rows are brought up to date when
    read.  Every reader sees the counts of the current z: the joint
    log-likelihood (1e-9), the FP32 head and the counts (bit-exact), also
    after a second round of sweeps."""
    V, K = 3000, 1024
    rng = np.random.default_rng(21)
    docs = random_docs(rng, V, 100, 60)
    arrays = chunks_from_docs(V, docs, 1)
    alpha, beta = 0.05, 0.01
    s, om = _pair(V, K, arrays, alpha, beta, 6)
    s.init_z()
    s.rebuild_all()
    for rnd in range(2):
        for it in range(1 + 3 * rnd, 4 + 3 * rnd):
            s.iterate(it)
        om.set_z([s.get_z(0)])
        om.rebuild()
        ref = om.loglik_joint(alpha, beta)
        got = s.loglik_joint()
        assert abs(got - ref) <= 1e-9 * abs(ref), (rnd, got, ref)
        assert np.array_equal(s.get_phi_head(), om.phi_head(beta)), f"round {rnd}: FP32 head differs"
        nwk, nk = s.get_phi()
        assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
    s.close()


@pytest.mark.parametrize("seed", range(6))
def test_random_configurations_bit_exact(seed):
    """Seeded mixes of K (multiples of 16 up to 1024), chunk counts, word
    shuffles, empty and long docs (counts above 63 split theta entries,
    rare words take the tiled path): z, n_wk, n_k and theta equal the
    oracle's after each of three sweeps, counts read only at the end and in
    between."""
    rng = np.random.default_rng(1000 + seed)
    K = int(rng.choice([16, 48, 96, 256, 640, 1024]))
    V = int(rng.integers(40, 2500))
    D = int(rng.integers(5, 120))
    chunks = int(rng.integers(1, 4))
    perm = rng.permutation(V)
    docs = [perm[d] for d in random_docs(rng, V, D, float(rng.uniform(5, 80)), empty_docs=int(rng.integers(0, 3)),
                                         zipf=float(rng.uniform(0.6, 1.3)))]
    if seed % 2:
        docs.append(perm[rng.integers(0, min(V, 30), int(rng.integers(2000, 9000)))])  # one long doc
    arrays = chunks_from_docs(V, docs, min(chunks, len(docs)))
    alpha, beta, s_seed = float(rng.uniform(0.01, 0.5)), float(rng.uniform(0.005, 0.1)), int(rng.integers(1, 1 << 30))
    s, om = _pair(V, K, arrays, alpha, beta, s_seed)
    nch = len(arrays[1])
    s.init_z()
    s.rebuild_all()
    om.init_z(s_seed)
    om.rebuild()
    for it in range(1, 4):
        ll, _ = s.iterate(it)
        out, _, _ = om.sweep(alpha, beta, s_seed, it)
        om.rebuild()
        for c in range(nch):
            bad = np.flatnonzero(s.get_z(c) != om.z[c])
            assert bad.size == 0, f"K={K} V={V} chunks={nch} sweep {it}: {bad.size} tokens differ in chunk {c}"
        ref = out.ll_sum / om.num_tokens
        assert abs(ll - ref) <= LL_RTOL * abs(ref)
        if it == 2:
            nwk, nk = s.get_phi()
            assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
    nwk, nk = s.get_phi()
    assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)
    _theta_equal(s, om, arrays)
    s.close()


@pytest.mark.parametrize("chunks", [1, 2, 4])
def test_back_to_back_iterations_equal_synchronised_ones(chunks):
    """culda_iterate returns with the last chunks' theta rebuilds still on the
    second stream; the next iteration orders each chunk's sweep after that
    chunk's rebuild.  Back-to-back iterations (nothing in between that waits)
    must give the chain of iterations each followed by a full sync, and the
    per-iteration sampler bytes must be the ones of the rebuild each sweep
    read; iterate_n (iterations enqueued back to back in one call) must give
    the same chain, LLs and summed bytes (a one-chunk corpus replays a
    captured graph there, with no per-sweep timing).  The last chunk's
    rebuild starts after the count update; two chunks is the case where the
    iteration's final wait is for the first chunk's rebuild.  The chain ends
    at the oracle's z."""
    V, K, arrays = _setup(31, V=500, D=400, mean_len=90, chunks=chunks, K=128)
    alpha, beta, seed = 0.2, 0.01, 17
    live, om = _pair(V, K, arrays, alpha, beta, seed, profile=True)
    synced, _ = _pair(V, K, arrays, alpha, beta, seed, profile=True)
    batched, _ = _pair(V, K, arrays, alpha, beta, seed, profile=True)
    for s in (live, synced, batched):
        s.init_z()
        s.rebuild_all()
    ll_live, ll_synced, b_live, b_synced = [], [], [], []
    for it in range(1, 7):
        ll_live.append(live.iterate(it)[0])
        ms, b, n = live.last_sweeps()
        assert n == chunks and ms > 0 and b == live.sampler_bytes()
        b_live.append(b)
        ll_synced.append(synced.iterate(it)[0])
        synced.sync()
        b_synced.append(synced.sampler_bytes())
    assert ll_live == ll_synced and b_live == b_synced
    # the same six iterations as one call (no host round trip between them)
    ll_n, _ = batched.iterate_n(1, 2)
    ll_m, _ = batched.iterate_n(3, 4)
    ms, b, n = batched.last_sweeps()
    assert ll_n + ll_m == ll_live
    assert n == 4 * chunks and b == sum(b_live[2:])
    assert (ms == 0) if chunks == 1 else (ms > 0)
    for c in range(chunks):
        assert np.array_equal(live.get_z(c), synced.get_z(c))
        assert np.array_equal(batched.get_z(c), synced.get_z(c))
        assert all(np.array_equal(x, y) for x, y in zip(live.get_theta(c), synced.get_theta(c)))
        assert all(np.array_equal(x, y) for x, y in zip(batched.get_theta(c), synced.get_theta(c)))
    assert live.stats().sum_token_nnz == synced.stats().sum_token_nnz == batched.stats().sum_token_nnz
    om.init_z(seed)
    om.rebuild()
    for it in range(1, 7):
        om.sweep(alpha, beta, seed, it)
        om.rebuild()
    for c in range(chunks):
        assert np.array_equal(synced.get_z(c), om.z[c]), f"chunk {c}"
    import culda_cgs_amd as cg
    for first, count in ((7, 0), (7, 1025), (0, 1)):
        with pytest.raises(cg.CuldaError):
            batched.iterate_n(first, count)
    for s in (live, synced, batched):
        s.close()
