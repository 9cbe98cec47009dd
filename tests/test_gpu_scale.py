"""GPU tests at the benchmark configurations' real sizes and end to end.

* the 1M-token configuration (K=1024, V=10k, D=10k Poisson(100)): several
  sweeps bit-exact against the oracle;
* the ~100M-token configuration (K=1024, V=100k, D=300k lognormal, 4
  chunks): size-independent properties -- count invariants, 4 chunks == 1
  chunk bit for bit, determinism across runs, LL finite and rising;
* checkpoint / resume, save hooks and the CLI.
"""
import hashlib
import re
import subprocess

import numpy as np
import pytest

import culda_cgs_amd as cg

pytestmark = pytest.mark.gpu

ALPHA = 50.0 / 1024.0
BETA = 0.01


def _c1():
    return cg.Corpus.generate(10_000, 10_000, 100, 1, len_mean=100.0, doc_topic_prior=0.1, topic_word_prior=0.05,
                              zipf_s=1.07, seed=1, num_threads=16)


def _c2(chunks):
    return cg.Corpus.generate(100_000, 300_000, 500, chunks, len_kind=1, len_mu=5.6, len_sigma=0.6,
                              doc_topic_prior=0.1, topic_word_prior=0.01, zipf_s=1.0, seed=2, num_threads=16)


@pytest.mark.parametrize("slot_tokens", [0, 2048])
def test_c1_full_size_bit_exact_vs_oracle(slot_tokens):
    """Default slots (128 tokens at this size) and 2048-token slots, whose
    batches run past the 256 tokens one set of Philox blocks serves."""
    import oracle
    corpus = _c1()
    info = corpus.info()
    assert 950_000 < info.num_tokens < 1_050_000
    _, end, wi, wt = corpus.chunk(0)
    assert np.diff(wi).max() > 2048
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=1, slot_tokens=slot_tokens) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        om = oracle.OracleModel(info.num_words, 1024, corpus.doc_cum(), [end], [wi], [wt])
        om.init_z(1)
        om.rebuild()
        assert np.array_equal(s.get_z(0), om.z[0])
        for it in range(1, 4):
            ll, _ = s.iterate(it)
            ref = om.sweep_fast(ALPHA, BETA, 1, it) / om.num_tokens
            om.rebuild()
            z = s.get_z(0)
            bad = np.count_nonzero(z != om.z[0])
            assert bad == 0, f"sweep {it}: {bad} of {z.size} tokens differ"
            assert abs(ll - ref) <= 1e-5 * abs(ref)
        nwk, nk = s.get_phi()
        assert np.array_equal(nwk.ravel(), om.nwk) and np.array_equal(nk, om.nk)


def _digest(s, chunks):
    h = hashlib.sha256()
    for c in range(chunks):
        h.update(s.get_z(c).tobytes())
    return h.hexdigest()


def test_c2_full_size_properties():
    c4 = _c2(4)
    info = c4.info()
    N, V, D = info.num_tokens, info.num_words, info.num_docs
    assert 9e7 < N < 1.05e8 and info.num_chunks == 4
    lls = []
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=3) as s4:
        s4.upload(c4)
        s4.init_z()
        s4.rebuild_all()
        for it in range(1, 4):
            ll, _ = s4.iterate(it)
            lls.append(ll)
        d4 = _digest(s4, 4)
        # count invariants (validPhi / validTheta, model_phi.cpp:194-280,
        # model_theta_chunk.cpp:142-273)
        nwk, nk = s4.get_phi()
        assert int(nk.sum()) == N
        assert np.array_equal(nwk.sum(axis=0, dtype=np.int64), nk.astype(np.int64))
        word_cnt = sum(np.diff(c4.chunk(c)[2]) for c in range(4))
        assert np.array_equal(nwk.sum(axis=1, dtype=np.int64), word_cnt)
        doc_len = np.diff(np.concatenate([[0], c4.doc_cum()]))
        for c in range(4):
            start, end, _, _ = c4.chunk(c)
            off, cnt, top = s4.get_theta(c)
            sums = np.add.reduceat(cnt.astype(np.int64), off[:-1]) if cnt.size else np.zeros(0)
            nnz = np.diff(off)
            assert np.all(nnz >= 1) and np.all(nnz <= np.minimum(doc_len[start:end], 1024))
            assert np.array_equal(sums, doc_len[start:end])
            # ascending, distinct topics inside every row
            inc = np.diff(top.astype(np.int32))
            row_start = np.zeros(top.size, bool)
            row_start[off[:-1]] = True
            assert np.all(inc[~row_start[1:]] > 0)
        assert all(np.isfinite(lls)) and lls[-1] > lls[0]
    # one chunk, same seed: identical z (chunking invariance at full size)
    c1 = c4.rechunk(1)
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=3) as s1:
        s1.upload(c1)
        s1.init_z()
        s1.rebuild_all()
        for it in range(1, 4):
            ll1, _ = s1.iterate(it)
        z1 = s1.get_z(0)
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=3, fused_phi=False) as s4b:
        s4b.upload(c4)
        s4b.init_z()
        s4b.rebuild_all()
        for it in range(1, 4):
            s4b.iterate(it)
        assert _digest(s4b, 4) == d4, "chunks-in-turn phi accumulation changed the chain"
        # word-major z of 4 chunks, interleaved per word, equals the 1-chunk z
        parts = [(c4.chunk(c)[2], s4b.get_z(c)) for c in range(4)]
    wi1 = c1.chunk(0)[2]
    rebuilt = np.empty_like(z1)
    pos = wi1[:-1].copy()
    for wi, z in parts:
        cnt = np.diff(wi)
        # place each chunk's word ranges after the previous chunks' tokens of that word
        idx = np.repeat(pos - wi[:-1], cnt) + np.arange(z.size)
        rebuilt[idx] = z
        pos += cnt
    assert np.array_equal(rebuilt, z1)
    assert abs(ll1 - lls[-1]) <= 1e-12 * abs(ll1)


def test_determinism_across_runs():
    corpus = _c1()
    digests = []
    for _ in range(2):
        with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=9, slot_tokens=512) as s:
            s.upload(corpus)
            s.init_z()
            s.rebuild_all()
            lls = [s.iterate(it)[0] for it in range(1, 4)]
            digests.append((_digest(s, 1), lls))
    assert digests[0] == digests[1]
    # a different slot size changes the launch shape but not the result
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=9, slot_tokens=4096) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        for it in range(1, 4):
            s.iterate(it)
        assert _digest(s, 1) == digests[0][0]


def test_checkpoint_resume_bit_exact(tmp_path):
    corpus = _c1()
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=4) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        for it in range(1, 5):
            s.iterate(it)
        ref = _digest(s, 1)
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=4) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        for it in range(1, 3):
            s.iterate(it)
        s.save_checkpoint(tmp_path / "ck", 3)
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=4) as s:
        s.upload(corpus)
        nxt = s.load_checkpoint(tmp_path / "ck")
        assert nxt == 3
        s.rebuild_all()
        for it in range(nxt, 5):
            s.iterate(it)
        assert _digest(s, 1) == ref
    with cg.Sampler(k=1024, alpha=ALPHA, beta=BETA, seed=5) as s:  # other seed: refused
        s.upload(corpus)
        with pytest.raises(cg.CuldaError):
            s.load_checkpoint(tmp_path / "ck")


def test_save_hooks_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    corpus = cg.Corpus.generate(300, 80, 5, 2, len_mean=30.0, seed=11)
    with cg.Sampler(k=64, alpha=0.1, beta=0.01, seed=2) as s:
        s.upload(corpus)
        s.init_z()
        s.rebuild_all()
        s.iterate(1)
        s.save_phi(tmp_path / "m")
        s.save_theta(tmp_path / "m")
        nwk, nk = s.get_phi()
        raw = np.fromfile(tmp_path / "m.phi.bin", dtype="<i4")
        assert raw[0] == 300 and raw[1] == 64
        assert np.array_equal(raw[2:2 + 300 * 64].reshape(300, 64), nwk)
        assert np.array_equal(raw[2 + 300 * 64:], nk)
        txt = (tmp_path / "m.phi.txt").read_text()
        assert txt.count("Word id:") == 300
        b = (tmp_path / "m.theta.bin").read_bytes()
        D, K = np.frombuffer(b[:8], "<i4")
        off = np.frombuffer(b[8:8 + 8 * (D + 1)], "<i8")
        n = int(off[-1])
        cnt = np.frombuffer(b[8 + 8 * (D + 1):8 + 8 * (D + 1) + 4 * n], "<i4")
        top = np.frombuffer(b[8 + 8 * (D + 1) + 4 * n:], "<i2")
        assert D == 80 and K == 64 and top.size == n
        lens = np.diff(np.concatenate([[0], corpus.doc_cum()]))
        assert np.array_equal(np.add.reduceat(cnt, off[:-1]), lens)
        ref = np.concatenate([s.get_theta(c)[1] for c in range(2)])
        assert np.array_equal(cnt, ref)


def test_cli_end_to_end(tmp_path):
    rng = np.random.default_rng(1)
    words = [f"w{i}" for i in range(200)]
    lines = []
    for d in range(120):
        n = rng.integers(5, 60)
        lines.append(f"doc{d} " + " ".join(rng.choice(words, size=n)))
    (tmp_path / "in.txt").write_text("\n".join(lines) + "\n")
    cg.format_text(tmp_path / "in.txt", tmp_path / "c", 2)
    r = subprocess.run([str(cg.CLI_PATH), "-k", "64", "-t", "4", "-a", "50/64", "-b", "0.01", "-c", "2",
                        "-i", str(tmp_path / "c"), "-o", str(tmp_path / "out"), "--checkpoint",
                        str(tmp_path / "ck")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    it_lines = [l for l in r.stdout.splitlines() if l.startswith("Iteration")]
    assert len(it_lines) == 4
    pat = re.compile(r"Iteration\s+(\d+):\s+([\d.]+) sec, ([\d.]+) sec, logLikelyhood = (-?[\d.]+), +([\d.]+) M")
    lls = [float(pat.match(l).group(4)) for l in it_lines]
    assert all(np.isfinite(lls))
    assert (tmp_path / "out.phi.bin").exists() and (tmp_path / "out.theta.txt").exists()
    # resume two more iterations from the checkpoint
    r2 = subprocess.run([str(cg.CLI_PATH), "-k", "64", "-t", "2", "-a", "50/64", "-c", "2", "-i", str(tmp_path / "c"),
                         "--resume", str(tmp_path / "ck"), "--json"], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert '"iteration": 5' in r2.stdout and '"iteration": 6' in r2.stdout
