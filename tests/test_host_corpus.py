"""Host side of the boundary (no GPU): data_chunk formatter, writer, loader,
re-chunking and the synthetic generator, through libculda.so's C ABI."""
import json
from pathlib import Path

import numpy as np
import pytest

import culda_cgs_amd as cg
from helpers import chunks_from_docs, random_docs

GOLD = json.loads((Path(__file__).parent / "golden" / "format_tiny.json").read_text())


def _chunk_name(prefix, what, c, C):
    return Path(f"{prefix}{what}{c if C > 1 else ''}")


@pytest.mark.parametrize("C", [1, 2, 3])
def test_format_text_matches_reference_layout(tmp_path, C):
    inp = tmp_path / "in.txt"
    inp.write_text(GOLD["input"])
    prefix = tmp_path / "out"
    cg.format_text(inp, prefix, C)
    exp = GOLD["chunks"][str(C)]
    assert Path(f"{prefix}.vocab").read_text() == GOLD["vocab"]
    assert Path(f"{prefix}.doc.idx").read_text() == GOLD["doc_idx"]
    assert Path(f"{prefix}.chunk").read_text() == exp["chunk"]
    for c in range(C):
        assert _chunk_name(prefix, ".word.idx", c, C).read_text() == exp["word_idx"][c]
        data = np.fromfile(_chunk_name(prefix, ".word.data", c, C), dtype="<i4")
        assert data.tolist() == exp["word_data"][c]
    # the loader reads it back
    corpus = cg.Corpus.load(prefix, C)
    info = corpus.info()
    assert (info.num_words, info.num_docs, info.num_tokens, info.num_chunks) == (4, 5, 10, C)
    assert corpus.doc_cum().tolist() == [3, 4, 8, 8, 10]


def test_loader_errors(tmp_path):
    inp = tmp_path / "in.txt"
    inp.write_text(GOLD["input"])
    prefix = tmp_path / "out"
    cg.format_text(inp, prefix, 2)
    with pytest.raises(cg.CuldaError) as e:  # chunk count does not match .chunk
        cg.Corpus.load(prefix, 3)
    assert e.value.status == -3
    with pytest.raises(cg.CuldaError) as e:  # missing files
        cg.Corpus.load(tmp_path / "nothing", 1)
    assert e.value.status == -2
    # truncated data file
    d1 = Path(f"{prefix}.word.data1")
    d1.write_bytes(d1.read_bytes()[:-4])
    with pytest.raises(cg.CuldaError) as e:
        cg.Corpus.load(prefix, 2)
    assert e.value.status == -3
    # a doc id outside its chunk
    cg.format_text(inp, prefix, 2)
    d0 = Path(f"{prefix}.word.data0")
    a = np.fromfile(d0, dtype="<i4")
    a[0] = 4
    a.tofile(d0)
    with pytest.raises(cg.CuldaError):
        cg.Corpus.load(prefix, 2)


def test_write_load_roundtrip_and_rechunk(tmp_path):
    rng = np.random.default_rng(0)
    V = 50
    docs = random_docs(rng, V, 40, 12, empty_docs=2)
    doc_cum, ends, wi, wt = chunks_from_docs(V, docs, 1)
    corpus = cg.Corpus.from_arrays(V, doc_cum, ends, wi, wt)
    corpus.write(tmp_path / "c")
    back = cg.Corpus.load(tmp_path / "c", 1)
    _, e, wi2, wt2 = back.chunk(0)
    assert e == len(docs) and np.array_equal(wi2, wi[0]) and np.array_equal(wt2, wt[0])
    for C in (2, 3, 5):
        r = corpus.rechunk(C)
        dc, en, wic, wtc = chunks_from_docs(V, docs, C)
        for c in range(C):
            s, e, a, b = r.chunk(c)
            assert e == en[c]
            assert np.array_equal(a, wic[c]) and np.array_equal(b, wtc[c])
        # concatenating the chunks' word lists gives the one-chunk lists
        for w in range(V):
            cat = np.concatenate([r.chunk(c)[3][r.chunk(c)[2][w]:r.chunk(c)[2][w + 1]] for c in range(C)])
            assert np.array_equal(cat, wt[0][wi[0][w]:wi[0][w + 1]])


def test_from_arrays_validation():
    V = 3
    with pytest.raises(cg.CuldaError):  # doc length disagrees with tokens
        cg.Corpus.from_arrays(V, [2, 3], [2], [np.array([0, 1, 2, 2])], [np.array([0, 1], np.int32)])
    with pytest.raises(cg.CuldaError):  # offsets decrease
        cg.Corpus.from_arrays(V, [1], [1], [np.array([0, 1, 0, 1])], [np.array([0], np.int32)])


def test_generator_deterministic_and_shaped():
    kw = dict(len_mean=80.0, doc_topic_prior=0.1, topic_word_prior=0.05, zipf_s=1.07, seed=1)
    a = cg.Corpus.generate(2000, 3000, 50, 1, num_threads=1, **kw)
    b = cg.Corpus.generate(2000, 3000, 50, 1, num_threads=7, **kw)
    assert np.array_equal(a.chunk(0)[3], b.chunk(0)[3])
    assert np.array_equal(a.doc_cum(), b.doc_cum())
    lens = np.diff(np.concatenate([[0], a.doc_cum()]))
    assert lens.min() >= 1
    assert abs(lens.mean() - 80.0) < 1.0  # Poisson(80): sd of the mean ~0.16
    assert abs(lens.var() / 80.0 - 1.0) < 0.15
    # word frequencies follow the Zipf reweighting: log-log slope near -1.07
    wi = a.chunk(0)[2]
    cnt = np.diff(wi).astype(float)
    r = np.arange(1, 2001)
    m = (r >= 5) & (r <= 500) & (cnt > 0)
    slope = np.polyfit(np.log(r[m]), np.log(cnt[m]), 1)[0]
    assert -1.6 < slope < -0.6
    c4 = cg.Corpus.generate(2000, 3000, 50, 4, num_threads=3, **kw)
    assert c4.info().num_chunks == 4 and c4.info().num_tokens == a.info().num_tokens


def test_generator_lognormal_lengths():
    c = cg.Corpus.generate(500, 4000, 20, 1, len_kind=1, len_mu=3.0, len_sigma=0.5, seed=5)
    lens = np.diff(np.concatenate([[0], c.doc_cum()]))
    ll = np.log(lens)
    assert abs(ll.mean() - 3.0) < 0.05 and abs(ll.std() - 0.5) < 0.05


def _doc_words(corpus):
    """Per doc: its word ids in ascending order, from the word-major chunks."""
    info = corpus.info()
    docs = [[] for _ in range(info.num_docs)]
    for c in range(info.num_chunks):
        _, _, wi, wt = corpus.chunk(c)
        for w in range(info.num_words):
            for d in wt[wi[w]:wi[w + 1]]:
                docs[d].append(w)
    return [np.array(sorted(x), np.int64) for x in docs]


@pytest.mark.parametrize("every,chunks", [(10, 1), (3, 2), (2, 3)])
def test_heldout_split(every, chunks):
    rng = np.random.default_rng(7)
    V = 60
    docs = random_docs(rng, V, 47, 9, empty_docs=2)
    full = cg.Corpus.from_arrays(V, *chunks_from_docs(V, docs, 2))
    train, obs, ev = full.heldout_split(every, chunks)
    assert train.info().num_chunks == chunks and obs.info().num_chunks == 1 and ev.info().num_chunks == 1
    assert train.info().num_words == V and obs.info().num_words == V and ev.info().num_words == V
    fw, tw, ow, ew = _doc_words(full), _doc_words(train), _doc_words(obs), _doc_words(ev)
    held = [d for d in range(len(fw)) if d % every == every - 1]
    kept = [d for d in range(len(fw)) if d % every != every - 1]
    assert len(tw) == len(kept) and len(ow) == len(held) == len(ev.doc_cum())
    for i, d in enumerate(kept):
        assert np.array_equal(tw[i], fw[d])
    for i, d in enumerate(held):
        assert np.array_equal(ow[i], fw[d][0::2]), f"observed half of doc {d}"
        assert np.array_equal(ew[i], fw[d][1::2]), f"evaluated half of doc {d}"
    n = full.info().num_tokens
    assert train.info().num_tokens + obs.info().num_tokens + ev.info().num_tokens == n


def test_heldout_split_errors():
    V = 10
    docs = [np.array([1, 2, 3]), np.array([4])]
    full = cg.Corpus.from_arrays(V, *chunks_from_docs(V, docs, 1))
    with pytest.raises(cg.CuldaError):
        full.heldout_split(1)
    with pytest.raises(cg.CuldaError):
        full.heldout_split(10)  # no doc held out: the held-out corpora would be empty


def test_write_failure_reports_io_error_and_leaves_no_partial_file(tmp_path):
    """A save whose bytes cannot reach the disk (the temporary file is
    /dev/full, so the close reports ENOSPC) fails with CULDA_ERR_IO, and
    neither the temporary nor a truncated target file is left behind; an
    existing good file survives a failed rewrite."""
    import os
    c = cg.Corpus.generate(50, 20, 3, 1, len_mean=10.0, seed=1, num_threads=1)
    prefix = tmp_path / "x"
    c.write(prefix)
    good = (tmp_path / "x.vocab").read_text()
    os.symlink("/dev/full", tmp_path / "x.vocab.tmp")
    with pytest.raises(cg.CuldaError) as e:
        c.write(prefix)
    assert e.value.status == -2
    assert not (tmp_path / "x.vocab.tmp").exists() and not os.path.islink(tmp_path / "x.vocab.tmp")
    assert (tmp_path / "x.vocab").read_text() == good
    with pytest.raises(cg.CuldaError) as e:
        c.write(tmp_path / "missing_dir" / "x")
    assert e.value.status == -2


@pytest.mark.parametrize("alpha", [0.1, 0.5, 2.0])
def test_generator_doc_topics_are_dirichlet_multinomial(alpha):
    """Two topics whose word distributions (Dirichlet(1e-4) over two words,
    seed 2) put nearly all their mass on different words: a doc's share of
    word 0 is then its topic-0 share, which under theta_d ~ Dirichlet(alpha)
    is Beta-binomial: mean 1/2 and variance (1 + (n-1)/(2 alpha + 1)) / (4n)
    for a doc of n tokens.  The measured variance matches to a few percent
    (a per-topic alpha taken for the total, or topics drawn iid per token,
    misses by 30% or more)."""
    c = cg.Corpus.generate(2, 4000, 2, 1, len_mean=50.0, doc_topic_prior=alpha, topic_word_prior=1e-4,
                           zipf_s=0.0, seed=2, num_threads=3)
    _, _, wi, wt = c.chunk(0)
    lens = np.diff(np.concatenate([[0], c.doc_cum()]))
    f = np.bincount(wt[:wi[1]], minlength=lens.size) / lens
    assert 0.45 < wi[1] / wi[2] < 0.55  # both words in use: the topics are separated
    pred = 0.25 * (1 + (lens - 1) / (2 * alpha + 1)) / lens
    ratio = np.mean((f - 0.5) ** 2) / np.mean(pred)
    assert abs(ratio - 1.0) < 0.06, ratio
    assert abs(f.mean() - 0.5) < 0.02
