"""Small corpora in the data_chunk layout, built with numpy (independent of
the product's generator), and glue shared by the tests."""
import numpy as np


def greedy_ends(lens, num_chunks):
    """Doc -> chunk split of src_format/format.cpp:107-132."""
    n = int(np.sum(lens))
    per = (n + num_chunks - 1) // num_chunks
    ends, d = [], 0
    for _ in range(num_chunks):
        acc = 0
        while d < len(lens):
            acc += int(lens[d])
            d += 1
            if acc >= per:
                break
        ends.append(d)
    return np.array(ends, np.int32)


def chunks_from_docs(num_words, docs, num_chunks):
    """docs: list of arrays of word ids.  Returns (doc_cum, ends, word_idx, tokens)."""
    lens = np.array([len(d) for d in docs], np.int64)
    doc_cum = np.cumsum(lens).astype(np.int64)
    ends = greedy_ends(lens, num_chunks)
    word_idx, tokens = [], []
    start = 0
    for e in ends:
        ws = [np.asarray(docs[d], np.int64) for d in range(start, e)]
        ds = [np.full(len(docs[d]), d, np.int64) for d in range(start, e)]
        w = np.concatenate(ws) if ws else np.zeros(0, np.int64)
        dd = np.concatenate(ds) if ds else np.zeros(0, np.int64)
        order = np.lexsort((dd, w))  # word-major, docs ascending inside a word
        counts = np.bincount(w, minlength=num_words)
        word_idx.append(np.concatenate([[0], np.cumsum(counts)]).astype(np.int64))
        tokens.append(dd[order].astype(np.int32))
        start = e
    return doc_cum, ends, word_idx, tokens


def random_docs(rng, num_words, num_docs, mean_len, empty_docs=0, zipf=1.0):
    p = 1.0 / np.arange(1, num_words + 1) ** zipf
    p /= p.sum()
    docs = []
    for d in range(num_docs):
        n = 0 if d < empty_docs else max(1, rng.poisson(mean_len))
        docs.append(rng.choice(num_words, size=n, p=p))
    rng.shuffle(docs)
    return docs


def oracle_model(num_words, k, doc_cum, ends, word_idx, tokens):
    import oracle
    return oracle.OracleModel(num_words, k, doc_cum, ends, word_idx, tokens)


def product_corpus(num_words, doc_cum, ends, word_idx, tokens):
    import culda_cgs_amd as cg
    return cg.Corpus.from_arrays(num_words, doc_cum, ends, word_idx, tokens)


def theta_rows_global(om):
    """Oracle theta as (offsets, counts, topics) in global doc order."""
    return om.th_off, om.th_cnt[: om.th_off[-1]], om.th_top[: om.th_off[-1]]
