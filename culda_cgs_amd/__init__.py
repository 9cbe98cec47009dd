"""Python binding of libculda.so, the MI355X collapsed-Gibbs LDA sampler.

Thin ctypes layer over the C ABI declared in include/culda.h.  The compute
path is the gfx950 HIP code inside libculda.so; there is no Python or CPU
fallback: if the library is missing, every entry point raises.

Roles mirror the reference host classes: ``Corpus`` is Document/DocChunk
(src_culda/model/doc.h:25, doc_chunk.h:19) plus the src_format writer;
``Sampler`` owns the device model (ModelPhi / ModelTheta,
model/model_phi.h:22, model/model_theta.h:22) and runs the iteration of
train/SingleChunkSingleGPU.h:70-127.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libculda.so"
CLI_PATH = _HERE / "bin" / "culda"

_lib = None


class CuldaError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"culda error {status}: {message}")
        self.status = status


class CorpusInfo(C.Structure):
    _fields_ = [("num_words", C.c_int32), ("num_docs", C.c_int32), ("num_chunks", C.c_int32),
                ("max_doc_len", C.c_int32), ("num_tokens", C.c_int64)]


class ChunkView(C.Structure):
    _fields_ = [("docid_start", C.c_int32), ("docid_end", C.c_int32), ("num_tokens", C.c_int64),
                ("word_indices", C.POINTER(C.c_int64)), ("word_tokens", C.POINTER(C.c_int32))]


class GenParams(C.Structure):
    _fields_ = [("num_words", C.c_int32), ("num_docs", C.c_int32), ("true_topics", C.c_int32),
                ("num_chunks", C.c_int32), ("len_kind", C.c_int32), ("num_threads", C.c_int32),
                ("len_mean", C.c_double), ("len_mu", C.c_double), ("len_sigma", C.c_double),
                ("doc_topic_prior", C.c_double), ("topic_word_prior", C.c_double),
                ("zipf_s", C.c_double), ("seed", C.c_uint64)]


class Params(C.Structure):
    _fields_ = [("k", C.c_int32), ("alpha", C.c_float), ("beta", C.c_float), ("seed", C.c_uint64),
                ("device", C.c_int32), ("slot_tokens", C.c_int32), ("profile", C.c_int32),
                ("fused_phi", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("num_tokens", C.c_int64), ("num_slots", C.c_int64), ("sum_doc_nnz", C.c_int64),
                ("sum_token_nnz", C.c_int64), ("theta_capacity", C.c_int64), ("sampler_ms", C.c_double),
                ("phi_ms", C.c_double), ("theta_ms", C.c_double), ("other_ms", C.c_double),
                ("sampler_launches", C.c_int32), ("pad", C.c_int32)]


class MHParams(C.Structure):
    _fields_ = [("batch", C.c_int32), ("infer_batch", C.c_int32), ("max_snapshots", C.c_int32), ("pad", C.c_int32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("init_seed", C.c_uint64)]


class Tune(C.Structure):
    _fields_ = [("slot_tokens", C.c_int32), ("persistent_groups", C.c_int32), ("num_tried", C.c_int32),
                ("theta_groups", C.c_int32), ("best_ms", C.c_double), ("tried_slot_tokens", C.c_int32 * 32),
                ("tried_ms", C.c_double * 32), ("theta_tried_ms", C.c_double * 8)]

    def tried(self):
        return [(self.tried_slot_tokens[i], self.tried_ms[i]) for i in range(self.num_tried)]

    def theta_tried(self):
        """(theta workgroups per CU, ms of two iterations) of the timed candidates."""
        return [(g, self.theta_tried_ms[i]) for i, g in enumerate((1, 2, 3, 4, 6, 8)) if self.theta_tried_ms[i] > 0]


# name -> (restype, argtypes); every int-returning call is checked.
_P = C.c_void_p
_SIGS = {
    "culda_last_error": (C.c_char_p, []),
    "culda_abi_version": (C.c_int, []),
    "culda_corpus_load": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(_P)]),
    "culda_corpus_from_arrays": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.c_int32,
                                           C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_int64)),
                                           C.POINTER(C.POINTER(C.c_int32)), C.POINTER(_P)]),
    "culda_corpus_write": (C.c_int, [_P, C.c_char_p]),
    "culda_format_text": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32]),
    "culda_corpus_rechunk": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "culda_corpus_generate": (C.c_int, [C.POINTER(GenParams), C.POINTER(_P)]),
    "culda_corpus_info_get": (C.c_int, [_P, C.POINTER(CorpusInfo)]),
    "culda_corpus_doc_cum": (C.c_int, [_P, C.POINTER(C.POINTER(C.c_int64))]),
    "culda_corpus_chunk": (C.c_int, [_P, C.c_int32, C.POINTER(ChunkView)]),
    "culda_corpus_free": (None, [_P]),
    "culda_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    "culda_destroy": (None, [_P]),
    "culda_upload": (C.c_int, [_P, _P]),
    "culda_check_corpus": (C.c_int, [_P, C.c_int32]),
    "culda_init_z": (C.c_int, [_P]),
    "culda_set_z": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint16), C.c_int64]),
    "culda_get_z": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint16), C.c_int64]),
    "culda_rebuild_phi": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "culda_phi_head": (C.c_int, [_P]),
    "culda_rebuild_theta": (C.c_int, [_P, C.c_int32]),
    "culda_sample": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "culda_loglik": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "culda_iterate": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "culda_rebuild_all": (C.c_int, [_P]),
    "culda_loglik_joint": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "culda_loglik_joint_noalloc": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "culda_sync": (C.c_int, [_P]),
    "culda_get_phi": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "culda_get_phi_head": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "culda_get_theta": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int16)]),
    "culda_save_phi": (C.c_int, [_P, C.c_char_p]),
    "culda_save_theta": (C.c_int, [_P, C.c_char_p]),
    "culda_save_checkpoint": (C.c_int, [_P, C.c_char_p, C.c_int32]),
    "culda_load_checkpoint": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_int32)]),
    "culda_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "culda_sampler_bytes": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "culda_last_sweeps": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "culda_iterate_n": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "culda_stream_copy": (C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "culda_autotune": (C.c_int, [_P, C.POINTER(Tune)]),
    "culda_mh_enable": (C.c_int, [_P, C.POINTER(MHParams)]),
    "culda_mh_collect": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "culda_mh_train": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "culda_mh_iterate": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "culda_mh_get_weights": (C.c_int, [_P] + [C.POINTER(C.c_float)] * 5),
    "culda_mh_set_weights": (C.c_int, [_P] + [C.POINTER(C.c_float)] * 5),
    "culda_mh_logits": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    "culda_mh_gradients": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_double)]
                           + [C.POINTER(C.c_float)] * 5),
    "culda_mh_tables": (C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
    "culda_mh_sketch": (C.c_int, [_P, C.POINTER(C.c_uint16)]),
    "culda_corpus_heldout_split": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "culda_heldout_set_eval": (C.c_int, [_P, _P]),
    "culda_heldout_ll": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_double)]),
    "culda_batch_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    "culda_batch_destroy": (None, [_P]),
    "culda_batch_add": (C.c_int, [_P, _P, C.c_float, C.c_float, C.c_uint64, C.POINTER(C.c_int32)]),
    "culda_batch_upload": (C.c_int, [_P]),
    "culda_batch_init": (C.c_int, [_P]),
    "culda_batch_set_persistent": (C.c_int, [_P, C.c_int32]),
    "culda_batch_iterate": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "culda_batch_autotune": (C.c_int, [_P, C.POINTER(Tune)]),
    "culda_batch_get_z": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_uint16), C.c_int64]),
    "culda_batch_get_phi": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "culda_batch_info": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "culda_batch_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "culda_batch_sampler_bytes": (C.c_int, [_P, C.POINTER(C.c_double)]),
}


def lib():
    """Load libculda.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        # CULDA_LIBRARY: another build of the same library (the sanitized host build)
        path = Path(os.environ.get("CULDA_LIBRARY", LIB_PATH))
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build it first (python -c "
                               "'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise CuldaError(st, lib().culda_last_error().decode(errors="replace"))


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Corpus:
    """Host corpus in the src_format data_chunk layout."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.culda_corpus_free(self._h)
            self._h = C.c_void_p()

    @classmethod
    def load(cls, prefix, num_chunks):
        h = C.c_void_p()
        _check(lib().culda_corpus_load(str(prefix).encode(), num_chunks, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, num_words, doc_cum, chunk_doc_end, word_indices, word_tokens):
        doc_cum = np.ascontiguousarray(doc_cum, dtype=np.int64)
        ends = np.ascontiguousarray(chunk_doc_end, dtype=np.int32)
        wi = [np.ascontiguousarray(x, dtype=np.int64) for x in word_indices]
        wt = [np.ascontiguousarray(x, dtype=np.int32) for x in word_tokens]
        nc = len(wi)
        wi_p = (C.POINTER(C.c_int64) * nc)(*[_ptr(x, C.c_int64) for x in wi])
        wt_p = (C.POINTER(C.c_int32) * nc)(*[_ptr(x, C.c_int32) for x in wt])
        h = C.c_void_p()
        _check(lib().culda_corpus_from_arrays(num_words, len(doc_cum), _ptr(doc_cum, C.c_int64), nc,
                                              _ptr(ends, C.c_int32), wi_p, wt_p, C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, num_words, num_docs, true_topics, num_chunks=1, *, len_kind=0, len_mean=100.0,
                 len_mu=0.0, len_sigma=0.0, doc_topic_prior=0.1, topic_word_prior=0.05, zipf_s=1.0,
                 seed=1, num_threads=0):
        p = GenParams(num_words, num_docs, true_topics, num_chunks, len_kind, num_threads, len_mean, len_mu,
                      len_sigma, doc_topic_prior, topic_word_prior, zipf_s, seed)
        h = C.c_void_p()
        _check(lib().culda_corpus_generate(C.byref(p), C.byref(h)))
        return cls(h.value)

    def heldout_split(self, every=10, train_chunks=1):
        """(train, observed, eval) corpora for document completion: docs
        d % every == every - 1 held out, their tokens alternating observed /
        evaluated in ascending word order."""
        t, o, e = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().culda_corpus_heldout_split(self._h, every, train_chunks, C.byref(t), C.byref(o), C.byref(e)))
        return Corpus(t.value), Corpus(o.value), Corpus(e.value)

    def write(self, prefix):
        _check(lib().culda_corpus_write(self._h, str(prefix).encode()))

    def rechunk(self, num_chunks):
        h = C.c_void_p()
        _check(lib().culda_corpus_rechunk(self._h, num_chunks, C.byref(h)))
        return Corpus(h.value)

    def info(self):
        i = CorpusInfo()
        _check(lib().culda_corpus_info_get(self._h, C.byref(i)))
        return i

    def doc_cum(self):
        p = C.POINTER(C.c_int64)()
        _check(lib().culda_corpus_doc_cum(self._h, C.byref(p)))
        n = self.info().num_docs
        return np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int64)

    def check(self, k):
        """Raise CuldaError if the corpus exceeds a documented limit for k topics
        (no device needed; upload runs the same checks)."""
        _check(lib().culda_check_corpus(self._h, k))

    def chunk_tokens(self, c):
        """Number of tokens of chunk c (nothing copied)."""
        v = ChunkView()
        _check(lib().culda_corpus_chunk(self._h, c, C.byref(v)))
        return int(v.num_tokens)

    def chunk(self, c):
        """(docid_start, docid_end, word_indices[V+1], word_tokens[N_c]) as copies."""
        v = ChunkView()
        _check(lib().culda_corpus_chunk(self._h, c, C.byref(v)))
        V = self.info().num_words
        wi = np.ctypeslib.as_array(v.word_indices, shape=(V + 1,)).copy()
        wt = (np.ctypeslib.as_array(v.word_tokens, shape=(v.num_tokens,)).copy()
              if v.num_tokens else np.zeros(0, np.int32))
        return v.docid_start, v.docid_end, wi, wt


def format_text(input_path, prefix, num_chunks=1):
    """./format input output_prefix numChunks (src_format/format.cpp:19-168)."""
    _check(lib().culda_format_text(str(input_path).encode(), str(prefix).encode(), num_chunks))


class Sampler:
    """Device model + iteration driver on one GPU."""

    def __init__(self, k=1024, alpha=50.0 / 1024.0, beta=0.01, seed=1, device=0, slot_tokens=0,
                 profile=False, fused_phi=True):
        self.params = Params(k, alpha, beta, seed, device, slot_tokens, 1 if profile else 0,
                             1 if fused_phi else 0)
        h = C.c_void_p()
        _check(lib().culda_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.k = k
        self.corpus = None

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().culda_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- lifecycle -------------------------------------------------------
    def upload(self, corpus: Corpus):
        _check(lib().culda_upload(self._h, corpus._h))
        self.corpus = corpus
        self._info = corpus.info()
        self._chunk_tokens = [corpus.chunk_tokens(c) for c in range(self._info.num_chunks)]

    def init_z(self):
        _check(lib().culda_init_z(self._h))

    def rebuild_all(self):
        _check(lib().culda_rebuild_all(self._h))

    def rebuild_phi(self, chunk=-1, accumulate=False):
        _check(lib().culda_rebuild_phi(self._h, chunk, 1 if accumulate else 0))

    def phi_head(self):
        _check(lib().culda_phi_head(self._h))

    def rebuild_theta(self, chunk):
        _check(lib().culda_rebuild_theta(self._h, chunk))

    def sample(self, chunk, iteration):
        _check(lib().culda_sample(self._h, chunk, iteration))

    def loglik(self):
        v = C.c_double()
        _check(lib().culda_loglik(self._h, C.byref(v)))
        return v.value

    def iterate(self, iteration):
        ll, sec = C.c_double(), C.c_double()
        _check(lib().culda_iterate(self._h, iteration, C.byref(ll), C.byref(sec)))
        return ll.value, sec.value

    def iterate_n(self, first, count):
        """count iterations from sweep `first` with no host round trip between
        them: (LL/token per iteration, seconds)."""
        ll, sec = (C.c_double * count)(), C.c_double()
        _check(lib().culda_iterate_n(self._h, first, count, ll, C.byref(sec)))
        return list(ll), sec.value

    def sync(self):
        _check(lib().culda_sync(self._h))

    def loglik_joint(self):
        v = C.c_double()
        _check(lib().culda_loglik_joint(self._h, C.byref(v)))
        return v.value

    def loglik_joint_noalloc(self):
        """The same value without any device allocation (static scratch)."""
        v = C.c_double()
        _check(lib().culda_loglik_joint_noalloc(self._h, C.byref(v)))
        return v.value

    # -- state transfer ---------------------------------------------------
    def get_z(self, chunk):
        z = np.empty(self._chunk_tokens[chunk], np.uint16)
        _check(lib().culda_get_z(self._h, chunk, _ptr(z, C.c_uint16), z.size))
        return z

    def set_z(self, chunk, z):
        z = np.ascontiguousarray(z, dtype=np.uint16)
        _check(lib().culda_set_z(self._h, chunk, _ptr(z, C.c_uint16), z.size))

    def get_phi(self):
        V = self._info.num_words
        nwk = np.empty(V * self.k, np.int32)
        nk = np.empty(self.k, np.int32)
        _check(lib().culda_get_phi(self._h, _ptr(nwk, C.c_int32), _ptr(nk, C.c_int32)))
        return nwk.reshape(V, self.k), nk

    def get_phi_head(self):
        V = self._info.num_words
        h = np.empty(V * self.k, np.float32)
        _check(lib().culda_get_phi_head(self._h, _ptr(h, C.c_float)))
        return h.reshape(V, self.k)

    def get_theta(self, chunk):
        """Compact CSR over the chunk's docs: (offsets, counts, topics)."""
        v = ChunkView()
        _check(lib().culda_corpus_chunk(self.corpus._h, chunk, C.byref(v)))
        off = np.empty(v.docid_end - v.docid_start + 1, np.int64)
        _check(lib().culda_get_theta(self._h, chunk, _ptr(off, C.c_int64), None, None))
        n = int(off[-1])
        cnt = np.empty(max(n, 1), np.int32)
        top = np.empty(max(n, 1), np.int16)
        _check(lib().culda_get_theta(self._h, chunk, _ptr(off, C.c_int64), _ptr(cnt, C.c_int32),
                                     _ptr(top, C.c_int16)))
        return off, cnt[:n], top[:n]

    def save_phi(self, prefix):
        _check(lib().culda_save_phi(self._h, str(prefix).encode()))

    def save_theta(self, prefix):
        _check(lib().culda_save_theta(self._h, str(prefix).encode()))

    def save_checkpoint(self, path, next_iteration):
        _check(lib().culda_save_checkpoint(self._h, str(path).encode(), next_iteration))

    def load_checkpoint(self, path):
        it = C.c_int32()
        _check(lib().culda_load_checkpoint(self._h, str(path).encode(), C.byref(it)))
        return it.value

    def stats(self):
        s = Stats()
        _check(lib().culda_get_stats(self._h, C.byref(s)))
        return s

    def sampler_bytes(self):
        v = C.c_double()
        _check(lib().culda_sampler_bytes(self._h, C.byref(v)))
        return v.value

    def last_sweeps(self):
        """(sampler ms, algorithmic bytes, launches) of the last iterate(),
        without waiting for the theta rebuilds it left running."""
        ms, b, n = C.c_double(), C.c_double(), C.c_int32()
        _check(lib().culda_last_sweeps(self._h, C.byref(ms), C.byref(b), C.byref(n)))
        return ms.value, b.value, n.value

    def stream_copy(self, n_bytes, repeats=10):
        v = C.c_double()
        _check(lib().culda_stream_copy(self._h, n_bytes, repeats, C.byref(v)))
        return v.value

    def set_eval(self, eval_corpus: Corpus):
        """This sampler holds held-out docs' observed halves; give it the
        evaluated halves."""
        _check(lib().culda_heldout_set_eval(self._h, eval_corpus._h))
        self.eval_corpus = eval_corpus

    def heldout_ll(self, train: "Sampler", fold_in_sweeps=10):
        """Document-completion LL/token of the evaluated tokens under the
        training sampler's current counts."""
        v = C.c_double()
        _check(lib().culda_heldout_ll(train._h, self._h, fold_in_sweeps, C.byref(v)))
        return v.value

    # -- learned-MH proposal ---------------------------------------------
    MH_IN, MH_HIDDEN, MH_EMB = 384, 512, 128

    def mh_enable(self, batch=0, infer_batch=0, max_snapshots=0, lr=0.0, beta1=0.0, beta2=0.0, eps=0.0,
                  init_seed=1):
        self.mh_params = MHParams(batch, infer_batch, max_snapshots, 0, lr, beta1, beta2, eps, init_seed)
        _check(lib().culda_mh_enable(self._h, C.byref(self.mh_params)))

    def mh_collect(self):
        n = C.c_int32()
        _check(lib().culda_mh_collect(self._h, C.byref(n)))
        return n.value

    def mh_train(self, epochs=1):
        loss, sec = C.c_double(), C.c_double()
        _check(lib().culda_mh_train(self._h, epochs, C.byref(loss), C.byref(sec)))
        return loss.value, sec.value

    def mh_iterate(self, iteration, proposals=2):
        rate, sec = C.c_double(), C.c_double()
        _check(lib().culda_mh_iterate(self._h, iteration, proposals, C.byref(rate), C.byref(sec)))
        return rate.value, sec.value

    def _mh_shapes(self):
        K, V = self.k, self._info.num_words
        return {"w1": (self.MH_IN, self.MH_HIDDEN), "b1": (self.MH_HIDDEN,), "w2": (self.MH_HIDDEN, K), "b2": (K,),
                "emb": (V, self.MH_EMB)}

    def mh_get_weights(self):
        out = {k: np.empty(v, np.float32) for k, v in self._mh_shapes().items()}
        _check(lib().culda_mh_get_weights(self._h, *[_ptr(out[k], C.c_float) for k in ("w1", "b1", "w2", "b2", "emb")]))
        return out

    def mh_set_weights(self, w):
        arr = {k: np.ascontiguousarray(w[k], np.float32).reshape(v) for k, v in self._mh_shapes().items()}
        _check(lib().culda_mh_set_weights(self._h, *[_ptr(arr[k], C.c_float) for k in ("w1", "b1", "w2", "b2", "emb")]))

    def mh_logits(self, docs, words):
        d = np.ascontiguousarray(docs, np.int32)
        w = np.ascontiguousarray(words, np.int32)
        out = np.empty((d.size, self.k), np.float32)
        _check(lib().culda_mh_logits(self._h, d.size, _ptr(d, C.c_int32), _ptr(w, C.c_int32), _ptr(out, C.c_float)))
        return out

    def mh_gradients(self, docs, words, targets, train_step=False):
        d = np.ascontiguousarray(docs, np.int32)
        w = np.ascontiguousarray(words, np.int32)
        t = np.ascontiguousarray(targets, np.int32)
        g = {k: np.empty(v, np.float32) for k, v in self._mh_shapes().items()}
        loss = C.c_double()
        _check(lib().culda_mh_gradients(self._h, d.size, _ptr(d, C.c_int32), _ptr(w, C.c_int32), _ptr(t, C.c_int32),
                                        1 if train_step else 0, C.byref(loss),
                                        *[_ptr(g[k], C.c_float) for k in ("w1", "b1", "w2", "b2", "emb")]))
        return loss.value, g

    def mh_tables(self, first, n):
        W = np.empty((n, self.k), np.int32)
        F = np.empty((n, self.k), np.uint32)
        A = np.empty((n, self.k), np.uint32)
        _check(lib().culda_mh_tables(self._h, first, n, _ptr(W, C.c_int32), _ptr(F, C.c_uint32), _ptr(A, C.c_uint32)))
        return W, F, A

    def mh_sketch(self):
        out = np.empty((self._info.num_docs, 256), np.uint16)
        _check(lib().culda_mh_sketch(self._h, _ptr(out, C.c_uint16)))
        return out

    def autotune(self):
        """Pick the sampler's tokens per slot by timing one sweep per size and,
        with several chunks, the theta rebuild's workgroups per CU by timing
        whole iterations (``theta_groups``); the chain's state is kept."""
        t = Tune()
        _check(lib().culda_autotune(self._h, C.byref(t)))
        return t


class Batch:
    """Many independent corpora sampled together, one launch per step.

    Each added corpus is its own model with its own alpha, beta and seed;
    its chain is bit-identical to that corpus run alone in a ``Sampler``.
    """

    def __init__(self, k=1024, device=0, slot_tokens=0, profile=False):
        self.params = Params(k, 0.0, 0.0, 0, device, slot_tokens, 1 if profile else 0, 1)
        h = C.c_void_p()
        _check(lib().culda_batch_create(C.byref(self.params), C.byref(h)))
        self._h = h
        self.k = k
        self.corpora = []

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().culda_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, corpus: Corpus, alpha, beta, seed):
        m = C.c_int32()
        _check(lib().culda_batch_add(self._h, corpus._h, alpha, beta, seed, C.byref(m)))
        self.corpora.append(corpus)
        return m.value

    def upload(self):
        _check(lib().culda_batch_upload(self._h))

    def init(self):
        _check(lib().culda_batch_init(self._h))

    def set_persistent(self, groups):
        _check(lib().culda_batch_set_persistent(self._h, groups))

    def iterate(self, first_iteration, count=1, use_graph=False):
        """Run ``count`` iterations; returns (ll[count, M], seconds)."""
        M = len(self.corpora)
        ll = np.empty(count * M, np.float64)
        sec = C.c_double()
        _check(lib().culda_batch_iterate(self._h, first_iteration, count, 1 if use_graph else 0,
                                         _ptr(ll, C.c_double), C.byref(sec)))
        return ll.reshape(count, M), sec.value

    def autotune(self):
        t = Tune()
        _check(lib().culda_batch_autotune(self._h, C.byref(t)))
        return t

    def info(self):
        m, n, s = C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().culda_batch_info(self._h, C.byref(m), C.byref(n), C.byref(s)))
        return m.value, n.value, s.value

    def stats(self):
        s = Stats()
        _check(lib().culda_batch_get_stats(self._h, C.byref(s)))
        return s

    def sampler_bytes(self):
        v = C.c_double()
        _check(lib().culda_batch_sampler_bytes(self._h, C.byref(v)))
        return v.value

    def get_z(self, model, chunk):
        z = np.empty(self.corpora[model].chunk_tokens(chunk), np.uint16)
        _check(lib().culda_batch_get_z(self._h, model, chunk, _ptr(z, C.c_uint16), z.size))
        return z

    def get_nk(self, model):
        nk = np.empty(self.k, np.int32)
        _check(lib().culda_batch_get_phi(self._h, model, None, _ptr(nk, C.c_int32)))
        return nk

    def get_phi(self, model):
        V = self.corpora[model].info().num_words
        nwk = np.empty(V * self.k, np.int32)
        nk = np.empty(self.k, np.int32)
        _check(lib().culda_batch_get_phi(self._h, model, _ptr(nwk, C.c_int32), _ptr(nk, C.c_int32)))
        return nwk.reshape(V, self.k), nk


__all__ = ["Corpus", "Sampler", "Batch", "Tune", "MHParams", "CuldaError", "format_text", "lib", "LIB_PATH", "CLI_PATH"]
