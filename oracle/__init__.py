"""CPU oracle for the plain CGS iteration -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the timed CPU baseline; the
product (culda_cgs_amd / libculda.so) never uses it.

ctypes wrapper over liboracle.so, built from culda_oracle.c (a plain-C
restatement of the reference schedule; each function cites the reference
file:line it follows).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"
_lib = None


class _OcCorpus(C.Structure):
    _fields_ = [("V", C.c_int32), ("D", C.c_int32), ("C", C.c_int32), ("K", C.c_int32),
                ("doc_cum", C.POINTER(C.c_int64)), ("chunk_doc_end", C.POINTER(C.c_int32)),
                ("word_idx", C.POINTER(C.POINTER(C.c_int64))), ("tokens", C.POINTER(C.POINTER(C.c_int32))),
                ("z", C.POINTER(C.POINTER(C.c_uint16)))]


class SweepOut(C.Structure):
    _fields_ = [("ll_sum", C.c_double), ("ll_sum_exact", C.c_double), ("boundary", C.c_int64),
                ("mismatch", C.c_int64), ("s_bucket", C.c_int64), ("tokens", C.c_int64)]


THETA_MAX_COUNT = 63  # largest count of one theta entry


def lib():
    global _lib
    if _lib is None:
        # CULDA_ORACLE_LIBRARY: another build of the oracle (the sanitized one)
        path = Path(os.environ.get("CULDA_ORACLE_LIBRARY", LIB_PATH))
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(str(path))
        P = C.POINTER
        L.oc_philox4x64_10.argtypes = [P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]
        L.oc_philox4x64_10.restype = None
        L.oc_init_z.argtypes = [P(_OcCorpus), C.c_uint64]
        L.oc_init_z.restype = None
        L.oc_rebuild_phi.argtypes = [P(_OcCorpus), P(C.c_int32), P(C.c_int32)]
        L.oc_rebuild_phi.restype = None
        L.oc_rebuild_theta.argtypes = [P(_OcCorpus), P(C.c_int64), P(C.c_int32), P(C.c_int16)]
        L.oc_rebuild_theta.restype = C.c_int64
        L.oc_phi_head.argtypes = [C.c_int32, C.c_int32, C.c_float, P(C.c_int32), P(C.c_int32), P(C.c_float)]
        L.oc_phi_head.restype = None
        L.oc_sweep.argtypes = [P(_OcCorpus), C.c_float, C.c_float, C.c_uint64, C.c_uint32, P(C.c_int32),
                               P(C.c_int32), P(C.c_int64), P(C.c_int32), P(C.c_int16), P(P(C.c_uint16)),
                               P(P(C.c_uint8)), P(SweepOut)]
        L.oc_sweep.restype = None
        L.oc_sweep_range.argtypes = [P(_OcCorpus), C.c_float, C.c_float, C.c_uint64, C.c_uint32, P(C.c_int32),
                                     P(C.c_int32), P(C.c_int64), P(C.c_int32), P(C.c_int16), C.c_int32, C.c_int32,
                                     C.c_int, P(P(C.c_uint16)), P(P(C.c_uint8)), P(SweepOut)]
        L.oc_sweep_range.restype = None
        L.oc_sweep_fast.argtypes = [P(_OcCorpus), C.c_float, C.c_float, C.c_uint64, C.c_uint32, P(C.c_int32),
                                    P(C.c_int32), P(C.c_int64), P(C.c_int32), P(C.c_int16), P(C.c_double)]
        L.oc_sweep_fast.restype = None
        L.oc_loglik_joint.argtypes = [P(_OcCorpus), C.c_float, C.c_float, P(C.c_int32), P(C.c_int32),
                                      P(C.c_int64), P(C.c_int32), P(C.c_int16)]
        L.oc_loglik_joint.restype = C.c_double
        L.oc_max_threads.restype = C.c_int
        L.oc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def philox4x64_10(counter, key):
    c = np.ascontiguousarray(counter, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    o = np.zeros(4, np.uint64)
    lib().oc_philox4x64_10(_p(c, C.c_uint64), _p(k, C.c_uint64), _p(o, C.c_uint64))
    return o


def set_threads(n):
    lib().oc_set_threads(int(n))


def max_threads():
    return lib().oc_max_threads()


class OracleModel:
    """Corpus arrays (data_chunk layout) + the oracle's model state.

    doc_cum[D], chunk_doc_end[C], word_idx[c][V+1], tokens[c][N_c] (global doc
    ids); z[c][N_c] is owned here.
    """

    def __init__(self, num_words, k, doc_cum, chunk_doc_end, word_idx, tokens, z=None):
        self.V, self.K = int(num_words), int(k)
        self.doc_cum = np.ascontiguousarray(doc_cum, dtype=np.int64)
        self.D = self.doc_cum.size
        self.chunk_doc_end = np.ascontiguousarray(chunk_doc_end, dtype=np.int32)
        self.word_idx = [np.ascontiguousarray(w, dtype=np.int64) for w in word_idx]
        self.tokens = [np.ascontiguousarray(t, dtype=np.int32) for t in tokens]
        self.C = len(self.word_idx)
        if z is None:
            self.z = [np.zeros(t.size, np.uint16) for t in self.tokens]
        else:
            self.z = [np.array(x, dtype=np.uint16, copy=True) for x in z]
        nc = self.C
        self._wi = (C.POINTER(C.c_int64) * nc)(*[_p(x, C.c_int64) for x in self.word_idx])
        self._tk = (C.POINTER(C.c_int32) * nc)(*[_p(x, C.c_int32) for x in self.tokens])
        self._z = (C.POINTER(C.c_uint16) * nc)(*[_p(x, C.c_uint16) for x in self.z])
        self._c = _OcCorpus(self.V, self.D, nc, self.K, _p(self.doc_cum, C.c_int64),
                            _p(self.chunk_doc_end, C.c_int32), self._wi, self._tk, self._z)
        lens = np.diff(np.concatenate([[0], self.doc_cum]))
        self.doc_len = lens
        # theta entries: a topic with count c takes ceil(c / THETA_MAX_COUNT)
        self.theta_cap = int(np.minimum(lens, self.K + (lens + THETA_MAX_COUNT - 1) // THETA_MAX_COUNT).sum())
        self.nwk = np.zeros(self.V * self.K, np.int32)
        self.nk = np.zeros(self.K, np.int32)
        self.th_off = np.zeros(self.D + 1, np.int64)
        self.th_cnt = np.zeros(max(self.theta_cap, 1), np.int32)
        self.th_top = np.zeros(max(self.theta_cap, 1), np.int16)
        self.token_nnz = 0

    @property
    def num_tokens(self):
        return int(self.doc_cum[-1]) if self.D else 0

    def set_z(self, zs):
        for dst, src in zip(self.z, zs):
            dst[:] = src

    def init_z(self, seed):
        lib().oc_init_z(C.byref(self._c), seed)

    def rebuild(self):
        L = lib()
        L.oc_rebuild_phi(C.byref(self._c), _p(self.nwk, C.c_int32), _p(self.nk, C.c_int32))
        self.token_nnz = L.oc_rebuild_theta(C.byref(self._c), _p(self.th_off, C.c_int64),
                                            _p(self.th_cnt, C.c_int32), _p(self.th_top, C.c_int16))

    def rebuild_phi(self):
        """n_wk / n_k only."""
        lib().oc_rebuild_phi(C.byref(self._c), _p(self.nwk, C.c_int32), _p(self.nk, C.c_int32))

    def rebuild_theta(self):
        """Theta only (n_wk / n_k left as they are: fold-in with phi fixed)."""
        self.token_nnz = lib().oc_rebuild_theta(C.byref(self._c), _p(self.th_off, C.c_int64),
                                                _p(self.th_cnt, C.c_int32), _p(self.th_top, C.c_int16))

    def phi_head(self, beta):
        h = np.zeros(self.V * self.K, np.float32)
        lib().oc_phi_head(self.V, self.K, beta, _p(self.nwk, C.c_int32), _p(self.nk, C.c_int32),
                          _p(h, C.c_float))
        return h.reshape(self.V, self.K)

    def theta_entries(self, d):
        """Doc d's theta entries (counts, topics), a topic with n_dk > 63
        split into ceil(n_dk / 63) entries -- the order S is summed in."""
        o0, o1 = self.th_off[d], self.th_off[d + 1]
        return self.th_cnt[o0:o1], self.th_top[o0:o1]

    def theta_rows(self, d):
        """Doc d's nonzero (n_dk, topic) pairs, ascending topic."""
        cnt, top = self.theta_entries(d)
        if top.size == 0:
            return cnt, top
        first = np.concatenate([[True], top[1:] != top[:-1]])
        starts = np.flatnonzero(first)
        return np.add.reduceat(cnt, starts).astype(np.int32), top[starts]

    def sweep(self, alpha, beta, seed, sweep, want_exact=False):
        """Faithful FP32 sweep (writes self.z); returns (SweepOut, exact_z, flags)."""
        out = SweepOut()
        ex = fl = None
        exp = flp = None
        if want_exact:
            ex = [np.zeros_like(z) for z in self.z]
            fl = [np.zeros(z.size, np.uint8) for z in self.z]
            exp = (C.POINTER(C.c_uint16) * self.C)(*[_p(x, C.c_uint16) for x in ex])
            flp = (C.POINTER(C.c_uint8) * self.C)(*[_p(x, C.c_uint8) for x in fl])
        lib().oc_sweep(C.byref(self._c), alpha, beta, seed, sweep, _p(self.nwk, C.c_int32), _p(self.nk, C.c_int32),
                       _p(self.th_off, C.c_int64), _p(self.th_cnt, C.c_int32), _p(self.th_top, C.c_int16),
                       exp, flp, C.byref(out))
        return out, ex, fl

    def sweep_fast(self, alpha, beta, seed, sweep):
        """The same FP32 sweep without the FP64 shadow (timing); returns the
        FP64 sum of the LL terms."""
        ll = C.c_double()
        lib().oc_sweep_fast(C.byref(self._c), alpha, beta, seed, sweep, _p(self.nwk, C.c_int32),
                            _p(self.nk, C.c_int32), _p(self.th_off, C.c_int64), _p(self.th_cnt, C.c_int32),
                            _p(self.th_top, C.c_int16), C.byref(ll))
        return ll.value

    def sweep_words(self, alpha, beta, seed, sweep, w_lo, w_hi, exact=False):
        """The sweep restricted to the tokens of words [w_lo, w_hi) (other
        tokens keep their z); exact adds the FP64 shadow (boundary draws
        counted).  Returns SweepOut (tokens = tokens swept)."""
        out = SweepOut()
        lib().oc_sweep_range(C.byref(self._c), alpha, beta, seed, sweep, _p(self.nwk, C.c_int32),
                             _p(self.nk, C.c_int32), _p(self.th_off, C.c_int64), _p(self.th_cnt, C.c_int32),
                             _p(self.th_top, C.c_int16), int(w_lo), int(w_hi), 1 if exact else 0, None, None,
                             C.byref(out))
        return out

    def loglik_joint(self, alpha, beta):
        return lib().oc_loglik_joint(C.byref(self._c), alpha, beta, _p(self.nwk, C.c_int32),
                                     _p(self.nk, C.c_int32), _p(self.th_off, C.c_int64),
                                     _p(self.th_cnt, C.c_int32), _p(self.th_top, C.c_int16))
