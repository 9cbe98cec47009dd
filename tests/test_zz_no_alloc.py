"""culda_loglik_joint_noalloc (static device scratch, no allocation) against
culda_loglik_joint (a buffer allocated per call), through the C ABI: two
calls in a row, the second on a smaller corpus, in one process; and after
more iterations of the first context.  The joint LL is FP64 with lgamma;
the two differ only in the order of the final sums (<= 1e-12 relative)."""
import numpy as np
import pytest

import culda_cgs_amd as cg
from helpers import chunks_from_docs, product_corpus, random_docs

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _ctx(seed, V, D, mean_len, chunks, K):
    rng = np.random.default_rng(seed)
    arrays = chunks_from_docs(V, random_docs(rng, V, D, mean_len, empty_docs=1), chunks)
    s = cg.Sampler(k=K, alpha=0.1, beta=0.02, seed=seed)
    s.upload(product_corpus(V, *arrays))
    s.init_z()
    s.rebuild_all()
    for it in (1, 2):
        s.iterate(it)
    return s


def test_noalloc_joint_ll_equals_allocating_one():
    big = _ctx(41, 3000, 400, 120, 5, 1024)
    a1, b1 = big.loglik_joint(), big.loglik_joint_noalloc()
    assert np.isfinite(a1) and abs(a1 - b1) <= RTOL * abs(a1), (a1, b1)
    small = _ctx(42, 200, 30, 20, 1, 64)  # fewer chunks, docs and topics: less of the scratch in use
    a2, b2 = small.loglik_joint(), small.loglik_joint_noalloc()
    assert abs(a2 - b2) <= RTOL * abs(a2), (a2, b2)
    assert abs(a2 - a1) > 1e-3 * abs(a1)  # a different value, not the first call's left over
    big.iterate(3)
    a3, b3 = big.loglik_joint(), big.loglik_joint_noalloc()
    assert abs(a3 - b3) <= RTOL * abs(a3) and a3 != a1
    assert big.loglik_joint_noalloc() == b3  # repeatable
    big.close()
    small.close()


def test_noalloc_joint_ll_state_and_limit_errors():
    s = cg.Sampler(k=64, alpha=0.1, beta=0.02, seed=1)
    with pytest.raises(cg.CuldaError) as e:
        s.loglik_joint_noalloc()
    assert e.value.status == -6
    rng = np.random.default_rng(3)
    docs = [rng.integers(0, 100, 10) for _ in range(640)]
    arrays = chunks_from_docs(100, docs, 64)  # 64 chunks of 10 docs: one chunk over the limit
    assert min(t.size for t in arrays[3]) == 100
    s.upload(product_corpus(100, *arrays))
    s.init_z()
    s.rebuild_all()
    with pytest.raises(cg.CuldaError) as e:
        s.loglik_joint_noalloc()
    assert e.value.status == -7
    assert np.isfinite(s.loglik_joint())
    s.close()
