"""CPU oracle of the learned-MH proposal step -- TEST INFRASTRUCTURE.

Only tests/ import this module, as the checker of libculda.so's learned-MH
path; the product never uses it.  The learned-MH proposal has no
counterpart in the reference (it is an extension beside the plain sampler
of kernel/lda_train_kernel.cu:20-271), so everything here is restated from
its mathematical definition:

* ``sketch``: signed count sketch of a doc's topic counts divided by n_d,
  rounded to BF16 (round to nearest even);
* ``alias_sweep``: Walker alias table from integer weights by the
  *sequential* sweep (lights in index order fill up from the current heavy;
  a heavy whose residual drops to M or below is closed and aliases the next
  heavy) -- an independent formulation of the table the device builds with
  prefix sums and binary searches;
* ``alias_mass``: the exact integer mass every outcome receives from a
  table (what an alias draw samples);
* ``mh_token``: the independence-MH steps of one token given its weights,
  its table, the frozen counts and the Philox stream;
* ``mh_transition``: the one-step MH transition matrix, for invariance
  checks by enumeration.
"""
from __future__ import annotations

import numpy as np

from . import philox4x64_10

ALIAS_BITS = 20
PURPOSE_MH = 2


def bf16_bits(x):
    """FP32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint16)


def bf16_to_f32(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32)


def sketch_bucket(k):
    k = np.asarray(k, np.uint64)
    return ((k * 2654435761) & 0xFFFFFFFF) >> 24


def sketch_negative(k):
    k = np.asarray(k, np.uint64)
    return (((k * 0x85EBCA6B) & 0xFFFFFFFF) >> 31) != 0


def sketch(topics, counts, nd, width=256):
    """BF16 bits of sum_k sign(k) n_dk [h(k) = j] / n_d (integer sums, one
    FP32 division per bucket)."""
    acc = np.zeros(width, np.int64)
    t = np.asarray(topics, np.int64)
    c = np.asarray(counts, np.int64)
    np.add.at(acc, sketch_bucket(t).astype(np.int64), np.where(sketch_negative(t), -c, c))
    if nd == 0:
        return np.zeros(width, np.uint16)
    return bf16_bits(acc.astype(np.float32) / np.float32(nd))


def alias_sweep(W, M=1 << ALIAS_BITS):
    """(thresholds F, aliases A) of integer weights W with sum W = K * M."""
    W = [int(x) for x in W]
    K = len(W)
    assert sum(W) == K * M and min(W) >= 1
    F, A = [0] * K, [0] * K
    lights = [k for k in range(K) if W[k] < M]
    heavies = [k for k in range(K) if W[k] >= M]
    i, j = 0, 0
    r = W[heavies[0]]
    while True:
        h = heavies[j]
        if j == len(heavies) - 1:
            while i < len(lights):
                lt = lights[i]
                F[lt], A[lt] = W[lt], h
                r -= M - W[lt]
                i += 1
            assert r == M
            F[h], A[h] = M, h
            break
        if r > M:
            assert i < len(lights), "excess left with no light to absorb it"
            lt = lights[i]
            F[lt], A[lt] = W[lt], h
            r -= M - W[lt]
            i += 1
        else:
            F[h], A[h] = r, heavies[j + 1]
            r = W[heavies[j + 1]] - (M - r)
            j += 1
    return np.array(F, np.int64), np.array(A, np.int64)


def alias_mass(F, A, M=1 << ALIAS_BITS):
    """Integer mass of every outcome: bucket j gives F_j to j and M - F_j to A_j."""
    F = np.asarray(F, np.int64)
    A = np.asarray(A, np.int64)
    mass = F.copy()
    np.add.at(mass, A, M - F)
    return mass


def alias_draw(F, A, x0, x1, K):
    """The outcome of one draw from 64-bit words x0 (bucket) and x1 (threshold)."""
    j = (int(x0) * K) >> 64
    rr = int(x1) >> (64 - ALIAS_BITS)
    return j if rr < int(F[j]) else int(A[j])


def target(ndk, nwk, nk, V, alpha, beta):
    """Collapsed conditional weight (n_dk + a)(n_wk + b) / (n_k + V b), FP64,
    alpha and beta being the FP32 hyper-parameters widened to FP64."""
    a = float(np.float32(alpha))
    b = float(np.float32(beta))
    return ((float(ndk) + a) * (float(nwk) + b)) / (float(nk) + float(V) * b)


def mh_token(W, F, A, p_of, z0, seed, sweep, gid, proposals):
    """z after `proposals` independence-MH steps of one token."""
    K = len(W)
    cur = int(z0)
    pc = p_of(cur)
    for r in range(proposals):
        x = philox4x64_10([gid, sweep, PURPOSE_MH, r], [seed, 0])
        kp = alias_draw(F, A, x[0], x[1], K)
        u = float(int(x[2]) >> 11) * 2.0 ** -53
        if kp != cur:
            pp = p_of(kp)
            if u * (pc * float(W[kp])) < pp * float(W[cur]):
                cur, pc = kp, pp
    return cur


def mh_transition(W, p):
    """One-step transition matrix of the independence MH chain with
    proposal q = W / sum W and target proportional to p (FP64)."""
    W = np.asarray(W, np.float64)
    p = np.asarray(p, np.float64)
    q = W / W.sum()
    K = len(W)
    T = np.zeros((K, K))
    for k in range(K):
        for kp in range(K):
            if kp != k:
                T[k, kp] = q[kp] * min(1.0, (p[kp] * q[k]) / (p[k] * q[kp]))
        T[k, k] = 1.0 - T[k].sum()
    return T
