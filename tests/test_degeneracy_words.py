"""Degeneracy words (host code of the product, no GPU needed).

For a found shot the residue v = e ^ e_hat lies in ker(H); the decoder reports
degenerate = 1 iff v is in the row space of the other CSS matrix
(grand_decoder.cu:1080-1087, in_rowspace at :480-493).  The search kernel tests
this as "XOR of the degeneracy words over supp(v) is 0".  These tests pin that
equivalence against a GF(2) rank test on every vector of ker(H) for random
pairs (CSS or not), on the BB72 pair (12 logical qubits -> 12 functionals), and
the fall-back return of -1 when more than 32 functionals would be needed.
"""
import ctypes

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes


def _lib():
    lib = ctypes.CDLL(G.DEFAULT_LIB)
    lib.grand_degeneracy_words.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p]
    lib.grand_degeneracy_words.restype = ctypes.c_int
    return lib


def _words(lib, H, O):
    H = np.ascontiguousarray(H, dtype=np.int32)
    O = np.ascontiguousarray(O, dtype=np.int32)
    n, m = H.shape
    out = np.zeros(m, dtype=np.uint32)
    t = lib.grand_degeneracy_words(H.ctypes.data, O.ctypes.data, n, m, out.ctypes.data)
    return t, out


def gf2_rank(M):
    M = (np.array(M, dtype=np.uint8) & 1).copy()
    r = 0
    rows, cols = M.shape
    for c in range(cols):
        p = next((i for i in range(r, rows) if M[i, c]), None)
        if p is None:
            continue
        M[[r, p]] = M[[p, r]]
        for i in range(rows):
            if i != r and M[i, c]:
                M[i] ^= M[r]
        r += 1
        if r == rows:
            break
    return r


def in_rowspace(v, O):
    return gf2_rank(O) == gf2_rank(np.vstack([O, v]))


def kernel_vectors(H):
    """Every vector of ker(H) over GF(2) (small m)."""
    n, m = H.shape
    allv = ((np.arange(1 << m)[:, None] >> np.arange(m)) & 1).astype(np.int64)
    return allv[((allv @ H.T.astype(np.int64)) % 2 == 0).all(axis=1)]


def _xor_words(words, v):
    acc = 0
    for c in np.nonzero(v)[0]:
        acc ^= int(words[c])
    return acc


@pytest.mark.parametrize("seed", range(12))
def test_words_decide_rowspace_on_kernel_of_h(seed):
    lib = _lib()
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    m = int(rng.integers(6, 13))
    H = (rng.random((n, m)) < 0.4).astype(np.int32)
    if seed % 3 == 0:
        # a CSS pair: O's rows orthogonal to H's rows
        K = kernel_vectors(H)
        O = K[rng.integers(0, len(K), size=n)].astype(np.int32)
    else:
        O = (rng.random((n, m)) < 0.4).astype(np.int32)
    t, words = _words(lib, H, O)
    assert 0 <= t <= m
    assert all(int(w) < (1 << t) for w in words)
    for v in kernel_vectors(H):
        assert (_xor_words(words, v) == 0) == in_rowspace(v, O), (seed, v)


def test_bb72_has_one_functional_per_logical_qubit():
    lib = _lib()
    hx, hz = codes.bb(72)
    t, words = _words(lib, hz, hx)
    assert t == 12                      # [[72, 12, 6]]
    rng = np.random.default_rng(3)
    # stabilizers (rows of hx) are degenerate residues
    for _ in range(50):
        v = (rng.integers(0, 2, size=hx.shape[0]) @ hx) % 2
        assert _xor_words(words, v) == 0
    # random vectors of ker(hz) (GF(2) null-space basis from the RREF): compare with the rank test
    basis = []
    R = hz.copy() % 2
    piv = []
    r = 0
    for c in range(R.shape[1]):
        p = next((i for i in range(r, R.shape[0]) if R[i, c]), None)
        if p is None:
            continue
        R[[r, p]] = R[[p, r]]
        for i in range(R.shape[0]):
            if i != r and R[i, c]:
                R[i] ^= R[r]
        piv.append(c)
        r += 1
    for f in (c for c in range(R.shape[1]) if c not in piv):
        k = np.zeros(R.shape[1], dtype=np.int64)
        k[f] = 1
        for i, pc in enumerate(piv):
            if R[i, f]:
                k[pc] = 1
        basis.append(k)
    basis = np.array(basis)
    assert ((basis @ hz.T) % 2 == 0).all()
    for _ in range(200):
        v = (rng.integers(0, 2, size=len(basis)) @ basis) % 2
        assert (_xor_words(words, v) == 0) == in_rowspace(v, hx)


def test_more_than_32_functionals_returns_minus_one():
    lib = _lib()
    m = 40
    H = np.zeros((1, m), np.int32)
    O = np.zeros((1, m), np.int32)
    O[0, 0] = 1
    t, _ = _words(lib, H, O)
    assert t == -1
