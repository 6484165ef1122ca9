"""CPU statement of the learned pass (no counterpart in the reference), used to
check the GPU's learned decoding shot by shot.

Per shot with a nonzero syndrome:
  * the first active list A (threshold max_soma; grand_decoder.cu:1284-1290,
    :406-416, :495-504); for the DEM family a list longer than max_size means
    the first iteration searches nothing (:1550-1554), and so does the pass;
  * positions sorted by descending score, ties in list order (P);
  * patterns tried in stages k = 0..tam-1: the patterns whose last position in P
    is k, weight 1..maxw, each weight lexicographic over P[0..k);
  * the first pattern whose columns sum to the syndrome wins, with count = the
    patterns tried including it; after `budget` patterns without a match, the
    plain search result follows with the pass's patterns added to its count.
"""
import itertools

import numpy as np


def first_active(H, s):
    H = H.astype(np.int64)
    colw = H.sum(0)
    max_soma = int(colw.max())
    score = H[s == 1].sum(0)
    if not (colw == colw[0]).all():
        score = np.where(colw == 0, 0, score + max_soma - colw)
    return [c for c in range(H.shape[1]) if score[c] >= max_soma]


def learned_pass(H, s, scores, per_column, maxw, budget, max_size=None):
    """Returns (hit columns or None, patterns tried, tam)."""
    A = first_active(H, s)
    tam = len(A)
    if max_size is not None and tam > max_size:
        return None, 0, 0
    if tam == 0 or (not per_column and tam > len(scores)):
        return None, 0, 0
    sc = np.array([scores[c] if per_column else scores[i] for i, c in enumerate(A)], dtype=np.float32)
    key = np.where(np.isnan(sc), np.inf, -sc).astype(np.float32)
    P = [A[i] for i in np.argsort(key, kind="stable")]
    # columns and syndrome as Python integers (bit r = row r): a pattern
    # matches when the XOR of its columns equals the syndrome
    colint = [int("".join("1" if H[r, c] else "0" for r in reversed(range(H.shape[0]))), 2) for c in P]
    sint = int("".join("1" if s[r] else "0" for r in reversed(range(H.shape[0]))), 2)
    tried = 0
    for k in range(tam):
        tk = sint ^ colint[k]
        for wp in range(min(maxw, k + 1)):
            for S in itertools.combinations(range(k), wp):
                if tried >= budget:
                    return None, tried, tam
                tried += 1
                x = tk
                for j in S:
                    x ^= colint[j]
                if x == 0:
                    return sorted([P[j] for j in S] + [P[k]]), tried, tam
    return None, tried, tam
