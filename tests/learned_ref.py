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


def _pack(bits):
    """0/1 rows (..., n) -> uint64 words (..., ceil(n / 64)), bit r of the row in word r // 64."""
    bits = np.asarray(bits, dtype=np.uint64)
    n = bits.shape[-1]
    nw = (n + 63) // 64
    out = np.zeros(bits.shape[:-1] + (nw,), dtype=np.uint64)
    for r in range(n):
        out[..., r // 64] |= bits[..., r] << np.uint64(r % 64)
    return out


def learned_pass(H, s, scores, per_column, maxw, budget, max_size=None, block=4096):
    """Returns (hit columns or None, patterns tried, tam).  The patterns of one
    (stage, weight) are tested in blocks of lexicographic order (numpy XOR of
    the packed columns); the first matching one in order is the hit."""
    A = first_active(H, s)
    tam = len(A)
    if max_size is not None and tam > max_size:
        return None, 0, 0
    if tam == 0 or (not per_column and tam > len(scores)):
        return None, 0, 0
    sc = np.array([scores[c] if per_column else scores[i] for i, c in enumerate(A)], dtype=np.float32)
    key = np.where(np.isnan(sc), np.inf, -sc).astype(np.float32)
    P = [A[i] for i in np.argsort(key, kind="stable")]
    cols = _pack(np.asarray(H)[:, P].T)              # [tam][words]
    sv = _pack(np.asarray(s))
    tried = 0
    for k in range(tam):
        tk = sv ^ cols[k]
        for wp in range(min(maxw, k + 1)):
            combos = itertools.combinations(range(k), wp)
            while True:
                left = budget - tried
                if left <= 0:
                    return None, tried, tam
                chunk = list(itertools.islice(combos, min(block, left)))
                if not chunk:
                    break
                idx = np.array(chunk, dtype=np.int64).reshape(len(chunk), wp)
                x = np.repeat(tk[None, :], len(chunk), axis=0)
                for j in range(wp):
                    x ^= cols[idx[:, j]]
                z = np.nonzero(~x.any(axis=1))[0]
                if z.size:
                    i = int(z[0])
                    return sorted([P[j] for j in idx[i]] + [P[k]]), tried + i + 1, tam
                tried += len(chunk)
    return None, tried, tam
