"""Independent brute-force statement of the reference decoder's per-shot
semantics, written from the definitions with itertools and numpy (no unranking,
no binomial tables), used to pin the oracle on small instances.

Reference semantics followed (grand_decoder.cu): syndrome and zero-syndrome
shortcut :1206-1241; column scores and the irregular offset :352-416; active-set
iterations and the DEM-family tam_old rule :1284-1290, :1546-1558; branch choice
and `hibrido` :1296-1343; first-hit enumeration :295-350, :534-588; ML selection
:590-666; phase masks :754-786; post-processing :1067-1092, :1618-1644.
"""
from __future__ import annotations

import itertools

import numpy as np

CC, CLN = 0, 1
UNIFORM, PHASE, ML = 0, 1, 2


def _branch(cfg, tam):
    mode, thr_h, mx = cfg["mode"], cfg["thr_h"], cfg["max_size"]
    if (mode == 1 and thr_h < tam <= mx) or mode == 0:
        return 1
    if (mode == 1 and tam <= thr_h) or mode == -1:
        return -1
    return 0


def _phase_weights(phase, tam):
    if phase == 0:
        return [w for w in range(1, 6) if w <= tam]
    if phase == 1:
        return [6] if tam >= 6 else []
    return []


def _in_rowspace(other, v):
    """v in rowspace(other) over GF(2), by rank comparison."""
    def rank(M):
        M = M.copy() % 2
        r = 0
        for c in range(M.shape[1]):
            piv = np.nonzero(M[r:, c])[0]
            if piv.size == 0:
                continue
            p = r + piv[0]
            M[[r, p]] = M[[p, r]]
            for i in range(M.shape[0]):
                if i != r and M[i, c]:
                    M[i] ^= M[r]
            r += 1
            if r == M.shape[0]:
                break
        return r
    return rank(np.vstack([other, v[None, :]])) == rank(other)


def decode_shot(H, s, cfg, family, kind, *, masks=None, num_phases=1, llr=None, e=None, other=None, L=None,
                lg=None):
    """Returns dict of the parity fields plus 'decoded' (m,) and 'logicals'."""
    n, m = H.shape
    H = H.astype(np.int64)
    s = np.asarray(s, dtype=np.int64)
    out = dict(found=0, found_weight=0, original_weight=0, degenerate=0, is_equal=0, tam=0,
               total_tested_combinations=0, hibrido=-2, iteration=0)
    decoded = np.zeros(m, dtype=np.int64)
    if family == CC and e is not None:
        out["original_weight"] = int((e >= 1).sum())
    nl = L.shape[1] if L is not None else 0
    if (s >= 1).sum() == 0:
        return out, decoded, np.zeros(nl, dtype=np.int64)
    colw = H.sum(0)
    max_soma = int(colw.max())
    regular = bool((colw == colw[0]).all())
    score = H[s == 1].sum(0)
    if not regular:
        score = np.where(colw == 0, 0, score + max_soma - colw)
    W = min(cfg["W"], 6)
    it, tam, tam_old, hib, count = 0, 0, 0, -2, 0
    hit = None
    while it < max_soma and it < 3:
        thr = max_soma - it
        it += 1
        A = [c for c in range(m) if score[c] >= thr]
        tam = len(A)
        if family == CLN:
            if tam > cfg["max_size"]:
                tam = tam_old
            else:
                tam_old = tam
        if tam <= 0:
            continue
        br = _branch(cfg, tam)
        if br == 0:
            continue
        hib = (1 if hib == -1 else 0) if br > 0 else -1
        pos = A[:tam]
        if kind == ML:
            P = sorted(pos, key=lambda c: np.float32(llr[c]))       # stable
            lv = [np.float32(llr[c]) for c in P]
            max_w = min(W, tam)
            prefix = [np.float32(0)]
            for w in range(1, max_w + 1):
                prefix.append(np.float32(prefix[-1] + lv[w - 1]))
            best = None
            for w in range(1, max_w + 1):
                if best is not None and prefix[w] >= best[0]:
                    break
                for combo in itertools.combinations(range(tam), w):
                    count += 1
                    cols = [P[j] for j in combo]
                    if ((H[:, cols].sum(1) % 2) == s).all():
                        cost = np.float32(0)
                        for j in combo:
                            cost = np.float32(cost + lv[j])
                        if best is None or cost < best[0]:
                            best = (cost, cols)
            if best is not None:
                hit = best[1]
        else:
            if kind == UNIFORM:
                wl = list(range(1, W + 1))
            else:
                mk = next((mm for mm in masks if tam <= mm[0]), masks[-1])
                wl = []
                for k in range(num_phases):
                    wl += _phase_weights(mk[1 + k], tam)
            for w in wl:
                if w > tam:
                    continue
                for combo in itertools.combinations(range(tam), w):
                    count += 1
                    cols = [pos[j] for j in combo]
                    if ((H[:, cols].sum(1) % 2) == s).all():
                        hit = cols
                        break
                if hit is not None:
                    break
        if hit is not None:
            break
    out.update(iteration=it, tam=tam, total_tested_combinations=count, hibrido=hib)
    logicals = np.zeros(nl, dtype=np.int64)
    if hit is None:
        return out, decoded, logicals
    decoded[hit] = 1
    out["found"] = 1
    out["found_weight"] = int(decoded.sum())
    if family == CC and e is not None:
        out["is_equal"] = int((decoded == e).all())
        out["degenerate"] = int(_in_rowspace(other.astype(np.int64), (decoded ^ e) % 2))
    elif family == CLN:
        logicals = L[decoded == 1].sum(0) % 2
        out["degenerate"] = int((logicals == lg).all())
    return out, decoded, logicals
