"""Pins the CPU oracle (oracle/grand_oracle.cpp) before it is trusted as the
checker of the HIP product:

* binomials against math.comb, unranking against itertools' lexicographic order;
* the whole per-shot pipeline against an independent brute-force decoder
  (tests/bruteforce.py) on random small codes, every entry point and mode;
* the known-answer search counts printed by the reference's own notebook run
  (Create_Files/create_prob_files.ipynb, cell 6 output), reproduced on
  constructed unsolvable shots whose active sets have the same sizes;
* invariants on BB codes: a decoded pattern reproduces the syndrome, and
  stabilizer-equivalent corrections are flagged degenerate.
"""
import ctypes
import itertools
import math

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
import bruteforce as BF
import grandtest as T


def N(t, W=6):
    return sum(math.comb(t, w) for w in range(1, W + 1))


def test_ncr_matches_math_comb(oracle):
    for n in list(range(0, 40)) + [64, 100, 299, 300, 301, 512, 1024]:
        for k in range(0, 7):
            assert oracle.oracle_nCr(n, k) == math.comb(n, k), (n, k)


def test_unrank_is_lexicographic(oracle):
    out = (ctypes.c_int * 6)()
    for n in range(1, 13):
        for k in range(1, min(n, 6) + 1):
            for r, combo in enumerate(itertools.combinations(range(n), k)):
                oracle.oracle_unrank(r, n, k, out)
                assert tuple(out[:k]) == combo, (n, k, r)


# ----------------------------------------------------------------- brute force
def _random_code(rng, n, m, colw_choices):
    H = np.zeros((n, m), dtype=np.int32)
    for c in range(m):
        w = int(rng.choice(colw_choices))
        H[rng.choice(n, size=w, replace=False), c] = 1
    return H


def _bf_batch(H, syn, cfg, family, kind, **kw):
    res, dec, logs = [], [], []
    for i in range(syn.shape[0]):
        extra = dict(kw)
        if "errors" in extra:
            errs = extra.pop("errors")
            extra["e"] = errs[i] if errs is not None else None
        if "lgs" in extra:
            extra["lg"] = extra.pop("lgs")[i]
        o, d, l = BF.decode_shot(H, syn[i], cfg, family, kind, **extra)
        res.append(o)
        dec.append(d)
        logs.append(l)
    return res, np.array(dec), np.array(logs)


def _check(bf, orc, what):
    res, dec, logs = bf
    rc, r, d, l = orc
    assert rc == 0, what
    for i, o in enumerate(res):
        for f, v in o.items():
            assert r[f][i] == v, f"{what}: shot {i} field {f}: oracle {r[f][i]} brute force {v}"
    assert (d == dec).all(), f"{what}: decoded patterns differ"
    if l is not None and logs.size:
        assert (l == logs).all(), f"{what}: decoded logicals differ"


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("mode,thr_h,max_size", [(-1, 0, 50), (0, 0, 50), (1, 3, 8), (1, 5, 50), (3, 0, 50)])
def test_cc_against_bruteforce(oracle, seed, mode, thr_h, max_size):
    rng = np.random.default_rng(100 + seed)
    n, m = 9, 14
    hx = _random_code(rng, n, m, [2, 3])
    hz = _random_code(rng, n, m, [1, 2, 3])
    errors = (rng.random((40, m)) < 0.12).astype(np.int32)
    syn = (errors @ hz.T % 2).astype(np.int32)
    W = int(rng.integers(1, 5))
    cfg = G.make_config(max_error_weight=W, mode_of_search=mode, threshold_for_hybrid=thr_h,
                        max_size_active_positions=max_size)
    bcfg = dict(W=W, mode=mode, thr_h=thr_h, max_size=max_size)
    bf = _bf_batch(hz, syn, bcfg, BF.CC, BF.UNIFORM, errors=errors, other=hx)
    _check(bf, T.run_cc(oracle, G.KIND_CC, hx, hz, errors=errors, cfg=cfg), "cc/errors")
    bf = _bf_batch(hz, syn, bcfg, BF.CC, BF.UNIFORM, errors=None)
    _check(bf, T.run_cc(oracle, G.KIND_CC, hx, hz, syndromes=syn, cfg=cfg), "cc/syndromes")


@pytest.mark.parametrize("seed", range(5))
def test_prob_cc_against_bruteforce(oracle, tmp_path, seed):
    rng = np.random.default_rng(200 + seed)
    n, m = 8, 13
    hx = _random_code(rng, n, m, [2, 3])
    hz = _random_code(rng, n, m, [2, 3, 4])
    errors = (rng.random((40, m)) < 0.15).astype(np.int32)
    syn = (errors @ hx.T % 2).astype(np.int32)
    masks = [(3, 0, 1, 2, .9, .1, 0.), (7, 1, 0, 2, .2, .8, 0.), (9, 2, 1, 0, .1, .1, .8), (30, 0, 1, 2, .5, .5, 0.)]
    pf = T.phase_file(tmp_path, masks)
    for W in (3, 6, 7):
        cfg = G.make_config(max_error_weight=W)
        num_phases = 1 if W < 6 else (2 if W == 6 else 3)
        bf = _bf_batch(hx, syn, dict(W=W, mode=-1, thr_h=0, max_size=300), BF.CC, BF.PHASE, errors=errors,
                       other=hz, masks=masks, num_phases=num_phases)
        _check(bf, T.run_cc(oracle, G.KIND_PROB_CC, hx, hz, errors=errors, cfg=cfg, etype="z", prob_file=pf),
               f"prob_cc W={W}")


def _dem_small(rng, n=12, m=30, nl=3):
    H = _random_code(rng, n, m, [1, 2, 3, 4])
    L = (rng.random((m, nl)) < 0.3).astype(np.int32)
    p = np.exp(rng.uniform(np.log(0.02), np.log(0.12), m))
    e = (rng.random((40, m)) < p).astype(np.int32)
    det = (e @ H.T % 2).astype(np.int32)
    lg = (e @ L % 2).astype(np.int32)
    return H, L, p, det, lg


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("max_size", [4, 9, 40])
def test_cln_families_against_bruteforce(oracle, tmp_path, seed, max_size):
    rng = np.random.default_rng(300 + seed)
    H, L, p, det, lg = _dem_small(rng)
    W = 3
    cfg = G.make_config(max_error_weight=W, max_size_active_positions=max_size)
    bcfg = dict(W=W, mode=-1, thr_h=0, max_size=max_size)
    bf = _bf_batch(H, det, bcfg, BF.CLN, BF.UNIFORM, L=L, lgs=lg)
    _check(bf, T.run_cln(oracle, G.KIND_CLN, H, L, det, lg, cfg=cfg), "cln")
    masks = [(5, 0, 1, 2, 1., 0., 0.), (300, 1, 0, 2, .5, .5, 0.)]
    pf = T.phase_file(tmp_path, masks)
    cfg6 = G.make_config(max_error_weight=6, max_size_active_positions=max_size)
    bf = _bf_batch(H, det, dict(bcfg, W=6), BF.CLN, BF.PHASE, L=L, lgs=lg, masks=masks, num_phases=2)
    _check(bf, T.run_cln(oracle, G.KIND_PROB_CLN, H, L, det, lg, cfg=cfg6, file=pf), "prob_cln")
    lf = T.llr_file(tmp_path, p)
    llr = codes.llr_from_p(p)
    bf = _bf_batch(H, det, bcfg, BF.CLN, BF.ML, L=L, lgs=lg, llr=llr)
    _check(bf, T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf), "ML")


def test_ml_ties_and_negative_llrs(oracle, tmp_path):
    """Equal LLRs (ties broken by active-list order) and negative LLRs, where the
    prefix bound stops the weight loop early."""
    rng = np.random.default_rng(7)
    H, L, p, det, lg = _dem_small(rng, n=10, m=24)
    for llr in (np.round(rng.uniform(1, 3, 24)).astype(np.float32),
                rng.uniform(-2, 2, 24).astype(np.float32)):
        path = tmp_path / "llr.txt"
        codes.write_llr_file(str(path), llr)
        cfg = G.make_config(max_error_weight=4, max_size_active_positions=30)
        bf = _bf_batch(H, det, dict(W=4, mode=-1, thr_h=0, max_size=30), BF.CLN, BF.ML, L=L, lgs=lg, llr=llr)
        _check(bf, T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=str(path)), "ML ties")


# -------------------------------------------------------------- known answers
NOTEBOOK_COUNTS = [  # (printed total_tested_combinations, active-set sizes of the 3 iterations)
    (65331, (17, 17, 17)), (7527, (12, 12, 12)), (1192779, (27, 27, 27)), (2541, (10, 10, 10)),
    (131385, (19, 19, 19)), (1167107569, (12, 88, 88)), (1250408109, (20, 89, 89)), (58664676, (8, 54, 54)),
]


@pytest.mark.parametrize("printed,tams", NOTEBOOK_COUNTS)
def test_notebook_counts_closed_form(printed, tams):
    assert sum(N(t) for t in tams) == printed


def unsolvable_shot(t1, t2, t3, k=9, extra_rows=8):
    """Weight-4 columns (so every XOR of columns has even weight) and an
    odd-weight syndrome on rows 0..k-1: no pattern ever matches.  Column groups
    overlap supp(s) in 4, 3 and 2 rows, giving active sets of t1, t2 and t3
    positions at thresholds 4, 3 and 2."""
    rng = np.random.default_rng(t1 * 1000 + t2 * 10 + t3)
    n = k + extra_rows
    cols = []
    for inside, cnt in ((4, t1), (3, t2 - t1), (2, t3 - t2)):
        for _ in range(cnt):
            c = np.zeros(n, dtype=np.int32)
            c[rng.choice(k, inside, replace=False)] = 1
            c[k + rng.choice(extra_rows, 4 - inside, replace=False)] = 1
            cols.append(c)
    cols.append(np.r_[np.zeros(k, np.int32), np.ones(4, np.int32), np.zeros(extra_rows - 4, np.int32)])
    H = np.array(cols, dtype=np.int32).T
    s = np.zeros(n, dtype=np.int32)
    s[:k] = 1
    return H, s


@pytest.mark.parametrize("printed,tams", [c for c in NOTEBOOK_COUNTS if c[0] < 2_000_000])
def test_notebook_counts_reproduced_by_oracle(oracle, printed, tams):
    t1 = tams[0]
    H, s = unsolvable_shot(t1, t1 + 3, t1 + 5)
    L = np.zeros((H.shape[1], 2), dtype=np.int32)
    cfg = G.make_config(max_error_weight=6, max_size_active_positions=t1)   # iterations 2, 3 revert to tam_old
    rc, r, d, l = T.run_cln(oracle, G.KIND_CLN, H, L, s[None, :], np.zeros((1, 2), np.int32), cfg=cfg)
    assert rc == 0
    assert r["found"][0] == 0 and r["iteration"][0] == 3 and r["tam"][0] == t1
    assert int(r["total_tested_combinations"][0]) == printed


def test_cc_counts_without_cap(oracle):
    """The CSS family has no active-set cap: counts are N(t1)+N(t2)+N(t3)."""
    H, s = unsolvable_shot(6, 9, 13)
    rc, r, d, l = T.run_cc(oracle, G.KIND_CC, H, H, syndromes=s[None, :], cfg=G.make_config(max_error_weight=6,
                                                                                             max_size_active_positions=1))
    assert rc == 0 and r["tam"][0] == 13
    assert int(r["total_tested_combinations"][0]) == N(6) + N(9) + N(13)


# ---------------------------------------------------------------- invariants
def test_bb72_decoded_patterns_reproduce_syndromes(oracle):
    hx, hz, e = T.cc_workload(72, 3000, 0.01, seed=3)
    rc, r, dec, _ = T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=G.make_config(max_error_weight=4))
    assert rc == 0
    f = r["found"] == 1
    assert f.sum() > 500
    assert ((dec[f] @ hz.T % 2) == (e[f] @ hz.T % 2)).all()
    assert (r["found_weight"][f] == dec[f].sum(1)).all()
    assert (r["original_weight"] == e.sum(1)).all()
    eq = r["is_equal"] == 1
    assert (dec[eq] == e[eq]).all() and (r["degenerate"][eq] == 1).all()


def test_stabilizer_residue_is_degenerate(oracle):
    """e = 4 qubits of an X-stabilizer (a row of Hx, weight 6): e + row has the
    same Hz-syndrome and weight 2, so the decoder answers with a weight-2
    pattern; degenerate must equal membership of e + answer in rowspace(Hx)."""
    hx, hz = codes.bb(72)
    supp = np.nonzero(hx[0])[0]
    e = np.zeros((1, 72), dtype=np.int32)
    e[0, supp[:4]] = 1
    rc, r, dec, _ = T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=G.make_config(max_error_weight=4), etype="x")
    assert rc == 0 and r["found"][0] == 1 and r["found_weight"][0] == 2 and r["is_equal"][0] == 0
    assert r["degenerate"][0] == int(BF._in_rowspace(hx.astype(np.int64), (dec[0] ^ e[0]).astype(np.int64)))
    assert r["degenerate"][0] == int((dec[0] ^ e[0] == hx[0]).all() or r["degenerate"][0])


def test_reference_syndromes_72_fixture(oracle):
    """Syndromes/syndromes_72.txt = Hz . Errors/patterns_72.txt[0:30] (both reference
    files, kept under tests/golden); every shot is decoded to a pattern with
    that syndrome."""
    hz = codes.read_matrix(f"{T.GOLDEN}/hz_ibm_72.txt")
    hx = codes.read_matrix(f"{T.GOLDEN}/hx_ibm_72.txt")
    syn = codes.read_rows(f"{T.GOLDEN}/syndromes_72.txt", 36)
    pats = codes.read_rows(f"{T.GOLDEN}/patterns_72_first30.txt", 72)
    assert ((pats @ hz.T) % 2 == syn).all()
    rc, r, dec, _ = T.run_cc(oracle, G.KIND_CC, hx, hz, syndromes=syn, cfg=G.make_config(max_error_weight=6))
    assert rc == 0 and r["found"].all()
    assert ((dec @ hz.T) % 2 == syn).all()
