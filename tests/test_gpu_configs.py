"""GPU parity at the benchmark configurations themselves, and on the paths the
smaller parity tests leave out:

* grand_decode_ML_cln on the 216 x 3000 synthetic DEM at max_error_weight 4,
  max_size_active_positions 64 (classes up to C(64, 4) = 635,376 patterns),
  with the built-in ML tuning and with every class shared / none shared;
* ML with mass cost ties (LLRs in {1, 2, 3} over a DEM with duplicate columns)
  and with negative and signed-zero LLRs: the lowest-rank tie-break of the
  reference's strict `<` (grand_decoder.cu:640-650), the prefix bound with
  negative prefixes (:615-620) and -0.0 == +0.0;
* uniform grand_decode_cc on BB144 where the second iteration's list has 16 or
  more columns and the third iteration is not shared (the search kernel builds
  the third list before it decides, and must search the second list again);
* the five reference-named entry points (no _ex) through the reference's
  Python wrappers (Examples/auxilary_functions.py:255-395), every parity field
  against the oracle.

Each comparison is field by field on GrandDecoderResult (timings aside), and
pattern by pattern / logical by logical where the entry point returns them.
"""
import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
import grandtest as T

pytestmark = pytest.mark.gpu


def _llr_path(tmp_path, llr, name):
    path = str(tmp_path / name)
    codes.write_llr_file(path, np.asarray(llr, dtype=np.float32))
    return path


@pytest.fixture(scope="module")
def dem_full():
    """216 x 3000 DEM (seed 11), 1000 shots: the ML benchmark's code and shots."""
    return T.dem_workload(1000, seed=31)


# ----------------------------------------------------------- ML at full config
@pytest.mark.parametrize("split", [0, 1, -1])
def test_ml_cln_dem216x3000_maxsize64(oracle, product, dem_full, tmp_path, split):
    H, L, pr, det, lg = dem_full
    lf = _llr_path(tmp_path, codes.llr_from_p(pr), "llr.txt")
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
    o = T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)
    p = T.run_cln(product, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf, ex=G.make_ex(split_threshold=split))
    T.assert_same(o, p, f"ML 216x3000 W=4 max_size=64 split={split}")
    res = o[1]
    assert res["tam"].max() == 64 and res["total_tested_combinations"].max() > 635_376


def _twin_ties(H, llr, res, dec):
    """Found shots whose decoded pattern has a column with an identical twin
    column of equal LLR outside the pattern: each such shot has at least two
    matching patterns of equal cost, so the tie-break decides its answer."""
    cols = {}
    for c in range(H.shape[1]):
        cols.setdefault((H[:, c].tobytes(), float(llr[c])), []).append(c)
    n = 0
    for i in np.nonzero(res["found"])[0]:
        pat = set(np.nonzero(dec[i])[0].tolist())
        if any(len(set(cols[(H[:, c].tobytes(), float(llr[c]))]) - pat) > 0 for c in pat):
            n += 1
    return n


@pytest.mark.parametrize("split", [1, -1])
@pytest.mark.parametrize("llr_kind", ["ties", "signed"])
def test_ml_cost_ties_and_negative_llrs(oracle, product, dem_full, tmp_path, split, llr_kind):
    H, L, pr, det, lg = dem_full
    rng = np.random.default_rng(17)
    if llr_kind == "ties":
        llr = rng.integers(1, 4, size=H.shape[1]).astype(np.float32)
    else:
        # uniform in [-2, 2], every tenth rounded to an integer (exact negative
        # ties), and the members of each group of identical columns given -0.0
        # and +0.0 in turn (equal costs that differ in sign)
        llr = rng.uniform(-2.0, 2.0, size=H.shape[1]).astype(np.float32)
        llr[::10] = np.round(llr[::10])
        groups = {}
        for c in range(H.shape[1]):
            groups.setdefault(H[:, c].tobytes(), []).append(c)
        for g in groups.values():
            if len(g) > 1:
                for j, c in enumerate(g):
                    llr[c] = -0.0 if j % 2 == 0 else 0.0
        assert (np.signbit(llr) & (llr == 0)).any() and (~np.signbit(llr) & (llr == 0)).any()
    lf = _llr_path(tmp_path, llr, f"llr_{llr_kind}.txt")
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
    o = T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)
    p = T.run_cln(product, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf, ex=G.make_ex(split_threshold=split))
    T.assert_same(o, p, f"ML {llr_kind} split={split}")
    assert _twin_ties(H, llr, o[1], o[2]) >= 20
    if llr_kind == "signed":
        neg = [(llr[np.nonzero(o[2][i])[0]] < 0).any() for i in np.nonzero(o[1]["found"])[0]]
        assert sum(neg) >= 20


# ------------------------------------------- second list after the third built
@pytest.mark.parametrize("W,p", [(3, 0.02), (4, 0.01)])
@pytest.mark.parametrize("use_errors", [True, False])
def test_cc_bb144_second_iteration_lists(oracle, product, W, p, use_errors):
    hx, hz, e = T.cc_workload(144, 3000, p, seed=21)
    syn = codes.syndromes(hz, e)
    cfg = G.make_config(max_error_weight=W)
    args = dict(errors=e if use_errors else None, syndromes=None if use_errors else syn, cfg=cfg, etype="x")
    o = T.run_cc(oracle, G.KIND_CC, hx, hz, **args)
    pr = T.run_cc(product, G.KIND_CC, hx, hz, **args)
    T.assert_same(o, pr, f"cc BB144 W={W} p={p}")
    # shots whose second iteration searched a list of 16 or more columns
    tam2 = ((syn @ hz) >= 2).sum(1)
    assert ((o[1]["iteration"] >= 2) & (tam2 >= 16)).sum() >= 10


def test_cc_bb144_hybrid_window(oracle, product):
    hx, hz, e = T.cc_workload(144, 3000, 0.015, seed=23)
    cfg = G.make_config(max_error_weight=4, mode_of_search=1, threshold_for_hybrid=12,
                        max_size_active_positions=40)
    o = T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=cfg)
    p = T.run_cc(product, G.KIND_CC, hx, hz, errors=e, cfg=cfg)
    T.assert_same(o, p, "cc BB144 mode 1")
    h = o[1]["hibrido"]
    assert (h == -1).any() and (h == 0).any() and (h == 1).any()


# ------------------------------------------------- reference entry points by name
def _abi_same(o, p, what):
    assert o[0] == p[0] == 0, (what, o[0], p[0])
    a, b = G.results_to_numpy(o[1]), G.results_to_numpy(p[1])
    for f in G.PARITY_FIELDS:
        bad = np.nonzero(a[f] != b[f])[0]
        assert bad.size == 0, f"{what}: {f} differs at {bad.size} shots, first {bad[:1]}"
    return a


@pytest.mark.parametrize("use_errors", [True, False])
def test_reference_entry_cc_bb72(oracle, product, use_errors):
    """grand_decode_cc by name: BB72, p = 0.005, W = 3, 'x'."""
    hx, hz, e = T.cc_workload(72, 10_000, 0.005, seed=41)
    syn = codes.syndromes(hz, e)
    cfg = G.make_config(max_error_weight=3)
    args = (hx, hz, e if use_errors else None, None if use_errors else syn, cfg, "x")
    a = _abi_same(G.run_grand_shots_cc(oracle, *args), G.run_grand_shots_cc(product, *args), "grand_decode_cc")
    assert a["found"].sum() > 1000


def test_reference_entry_prob_cc_bb144(oracle, product, tmp_path):
    """grand_decode_prob_cc by name: BB144, p_i log-uniform in [1e-3, 1e-2], W = 4."""
    hx, hz, e = T.cc_workload(144, 10_000, "loguniform", seed=42)
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4)
    args = (hx, hz, e, None, cfg, pf, "x")
    a = _abi_same(G.run_grand_prob_shots_cc(oracle, *args), G.run_grand_prob_shots_cc(product, *args),
                  "grand_decode_prob_cc")
    assert a["found"].sum() > 1000


@pytest.mark.parametrize("entry", ["cln", "prob_cln", "ML_cln"])
def test_reference_entry_cln_family_dem(oracle, product, tmp_path, entry):
    """grand_decode_cln / _prob_cln / _ML_cln by name on the 216 x 3000 DEM."""
    H, L, pr, det, lg = T.dem_workload(300 if entry == "prob_cln" else 1000, seed=43)
    nl = L.shape[1]
    if entry == "cln":
        cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
        run = lambda lib: G.run_grand_shots_cln(lib, H, L, det, lg, nl, cfg)
    elif entry == "prob_cln":
        cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
        pf = T.phase_file(tmp_path)
        run = lambda lib: G.run_grand_prob_shots_cln(lib, H, L, det, lg, nl, cfg, pf)
    else:
        cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
        lf = T.llr_file(tmp_path, pr)
        run = lambda lib: G.run_grand_ML_shots_cln(lib, H, L, det, lg, nl, cfg, lf)
    a = _abi_same(run(oracle), run(product), f"grand_decode_{entry}")
    assert a["found"].sum() > 100
