"""GPU parity: the HIP product against the CPU oracle on seeded workloads of
every entry point, bit-exact on every result field except the timings, on the
decoded patterns and on the decoded logicals."""
import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
import grandtest as T

pytestmark = pytest.mark.gpu


def _both(oracle, product, runner, *args, ex=None, **kw):
    o = runner(oracle, *args, **kw)
    p = runner(product, *args, ex=ex, **kw)
    return o, p


@pytest.mark.parametrize("use_errors", [True, False])
def test_cc_bb72_w3(oracle, product, use_errors):
    hx, hz, e = T.cc_workload(72, 4000, 0.005, seed=1)
    cfg = G.make_config(max_error_weight=3)
    syn = T.codes.syndromes(hz, e)
    args = dict(errors=e if use_errors else None, syndromes=None if use_errors else syn, cfg=cfg, etype="x")
    o, p = _both(oracle, product, T.run_cc, G.KIND_CC, hx, hz, **args)
    T.assert_same(o, p, "cc BB72")
    assert o[1]["found"].sum() > 100


def test_cc_z_errors_bb144(oracle, product):
    hx, hz, e = T.cc_workload(144, 2000, 0.004, seed=2)
    cfg = G.make_config(max_error_weight=4)
    o, p = _both(oracle, product, T.run_cc, G.KIND_CC, hx, hz, errors=e, cfg=cfg, etype="z")
    T.assert_same(o, p, "cc BB144 z")


def test_prob_cc_bb144_w4(oracle, product, tmp_path):
    hx, hz, e = T.cc_workload(144, 6000, "loguniform", seed=3)
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4)
    o, p = _both(oracle, product, T.run_cc, G.KIND_PROB_CC, hx, hz, errors=e, cfg=cfg, prob_file=pf)
    T.assert_same(o, p, "prob_cc BB144")


def test_prob_cc_phase_orders(oracle, product, tmp_path):
    # masks that put weight 6 first for mid-size active sets (W = 6 -> 2 phases)
    masks = [(5, 0, 1, 2, 1, 0, 0), (15, 1, 0, 2, .3, .7, 0), (300, 0, 1, 2, .6, .4, 0)]
    pf = T.phase_file(tmp_path, masks)
    hx, hz, e = T.cc_workload(72, 1500, 0.01, seed=4)
    cfg = G.make_config(max_error_weight=6)
    o, p = _both(oracle, product, T.run_cc, G.KIND_PROB_CC, hx, hz, errors=e, cfg=cfg, prob_file=pf)
    T.assert_same(o, p, "prob_cc phases")


@pytest.mark.parametrize("mode,thr", [(0, 0), (1, 8), (1, 40), (7, 0)])
def test_cc_modes(oracle, product, mode, thr):
    hx, hz, e = T.cc_workload(72, 1500, 0.008, seed=5)
    cfg = G.make_config(max_error_weight=4, mode_of_search=mode, threshold_for_hybrid=thr,
                        max_size_active_positions=30)
    o, p = _both(oracle, product, T.run_cc, G.KIND_CC, hx, hz, errors=e, cfg=cfg)
    T.assert_same(o, p, f"cc mode {mode}")


def test_cln_dem(oracle, product):
    H, L, pr, det, lg = T.dem_workload(1500, seed=6)
    cfg = G.make_config(max_error_weight=3, max_size_active_positions=48)
    o, p = _both(oracle, product, T.run_cln, G.KIND_CLN, H, L, det, lg, cfg=cfg)
    T.assert_same(o, p, "cln")


def test_prob_cln_dem(oracle, product, tmp_path):
    H, L, pr, det, lg = T.dem_workload(1500, seed=7)
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=40)
    o, p = _both(oracle, product, T.run_cln, G.KIND_PROB_CLN, H, L, det, lg, cfg=cfg, file=pf)
    T.assert_same(o, p, "prob_cln")


def test_ml_cln_dem(oracle, product, tmp_path):
    H, L, pr, det, lg = T.dem_workload(600, seed=8)
    lf = T.llr_file(tmp_path, pr)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=28)
    o, p = _both(oracle, product, T.run_cln, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)
    T.assert_same(o, p, "ML cln")


@pytest.mark.parametrize("kind", [G.KIND_CC, G.KIND_PROB_CC])
def test_plain_engine_matches(oracle, product, kind, tmp_path):
    hx, hz, e = T.cc_workload(144, 300, "loguniform", seed=9)
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4)
    o = T.run_cc(oracle, kind, hx, hz, errors=e, cfg=cfg, prob_file=pf)
    p = T.run_cc(product, kind, hx, hz, errors=e, cfg=cfg, prob_file=pf, ex=G.make_ex(engine=G.GRAND_ENGINE_PLAIN))
    T.assert_same(o, p, "plain cc")


def test_plain_engine_ml(oracle, product, tmp_path):
    H, L, pr, det, lg = T.dem_workload(120, seed=10)
    lf = T.llr_file(tmp_path, pr)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=28)
    o = T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)
    p = T.run_cln(product, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf, ex=G.make_ex(engine=G.GRAND_ENGINE_PLAIN))
    T.assert_same(o, p, "plain ML")


# ---------------------------------------------------------------- shared classes
# split_threshold = 1 shares every weight class through the task table, 200
# shares the larger ones, -1 never shares: all must equal the oracle.
@pytest.mark.parametrize("split", [1, 200, -1])
def test_shared_classes_cc(oracle, product, split, tmp_path):
    hx, hz, e = T.cc_workload(144, 3000, "loguniform", seed=12)
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4)
    ex = G.make_ex(split_threshold=split)
    o = T.run_cc(oracle, G.KIND_PROB_CC, hx, hz, errors=e, cfg=cfg, prob_file=pf)
    p = T.run_cc(product, G.KIND_PROB_CC, hx, hz, errors=e, cfg=cfg, prob_file=pf, ex=ex)
    T.assert_same(o, p, f"shared prob_cc split={split}")


@pytest.mark.parametrize("split", [1, 500, -1])
def test_shared_classes_ml(oracle, product, split, tmp_path):
    H, L, pr, det, lg = T.dem_workload(400, seed=13)
    lf = T.llr_file(tmp_path, pr)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=32)
    ex = G.make_ex(split_threshold=split)
    o = T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)
    p = T.run_cln(product, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf, ex=ex)
    T.assert_same(o, p, f"shared ML split={split}")


@pytest.mark.parametrize("split", [1, 300])
def test_shared_classes_cln_deep(oracle, product, split):
    H, L, pr, det, lg = T.dem_workload(800, seed=14)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=40)
    ex = G.make_ex(split_threshold=split)
    o = T.run_cln(oracle, G.KIND_CLN, H, L, det, lg, cfg=cfg)
    p = T.run_cln(product, G.KIND_CLN, H, L, det, lg, cfg=cfg, ex=ex)
    T.assert_same(o, p, f"shared cln split={split}")


# ---------------------------------------------------------------- triage search
# Shots with a small first iteration are searched one per lane in the triage
# kernel (fused engine, first-hit kinds, codes of at most 6 syndrome words).
# Results must be those of the oracle and of the search kernel alone
# (triage_search = -1), including the cap, hybrid-mode and phase-order paths.
def _triage_cases(tmp_path):
    pf = T.phase_file(tmp_path)
    masks6 = [(5, 0, 1, 2, 1, 0, 0), (15, 1, 0, 2, .3, .7, 0), (300, 0, 1, 2, .6, .4, 0)]
    pf6 = T.phase_file(tmp_path, masks6, name="phase6.txt")
    hx72, hz72, e72 = T.cc_workload(72, 3000, 0.006, seed=31)
    hx144, hz144, e144 = T.cc_workload(144, 4000, "loguniform", seed=32)
    hx288, hz288, e288 = T.cc_workload(288, 1500, 0.004, seed=33)
    Hs, Ls, ps, dets, lgs = T.dem_workload(2000, seed=34, dem_seed=5, n_det=72, n_mech=400)
    cases = [
        ("cc BB72 errors", T.run_cc, (G.KIND_CC, hx72, hz72), dict(errors=e72, cfg=G.make_config(max_error_weight=3))),
        ("cc BB72 syndromes", T.run_cc, (G.KIND_CC, hx72, hz72),
         dict(syndromes=T.codes.syndromes(hz72, e72), cfg=G.make_config(max_error_weight=3))),
        ("cc BB72 z W1", T.run_cc, (G.KIND_CC, hx72, hz72), dict(errors=e72, etype="z",
                                                                 cfg=G.make_config(max_error_weight=1))),
        ("prob_cc BB144", T.run_cc, (G.KIND_PROB_CC, hx144, hz144),
         dict(errors=e144, cfg=G.make_config(max_error_weight=4), prob_file=pf)),
        ("prob_cc BB72 phase orders", T.run_cc, (G.KIND_PROB_CC, hx72, hz72),
         dict(errors=e72, cfg=G.make_config(max_error_weight=6), prob_file=pf6)),
        ("cc BB288", T.run_cc, (G.KIND_CC, hx288, hz288), dict(errors=e288, cfg=G.make_config(max_error_weight=4))),
        ("cc BB72 hybrid", T.run_cc, (G.KIND_CC, hx72, hz72),
         dict(errors=e72, cfg=G.make_config(max_error_weight=4, mode_of_search=1, threshold_for_hybrid=2,
                                            max_size_active_positions=30))),
        ("cc BB72 cpu mode", T.run_cc, (G.KIND_CC, hx72, hz72),
         dict(errors=e72, cfg=G.make_config(max_error_weight=4, mode_of_search=-1))),
        ("cc BB72 no search", T.run_cc, (G.KIND_CC, hx72, hz72),
         dict(errors=e72, cfg=G.make_config(max_error_weight=4, mode_of_search=7))),
        ("cln small DEM", T.run_cln, (G.KIND_CLN, Hs, Ls, dets, lgs),
         dict(cfg=G.make_config(max_error_weight=3, max_size_active_positions=48))),
        ("prob_cln small DEM, cap 2", T.run_cln, (G.KIND_PROB_CLN, Hs, Ls, dets, lgs),
         dict(cfg=G.make_config(max_error_weight=4, max_size_active_positions=2), file=pf)),
    ]
    return cases


def test_triage_search_matches_oracle_and_search_kernel(oracle, product, tmp_path):
    for what, runner, args, kw in _triage_cases(tmp_path):
        o = runner(oracle, *args, **kw)
        on = runner(product, *args, ex=G.make_ex(triage_search=0), **kw)
        off = runner(product, *args, ex=G.make_ex(triage_search=-1), **kw)
        T.assert_same(o, on, what + " (triage search on)")
        T.assert_same(on, off, what + " (on vs off)")
        print(what, "found", int(o[1]["found"].sum()), flush=True)


def test_triage_second_iteration(oracle, product, tmp_path):
    """The lane search of the second iteration (CSS family): BB144 prob_cc shots of the bench's error
    model, about 1.3 % of them decoded in the second iteration, bit-exact with the oracle and with the
    search kernel alone.  Then a code with a weight-1 and a weight-2 column (a weight-1 column is active
    at every lower threshold: its count of the syndrome's checks, 0, is within one of its weight) and one
    whose columns all have weight 2 after the change, syndromes as input, hybrid mode included."""
    pf = T.phase_file(tmp_path)
    hx, hz, e = T.cc_workload(144, 20000, "loguniform", seed=35)
    kw = dict(errors=e, cfg=G.make_config(max_error_weight=4), prob_file=pf)
    o = T.run_cc(oracle, G.KIND_PROB_CC, hx, hz, **kw)
    it2 = int(((o[1]["iteration"] == 2) & (o[1]["found"] == 1)).sum())
    assert it2 >= 150, it2
    on = T.run_cc(product, G.KIND_PROB_CC, hx, hz, ex=G.make_ex(triage_search=0), **kw)
    off = T.run_cc(product, G.KIND_PROB_CC, hx, hz, ex=G.make_ex(triage_search=-1), **kw)
    T.assert_same(o, on, "BB144 second iteration (triage search on)")
    T.assert_same(on, off, "BB144 second iteration (on vs off)")
    hx72, hz72, e72 = T.cc_workload(72, 3000, 0.01, seed=36)
    low = hz72.copy()
    rows = np.nonzero(low[:, 5])[0]
    low[rows[1:], 5] = 0                          # column 5: weight 1
    rows = np.nonzero(low[:, 9])[0]
    low[rows[2:], 9] = 0                          # column 9: weight 2
    two = hz72.copy()
    for c in range(two.shape[1]):                 # every column: weight 2
        rows = np.nonzero(two[:, c])[0]
        two[rows[2:], c] = 0
    for name, h in (("weight-1 column", low), ("weight-2 columns", two)):
        syn = T.codes.syndromes(h, e72)
        for cfg in (G.make_config(max_error_weight=4),
                    G.make_config(max_error_weight=4, mode_of_search=1, threshold_for_hybrid=3,
                                  max_size_active_positions=30)):
            args = (G.KIND_CC, hx72, h)
            o = T.run_cc(oracle, *args, syndromes=syn, cfg=cfg)
            on = T.run_cc(product, *args, syndromes=syn, cfg=cfg, ex=G.make_ex(triage_search=0))
            T.assert_same(o, on, f"{name} mode {cfg.mode_of_search}")
            print(name, "mode", cfg.mode_of_search, "iterations", np.bincount(o[1]["iteration"]), flush=True)


# ---------------------------------------------------------------- pair table
# Deep first-hit classes (weight >= 3, many more prefixes than pairs) are walked
# by (w-2)-stems and the LDS pair table when it fits the wave's share of the
# CU's LDS: at 1 and 2 waves per CU for these lists, not at 8.  Shared (split
# 256) and unshared classes must give the oracle's results bit for bit.
_PAIR_ORACLE = {}


def _pair_cases(oracle, tmp_path):
    if not _PAIR_ORACLE:
        pf = T.phase_file(tmp_path)
        hx, hz, e = T.cc_workload(144, 400, 0.015, seed=41)          # tam up to 87, 3.9e7 patterns
        _PAIR_ORACLE["prob_cc BB144"] = (T.run_cc, (G.KIND_PROB_CC, hx, hz),
                                         dict(errors=e, cfg=G.make_config(max_error_weight=4), prob_file=pf))
        hx72, hz72, e72 = T.cc_workload(72, 600, 0.03, seed=42)     # weight-5 classes, tam up to 59
        _PAIR_ORACLE["cc BB72 W5"] = (T.run_cc, (G.KIND_CC, hx72, hz72),
                                      dict(errors=e72, cfg=G.make_config(max_error_weight=5)))
        H, L, pr, det, lg = T.dem_workload(300, seed=43)
        _PAIR_ORACLE["cln DEM"] = (T.run_cln, (G.KIND_CLN, H, L, det, lg),
                                   dict(cfg=G.make_config(max_error_weight=4, max_size_active_positions=60)))
        for k, (runner, args, kw) in list(_PAIR_ORACLE.items()):
            _PAIR_ORACLE[k] = (runner, args, kw, runner(oracle, *args, **kw))
    return _PAIR_ORACLE


@pytest.mark.parametrize("wpc", [1, 2, 8])
def test_pair_table_deep_classes(oracle, product, tmp_path, wpc):
    for what, (runner, args, kw, o) in _pair_cases(oracle, tmp_path).items():
        for split in (256, -1):
            p = runner(product, *args, ex=G.make_ex(waves_per_cu=wpc, split_threshold=split), **kw)
            T.assert_same(o, p, f"{what} wpc={wpc} split={split}")
        print(what, "wpc", wpc, "max count", int(o[1]["total_tested_combinations"].max()), flush=True)


@pytest.mark.parametrize("kb", [4, 16, -1])
def test_pair_table_lds_knob(oracle, product, tmp_path, kb):
    """The LDS-packing knob: a pair table capped at 4 KiB (512 pairs: the deep classes of long lists
    fall back to the prefix walk), at 16 KiB, and none at all decode exactly as the oracle."""
    for what, (runner, args, kw, o) in _pair_cases(oracle, tmp_path).items():
        for split in (256, -1):
            p = runner(product, *args, ex=G.make_ex(split_threshold=split, pair_table_kb=kb), **kw)
            T.assert_same(o, p, f"{what} pair_table_kb={kb} split={split}")
        print(what, "pair_table_kb", kb, flush=True)


def test_deep_shots_in_small_batches(oracle, product, tmp_path):
    """Shots of the BB144 prob_cc bench workload whose second active list has 16 or more columns (the
    third iteration's shared run is then published before the second is searched: joined when the
    third iteration is searched, withdrawn when an earlier one finds the pattern), plus the deep shot
    37877 (1.2e6 patterns): as one small batch, the deepest (2.7e7 patterns) alone, and 37877 alone.
    A batch smaller than the CU count still launches helper waves.  Bit-exact with the oracle."""
    hx, hz = codes.bb(144)
    seed = 20240501
    p = codes.log_uniform(np.random.default_rng(seed), 1e-3, 1e-2, hx.shape[1])
    errors = codes.sample_errors(np.random.default_rng(seed + 1), 100_000, p)
    syn = (errors.astype(np.int64) @ hz.T) % 2
    tam2 = ((syn @ hz) >= 2).sum(axis=1)          # second list: columns with two checks in supp(s)
    sel = sorted(set(np.nonzero(tam2 >= 16)[0].tolist()) | {37877})
    assert 79335 in sel and len(sel) >= 8
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(max_error_weight=4)
    want = T.run_cc(oracle, G.KIND_PROB_CC, hx, hz, errors=np.ascontiguousarray(errors[sel]), cfg=cfg, prob_file=pf)
    it = want[1]["iteration"]
    assert (it == 3).sum() >= 3 and (it == 2).sum() >= 3, it
    for sub in (list(range(len(sel))), [sel.index(79335)], [sel.index(37877)]):
        got = T.run_cc(product, G.KIND_PROB_CC, hx, hz, errors=np.ascontiguousarray(errors[[sel[i] for i in sub]]),
                       cfg=cfg, prob_file=pf)
        ref = (want[0], want[1][sub], None if want[2] is None else want[2][sub], None)
        T.assert_same(ref, got, f"shots {[sel[i] for i in sub]}")


# ------------------------------------------------------- lane search limits
def _random_code(rng, n, m, wmax):
    """n x m binary matrix whose column c has weight uniform in 1..wmax (distinct rows)."""
    H = np.zeros((n, m), dtype=np.int32)
    for c in range(m):
        H[rng.choice(n, size=int(rng.integers(1, wmax + 1)), replace=False), c] = 1
    return H


@pytest.mark.parametrize("m,wmax", [(320, 7), (320, 8), (321, 7)])
def test_lane_search_limits(oracle, product, tmp_path, m, wmax):
    """The lane search's bit-sliced column counts at their limits: 320 columns (10 words per count
    plane) with column weights up to 7 (three planes) searched by lane; a weight-8 column or a 321st
    column turns the lane search off for the code.  Both families, uniform and phase-ordered, errors and
    syndromes as input: bit-exact with the oracle and with the search kernel alone."""
    rng = np.random.default_rng(1000 + m + wmax)
    pf = T.phase_file(tmp_path)
    hz = _random_code(rng, 96, m, wmax)
    hx = _random_code(rng, 96, m, 3)
    e = codes.sample_errors(rng, 3000, 0.003, m)
    Hd = _random_code(rng, 64, m, wmax)
    Ld = (rng.random((m, 8)) < 0.1).astype(np.int32)
    ed = codes.sample_errors(rng, 3000, 0.003, m)
    det = codes.syndromes(Hd, ed)
    lg = ((ed.astype(np.int64) @ Ld) % 2).astype(np.int32)
    cases = [
        ("cc errors", T.run_cc, (G.KIND_CC, hx, hz), dict(errors=e, cfg=G.make_config(max_error_weight=3))),
        ("cc syndromes", T.run_cc, (G.KIND_CC, hx, hz),
         dict(syndromes=codes.syndromes(hz, e), cfg=G.make_config(max_error_weight=3))),
        # (shots with at most two errors: weights 1..5 in three iterations over lists of 100+ columns
        # would cost the oracle minutes on a few deep shots)
        ("prob_cc errors", T.run_cc, (G.KIND_PROB_CC, hx, hz),
         dict(errors=e[e.sum(axis=1) <= 2][:600], cfg=G.make_config(max_error_weight=4), prob_file=pf)),
        ("cln", T.run_cln, (G.KIND_CLN, Hd, Ld, det, lg),
         dict(cfg=G.make_config(max_error_weight=3, max_size_active_positions=40))),
        ("prob_cln", T.run_cln, (G.KIND_PROB_CLN, Hd, Ld, det, lg),
         dict(cfg=G.make_config(max_error_weight=4, max_size_active_positions=40), file=pf)),
    ]
    for what, runner, args, kw in cases:
        o = runner(oracle, *args, **kw)
        on = runner(product, *args, ex=G.make_ex(triage_search=0), **kw)
        off = runner(product, *args, ex=G.make_ex(triage_search=-1), **kw)
        name = f"m={m} colw<={wmax} {what}"
        T.assert_same(o, on, name + " (lane search on)")
        T.assert_same(on, off, name + " (on vs off)")
        print(name, "found", int(o[1]["found"].sum()), "iterations", np.bincount(o[1]["iteration"]), flush=True)
