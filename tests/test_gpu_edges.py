"""GPU edge cases: committed golden fixtures from the reference data set, empty
and all-zero batches, the max-weight clamp, large check matrices, error paths
that the kernels themselves detect (non-binary inputs, oversized active sets)."""
import os

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
import grandtest as T

pytestmark = pytest.mark.gpu


def test_golden_bb72_fixtures(product):
    """Reference Syndromes/syndromes_72.txt and Errors/patterns_72.txt[0:30]
    against the committed oracle outputs (tests/golden/oracle_bb72_fixtures.npz)."""
    g = np.load(os.path.join(T.GOLDEN, "oracle_bb72_fixtures.npz"))
    hx = codes.read_matrix(os.path.join(T.GOLDEN, "hx_ibm_72.txt"))
    hz = codes.read_matrix(os.path.join(T.GOLDEN, "hz_ibm_72.txt"))
    syn = codes.read_rows(os.path.join(T.GOLDEN, "syndromes_72.txt"), 36)
    pats = codes.read_rows(os.path.join(T.GOLDEN, "patterns_72_first30.txt"), 72)
    pf = os.path.join(T.GOLDEN, "phase_masks_notebook.txt")
    for W in (3, 6):
        for kind, name in ((G.KIND_CC, "cc"), (G.KIND_PROB_CC, "prob_cc")):
            for src in ("syndromes", "errors"):
                kw = dict(syndromes=syn) if src == "syndromes" else dict(errors=pats)
                rc, r, dec, _ = T.run_cc(product, kind, hx, hz, cfg=G.make_config(max_error_weight=W), prob_file=pf, **kw)
                assert rc == 0
                want = g[f"{name}_W{W}_{src}_results"]
                for f in G.PARITY_FIELDS:
                    assert (r[f] == want[f]).all(), (name, W, src, f)
                assert (dec == g[f"{name}_W{W}_{src}_decoded"]).all()


def test_all_zero_and_single_shot_batches(oracle, product):
    hx, hz = codes.bb(144)
    cfg = G.make_config(max_error_weight=4)
    for e in (np.zeros((257, 144), np.int32), np.zeros((1, 144), np.int32)):
        o, p = T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=cfg), T.run_cc(product, G.KIND_CC, hx, hz, errors=e, cfg=cfg)
        T.assert_same(o, p, "zero batch")
        assert (p[1]["hibrido"] == -2).all() and (p[1]["iteration"] == 0).all()
    e = np.zeros((1, 144), np.int32)
    e[0, [5, 77]] = 1
    T.assert_same(T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=cfg),
                  T.run_cc(product, G.KIND_CC, hx, hz, errors=e, cfg=cfg), "single shot")


def test_weight_seven_is_searched_as_six(oracle, product):
    hx, hz, e = T.cc_workload(72, 400, 0.02, seed=21)
    for W in (6, 7):
        cfg = G.make_config(max_error_weight=W)
        T.assert_same(T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=cfg),
                      T.run_cc(product, G.KIND_CC, hx, hz, errors=e, cfg=cfg), f"W={W}")


def test_bb288_and_large_n(oracle, product):
    """BB288 (n = 144 checks) and a 392-check code (16-word columns)."""
    hx, hz, e = T.cc_workload(288, 600, 0.003, seed=22)
    cfg = G.make_config(max_error_weight=3)
    T.assert_same(T.run_cc(oracle, G.KIND_CC, hx, hz, errors=e, cfg=cfg),
                  T.run_cc(product, G.KIND_CC, hx, hz, errors=e, cfg=cfg), "BB288")
    rng = np.random.default_rng(23)
    n, m = 392, 784
    H = np.zeros((n, m), np.int32)
    for c in range(m):
        H[rng.choice(n, 3, replace=False), c] = 1
    e = (rng.random((300, m)) < 0.002).astype(np.int32)
    T.assert_same(T.run_cc(oracle, G.KIND_CC, H, H, errors=e, cfg=cfg),
                  T.run_cc(product, G.KIND_CC, H, H, errors=e, cfg=cfg), "n=392")


def test_dem_without_logicals(oracle, product):
    H, L, p = T.dem(11, 72, 400)
    L0 = np.zeros((H.shape[1], 0), np.int32)
    e = (np.random.default_rng(24).random((300, H.shape[1])) < 0.004).astype(np.int32)
    det = (e @ H.T % 2).astype(np.int32)
    lg = np.zeros((300, 0), np.int32)
    cfg = G.make_config(max_error_weight=3, max_size_active_positions=30)
    T.assert_same(T.run_cln(oracle, G.KIND_CLN, H, L0, det, lg, cfg=cfg),
                  T.run_cln(product, G.KIND_CLN, H, L0, det, lg, cfg=cfg), "n_logicos=0")


def test_non_binary_input_rejected(capfd, product):
    hx, hz = codes.bb(72)
    syn = np.zeros((3, 36), np.int32)
    syn[1, 4] = 2
    rc = T.run_cc(product, G.KIND_CC, hx, hz, syndromes=syn, cfg=G.make_config(max_error_weight=3))[0]
    assert rc == -1
    assert "input rows must hold only 0 and 1" in capfd.readouterr().err


def test_active_set_overflow_reported(capfd, product):
    """A code with more than 1024 columns all touching the flipped check: the
    active set exceeds the supported 1024 positions and the call returns -1."""
    n, m = 4, 1100
    H = np.zeros((n, m), np.int32)
    H[0, :] = 1
    syn = np.zeros((1, n), np.int32)
    syn[0, 0] = 1
    rc = T.run_cc(product, G.KIND_CC, H, H, syndromes=syn, cfg=G.make_config(max_error_weight=1))[0]
    assert rc == -1
    assert "exceeds the supported 1024" in capfd.readouterr().err
