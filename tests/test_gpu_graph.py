"""HIP-graph replay of a decode and the autotuner.

A captured decode is replayed several times, with new inputs written into the
same buffers and eager work between the replays; every replay must equal the
eager decode of the same inputs, field by field and pattern by pattern.  The
autotuner must leave its best configuration applied and change no result.
"""
import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import autotune, codes
from library_s_grand_amd.session import Session, results_from_bytes
import grandtest as T

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _same(a, b, what):
    for f in G.PARITY_FIELDS:
        assert (a[f] == b[f]).all(), (what, f)


def test_graph_replays_match_eager_cc(tmp_path):
    hx, hz = codes.bb(144)
    cfg = G.make_config(max_error_weight=4)
    pf = T.phase_file(tmp_path)
    shots = 3000
    sess = Session(G.KIND_PROB_CC, hx, hz, cfg, type_of_error="x", file=pf)
    ref = Session(G.KIND_PROB_CC, hx, hz, cfg, type_of_error="x", file=pf)
    stream = torch.cuda.Stream()
    d_err = torch.zeros(shots, 144, dtype=torch.int32, device="cuda")
    d_res = torch.zeros(shots * 72, dtype=torch.uint8, device="cuda")
    d_dec = torch.zeros(shots * sess.pattern_words, dtype=torch.int32, device="cuda")
    e_res = torch.zeros_like(d_res)
    e_dec = torch.zeros_like(d_dec)
    graph = sess.capture(d_err, True, d_res, stream.cuda_stream, d_decoded=d_dec)
    for rep in range(4):
        _, _, errors = T.cc_workload(144, shots, "loguniform", seed=100 + rep)
        d_err.copy_(torch.from_numpy(errors))
        torch.cuda.synchronize()
        graph.launch(stream.cuda_stream)
        stream.synchronize()
        ref.decode(d_err, True, e_res, d_decoded=e_dec)          # eager, default stream
        torch.cuda.synchronize()
        _same(results_from_bytes(d_res.cpu().numpy()), results_from_bytes(e_res.cpu().numpy()), rep)
        assert torch.equal(d_dec, e_dec), rep
    graph.close()
    sess.close()
    ref.close()


def test_graph_replays_match_oracle_dem(oracle, tmp_path):
    H, L, p, det, lg = T.dem_workload(1500, seed=41, dem_seed=5, n_det=72, n_mech=400)
    cfg = G.make_config(max_error_weight=4, max_size_active_positions=64)
    lf = T.llr_file(tmp_path, p)
    sess = Session(G.KIND_ML_CLN, H, L, cfg, n_logicos=L.shape[1], file=lf)
    stream = torch.cuda.Stream()
    d_det = torch.from_numpy(det).cuda()
    d_lg = torch.from_numpy(lg).cuda()
    d_res = torch.zeros(det.shape[0] * 72, dtype=torch.uint8, device="cuda")
    graph = sess.capture(d_det, False, d_res, stream.cuda_stream, d_logicals=d_lg)
    rc, o = T.run_cln(oracle, G.KIND_ML_CLN, H, L, det, lg, cfg=cfg, file=lf)[:2]
    assert rc == 0
    for rep in range(3):
        d_res.zero_()
        torch.cuda.synchronize()
        graph.launch(stream.cuda_stream)
        stream.synchronize()
        _same(results_from_bytes(d_res.cpu().numpy()), o, rep)
    graph.close()
    sess.close()


def test_autotuner_applies_best_and_keeps_results(tmp_path):
    lib = G.load_library()
    before = autotune.get(lib)
    hx, hz, errors = T.cc_workload(144, 20000, "loguniform", seed=7)
    cfg = G.make_config(max_error_weight=4)
    pf = T.phase_file(tmp_path)
    sess = Session(G.KIND_PROB_CC, hx, hz, cfg, type_of_error="x", file=pf, lib=lib)
    sess.enable_timing(True)
    d_err = torch.from_numpy(errors).cuda()
    d_res = torch.zeros(errors.shape[0] * 72, dtype=torch.uint8, device="cuda")
    sess.decode(d_err, True, d_res)
    first = results_from_bytes(d_res.cpu().numpy())

    def run():
        sess.decode(d_err, True, d_res)
        return sess.last_timing()

    space = {"waves_per_cu": (4, 8), "claim_batch": (1, 4), "rmax": (16, 32), "split_threshold": (1024, 2048),
             "pair_table_kb": (8, -1)}
    try:
        best, ms, table = autotune.tune(lib, run, reps=2, space=space, max_rounds=1)
        assert autotune.get(lib) == best and ms > 0 and len(table) >= 2
        autotune.save(str(tmp_path / "tune.json"), best)
        assert autotune.load(str(tmp_path / "tune.json")) == best
        sess.decode(d_err, True, d_res)
        _same(results_from_bytes(d_res.cpu().numpy()), first, "tuned")
    finally:
        autotune.reset(lib)                       # the built-in defaults of every search kind
        sess.close()
