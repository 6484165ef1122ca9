"""Serving loop on the GPU: requests of three codes arriving together are
batched (several requests per launch) and every request's results equal those
of decoding its rows alone through the ABI, and the oracle's; admission by HBM
footprint and by queue depth refuses what does not fit."""
import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes, serving
import grandtest as T

pytestmark = pytest.mark.gpu

SIZES = (72, 144, 288)


def _alone(lib, q, errors, W=4):
    hx, hz = codes.bb(q)
    rc, res, dec = G.run_grand_shots_cc_ex(lib, hx, hz, errors, None, G.make_config(max_error_weight=W), "x")
    assert rc == 0
    return G.results_to_numpy(res), dec


def test_mixed_requests_equal_alone_and_oracle(oracle):
    reqs = serving.mixed_bb_workload(48, rate_hz=1e6, seed=3, max_shots=300)
    srv, ids = serving.setup_bb_server(max_batch_shots=4096, lanes_per_code=2)
    try:
        before = srv.counters()
        out, _ = serving.serve_open_loop(srv, ids, reqs, want_decoded=True)
        after = srv.counters()
    finally:
        srv.close()
    batches = after["batches"] - before["batches"]
    assert after["requests"] - before["requests"] == len(reqs)
    assert batches < len(reqs), "requests that arrive together must share batches"
    prod = G.load_library()
    for i, (r, (_, lat, res, dec)) in enumerate(zip(reqs, out)):
        q = SIZES[r.code]
        a_res, a_dec = _alone(prod, q, r.errors)
        for f in G.PARITY_FIELDS:
            assert (res[f] == a_res[f]).all(), (i, q, f)
        assert (dec == a_dec).all(), (i, q)
        assert lat > 0
        if i % 8 == 0:
            o_res, o_dec = _alone(oracle, q, r.errors)
            for f in G.PARITY_FIELDS:
                assert (res[f] == o_res[f]).all(), (i, q, f)
            assert (dec == o_dec).all()


def test_hbm_budget_admission():
    hx, hz = codes.bb(144)
    cfg = G.make_config(max_error_weight=3)
    with serving.Server(max_batch_shots=1024, lanes_per_code=1, hbm_budget_bytes=1 << 34) as srv:
        cid = srv.add_code(G.KIND_CC, hx, hz, cfg)
        one = srv.footprint(cid)
        assert one > 0 and srv.footprint() == one
    with serving.Server(max_batch_shots=1024, lanes_per_code=1, hbm_budget_bytes=one - 1) as srv:
        with pytest.raises(MemoryError):
            srv.add_code(G.KIND_CC, hx, hz, cfg)
    with serving.Server(max_batch_shots=1024, lanes_per_code=1, hbm_budget_bytes=one) as srv:
        srv.add_code(G.KIND_CC, hx, hz, cfg)
    with serving.Server(max_batch_shots=1024, lanes_per_code=1, hbm_budget_bytes=one + one // 2) as srv:
        srv.add_code(G.KIND_CC, hx, hz, cfg)
        with pytest.raises(MemoryError):
            srv.add_code(G.KIND_CC, hx, hz, cfg)


def test_queue_and_size_limits():
    hx, hz = codes.bb(72)
    cfg = G.make_config(max_error_weight=3)
    with serving.Server(max_batch_shots=64, lanes_per_code=1, max_queued_shots=32) as srv:
        cid = srv.add_code(G.KIND_CC, hx, hz, cfg)
        with pytest.raises(BlockingIOError):
            srv.submit(cid, np.zeros((33, 72), np.int32))
        with pytest.raises(ValueError):
            srv.submit(cid, np.zeros((65, 72), np.int32))
        e = np.zeros((8, 72), np.int32)
        e[1, 5] = 1
        res, _, _ = srv.wait(srv.submit(cid, e))
        assert res["found"][1] == 1 and res["found"].sum() == 1


def test_dem_code_with_logicals(oracle):
    H, L, p, det, lg = T.dem_workload(200, seed=9, dem_seed=5, n_det=36, n_mech=120)
    cfg = G.make_config(max_error_weight=3, max_size_active_positions=64)
    with serving.Server(max_batch_shots=256, lanes_per_code=1) as srv:
        cid = srv.add_code(G.KIND_CLN, H, L, cfg, n_logicos=L.shape[1])
        t1 = srv.submit(cid, det[:120], input_is_errors=False, logicals=lg[:120])
        t2 = srv.submit(cid, det[120:], input_is_errors=False, logicals=lg[120:])
        r1, _, _ = srv.wait(t1)
        r2, _, _ = srv.wait(t2)
    res = np.concatenate([r1, r2])
    rc, o = T.run_cln(oracle, G.KIND_CLN, H, L, det, lg, cfg=cfg)[:2]
    assert rc == 0 and res["found"].sum() > 0
    for f in G.PARITY_FIELDS:
        assert (res[f] == o[f]).all(), f
