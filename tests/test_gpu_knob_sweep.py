"""GPU parity over seeded random combinations of codes, inputs and tuning knobs.

Each case draws a code (BB72, BB144, BB288 or a synthetic DEM), an error rate,
max_error_weight, the input form (error rows or syndromes, X or Z), the
hybrid window (mode_of_search 1 with max_size_active_positions 40) and the
session knobs that change how the search kernel divides its work but not its
results: split_threshold (0 built-in, 1 share every run, 2048, -1 share
none), waves_per_cu, pair_table_kb (built-in, capped at 8 or 16 KiB, -1 no
pair table) and triage_search (lane search on or off).  Every decode is
compared with the oracle field by field, pattern by pattern and logical by
logical.  The draws are fixed by the seed, so a failure names a case that
reruns as it failed.
"""
import numpy as np
import pytest

import library_s_grand_amd as G
import grandtest as T

pytestmark = pytest.mark.gpu


def _cc_cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nq = int(rng.choice([72, 144, 288]))
        p = float(rng.choice([0.002, 0.005, 0.01] if nq < 288 else [0.002, 0.004]))
        out.append(dict(nq=nq, p=p, W=int(rng.integers(2, 6 if p <= 0.005 else 5)),
                        shots=800 if nq < 288 else 400, seed=int(rng.integers(1 << 30)),
                        et=str(rng.choice(["x", "z"])), errors=bool(rng.integers(2)),
                        split=int(rng.choice([0, 1, 2048, -1])), wpc=int(rng.choice([0, 1, 2, 4])),
                        kb=int(rng.choice([0, 0, 8, -1])), tri=int(rng.choice([0, -1])),
                        window=bool(rng.integers(3) == 0)))
    return out


def _cln_cases(n, seed):
    rng = np.random.default_rng(seed)
    return [dict(n_det=int(rng.choice([48, 96, 216])), n_mech=int(rng.choice([300, 800, 1500])),
                 W=int(rng.integers(2, 5)), ms=int(rng.choice([24, 40, 64])), shots=500,
                 seed=int(rng.integers(1 << 30)), dem_seed=int(rng.integers(100)),
                 split=int(rng.choice([0, 1, -1])), wpc=int(rng.choice([0, 2, 4])),
                 kb=int(rng.choice([0, 0, 16]))) for _ in range(n)]


def _id(c):
    return "-".join(f"{k}{v}" for k, v in c.items())


@pytest.mark.parametrize("c", _cc_cases(8, 20261204), ids=_id)
def test_css_family_knob_sweep(oracle, product, tmp_path, c):
    hx, hz, e = T.cc_workload(c["nq"], c["shots"], c["p"], seed=c["seed"])
    if c["errors"]:
        kw = dict(errors=e)
    else:
        H = hz if c["et"] == "x" else hx
        kw = dict(syndromes=(e @ H.T % 2).astype(np.int32))
    win = dict(mode_of_search=1, max_size_active_positions=40) if c["window"] else {}
    cfg = G.make_config(max_error_weight=c["W"], **win)
    ex = G.make_ex(split_threshold=c["split"], waves_per_cu=c["wpc"], pair_table_kb=c["kb"],
                   triage_search=c["tri"])
    pf = T.phase_file(tmp_path)
    for kind in (G.KIND_CC, G.KIND_PROB_CC):
        o = T.run_cc(oracle, kind, hx, hz, cfg=cfg, etype=c["et"], prob_file=pf, **kw)
        p = T.run_cc(product, kind, hx, hz, cfg=cfg, etype=c["et"], prob_file=pf, ex=ex, **kw)
        T.assert_same(o, p, f"kind {kind}")


@pytest.mark.parametrize("c", _cln_cases(5, 41), ids=_id)
def test_dem_family_knob_sweep(oracle, product, tmp_path, c):
    H, L, p, det, lg = T.dem_workload(c["shots"], c["seed"], dem_seed=c["dem_seed"], n_det=c["n_det"],
                                      n_mech=c["n_mech"])
    cfg = G.make_config(max_error_weight=c["W"], max_size_active_positions=c["ms"])
    ex = G.make_ex(split_threshold=c["split"], waves_per_cu=c["wpc"], pair_table_kb=c["kb"])
    files = {G.KIND_CLN: None, G.KIND_PROB_CLN: T.phase_file(tmp_path), G.KIND_ML_CLN: T.llr_file(tmp_path, p)}
    for kind, f in files.items():
        o = T.run_cln(oracle, kind, H, L, det, lg, cfg=cfg, file=f)
        q = T.run_cln(product, kind, H, L, det, lg, cfg=cfg, file=f, ex=ex)
        T.assert_same(o, q, f"kind {kind}")
