"""CPU checks of the drop-in boundary (no GPU needed):

* the C header is valid C and its struct layouts match the reference's ctypes
  mirrors (grand_decoder.h:37-68, Examples/auxilary_functions.py:9-35);
* libgranddecoder.so loads and exports every function include/*.h declares;
* invalid configurations are rejected exactly like the reference drivers
  (return -1, same stderr message), by the product and by the oracle alike;
* the BB-code generator reproduces the reference's H_Matrix files.
"""
import ctypes
import glob
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
import grandtest as T

HEADERS = sorted(glob.glob(os.path.join(T.ROOT, "include", "*.h")))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(grand_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_compiles_as_c_and_layout(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "grand_decoder.h"
#include "grand_session.h"
#include "grand_server.h"
#include "grand_mlp.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(GrandDecoderConfig), sizeof(GrandDecoderResult),
         offsetof(GrandDecoderResult, tam), offsetof(GrandDecoderResult, total_tested_combinations),
         offsetof(GrandDecoderResult, hibrido), offsetof(GrandDecoderResult, iteration),
         offsetof(GrandDecoderResult, time_sum_ms), sizeof(GrandDecoderExConfig), sizeof(GrandServerConfig),
         offsetof(GrandServerConfig, hbm_budget_bytes));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(T.ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [24, 72, 20, 24, 32, 36, 40, 32, 24, 16]
    from library_s_grand_amd.serving import GrandServerConfig
    assert ctypes.sizeof(GrandServerConfig) == 24 and GrandServerConfig.hbm_budget_bytes.offset == 16
    assert ctypes.sizeof(G.GrandDecoderConfig) == 24 and ctypes.sizeof(G.GrandDecoderResult) == 72
    assert G.GrandDecoderResult.total_tested_combinations.offset == 24
    assert G.GrandDecoderResult.time_sum_ms.offset == 40


def test_product_exports_every_declared_function():
    lib = ctypes.CDLL(G.DEFAULT_LIB)
    names = declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_tuning_knobs_set_get_and_reset():
    """Host-side state only (no GPU): a set replaces the knobs, values <= 0 keep them, all four negative
    restores the built-in first-hit defaults (1 wave per CU, claim batch 2, 64 prefixes, split 65536,
    the pair table filling the wave's LDS share)."""
    from library_s_grand_amd import autotune
    lib = ctypes.CDLL(G.DEFAULT_LIB)
    try:
        autotune.apply(lib, {"waves_per_cu": 3, "claim_batch": 4, "rmax": 16, "split_threshold": 100,
                             "pair_table_kb": 16})
        assert autotune.get(lib) == {"waves_per_cu": 3, "claim_batch": 4, "rmax": 16, "split_threshold": 100,
                                     "pair_table_kb": 16}
        lib.grand_set_tuning(0, 0, 8, 0)
        assert autotune.get(lib) == {"waves_per_cu": 3, "claim_batch": 4, "rmax": 8, "split_threshold": 100,
                                     "pair_table_kb": 16}
        lib.grand_set_tuning(0, 99, 0, 0)                     # claim batch is capped at 8
        assert autotune.get(lib)["claim_batch"] == 8
        lib.grand_set_pair_table_kb(-1)                      # no pair table
        assert autotune.get(lib)["pair_table_kb"] == -1
    finally:
        autotune.reset(lib)
    assert autotune.get(lib) == {"waves_per_cu": 1, "claim_batch": 2, "rmax": 64, "split_threshold": 65536,
                                 "pair_table_kb": 0}


def test_every_exported_function_is_declared():
    out = subprocess.run(["nm", "-D", "--defined-only", G.DEFAULT_LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if ln.split()[-2:-1] == ["T"] and
                       ln.split()[-1].startswith("grand_")})
    declared = set(declared_functions())
    assert exported and not [n for n in exported if n not in declared]


def test_oracle_exports_abi():
    lib = T.oracle_lib()
    for n in G.ABI_SYMBOLS + G.EX_SYMBOLS:
        assert hasattr(lib, n)


# -------------------------------------------------------------- validation
def _cc_call(lib, kind, cfg, etype="x", prob=None, errors=True):
    hx, hz = codes.bb(72)
    e = np.zeros((2, 72), np.int32)
    e[0, 3] = 1
    if kind == G.KIND_CC:
        return T.run_cc(lib, kind, hx, hz, errors=e, cfg=cfg, etype=etype)[0]
    return T.run_cc(lib, kind, hx, hz, errors=e, cfg=cfg, etype=etype, prob_file=prob)[0]


INVALID_CC = [
    ("bad type", G.KIND_CC, dict(max_error_weight=3), "q", "Invalid type_of_error: q. Must be 'x' or 'z'."),
    ("W=0", G.KIND_CC, dict(max_error_weight=0), "x", "Invalid max_error_weight: 0. Must be between 1 and 7."),
    ("W=8", G.KIND_CC, dict(max_error_weight=8), "z", "Invalid max_error_weight: 8. Must be between 1 and 7."),
    ("prob W=9", G.KIND_PROB_CC, dict(max_error_weight=9), "x", "Invalid max_error_weight: 9. Must be 6 or 7."),
    ("prob bad type", G.KIND_PROB_CC, dict(max_error_weight=3), "y", "Invalid type_of_error: y"),
]


@pytest.mark.parametrize("what,kind,cfgkw,etype,msg", INVALID_CC)
def test_invalid_cc_configs_rejected_like_reference(capfd, oracle, tmp_path, what, kind, cfgkw, etype, msg):
    pf = T.phase_file(tmp_path)
    cfg = G.make_config(**cfgkw)
    rc_o = _cc_call(oracle, kind, cfg, etype, pf)
    err_o = capfd.readouterr().err
    rc_p = _cc_call(T.product_lib(), kind, cfg, etype, pf)
    err_p = capfd.readouterr().err
    assert rc_o == -1 and rc_p == -1, what
    assert msg in err_o and msg in err_p, (err_o, err_p)


def test_missing_phase_file_rejected(capfd, oracle, tmp_path):
    cfg = G.make_config(max_error_weight=3)
    missing = str(tmp_path / "nope.txt")
    for lib in (oracle, T.product_lib()):
        assert _cc_call(lib, G.KIND_PROB_CC, cfg, "x", missing) == -1
        err = capfd.readouterr().err
        assert "Error opening probability file" in err and "No phase order masks read from" in err


def _dem_call(lib, kind, cfg, file=None):
    H, L, p = T.dem(11, 36, 60)
    det = np.zeros((2, 36), np.int32)
    det[0, :2] = 1
    lg = np.zeros((2, L.shape[1]), np.int32)
    return T.run_cln(lib, kind, H, L, det, lg, cfg=cfg, file=file)[0]


def test_invalid_dem_configs_rejected(capfd, oracle, tmp_path):
    short = tmp_path / "short_llr.txt"
    short.write_text("1.0\n2.0\n")
    empty = tmp_path / "empty_phase.txt"
    empty.write_text("")
    cases = [
        (G.KIND_CLN, G.make_config(max_error_weight=0), None, "Invalid max_error_weight: 0"),
        (G.KIND_ML_CLN, G.make_config(max_error_weight=3), str(tmp_path / "none.txt"), "Error opening LLR file"),
        (G.KIND_ML_CLN, G.make_config(max_error_weight=3), str(short), "Error reading LLR file (expected 60 values)"),
        (G.KIND_PROB_CLN, G.make_config(max_error_weight=3), str(empty), "No phase order masks read from"),
        (G.KIND_PROB_CLN, G.make_config(max_error_weight=8), str(empty), "Invalid max_error_weight: 8"),
    ]
    for kind, cfg, f, msg in cases:
        for lib in (oracle, T.product_lib()):
            assert _dem_call(lib, kind, cfg, f) == -1, msg
            assert msg in capfd.readouterr().err, msg


def test_zero_shots_is_a_successful_noop(oracle):
    hx, hz = codes.bb(72)
    cfg = G.make_config(max_error_weight=3)
    res = (G.GrandDecoderResult * 1)()
    for lib in (oracle,):
        rc = lib.grand_decode_cc(hx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 hz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 36, 72, None, None, 0,
                                 ctypes.byref(cfg), res, ctypes.c_char(b"x"))
        assert rc == 0


# -------------------------------------------------------------- generators
def test_bb_generator_matches_reference_files():
    sha = json.load(open(os.path.join(T.GOLDEN, "reference_h_sha256.json")))
    for nq in (72, 144, 288):
        hx, hz = codes.bb(nq)
        for name, H in (("hx", hx), ("hz", hz)):
            digest = hashlib.sha256(codes.matrix_to_text(H).encode()).hexdigest()
            assert digest == sha[f"{name}_ibm_{nq}.txt"], (name, nq)
        assert not ((hx.astype(np.int64) @ hz.T) % 2).any()


def test_reference_text_formats_roundtrip(tmp_path):
    hz = codes.read_matrix(os.path.join(T.GOLDEN, "hz_ibm_72.txt"))
    assert hz.shape == (36, 72)
    pats = codes.read_rows(os.path.join(T.GOLDEN, "patterns_72_first30.txt"), 72)   # bitstring rows
    syn = codes.read_rows(os.path.join(T.GOLDEN, "syndromes_72.txt"), 36)           # int rows
    assert pats.shape == (30, 72) and syn.shape == (30, 36)
    p = tmp_path / "llr.txt"
    llr = np.array([1.5, 2.25, 7.123456789], np.float32)
    codes.write_llr_file(str(p), llr)
    assert (np.loadtxt(p, dtype=np.float32) == llr).all()


def test_csv_writers_column_order(tmp_path):
    res = (G.GrandDecoderResult * 2)()
    res[1].found, res[1].tam, res[1].total_tested_combinations = 1, 5, 12
    f = tmp_path / "cc.csv"
    G.save_results_cc_csv(str(f), res, 2, using_errors=False)
    lines = f.read_text().splitlines()
    assert lines[0].startswith("shot,found,found_weight,size_active_positions,total_tested_combinations")
    assert lines[2].startswith("1,1,0,5,12")


def test_cln_csv_one_row_per_shot(tmp_path):
    res = (G.GrandDecoderResult * 3)()
    res[2].found, res[2].degenerate = 1, 1
    f = tmp_path / "cln.csv"
    G.save_results_cln_csv(str(f), res, 3, True)
    lines = f.read_text().splitlines()
    assert len(lines) == 4 and lines[3].startswith("2,1,0,0,1,0")


def test_reference_file_and_simulation_helpers(tmp_path):
    hz = G.read_matrix_from_file(os.path.join(T.GOLDEN, "hz_ibm_72.txt"))
    assert hz.dtype == np.int8 and hz.shape == (36, 72)
    pats = G.read_errors_or_syndromes_from_file(os.path.join(T.GOLDEN, "patterns_72_first30.txt"), 72, max_shots=7)
    assert pats.shape == (7, 72)
    syn = np.array([G.compute_syndrome(hz.astype(np.int64), e) for e in pats])
    ref = G.read_errors_or_syndromes_from_file(os.path.join(T.GOLDEN, "syndromes_72.txt"), 36)[:7]
    assert (syn == ref).all()
    out = tmp_path / "syn.txt"
    G.save_syndromes_to_file(syn, str(out))
    assert (G.read_errors_or_syndromes_from_file(str(out), 36) == syn).all()
    np.random.seed(3)
    e = G.simulate_random_error(72, 0.01)
    assert e.any() and e.shape == (72,)
    pos = G.error_positions_random(72, 5)
    assert len(set(pos.tolist())) == 5 and pos.max() < 72
