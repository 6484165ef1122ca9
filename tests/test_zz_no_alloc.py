"""Decode without allocation (grand_session_decode_device_noalloc) against the
allocating entry point (grand_session_decode_device), through the C ABI:
every result field (timings aside), the decoded patterns and the decoded
logicals bit for bit, on two calls in a row, the second with fewer shots (the
fixed scratch keeps the first call's contents).  Also: a batch larger than the
fixed scratch is rejected before any device work."""
import ctypes

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import codes
from library_s_grand_amd.session import Session, results_from_bytes
import grandtest as T

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

_VP = ctypes.c_void_p


def _bind(lib):
    lib.grand_session_decode_device_noalloc.argtypes = [_VP, _VP, ctypes.c_int, _VP, ctypes.c_int, _VP, _VP, _VP,
                                                        _VP, _VP]
    lib.grand_session_decode_device_noalloc.restype = ctypes.c_int
    return lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _decode(sess, noalloc, d_in, is_err, d_lg, shots):
    res = torch.full((shots * 72,), 0xAB, dtype=torch.uint8, device="cuda")
    dec = torch.full((shots, sess.pattern_words), -1, dtype=torch.int32, device="cuda")
    dlg = torch.full((shots, max(sess.logical_words, 1)), -1, dtype=torch.int32, device="cuda")
    fn = sess.lib.grand_session_decode_device_noalloc if noalloc else sess.lib.grand_session_decode_device
    rc = fn(sess.h, _p(d_in), int(is_err), _p(d_lg), shots, _p(res), _p(dec), _p(dlg) if d_lg is not None else None,
            None, None)
    assert rc == 0
    torch.cuda.synchronize()
    assert sess.error_bits() == 0
    return results_from_bytes(res.cpu().numpy()), dec.cpu().numpy(), dlg.cpu().numpy()


def _same(a, b, what):
    for f in G.PARITY_FIELDS:
        bad = np.nonzero(a[0][f] != b[0][f])[0]
        assert bad.size == 0, f"{what}: {f} differs at {bad.size} shots, first {bad[:1]}"
    assert (a[1] == b[1]).all(), f"{what}: decoded patterns differ"
    assert (a[2] == b[2]).all(), f"{what}: decoded logicals differ"


@pytest.mark.parametrize("split", [0, 1])
def test_noalloc_matches_allocating_decode_bb144(tmp_path, split):
    lib = _bind(G.load_library())
    hx, hz, e = T.cc_workload(144, 30_000, "loguniform", seed=61)
    pf = T.phase_file(tmp_path)
    sess = Session(G.KIND_PROB_CC, hx, hz, G.make_config(max_error_weight=4), type_of_error="x", file=pf,
                   ex=G.make_ex(split_threshold=split), lib=lib)
    d_e = torch.from_numpy(e).cuda()
    # two calls in a row, the second smaller and on other shots
    for lo, hi in ((0, 20_000), (20_000, 25_000)):
        d_in = d_e[lo:hi].contiguous()
        ref = _decode(sess, False, d_in, True, None, hi - lo)
        got = _decode(sess, True, d_in, True, None, hi - lo)
        _same(ref, got, f"BB144 prob_cc shots {lo}..{hi} split={split}")
        assert ref[0]["found"].sum() > 1000
    sess.close()


def test_noalloc_matches_allocating_decode_dem_ml(tmp_path):
    lib = _bind(G.load_library())
    H, L, pr, det, lg = T.dem_workload(3000, seed=62)
    lf = T.llr_file(tmp_path, pr)
    sess = Session(G.KIND_ML_CLN, H, L, G.make_config(max_error_weight=4, max_size_active_positions=64),
                   n_logicos=L.shape[1], file=lf, lib=lib)
    d_det, d_lg = torch.from_numpy(det).cuda(), torch.from_numpy(lg).cuda()
    for lo, hi in ((0, 2000), (2000, 2700)):
        a, b = d_det[lo:hi].contiguous(), d_lg[lo:hi].contiguous()
        ref = _decode(sess, False, a, False, b, hi - lo)
        got = _decode(sess, True, a, False, b, hi - lo)
        _same(ref, got, f"DEM ML shots {lo}..{hi}")
    sess.close()


def test_noalloc_rejects_batches_beyond_the_scratch():
    lib = _bind(G.load_library())
    hx, hz, e = T.cc_workload(144, 64, 0.01, seed=63)
    sess = Session(G.KIND_CC, hx, hz, G.make_config(max_error_weight=3), type_of_error="x", lib=lib)
    d_e = torch.from_numpy(e).cuda()
    res = torch.empty(64 * 72, dtype=torch.uint8, device="cuda")
    # 10^6 BB144 shots need about 76 MB of scratch: rejected on the host, nothing is launched
    rc = lib.grand_session_decode_device_noalloc(sess.h, _p(d_e), 1, None, 1_000_000, _p(res), None, None, None, None)
    assert rc == -1
    # a small batch still decodes, equal to the allocating entry point
    _same(_decode(sess, False, d_e, True, None, 64), _decode(sess, True, d_e, True, None, 64), "BB144 cc 64 shots")
    sess.close()
