"""MI355X GRAND decoder: ctypes binding of libgranddecoder.so.

The names and argument meaning mirror the reference's Python layer
(Examples/auxilary_functions.py): the two ctypes structs (:9-35),
`load_library` with the argtypes of all five entry points (:172-253) and the
`run_grand_*` batch wrappers (:255-395), so scripts written against the
reference run unchanged.  `run_grand_*_ex` additionally return the decoded
patterns (and decoded logicals for the DEM family) from the `_ex` entry points.

The decoder itself is the HIP library built by `build.py`; nothing here falls
back to a CPU implementation: loading fails loudly if the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libgranddecoder.so"
DEFAULT_LIB = os.path.join(HERE, LIB_NAME)

GRAND_MAX_W = 6
GRAND_ENGINE_FUSED = 0
GRAND_ENGINE_PLAIN = 1

KIND_CC, KIND_PROB_CC, KIND_CLN, KIND_PROB_CLN, KIND_ML_CLN = range(5)


class GrandDecoderConfig(ctypes.Structure):
    _fields_ = [
        ("max_error_weight", ctypes.c_int),
        ("threads_per_block", ctypes.c_int),
        ("use_gpu_sum", ctypes.c_int),
        ("max_size_active_positions", ctypes.c_int),
        ("threshold_for_hybrid", ctypes.c_int),
        ("mode_of_search", ctypes.c_int),
    ]


class GrandDecoderResult(ctypes.Structure):
    _fields_ = [
        ("found", ctypes.c_int),
        ("found_weight", ctypes.c_int),
        ("original_weight", ctypes.c_int),
        ("degenerate", ctypes.c_int),
        ("is_equal", ctypes.c_int),
        ("tam", ctypes.c_int),
        ("total_tested_combinations", ctypes.c_ulonglong),
        ("hibrido", ctypes.c_int),
        ("iteration", ctypes.c_int),
        ("time_sum_ms", ctypes.c_double),
        ("total_time_search_ms", ctypes.c_double),
        ("transfer_time_ms", ctypes.c_double),
        ("total_to_find_ms", ctypes.c_double),
    ]


class GrandDecoderExConfig(ctypes.Structure):
    _fields_ = [
        ("engine", ctypes.c_int),
        ("waves_per_cu", ctypes.c_int),
        ("claim_batch", ctypes.c_int),
        ("split_threshold", ctypes.c_int),
        ("triage_search", ctypes.c_int),
        ("pair_table_kb", ctypes.c_int),
        ("reserved", ctypes.c_int * 2),
    ]


# numpy view of GrandDecoderResult (72 bytes)
RESULT_DTYPE = np.dtype([
    ("found", np.int32), ("found_weight", np.int32), ("original_weight", np.int32), ("degenerate", np.int32),
    ("is_equal", np.int32), ("tam", np.int32), ("total_tested_combinations", np.uint64), ("hibrido", np.int32),
    ("iteration", np.int32), ("time_sum_ms", np.float64), ("total_time_search_ms", np.float64),
    ("transfer_time_ms", np.float64), ("total_to_find_ms", np.float64),
])
assert RESULT_DTYPE.itemsize == ctypes.sizeof(GrandDecoderResult) == 72

PARITY_FIELDS = ("found", "found_weight", "original_weight", "degenerate", "is_equal", "tam",
                 "total_tested_combinations", "hibrido", "iteration")

_IP = ctypes.POINTER(ctypes.c_int)
_CFG = ctypes.POINTER(GrandDecoderConfig)
_RES = ctypes.POINTER(GrandDecoderResult)
_EX = ctypes.POINTER(GrandDecoderExConfig)

_CC_ARGS = [_IP, _IP, ctypes.c_int, ctypes.c_int, _IP, _IP, ctypes.c_int, _CFG, _RES, ctypes.c_char]
_CLN_ARGS = [_IP, _IP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _IP, _IP, ctypes.c_int, _CFG, _RES]

ABI_SYMBOLS = ("grand_decode_cc", "grand_decode_prob_cc", "grand_decode_cln", "grand_decode_prob_cln",
               "grand_decode_ML_cln")
EX_SYMBOLS = tuple(s + "_ex" for s in ABI_SYMBOLS)


def bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Set argtypes/restype of the ABI and _ex entry points on a loaded library
    (the reference's load_library, auxilary_functions.py:172-253)."""
    lib.grand_decode_cc.argtypes = _CC_ARGS
    lib.grand_decode_prob_cc.argtypes = _CC_ARGS + [ctypes.c_char_p]
    lib.grand_decode_cln.argtypes = _CLN_ARGS
    lib.grand_decode_prob_cln.argtypes = _CLN_ARGS + [ctypes.c_char_p]
    lib.grand_decode_ML_cln.argtypes = _CLN_ARGS + [ctypes.c_char_p]
    lib.grand_decode_cc_ex.argtypes = _CC_ARGS + [_EX, _IP]
    lib.grand_decode_prob_cc_ex.argtypes = _CC_ARGS + [ctypes.c_char_p, _EX, _IP]
    lib.grand_decode_cln_ex.argtypes = _CLN_ARGS + [_EX, _IP, _IP]
    lib.grand_decode_prob_cln_ex.argtypes = _CLN_ARGS + [ctypes.c_char_p, _EX, _IP, _IP]
    lib.grand_decode_ML_cln_ex.argtypes = _CLN_ARGS + [ctypes.c_char_p, _EX, _IP, _IP]
    for name in ABI_SYMBOLS + EX_SYMBOLS:
        getattr(lib, name).restype = ctypes.c_int
    return lib


def load_library(lib_path: str | None = None) -> ctypes.CDLL:
    """Load libgranddecoder.so (the HIP product) and bind its entry points."""
    path = lib_path or os.environ.get("GRAND_LIB", DEFAULT_LIB)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run library_s_grand_amd/build.py (or __graft_entry__.build())")
    return bind(ctypes.CDLL(path))


def _ptr(a):
    return a.ctypes.data_as(_IP) if a is not None else _IP()


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def results_to_numpy(results) -> np.ndarray:
    """Copy a ctypes GrandDecoderResult array into a structured numpy array."""
    n = len(results)
    buf = (ctypes.c_char * (72 * n)).from_buffer(results) if n else b""
    return np.frombuffer(bytes(buf), dtype=RESULT_DTYPE, count=n).copy()


def make_config(max_error_weight=4, threads_per_block=64, use_gpu_sum=1, max_size_active_positions=300,
                threshold_for_hybrid=0, mode_of_search=-1) -> GrandDecoderConfig:
    return GrandDecoderConfig(max_error_weight, threads_per_block, use_gpu_sum, max_size_active_positions,
                              threshold_for_hybrid, mode_of_search)


def make_ex(engine=GRAND_ENGINE_FUSED, waves_per_cu=0, claim_batch=0, split_threshold=0,
            triage_search=0, pair_table_kb=0) -> GrandDecoderExConfig:
    ex = GrandDecoderExConfig()
    ex.engine, ex.waves_per_cu, ex.claim_batch, ex.split_threshold = engine, waves_per_cu, claim_batch, split_threshold
    ex.triage_search, ex.pair_table_kb = triage_search, pair_table_kb
    return ex


# --------------------------------------------------------------- batch wrappers
def run_grand_shots_cc(lib, Hx, Hz, errors, syndromes, cfg, error_type):
    """auxilary_functions.py:255-284: errors take precedence over syndromes."""
    rc, res, _ = run_grand_shots_cc_ex(lib, Hx, Hz, errors, syndromes, cfg, error_type, ex=None, want_decoded=False,
                                       _abi=True)
    return rc, res


def run_grand_prob_shots_cc(lib, Hx, Hz, errors, syndromes, cfg, prob_file_path, error_type):
    """auxilary_functions.py:286-315."""
    rc, res, _ = run_grand_prob_shots_cc_ex(lib, Hx, Hz, errors, syndromes, cfg, prob_file_path, error_type, ex=None,
                                            want_decoded=False, _abi=True)
    return rc, res


def run_grand_shots_cln(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg):
    """auxilary_functions.py:317-343."""
    rc, res, _, _ = run_grand_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, ex=None,
                                           want_decoded=False, _abi=True)
    return rc, res


def run_grand_prob_shots_cln(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, prob_file_path):
    """auxilary_functions.py:345-370."""
    rc, res, _, _ = run_grand_prob_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg,
                                                prob_file_path, ex=None, want_decoded=False, _abi=True)
    return rc, res


def run_grand_ML_shots_cln(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, llr_file_path):
    """auxilary_functions.py:372-395."""
    rc, res, _, _ = run_grand_ML_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, llr_file_path,
                                              ex=None, want_decoded=False, _abi=True)
    return rc, res


def _cc_common(Hx, Hz, errors, syndromes):
    n, m = Hx.shape
    Hx, Hz = _i32(Hx), _i32(Hz)
    errors, syndromes = _i32(errors), _i32(syndromes)
    if errors is None and syndromes is None:
        raise ValueError("either errors or syndromes is required")
    shots = errors.shape[0] if errors is not None else syndromes.shape[0]
    return n, m, Hx, Hz, errors, (None if errors is not None else syndromes), shots


def _char(t):
    if isinstance(t, (bytes, bytearray)):
        return ctypes.c_char(bytes(t[:1]))
    return ctypes.c_char(str(t).encode()[:1])


def run_grand_shots_cc_ex(lib, Hx, Hz, errors, syndromes, cfg, error_type, ex=None, want_decoded=True, _abi=False):
    n, m, Hx, Hz, errors, syndromes, shots = _cc_common(Hx, Hz, errors, syndromes)
    results = (GrandDecoderResult * shots)()
    dec = np.zeros((shots, m), dtype=np.int32) if want_decoded else None
    if _abi:
        rc = lib.grand_decode_cc(_ptr(Hx), _ptr(Hz), n, m, _ptr(errors), _ptr(syndromes), shots, ctypes.byref(cfg),
                                 results, _char(error_type))
    else:
        rc = lib.grand_decode_cc_ex(_ptr(Hx), _ptr(Hz), n, m, _ptr(errors), _ptr(syndromes), shots,
                                    ctypes.byref(cfg), results, _char(error_type),
                                    ctypes.byref(ex) if ex is not None else _EX(), _ptr(dec))
    return rc, results, dec


def run_grand_prob_shots_cc_ex(lib, Hx, Hz, errors, syndromes, cfg, prob_file_path, error_type, ex=None,
                               want_decoded=True, _abi=False):
    n, m, Hx, Hz, errors, syndromes, shots = _cc_common(Hx, Hz, errors, syndromes)
    results = (GrandDecoderResult * shots)()
    dec = np.zeros((shots, m), dtype=np.int32) if want_decoded else None
    path = prob_file_path.encode("utf-8")
    if _abi:
        rc = lib.grand_decode_prob_cc(_ptr(Hx), _ptr(Hz), n, m, _ptr(errors), _ptr(syndromes), shots,
                                      ctypes.byref(cfg), results, _char(error_type), path)
    else:
        rc = lib.grand_decode_prob_cc_ex(_ptr(Hx), _ptr(Hz), n, m, _ptr(errors), _ptr(syndromes), shots,
                                         ctypes.byref(cfg), results, _char(error_type), path,
                                         ctypes.byref(ex) if ex is not None else _EX(), _ptr(dec))
    return rc, results, dec


def _cln(lib, sym, H_dem, L_dem, detectors, logicals, n_logicos, cfg, path, ex, want_decoded, abi):
    n, m = H_dem.shape
    H_dem, L_dem = _i32(H_dem), _i32(L_dem)
    detectors, logicals = _i32(detectors), _i32(logicals)
    shots = detectors.shape[0]
    results = (GrandDecoderResult * shots)()
    dec = np.zeros((shots, m), dtype=np.int32) if want_decoded else None
    dlog = np.zeros((shots, n_logicos), dtype=np.int32) if want_decoded else None
    args = [_ptr(H_dem), _ptr(L_dem), n, m, n_logicos, _ptr(detectors), _ptr(logicals), shots, ctypes.byref(cfg),
            results]
    if path is not None:
        args.append(path.encode("utf-8"))
    if abi:
        rc = getattr(lib, sym)(*args)
    else:
        rc = getattr(lib, sym + "_ex")(*args, ctypes.byref(ex) if ex is not None else _EX(), _ptr(dec), _ptr(dlog))
    return rc, results, dec, dlog


def run_grand_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, ex=None, want_decoded=True,
                           _abi=False):
    return _cln(lib, "grand_decode_cln", H_dem, L_dem, detectors, logicals, n_logicos, cfg, None, ex, want_decoded,
                _abi)


def run_grand_prob_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, prob_file_path, ex=None,
                                want_decoded=True, _abi=False):
    return _cln(lib, "grand_decode_prob_cln", H_dem, L_dem, detectors, logicals, n_logicos, cfg, prob_file_path, ex,
                want_decoded, _abi)


def run_grand_ML_shots_cln_ex(lib, H_dem, L_dem, detectors, logicals, n_logicos, cfg, llr_file_path, ex=None,
                              want_decoded=True, _abi=False):
    return _cln(lib, "grand_decode_ML_cln", H_dem, L_dem, detectors, logicals, n_logicos, cfg, llr_file_path, ex,
                want_decoded, _abi)


# ------------------------------------------------------------- CSV writers
def save_results_cc_csv(output_file, results, num_shots, using_errors):
    """Column order of auxilary_functions.py:75-95."""
    import csv
    with open(output_file, "w", newline="", encoding="utf-8") as f:
        w = csv.writer(f)
        if using_errors:
            w.writerow(["shot", "found", "found_weight", "original_weight", "degenerate", "is_equal",
                        "size_active_positions", "total_tested_combinations", "hybrid", "iteration", "time_sum_ms",
                        "total_time_search_ms", "transfer_time_ms", "total_to_find_ms"])
        else:
            w.writerow(["shot", "found", "found_weight", "size_active_positions", "total_tested_combinations",
                        "hybrid", "iteration", "time_sum_ms", "total_time_search_ms", "transfer_time_ms",
                        "total_to_find_ms"])
        for i in range(num_shots):
            r = results[i]
            if using_errors:
                w.writerow([i, r.found, r.found_weight, r.original_weight, r.degenerate, r.is_equal, r.tam,
                            r.total_tested_combinations, r.hibrido, r.iteration, r.time_sum_ms,
                            r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms])
            else:
                w.writerow([i, r.found, r.found_weight, r.tam, r.total_tested_combinations, r.hibrido, r.iteration,
                            r.time_sum_ms, r.total_time_search_ms, r.transfer_time_ms, r.total_to_find_ms])


def save_results_cln_csv(output_file, results, num_shots, using_errors=True):
    """Column order of auxilary_functions.py:97-106."""
    import csv
    with open(output_file, "w", newline="", encoding="utf-8") as f:
        w = csv.writer(f)
        w.writerow(["shot", "found", "found_weight", "original_weight", "degenerate", "is_equal",
                    "size_active_positions", "total_tested_combinations", "hybrid", "iteration", "time_sum_ms",
                    "total_time_search_ms", "transfer_time_ms", "total_to_find_ms"])
        for i in range(num_shots):
            r = results[i]
            w.writerow([i, r.found, r.found_weight, r.original_weight, r.degenerate, r.is_equal, r.tam,
                        r.total_tested_combinations, r.hibrido, r.iteration, r.time_sum_ms, r.total_time_search_ms,
                        r.transfer_time_ms, r.total_to_find_ms])


# ------------------------------------------------------ file and simulation helpers
# Same names and behaviour as auxilary_functions.py:37-73, :109-131, :168-170.
def save_syndromes_to_file(syndromes, output_file):
    """One row per line, values separated by single spaces (auxilary_functions.py:37-41)."""
    import csv
    with open(output_file, "w", newline="", encoding="utf-8") as f:
        w = csv.writer(f, delimiter=" ")
        for row in syndromes:
            w.writerow([int(v) for v in row])


def read_errors_or_syndromes_from_file(file_path, n_cols, max_shots=None):
    """Rows of n_cols ints or one n_cols-character bitstring (auxilary_functions.py:44-73)."""
    from .codes import read_rows
    return read_rows(file_path, n_cols, max_shots)


def read_matrix_from_file(file_path):
    """Whitespace-separated rows; int8 like auxilary_functions.py:109-121."""
    from .codes import read_matrix
    return read_matrix(file_path).astype(np.int8)


def compute_syndrome(H, error):
    """(H @ e) mod 2 (auxilary_functions.py:124-126)."""
    return (np.asarray(H) @ np.asarray(error)) % 2


def simulate_random_error(n, p_error):
    """Nonzero i.i.d. Bernoulli(p_error) vector, redrawn until nonzero (auxilary_functions.py:128-132)."""
    while True:
        e = (np.random.rand(n) < p_error).astype(int)
        if np.any(e):
            return e


def error_positions_random(n, error_weight):
    """error_weight distinct positions in [0, n) (auxilary_functions.py:168-170)."""
    return np.random.choice(n, size=error_weight, replace=False)
