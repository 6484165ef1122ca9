"""Device-resident decoding sessions (include/grand_session.h) over torch
device memory: the benchmark, the serving loop and graph replay use these."""
from __future__ import annotations

import ctypes

import numpy as np

from . import (GrandDecoderConfig, GrandDecoderExConfig, RESULT_DTYPE, load_library, KIND_CC, KIND_PROB_CC,
               KIND_CLN, KIND_PROB_CLN, KIND_ML_CLN)

_VP = ctypes.c_void_p


def bind_session(lib):
    if getattr(lib, "_grand_session_bound", False):
        return lib
    lib.grand_session_create.argtypes = [ctypes.c_int, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_char, ctypes.POINTER(GrandDecoderConfig), ctypes.c_char_p,
                                         ctypes.POINTER(GrandDecoderExConfig)]
    lib.grand_session_create.restype = _VP
    lib.grand_session_decode_device.argtypes = [_VP, _VP, ctypes.c_int, _VP, ctypes.c_int, _VP, _VP, _VP, _VP, _VP]
    lib.grand_session_decode_device.restype = ctypes.c_int
    lib.grand_session_decode_learned.argtypes = [_VP, _VP, ctypes.c_int, _VP, ctypes.c_int, _VP, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _VP, _VP, _VP, _VP,
                                                 _VP]
    lib.grand_session_decode_learned.restype = ctypes.c_int
    lib.grand_session_capture.argtypes = [_VP, _VP, ctypes.c_int, _VP, ctypes.c_int, _VP, _VP, _VP, _VP, _VP]
    lib.grand_session_capture.restype = _VP
    lib.grand_graph_launch.argtypes = [_VP, _VP]
    lib.grand_graph_launch.restype = ctypes.c_int
    lib.grand_graph_destroy.argtypes = [_VP]
    lib.grand_graph_destroy.restype = None
    lib.grand_session_error_bits.argtypes = [_VP, _VP]
    lib.grand_session_enable_timing.argtypes = [_VP, ctypes.c_int]
    lib.grand_session_last_timing.argtypes = [_VP, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.grand_session_timing_totals.argtypes = [_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_longlong)]
    lib.grand_session_timing_reset.argtypes = [_VP]
    lib.grand_session_set_work_counters.argtypes = [_VP, _VP]
    lib.grand_session_ms_per_tick.argtypes = [_VP]
    lib.grand_session_ms_per_tick.restype = ctypes.c_double
    lib.grand_session_pattern_words.argtypes = [_VP]
    lib.grand_session_logical_words.argtypes = [_VP]
    lib.grand_session_footprint.argtypes = [_VP, ctypes.c_int]
    lib.grand_session_footprint.restype = ctypes.c_size_t
    lib.grand_session_destroy.argtypes = [_VP]
    lib.grand_device_info.argtypes = [ctypes.POINTER(ctypes.c_int)] * 3
    lib.grand_set_tuning.argtypes = [ctypes.c_int] * 4
    lib.grand_get_tuning.argtypes = [ctypes.POINTER(ctypes.c_int)] * 4
    lib.grand_set_pair_table_kb.argtypes = [ctypes.c_int]
    lib.grand_get_pair_table_kb.argtypes = []
    lib.grand_get_pair_table_kb.restype = ctypes.c_int
    lib._grand_session_bound = True
    return lib


def device_info(lib):
    bind_session(lib)
    cu, clk, wclk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib.grand_device_info(ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(wclk)) != 0:
        raise RuntimeError("no GPU")
    return {"cus": cu.value, "clock_khz": clk.value, "wall_clock_khz": wclk.value}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Session:
    """One uploaded code; decode() takes torch CUDA tensors."""

    def __init__(self, kind, H, H2, cfg, n_logicos=0, type_of_error="x", file=None, ex=None, lib=None):
        self.lib = bind_session(lib or load_library())
        self.kind = kind
        H = np.ascontiguousarray(H, dtype=np.int32)
        H2 = np.ascontiguousarray(H2, dtype=np.int32) if H2 is not None else None
        self.n, self.m = H.shape
        self.n_logicos = n_logicos
        self._keep = (H, H2)
        self.h = self.lib.grand_session_create(
            kind, H.ctypes.data, H2.ctypes.data if H2 is not None else None, self.n, self.m, n_logicos,
            ctypes.c_char(type_of_error.encode()[:1]), ctypes.byref(cfg),
            file.encode() if file else None, ctypes.byref(ex) if ex is not None else None)
        if not self.h:
            raise ValueError("grand_session_create rejected the code/config (see stderr)")
        self.pattern_words = self.lib.grand_session_pattern_words(self.h)
        self.logical_words = self.lib.grand_session_logical_words(self.h)
        self.ms_per_tick = self.lib.grand_session_ms_per_tick(self.h)

    def decode(self, d_input, input_is_errors, d_results, d_logicals=None, d_decoded=None, d_decoded_logicals=None,
               d_stats=None, stream=None, num_shots=None):
        shots = num_shots if num_shots is not None else d_input.shape[0]
        st = None if stream is None else ctypes.c_void_p(stream)
        rc = self.lib.grand_session_decode_device(self.h, _ptr(d_input), int(bool(input_is_errors)), _ptr(d_logicals),
                                                  shots, _ptr(d_results), _ptr(d_decoded),
                                                  _ptr(d_decoded_logicals), _ptr(d_stats), st)
        if rc != 0:
            raise RuntimeError(f"grand_session_decode_device returned {rc}")

    def decode_learned(self, d_input, input_is_errors, d_results, d_scores, per_column, max_weight=6,
                       budget=100_000, d_logicals=None, d_decoded=None, d_decoded_logicals=None, d_stats=None,
                       stream=None, num_shots=None):
        """Learned pass (scores: float32 CUDA tensor [shots][stride]) before the plain search."""
        shots = num_shots if num_shots is not None else d_input.shape[0]
        st = None if stream is None else ctypes.c_void_p(stream)
        rc = self.lib.grand_session_decode_learned(self.h, _ptr(d_input), int(bool(input_is_errors)), _ptr(d_logicals),
                                                   shots, _ptr(d_scores), d_scores.shape[1], int(bool(per_column)),
                                                   max_weight, budget, _ptr(d_results), _ptr(d_decoded),
                                                   _ptr(d_decoded_logicals), _ptr(d_stats), st)
        if rc != 0:
            raise RuntimeError(f"grand_session_decode_learned returned {rc}")

    def capture(self, d_input, input_is_errors, d_results, stream, d_logicals=None, d_decoded=None,
                d_decoded_logicals=None, d_stats=None, num_shots=None):
        """HIP graph of one decode on `stream` (a non-default stream handle);
        launch() decodes the buffers' current contents."""
        shots = num_shots if num_shots is not None else d_input.shape[0]
        h = self.lib.grand_session_capture(self.h, _ptr(d_input), int(bool(input_is_errors)), _ptr(d_logicals), shots,
                                           _ptr(d_results), _ptr(d_decoded), _ptr(d_decoded_logicals), _ptr(d_stats),
                                           ctypes.c_void_p(stream))
        if not h:
            raise RuntimeError("grand_session_capture failed")
        return Graph(self.lib, h)

    def error_bits(self, stream=None):
        return self.lib.grand_session_error_bits(self.h, None if stream is None else ctypes.c_void_p(stream))

    def enable_timing(self, on=True):
        self.lib.grand_session_enable_timing(self.h, int(on))

    def last_timing(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        if self.lib.grand_session_last_timing(self.h, ctypes.byref(a), ctypes.byref(b)) != 0:
            raise RuntimeError("timing not enabled")
        return a.value, b.value

    def timing_totals(self):
        """(triage ms, search ms, decodes) summed over the timed decodes since the
        last timing_reset; waits for all of them."""
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        if self.lib.grand_session_timing_totals(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)) != 0:
            raise RuntimeError("timing not enabled")
        return a.value, b.value, n.value

    def timing_reset(self):
        if self.lib.grand_session_timing_reset(self.h) != 0:
            raise RuntimeError("timing reset failed")

    def set_work_counters(self, d_work):
        """d_work: int64 CUDA tensor of 3 (zeroed by the caller), or None to stop counting."""
        self.lib.grand_session_set_work_counters(self.h, _ptr(d_work))

    def footprint(self, shots):
        return self.lib.grand_session_footprint(self.h, shots)

    def close(self):
        if self.h:
            self.lib.grand_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured decode (grand_session_capture)."""

    def __init__(self, lib, h):
        self.lib, self.h = lib, h

    def launch(self, stream):
        if self.lib.grand_graph_launch(self.h, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("grand_graph_launch failed")

    def close(self):
        if self.h:
            self.lib.grand_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def results_from_bytes(buf: np.ndarray) -> np.ndarray:
    """uint8 buffer of 72-byte records -> structured array."""
    return np.frombuffer(np.ascontiguousarray(buf).tobytes(), dtype=RESULT_DTYPE)


__all__ = ["Session", "bind_session", "device_info", "results_from_bytes", "KIND_CC", "KIND_PROB_CC", "KIND_CLN",
           "KIND_PROB_CLN", "KIND_ML_CLN"]
