"""Serving loop: continuous batching of decode requests over resident codes.

The server itself is native (``grand_server_*`` in libgranddecoder.so, declared
in include/grand_server.h): a worker thread keeps per-code lanes busy, each lane
being a device session with its own stream, reserved workspace and pinned
staging.  This module binds it and provides the open-loop driver used to
measure it: requests of several codes arrive as a Poisson process, each with a
log-uniform number of shots, and every request's latency (submit to results
copied out) is recorded.

    python -m library_s_grand_amd.serving --rate 2000 --requests 4000

The reference has no serving path; it decodes one synchronous batch per call
(grand_decoder.cu:790-2277).  Results of a request are those of decoding its
rows alone (shots are independent), which the tests check against the ABI.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import GrandDecoderConfig, KIND_CC, RESULT_DTYPE, load_library, make_config
from . import codes as C

REQ_DONE, REQ_FAILED, REQ_PENDING = 0, -1, 1


class GrandServerConfig(ctypes.Structure):
    _fields_ = [("max_batch_shots", ctypes.c_int), ("lanes_per_code", ctypes.c_int),
                ("max_queued_shots", ctypes.c_int), ("reserved", ctypes.c_int),
                ("hbm_budget_bytes", ctypes.c_longlong)]


def bind_server(lib):
    IP = ctypes.POINTER(ctypes.c_int)
    lib.grand_server_create.restype = ctypes.c_void_p
    lib.grand_server_create.argtypes = [ctypes.POINTER(GrandServerConfig)]
    lib.grand_server_add_code.restype = ctypes.c_int
    lib.grand_server_add_code.argtypes = [ctypes.c_void_p, ctypes.c_int, IP, IP, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_char, ctypes.POINTER(GrandDecoderConfig),
                                          ctypes.c_char_p]
    lib.grand_server_footprint.restype = ctypes.c_longlong
    lib.grand_server_footprint.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.grand_server_submit.restype = ctypes.c_longlong
    lib.grand_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.grand_server_poll.restype = ctypes.c_int
    lib.grand_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
    lib.grand_server_wait.restype = ctypes.c_int
    lib.grand_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
    lib.grand_server_counters.restype = None
    lib.grand_server_counters.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_longlong)] * 3
    lib.grand_server_destroy.restype = None
    lib.grand_server_destroy.argtypes = [ctypes.c_void_p]
    return lib


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Server:
    """Owns a native server.  Arrays handed to submit() are kept alive by the
    server object until their request is waited for."""

    def __init__(self, max_batch_shots=16384, lanes_per_code=2, max_queued_shots=1 << 20, hbm_budget_bytes=0,
                 lib=None):
        self.lib = bind_server(lib or load_library())
        cfg = GrandServerConfig(max_batch_shots, lanes_per_code, max_queued_shots, 0, hbm_budget_bytes)
        self.h = self.lib.grand_server_create(ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("grand_server_create failed")
        self._live = {}
        self._codes = []

    def add_code(self, kind, H, H2, cfg: GrandDecoderConfig, n_logicos=0, type_of_error="x", file=None):
        """Registers a code; returns its id.  Raises MemoryError when the HBM
        budget refuses it, ValueError for invalid input."""
        H = np.ascontiguousarray(H, dtype=np.int32)
        H2 = None if H2 is None else np.ascontiguousarray(H2, dtype=np.int32)
        n, m = H.shape
        rc = self.lib.grand_server_add_code(self.h, kind, _ip(H), _ip(H2), n, m, n_logicos,
                                            ctypes.c_char(type_of_error.encode()), ctypes.byref(cfg),
                                            file.encode() if file else None)
        if rc == -2:
            raise MemoryError("code refused by the server's HBM budget")
        if rc < 0:
            raise ValueError("invalid code")
        self._codes.append(dict(kind=kind, n=n, m=m, n_logicos=n_logicos))
        return rc

    def footprint(self, code_id=-1):
        return int(self.lib.grand_server_footprint(self.h, code_id))

    def submit(self, code_id, rows, input_is_errors=True, logicals=None, want_decoded=False):
        """Queues a request; returns its ticket.  Raises BlockingIOError when
        the queue is full."""
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        lg = None if logicals is None else np.ascontiguousarray(logicals, dtype=np.int32)
        shots = rows.shape[0]
        res = np.empty(shots, dtype=RESULT_DTYPE)          # the native records, written in place
        dec = np.zeros((shots, self._codes[code_id]["m"]), np.int32) if want_decoded else None
        t = self.lib.grand_server_submit(self.h, code_id, rows.ctypes.data, int(bool(input_is_errors)),
                                         None if lg is None else lg.ctypes.data, shots, res.ctypes.data,
                                         None if dec is None else dec.ctypes.data)
        if t == -2:
            raise BlockingIOError("server queue full")
        if t < 0:
            raise ValueError("request refused")
        self._live[t] = (rows, lg, res, dec)
        return t

    def wait(self, ticket):
        """Returns (results structured array, decoded or None, latency_ms)."""
        lat = ctypes.c_double(0.0)
        st = self.lib.grand_server_wait(self.h, ticket, ctypes.byref(lat))
        rows, lg, res, dec = self._live.pop(ticket)
        if st != REQ_DONE:
            raise RuntimeError(f"request {ticket} failed")
        return res, dec, lat.value

    def counters(self):
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self.lib.grand_server_counters(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return dict(requests=a.value, shots=b.value, batches=c.value)

    def close(self):
        if self.h:
            self.lib.grand_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ------------------------------------------------------------------ workload
@dataclass
class Request:
    t: float              # arrival time (s from start)
    code: int             # index into the workload's code list
    errors: np.ndarray    # shots x m int32


def mixed_bb_workload(n_requests, rate_hz, seed=7, code_sizes=(72, 144, 288), max_shots=4096,
                      p_range=(0.002, 0.008)):
    """Poisson arrivals over BB codes; each request picks a code uniformly, a
    shot count log-uniform in [1, max_shots] and a physical error rate
    uniform in p_range (i.i.d. X errors on every qubit)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, n_requests)
    t = np.cumsum(gaps) - gaps[0]
    m_of = {q: q for q in code_sizes}
    reqs = []
    for i in range(n_requests):
        ci = int(rng.integers(len(code_sizes)))
        shots = int(np.clip(np.round(np.exp(rng.uniform(0.0, np.log(max_shots)))), 1, max_shots))
        p = float(rng.uniform(*p_range))
        e = (rng.random((shots, m_of[code_sizes[ci]])) < p).astype(np.int32)
        reqs.append(Request(float(t[i]), ci, e))
    return reqs


def serve_open_loop(server, code_ids, reqs, want_decoded=False):
    """Submits each request at its arrival time from a feeder thread and waits
    for all of them.  Returns per-request (ticket, latency_ms, results, decoded)
    in submission order and the wall time of the run."""
    tickets = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]
    err = []

    def feeder():
        t0 = time.perf_counter()
        try:
            for i, r in enumerate(reqs):
                delay = r.t - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                while True:
                    try:
                        tickets[i] = server.submit(code_ids[r.code], r.errors, True, want_decoded=want_decoded)
                        break
                    except BlockingIOError:
                        time.sleep(1e-4)
                ready[i].set()
        except Exception as e:   # surfaced to the caller below
            err.append(e)
            for ev in ready:
                ev.set()

    start = time.perf_counter()
    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    out = []
    for i in range(len(reqs)):
        ready[i].wait()
        if err:
            break
        res, dec, lat = server.wait(tickets[i])
        out.append((tickets[i], lat, res, dec))
    th.join()
    if err:
        raise err[0]
    return out, time.perf_counter() - start


def setup_bb_server(code_sizes=(72, 144, 288), W=4, server=None, **kw):
    srv = server or Server(**kw)
    cfg = make_config(max_error_weight=W)
    ids = []
    for q in code_sizes:
        hx, hz = C.bb(q)
        ids.append(srv.add_code(KIND_CC, hx, hz, cfg, type_of_error="x"))
    return srv, ids


def _summary(reqs, out, wall, rate_hz, before, after, fp, lanes_per_code):
    lat = np.array([o[1] for o in out]) * 1e3
    shots = sum(r.errors.shape[0] for r in reqs)
    batches = after["batches"] - before["batches"]
    return {
        "workload": f"BB72/BB144/BB288 grand_decode_cc W=4, {len(reqs)} requests, Poisson {rate_hz:g}/s, "
                    f"shots per request log-uniform [1,4096], p ~ U[0.002,0.008]",
        "offered_requests_per_s": rate_hz, "requests": len(reqs), "shots": shots, "wall_s": wall,
        "offered_shots_per_s": shots / reqs[-1].t if reqs[-1].t > 0 else None,
        "served_shots_per_s": shots / wall, "requests_per_s": len(reqs) / wall,
        "p50_us_per_request": float(np.percentile(lat, 50)), "p99_us_per_request": float(np.percentile(lat, 99)),
        "max_us_per_request": float(lat.max()),
        "batches": batches, "mean_requests_per_batch": len(reqs) / max(batches, 1),
        "lanes_per_code": lanes_per_code, "hbm_footprint_bytes": fp,
    }


def measure(rate_hz=2000.0, n_requests=4000, seed=7, lanes_per_code=2, max_batch_shots=16384):
    """Open-loop run of the mixed BB72/BB144/BB288 workload; returns the
    summary dict (latency percentiles over requests, throughput)."""
    return sweep((rate_hz,), n_requests, seed, lanes_per_code, max_batch_shots)[0]


def sweep(rates, n_requests=4000, seed=7, lanes_per_code=2, max_batch_shots=16384):
    """The same request mix (one set of requests: codes, shot counts, errors)
    offered at each Poisson rate in turn on one server; one summary per rate.
    Past the rate the decoder sustains, served < offered and requests queue:
    batches then hold several requests each and the latency grows."""
    base = mixed_bb_workload(n_requests, 1.0, seed)
    srv, ids = setup_bb_server(lanes_per_code=lanes_per_code, max_batch_shots=max_batch_shots)
    out_rows = []
    try:
        # warm each lane's kernels once
        for cid, q in zip(ids, (72, 144, 288)):
            srv.wait(srv.submit(cid, np.zeros((1, q), np.int32)))
        fp = srv.footprint()
        for rate in rates:
            reqs = [Request(r.t / rate, r.code, r.errors) for r in base]
            before = srv.counters()
            out, wall = serve_open_loop(srv, ids, reqs)
            after = srv.counters()
            out_rows.append(_summary(reqs, out, wall, rate, before, after, fp, lanes_per_code))
    finally:
        srv.close()
    return out_rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--requests", type=int, default=4000)
    ap.add_argument("--lanes", type=int, default=2)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--sweep", default="", help="comma-separated request rates (per second) offered in turn")
    a = ap.parse_args()
    if a.sweep:
        rates = [float(x) for x in a.sweep.split(",") if x]
        for row in sweep(rates, a.requests, a.seed, a.lanes):
            print(json.dumps(row), flush=True)
    else:
        print(json.dumps(measure(a.rate, a.requests, a.seed, a.lanes)))


if __name__ == "__main__":
    main()
