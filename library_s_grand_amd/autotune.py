"""Autotuner of the persistent search kernel.

Knobs (grand_set_tuning, process-wide defaults used by every session that does
not override them in its GrandDecoderExConfig):

  waves_per_cu      persistent grid = CUs x waves_per_cu (residency is capped by
                    the kernel's registers and LDS: 2 waves per SIMD today)
  claim_batch       worklist shots taken per atomic claim (the kernel lowers it
                    for short worklists)
  rmax              prefixes per lane per chunk (chunk = 64 x rmax prefixes)
  split_threshold   weight classes with more prefixes than this are shared
                    between waves through the task table
  pair_table_kb     LDS packing (first-hit kinds): KiB of a wave's LDS share
                    given to the pair table of deep classes (0 = all that is
                    left, -1 = none); a smaller table is cheaper to build per
                    class and leaves more classes to the prefix walk

The tuner measures the search kernel with HIP events on the caller's workload
(device-resident inputs, a session with timing enabled) and walks the knobs one
at a time from the current defaults, keeping a change only when it is faster by
more than the noise margin; it repeats the sweep until nothing improves.  The
result can be saved as JSON and applied again (apply()).
"""
from __future__ import annotations

import ctypes
import json

import numpy as np

from .session import bind_session

SPACE = {
    "waves_per_cu": (1, 2, 4, 8, 16),
    "claim_batch": (1, 2, 4, 8),
    "rmax": (8, 16, 32, 64),
    "split_threshold": (256, 1024, 2048, 8192, 32768, 65536),
    "pair_table_kb": (0, 8, 16, -1),
}
KEYS = ("waves_per_cu", "claim_batch", "rmax", "split_threshold", "pair_table_kb")


def get(lib):
    bind_session(lib)
    v = [ctypes.c_int() for _ in KEYS[:4]]
    lib.grand_get_tuning(*[ctypes.byref(x) for x in v])
    cfg = {k: x.value for k, x in zip(KEYS, v)}
    cfg["pair_table_kb"] = int(lib.grand_get_pair_table_kb())
    return cfg


def reset(lib):
    """Restores the library's built-in defaults (one set for the first-hit kinds, one for ML)."""
    bind_session(lib)
    lib.grand_set_tuning(-1, -1, -1, -1)


def apply(lib, cfg):
    bind_session(lib)
    lib.grand_set_tuning(*[int(cfg[k]) for k in KEYS[:4]])
    if "pair_table_kb" in cfg:
        lib.grand_set_pair_table_kb(int(cfg["pair_table_kb"]))


def measure(run, reps=5):
    """Median search-kernel milliseconds of `run()` (which decodes once and
    returns (triage_ms, search_ms))."""
    run()   # warm the configuration
    return float(np.median([run()[1] for _ in range(reps)]))


def tune(lib, run, reps=5, margin=0.02, max_rounds=3, space=SPACE, log=None):
    """Coordinate search from the current defaults; returns (best config,
    its search ms, every measurement)."""
    best = get(lib)
    apply(lib, best)
    best_ms = measure(run, reps)
    table = [dict(best, search_ms=best_ms)]
    for _ in range(max_rounds):
        improved = False
        for key in KEYS:
            for val in space.get(key, ()):
                if val == best[key]:
                    continue
                cand = dict(best, **{key: val})
                apply(lib, cand)
                ms = measure(run, reps)
                table.append(dict(cand, search_ms=ms))
                if log:
                    log(f"{cand} -> {ms:.3f} ms")
                if ms < best_ms * (1.0 - margin):
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    apply(lib, best)
    return best, best_ms, table


def save(path, cfg):
    with open(path, "w", encoding="utf-8") as f:
        json.dump({k: int(cfg[k]) for k in KEYS if k in cfg}, f)


def load(path):
    with open(path, encoding="utf-8") as f:
        return json.load(f)
