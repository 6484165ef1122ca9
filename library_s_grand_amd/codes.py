"""Seeded generators for the decoder's inputs, in the reference's on-disk formats.

* Bivariate bicycle (BB) CSS codes: x = S_l (x) I_m, y = I_l (x) S_m with S_k the
  k x k cyclic shift, Hx = [A | B], Hz = [B^T | A^T].  BB72, BB144 and BB288 as
  named in the reference's data set (H_Matrix/hz_ibm_{72,144,288}.txt).
* A synthetic circuit-level detector error model (DEM): detectors grouped in
  rounds of 36, every mechanism flips 1..4 detectors inside one window, one of
  the logicals with probability 0.1.
* Shot samplers (iid Bernoulli errors / mechanisms) and the text formats read by
  the reference: 0/1 matrices as whitespace-separated ints per row
  (Examples/auxilary_functions.py:109-121), error/syndrome rows as bitstrings or
  int rows (auxilary_functions.py:44-73), phase-order mask lines
  "tam_upper order0 order1 order2 p0 p1 p2" (grand_decoder.cu:728-752), and
  LLR files of m floats (grand_decoder.cu:695-713).
"""
from __future__ import annotations

import numpy as np


def _shift(k: int) -> np.ndarray:
    return np.roll(np.eye(k, dtype=np.int64), 1, axis=1)


def _poly(terms, l: int, m: int) -> np.ndarray:
    """Sum over GF(2) of monomials given as ('x'|'y', power)."""
    x = np.kron(_shift(l), np.eye(m, dtype=np.int64))
    y = np.kron(np.eye(l, dtype=np.int64), _shift(m))
    out = np.zeros((l * m, l * m), dtype=np.int64)
    for var, p in terms:
        base = x if var == "x" else y
        out = (out + np.linalg.matrix_power(base, p)) % 2
    return out


def bb_code(l: int, m: int, a_terms, b_terms):
    """Return (Hx, Hz) int32 matrices of shape (l*m, 2*l*m)."""
    A = _poly(a_terms, l, m)
    B = _poly(b_terms, l, m)
    hx = np.concatenate([A, B], axis=1) % 2
    hz = np.concatenate([B.T, A.T], axis=1) % 2
    return hx.astype(np.int32), hz.astype(np.int32)


BB_CODES = {
    72: (6, 6, [("x", 3), ("y", 1), ("y", 2)], [("y", 3), ("x", 1), ("x", 2)]),
    144: (12, 6, [("x", 3), ("y", 1), ("y", 2)], [("y", 3), ("x", 1), ("x", 2)]),
    288: (12, 12, [("x", 3), ("y", 2), ("y", 7)], [("y", 3), ("x", 1), ("x", 2)]),
}


def bb(n_qubits: int):
    """BB72 [[72,12,6]], BB144 [[144,12,12]] or BB288 [[288,12,18]] as (Hx, Hz)."""
    l, m, a, b = BB_CODES[n_qubits]
    return bb_code(l, m, a, b)


def matrix_to_text(H: np.ndarray) -> str:
    """One row per line, entries separated by single spaces (H_Matrix/*.txt)."""
    return "".join(" ".join(str(int(v)) for v in row) + "\n" for row in H)


def read_matrix(path: str) -> np.ndarray:
    """Whitespace-separated 0/1 ints per row (auxilary_functions.py:109-121)."""
    rows = []
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            line = line.strip()
            if line:
                rows.append([int(x) for x in line.split()])
    return np.array(rows, dtype=np.int32)


def read_rows(path: str, n_cols: int, max_rows: int | None = None) -> np.ndarray:
    """Rows given either as n_cols ints or as one n_cols-character bitstring
    (auxilary_functions.py:44-73)."""
    out = []
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            parts = line.strip().split()
            if not parts:
                continue
            if len(parts) == n_cols:
                out.append([int(x) for x in parts])
            elif len(parts) == 1 and len(parts[0]) == n_cols:
                out.append([int(ch) for ch in parts[0]])
            else:
                raise ValueError(f"row has {len(parts)} tokens, expected {n_cols}")
            if max_rows is not None and len(out) >= max_rows:
                break
    return np.array(out, dtype=np.int32).reshape(-1, n_cols)


def sample_errors(rng: np.random.Generator, shots: int, p, n_pos: int | None = None) -> np.ndarray:
    """iid Bernoulli(p) errors, int32 shots x n_pos (p scalar, then n_pos is
    required, or one probability per position)."""
    p = np.asarray(p, dtype=np.float64)
    width = p.shape[-1] if p.ndim else n_pos
    return (rng.random((shots, width)) < p).astype(np.int32)


def syndromes(H: np.ndarray, errors: np.ndarray) -> np.ndarray:
    return ((errors.astype(np.int64) @ H.T.astype(np.int64)) % 2).astype(np.int32)


def log_uniform(rng: np.random.Generator, lo: float, hi: float, size) -> np.ndarray:
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def synthetic_dem(seed: int, n_det: int = 216, n_mech: int = 3000, window: int = 36, n_log: int = 12,
                  max_colw: int = 4, p_lo: float = 1e-4, p_hi: float = 5e-3, p_log: float = 0.1):
    """Synthetic circuit-level DEM: (H n_det x n_mech, L n_mech x n_log, p n_mech).

    Mechanism j lives in window j * n_windows // n_mech and flips a uniform
    1..max_colw distinct detectors inside it; it flips one random logical with
    probability p_log.  Probabilities are log-uniform in [p_lo, p_hi]."""
    rng = np.random.default_rng(seed)
    n_win = n_det // window
    H = np.zeros((n_det, n_mech), dtype=np.int32)
    for j in range(n_mech):
        w = (j * n_win) // n_mech
        k = int(rng.integers(1, max_colw + 1))
        rows = rng.choice(window, size=k, replace=False) + w * window
        H[rows, j] = 1
    L = np.zeros((n_mech, n_log), dtype=np.int32)
    hit = rng.random(n_mech) < p_log
    L[np.nonzero(hit)[0], rng.integers(0, n_log, size=int(hit.sum()))] = 1
    p = log_uniform(rng, p_lo, p_hi, n_mech)
    return H, L, p


def llr_from_p(p) -> np.ndarray:
    """LLR = log((1 - p) / p) in float32."""
    p = np.asarray(p, dtype=np.float64)
    return np.log((1.0 - p) / p).astype(np.float32)


def write_llr_file(path: str, llr) -> None:
    """m floats, one per line, printed with enough digits to round-trip fp32."""
    with open(path, "w", encoding="utf-8") as f:
        for v in np.asarray(llr, dtype=np.float32):
            f.write(f"{float(v):.9g}\n")


def write_phase_file(path: str, masks) -> None:
    """masks: iterable of (tam_upper, o0, o1, o2, p0, p1, p2)."""
    with open(path, "w", encoding="utf-8") as f:
        for t, o0, o1, o2, p0, p1, p2 in masks:
            f.write(f"{int(t)} {int(o0)} {int(o1)} {int(o2)} {p0:.6f} {p1:.6f} {p2:.6f}\n")


# Single-mask file searching phase 0 (weights 1..5) for every tam.
DEFAULT_PHASE_MASKS = [(300, 0, 1, 2, 1.0, 0.0, 0.0)]
