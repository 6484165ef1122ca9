"""Learned search order for the DEM family: an MLP scores the active positions.

    detector bits (216) -> 256 -> 256 -> scores, ReLU, BF16 weights, FP32 accumulation

The network runs on the project's own HIP kernels (grand_mlp.hip): BF16 MFMA
GEMMs with bias/ReLU epilogues for the forward pass, ReLU-masked and split-K
GEMMs for the backward pass, masked BCE, column sums and Adam.  torch only
provides device memory and random numbers.

Scores are either one per position of the shot's first active list (64
outputs, the list capacity: ``target="slot"``) or one per mechanism
(``target="column"``, m outputs, gathered at the active columns by the search
kernel).  Labels come from the plain decoder: a position is positive when the
plain search's decoded pattern contains it.

Decoding with the scores (Session.decode_learned) runs, per shot, a learned pass
over the first active list: positions in descending score order, patterns of
weight 1..8 tried so that a pattern comes as soon as its worst-scored position
does, every candidate checked exactly against the syndrome, at most ``budget``
patterns; then plain GRAND as usual.  Counts include both.

    python -m library_s_grand_amd.learned --train 2000000 --eval 1000000

No counterpart in the reference (grand_decoder.cu has no learned component).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import KIND_PROB_CLN, load_library, make_config
from . import codes as C

_VP = ctypes.c_void_p
A_MK, A_KM = 0, 1
B_NK, B_KN = 0, 1
EPI_STORE, EPI_MASK, EPI_ATOMIC = 0, 1, 2


def bind_mlp(lib):
    if getattr(lib, "_grand_mlp_bound", False):
        return lib
    i, f = ctypes.c_int, ctypes.c_float
    lib.grand_mlp_gemm.argtypes = [i, i, i, _VP, i, _VP, i, i, i, i, _VP, i, _VP, i, _VP, i, _VP, i, i, _VP]
    lib.grand_mlp_gemm.restype = ctypes.c_int
    lib.grand_mlp_bce.argtypes = [_VP, i, _VP, _VP, i, i, i, f, _VP, i, _VP, _VP]
    lib.grand_mlp_bce.restype = ctypes.c_int
    lib.grand_mlp_colsum.argtypes = [_VP, i, i, i, _VP, _VP]
    lib.grand_mlp_colsum.restype = ctypes.c_int
    lib.grand_mlp_adam.argtypes = [_VP, _VP, _VP, _VP, i, f, f, f, f, i, _VP, _VP]
    lib.grand_mlp_adam.restype = ctypes.c_int
    lib.grand_mlp_rows_to_bf16.argtypes = [_VP, i, i, _VP, i, _VP]
    lib.grand_mlp_rows_to_bf16.restype = ctypes.c_int
    lib._grand_mlp_bound = True
    return lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class MLP:
    """Fully connected ReLU network on the HIP kernels.  Weights: fp32 masters
    [out][in] plus their bf16 copies used by the GEMMs; Adam state in fp32."""

    def __init__(self, dims=(216, 256, 256, 64), seed=0, lib=None):
        import torch
        self.torch = torch
        self.lib = bind_mlp(lib or load_library())
        self.dims = tuple(int(d) for d in dims)
        for d in self.dims[:-1]:
            if d % 8:
                raise ValueError("hidden and input widths must be multiples of 8")
        g = torch.Generator().manual_seed(seed)
        dev = torch.device("cuda")
        self.W, self.b, self.Wb = [], [], []
        for din, dout in zip(self.dims[:-1], self.dims[1:]):
            bound = (6.0 / din) ** 0.5 if dout != self.dims[-1] else (1.0 / din) ** 0.5
            w = (torch.rand(dout, din, generator=g) * 2 - 1) * bound
            self.W.append(w.to(dev).contiguous())
            self.b.append(torch.zeros(dout, device=dev))
            self.Wb.append(self.W[-1].to(torch.bfloat16).contiguous())
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.steps = 0
        self.acts = None

    @property
    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def to_input(self, rows_i32):
        """int32 0/1 rows [B][dims[0]] (CUDA) -> bf16 rows."""
        B = rows_i32.shape[0]
        x = self.torch.empty(B, self.dims[0], dtype=self.torch.bfloat16, device=rows_i32.device)
        _check(self.lib.grand_mlp_rows_to_bf16(_p(rows_i32), rows_i32.shape[1], B, _p(x), self.dims[0], self.stream),
               "rows_to_bf16")
        return x

    def forward(self, x, keep=False):
        """x: bf16 [B][dims[0]] -> fp32 logits [B][dims[-1]]."""
        torch = self.torch
        B = x.shape[0]
        h = x
        acts = [x]
        L = len(self.W)
        out = None
        for l in range(L):
            dout, din = self.W[l].shape
            if l < L - 1:
                nxt = torch.empty(B, dout, dtype=torch.bfloat16, device=x.device)
                _check(self.lib.grand_mlp_gemm(A_MK, B_NK, EPI_STORE, _p(h), din, _p(self.Wb[l]), din, B, dout, din,
                                               None, 0, _p(nxt), dout, _p(self.b[l]), 1, None, 0, 1, self.stream),
                       "gemm fwd")
                h = nxt
                acts.append(h)
            else:
                out = torch.empty(B, dout, dtype=torch.float32, device=x.device)
                _check(self.lib.grand_mlp_gemm(A_MK, B_NK, EPI_STORE, _p(h), din, _p(self.Wb[l]), din, B, dout, din,
                                               _p(out), dout, None, 0, _p(self.b[l]), 0, None, 0, 1, self.stream),
                       "gemm fwd")
        self.acts = acts if keep else None
        return out

    def loss_grad(self, logits, labels_u8, mask_u8):
        """Masked BCE: returns (sum of losses, number of terms, dL/dz as bf16),
        the loss being averaged over the masked terms."""
        torch = self.torch
        B, N = logits.shape
        count = float(mask_u8.sum().item()) if mask_u8 is not None else float(B * N)
        dz = torch.empty(B, N, dtype=torch.bfloat16, device=logits.device)
        loss = torch.zeros(1, dtype=torch.float32, device=logits.device)
        _check(self.lib.grand_mlp_bce(_p(logits), N, _p(labels_u8), _p(mask_u8), N, B, N, 1.0 / max(count, 1.0),
                                      _p(dz), N, _p(loss), self.stream), "bce")
        return loss, count, dz

    def backward(self, dz):
        """Accumulates parameter gradients for dz = dL/dlogits (bf16 [B][out])
        using the activations kept by forward(keep=True)."""
        torch = self.torch
        acts = self.acts
        B = dz.shape[0]
        g = dz
        splits = max(1, min(64, B // 256))
        for l in reversed(range(len(self.W))):
            dout, din = self.W[l].shape
            h = acts[l]
            _check(self.lib.grand_mlp_colsum(_p(g), dout, B, dout, _p(self.gb[l]), self.stream), "colsum")
            _check(self.lib.grand_mlp_gemm(A_KM, B_KN, EPI_ATOMIC, _p(g), dout, _p(h), din, dout, din, B,
                                           _p(self.gW[l]), din, None, 0, None, 0, None, 0, splits, self.stream),
                   "gemm dW")
            if l > 0:
                gn = torch.empty(B, din, dtype=torch.bfloat16, device=dz.device)
                _check(self.lib.grand_mlp_gemm(A_MK, B_KN, EPI_MASK, _p(g), dout, _p(self.Wb[l]), din, B, din, dout,
                                               None, 0, _p(gn), din, None, 0, _p(h), din, 1, self.stream),
                       "gemm dX")
                g = gn

    def zero_grad(self):
        for t in self.gW + self.gb:
            t.zero_()

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.steps += 1
        for l in range(len(self.W)):
            _check(self.lib.grand_mlp_adam(_p(self.W[l]), _p(self.mW[l]), _p(self.vW[l]), _p(self.gW[l]),
                                           self.W[l].numel(), lr, b1, b2, eps, self.steps, _p(self.Wb[l]),
                                           self.stream), "adam")
            _check(self.lib.grand_mlp_adam(_p(self.b[l]), _p(self.mb[l]), _p(self.vb[l]), _p(self.gb[l]),
                                           self.b[l].numel(), lr, b1, b2, eps, self.steps, None, self.stream),
                   "adam")

    def predict(self, rows_i32, batch=65536):
        """Scores (fp32 [B][out]) for int32 0/1 input rows on the device."""
        outs = []
        for i in range(0, rows_i32.shape[0], batch):
            outs.append(self.forward(self.to_input(rows_i32[i:i + batch])))
        return self.torch.cat(outs) if outs else None

    def state(self):
        return {"dims": self.dims, "W": [w.cpu().numpy() for w in self.W], "b": [b.cpu().numpy() for b in self.b]}


# --------------------------------------------------------------------- data
class DemTask:
    """A DEM with its device copies: detectors and logical observables of
    sampled shots, plain-decoder labels and first-active-list positions."""

    def __init__(self, seed=11, n_det=216, n_mech=3000, max_size=64, W=4):
        import torch
        self.torch = torch
        self.H, self.L, self.p = C.synthetic_dem(seed, n_det=n_det, n_mech=n_mech)
        self.n, self.m = self.H.shape
        self.NL = self.L.shape[1]
        self.max_size = max_size
        self.W = W
        dev = torch.device("cuda")
        self.dH = torch.from_numpy(self.H.astype(np.float32)).to(dev)            # [n][m]
        self.dL = torch.from_numpy(self.L.astype(np.float32)).to(dev)            # [m][NL]
        self.dp = torch.from_numpy(self.p.astype(np.float32)).to(dev)
        self.colw = self.dH.sum(0)
        self.tmp = tempfile.mkdtemp(prefix="grand_learned_")
        self.phase_file = os.path.join(self.tmp, "phase.txt")
        C.write_phase_file(self.phase_file, C.DEFAULT_PHASE_MASKS)
        self.cfg = make_config(max_error_weight=W, max_size_active_positions=max_size)

    def sample(self, shots, seed, chunk=100_000):
        """Bernoulli mechanisms on the device: detectors and observables (int32)."""
        torch = self.torch
        g = torch.Generator(device="cuda").manual_seed(seed)
        dets, lgs = [], []
        for i in range(0, shots, chunk):
            k = min(chunk, shots - i)
            e = (torch.rand(k, self.m, generator=g, device="cuda") < self.dp).float()
            dets.append(torch.remainder(e @ self.dH.T, 2).to(torch.int32))
            lgs.append(torch.remainder(e @ self.dL, 2).to(torch.int32))
        return torch.cat(dets), torch.cat(lgs)

    def first_active(self, det):
        """Boolean [B][m] first active list (all of a column's detectors lit)
        and its length per shot (grand_decoder.cu:406-416, :495-504 at thr =
        max_soma for an irregular code)."""
        torch = self.torch
        score = det.float() @ self.dH
        act = (score >= self.colw) & (self.colw > 0)
        return act, act.sum(1)

    def session(self):
        from .session import Session
        return Session(KIND_PROB_CLN, self.H, self.L, self.cfg, n_logicos=self.NL, file=self.phase_file)

    def decode(self, sess, det, lg, scores=None, per_column=False, budget=100_000, max_weight=8, chunk=200_000):
        """Plain (scores None) or learned decode on the device; returns the
        result records (numpy), decoded pattern bits (torch bool [B][m]) and
        decoded logicals (torch int32 [B][NL])."""
        from .session import results_from_bytes
        torch = self.torch
        B = det.shape[0]
        mw = sess.pattern_words
        lw = sess.logical_words
        res = torch.empty(B * 72, dtype=torch.uint8, device="cuda")
        dec = torch.empty(B, mw, dtype=torch.int32, device="cuda")
        dlg = torch.empty(B, lw, dtype=torch.int32, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        for i in range(0, B, chunk):
            k = min(chunk, B - i)
            args = dict(d_logicals=lg[i:], d_decoded=dec[i:], d_decoded_logicals=dlg[i:], stream=stream,
                        num_shots=k)
            if scores is None:
                sess.decode(det[i:], False, res[i * 72:], **args)
            else:
                sess.decode_learned(det[i:], False, res[i * 72:], scores[i:], per_column, max_weight=max_weight,
                                    budget=budget, **args)
            bits_raised = sess.error_bits(stream)
            if bits_raised != 0:
                raise RuntimeError(f"decoder error bits {bits_raised} in shots [{i}, {i + k})")
        r = results_from_bytes(res.cpu().numpy())
        bits = ((dec.view(torch.int32).unsqueeze(2) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1)
        bits = bits.reshape(B, mw * 32)[:, :self.m].bool()
        lbits = ((dlg.unsqueeze(2) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1)
        lbits = lbits.reshape(B, lw * 32)[:, :self.NL].to(torch.int32)
        return r, bits, lbits

    def labels(self, det, dec_bits, found, target, chunk=200_000):
        """Targets and masks for the network from the plain decoder's patterns.
        slot: [B][max_size], position j of the first active list; column: [B][m].
        Shots the plain decoder did not solve, or whose first list is empty or
        longer than max_size, are masked out."""
        torch = self.torch
        ys, ms = [], []
        for i in range(0, det.shape[0], chunk):
            y, m = self._labels(det[i:i + chunk], dec_bits[i:i + chunk], found[i:i + chunk], target)
            ys.append(y)
            ms.append(m)
        return torch.cat(ys), torch.cat(ms)

    def _labels(self, det, dec_bits, found, target):
        torch = self.torch
        act, tam = self.first_active(det)
        usable = (torch.from_numpy(found.astype(np.bool_)).to(det.device)) & (tam > 0) & (tam <= self.max_size)
        if target == "column":
            y = (dec_bits & act).to(torch.uint8)
            mask = (act & usable[:, None]).to(torch.uint8)
            return y, mask
        B = det.shape[0]
        slot = torch.cumsum(act.int(), 1) - 1                                   # position of each active column
        y = torch.zeros(B, self.max_size, dtype=torch.uint8, device=det.device)
        mask = torch.zeros_like(y)
        rows, cols = torch.nonzero(act & usable[:, None], as_tuple=True)
        s = slot[rows, cols]
        keep = s < self.max_size
        rows, cols, s = rows[keep], cols[keep], s[keep]
        y[rows, s] = dec_bits[rows, cols].to(torch.uint8)
        mask[rows, s] = 1
        return y, mask


def logical_error_rate(res, lbits, lg, det):
    """Fraction of shots whose decoded logical flips differ from the observed
    ones; a shot with a nonzero syndrome and no pattern found counts as an
    error.  (The reference's `degenerate` flag is 0 for zero-syndrome shots, so
    it is not used: grand_decoder.cu:1489.)"""
    import torch
    wrong = (lbits != lg).any(1).cpu().numpy()
    nz = det.any(1).cpu().numpy()
    unsolved = nz & (res["found"] == 0)
    return float((wrong | unsolved).mean())


def train(task, model, det, y, mask, epochs=4, batch=4096, lr=2e-3, seed=0, log=None):
    torch = task.torch
    B = det.shape[0]
    g = torch.Generator(device="cuda").manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(B, generator=g, device="cuda")
        tot, cnt = 0.0, 0.0
        for i in range(0, B - batch + 1, batch):
            idx = perm[i:i + batch]
            x = model.to_input(det[idx].contiguous())
            z = model.forward(x, keep=True)
            loss, c, dz = model.loss_grad(z, y[idx].contiguous(), mask[idx].contiguous())
            model.zero_grad()
            model.backward(dz)
            model.step(lr * (0.5 ** ep))
            tot += float(loss.item())
            cnt += c
        hist.append(tot / max(cnt, 1.0))
        if log:
            log(f"epoch {ep}: mean BCE {hist[-1]:.5f}")
    return hist


def experiment(train_shots=2_000_000, eval_shots=1_000_000, heads=("column", "slot"), epochs=4, budget=100_000,
               max_weights=(5, 6, 8), seed=2024, log=print):
    """Trains each head on plain-decoder pairs and compares plain vs learned
    decoding on held-out shots: mean patterns per shot, logical error rate and
    found rate, with the learned pass limited to weights 1..max_weight for each
    of max_weights.  Plain grand_decode_prob_cln with this phase file searches
    weights 1..5, so max_weight 5 is the like-for-like row; 6 and 8 let the pass
    try heavier patterns than plain ever does (a wider search, not only a
    reordering)."""
    import torch
    t0 = time.time()
    task = DemTask()
    sess = task.session()
    det_tr, lg_tr = task.sample(train_shots, seed)
    r_tr, bits_tr, _ = task.decode(sess, det_tr, lg_tr)
    det_ev, lg_ev = task.sample(eval_shots, seed + 1)
    t_data = time.time() - t0
    t2 = time.time()
    r_pl, _, lb_pl = task.decode(sess, det_ev, lg_ev)
    torch.cuda.synchronize()
    t_plain = time.time() - t2
    cp = r_pl["total_tested_combinations"].astype(np.float64)
    ler_pl = logical_error_rate(r_pl, lb_pl, lg_ev, det_ev)
    rows, hists, t_train, t_learned = [], {}, 0.0, 0.0
    for head in heads:
        y, mask = task.labels(det_tr, bits_tr, r_tr["found"], head)
        out_dim = task.m if head == "column" else task.max_size
        model = MLP((task.n, 256, 256, out_dim), seed=seed)
        t1 = time.time()
        hists[head] = train(task, model, det_tr, y, mask, epochs=epochs, log=log)
        torch.cuda.synchronize()
        t_train += time.time() - t1
        del y, mask
        scores = model.predict(det_ev)
        for mw in max_weights:
            t3 = time.time()
            r_le, _, lb_le = task.decode(sess, det_ev, lg_ev, scores=scores, per_column=(head == "column"),
                                         budget=budget, max_weight=mw)
            torch.cuda.synchronize()
            t_learned += time.time() - t3
            cl = r_le["total_tested_combinations"].astype(np.float64)
            rows.append({"head": head, "model": list(model.dims), "learned_max_weight": mw,
                         "beyond_plain_weights": mw > 5,
                         "mean_patterns": float(cl.mean()),
                         "pattern_reduction": float(cp.mean() / max(cl.mean(), 1e-9)),
                         "median_patterns_nonzero": float(np.median(cl[cl > 0])) if (cl > 0).any() else 0.0,
                         "ler": logical_error_rate(r_le, lb_le, lg_ev, det_ev),
                         "found": float(r_le["found"].mean())})
        del scores, model
    best = max((r for r in rows if not r["beyond_plain_weights"]), key=lambda r: r["pattern_reduction"], default=None)
    out = {
        "dem": "synthetic 216 x 3000 (6 rounds x 36 detectors, column weight 1..4), 12 logicals, "
               "p log-uniform [1e-4, 5e-3]",
        "entry_point": "grand_decode_prob_cln", "max_error_weight": task.W, "max_size_active_positions": task.max_size,
        "plain_searched_weights": "1..5 per iteration (phase 0 of the phase file), up to 3 active-set iterations",
        "train_shots": train_shots, "eval_shots": eval_shots, "epochs": epochs, "train_bce": hists,
        "budget": budget,
        "plain": {"mean_patterns": float(cp.mean()),
                  "median_patterns_nonzero": float(np.median(cp[cp > 0])) if (cp > 0).any() else 0.0,
                  "ler": ler_pl, "found": float(r_pl["found"].mean())},
        "rows": rows,
        "like_for_like": best,
        "seconds": {"data_and_labels": t_data, "train": t_train, "plain_decode": t_plain,
                    "learned_infer_and_decode": t_learned},
    }
    sess.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2_000_000)
    ap.add_argument("--eval", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--heads", default="column,slot", help="comma-separated: column (one score per mechanism), "
                                                          "slot (one score per active-list position)")
    ap.add_argument("--budget", type=int, default=100_000)
    ap.add_argument("--max-weights", default="5,6,8")
    a = ap.parse_args()
    log = lambda msg: print(msg, file=sys.stderr, flush=True)   # noqa: E731
    heads = tuple(h for h in a.heads.split(",") if h)
    mws = tuple(int(x) for x in a.max_weights.split(",") if x)
    print(json.dumps(experiment(a.train, a.eval, heads, a.epochs, a.budget, mws, log=log)), flush=True)


if __name__ == "__main__":
    main()
