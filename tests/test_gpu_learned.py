"""Learned search order on the GPU.

* the BF16 MFMA GEMM, in every operand layout and epilogue the MLP uses, against
  an fp32 matmul of the same bf16 values;
* the MLP's forward pass, loss and every parameter gradient against the CPU
  fp32 reference (tests/mlp_ref.py) with the same bf16 rounding points;
* the learned pass, shot by shot, against its CPU statement (tests/learned_ref.py)
  with the oracle's plain results after the budget;
* budget 0 gives exactly the plain results;
* training on plain-decoder pairs lowers the loss, and the trained order needs
  fewer patterns than plain decoding at no higher logical error rate.
"""
import ctypes

import numpy as np
import pytest

import library_s_grand_amd as G
from library_s_grand_amd import learned
import grandtest as T
import learned_ref
import mlp_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _bf(x):
    return torch.from_numpy(mlp_ref.bf16(x)).to(torch.bfloat16).cuda()


def _gemm(lib, la, lb, epi, A, lda, B, ldb, M, N, K, C=None, ldc=0, Cb=None, ldcb=0, bias=None, relu=0, mask=None,
          ldm=0, splits=1):
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.grand_mlp_gemm(la, lb, epi, learned._p(A), lda, learned._p(B), ldb, M, N, K, learned._p(C), ldc,
                            learned._p(Cb), ldcb, learned._p(bias), relu, learned._p(mask), ldm, splits, st)
    assert rc == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("M,N,K", [(100, 70, 50), (257, 256, 216), (64, 3000, 256)])
def test_gemm_forward_layout(M, N, K):
    lib = learned.bind_mlp(G.load_library())
    rng = np.random.default_rng(M + N + K)
    a = mlp_ref.bf16(rng.standard_normal((M, K)).astype(np.float32))
    w = mlp_ref.bf16(rng.standard_normal((N, K)).astype(np.float32))
    bias = rng.standard_normal(N).astype(np.float32)
    C = torch.zeros(M, N, device="cuda")
    Cb = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda")
    _gemm(lib, 0, 0, 0, _bf(a), K, _bf(w), K, M, N, K, C=C, ldc=N, Cb=Cb, ldcb=N,
          bias=torch.from_numpy(bias).cuda(), relu=1)
    ref = np.maximum(a @ w.T + bias, 0)
    got = C.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())
    gotb = Cb.float().cpu().numpy()
    assert np.abs(gotb - ref).max() <= 2 ** -7 * max(1.0, np.abs(ref).max())


def test_gemm_backward_layouts():
    lib = learned.bind_mlp(G.load_library())
    rng = np.random.default_rng(5)
    B, dout, din = 300, 64, 256
    g = mlp_ref.bf16(rng.standard_normal((B, dout)).astype(np.float32))
    h = mlp_ref.bf16(np.maximum(rng.standard_normal((B, din)), 0).astype(np.float32))
    w = mlp_ref.bf16(rng.standard_normal((dout, din)).astype(np.float32))
    # dX = (g W) * [h > 0]
    out = torch.zeros(B, din, dtype=torch.bfloat16, device="cuda")
    _gemm(lib, 0, 1, 1, _bf(g), dout, _bf(w), din, B, din, dout, Cb=out, ldcb=din, mask=_bf(h), ldm=din)
    ref = (g @ w) * (h > 0)
    assert np.abs(out.float().cpu().numpy() - ref).max() <= 2 ** -7 * np.abs(ref).max()
    # dW = g^T h, split over the batch with atomics
    for splits in (1, 3):
        dW = torch.zeros(dout, din, device="cuda")
        _gemm(lib, 1, 1, 2, _bf(g), dout, _bf(h), din, dout, din, B, C=dW, ldc=din, splits=splits)
        refw = g.T @ h
        assert np.abs(dW.cpu().numpy() - refw).max() <= 1e-4 * np.abs(refw).max()


def test_mlp_forward_backward_match_cpu_reference():
    rng = np.random.default_rng(11)
    B = 512
    model = learned.MLP((216, 256, 256, 64), seed=3)
    # non-zero biases so that the bias paths are exercised
    for b in model.b:
        b.copy_(torch.from_numpy(rng.standard_normal(b.numel()).astype(np.float32) * 0.1))
    x = (rng.random((B, 216)) < 0.05).astype(np.int32)
    y = (rng.random((B, 64)) < 0.1).astype(np.uint8)
    mask = (rng.random((B, 64)) < 0.5).astype(np.uint8)
    z = model.forward(model.to_input(torch.from_numpy(x).cuda()), keep=True)
    loss, count, dz = model.loss_grad(z, torch.from_numpy(y).cuda(), torch.from_numpy(mask).cuda())
    model.zero_grad()
    model.backward(dz)
    torch.cuda.synchronize()
    Ws = [w.cpu().numpy() for w in model.W]
    bs = [b.cpu().numpy() for b in model.b]
    z_ref, acts = mlp_ref.forward(Ws, bs, x.astype(np.float32))
    assert np.abs(z.cpu().numpy() - z_ref).max() <= 2e-2 * np.abs(z_ref).max()
    l_ref, c_ref, dz_ref = mlp_ref.bce(z_ref, y.astype(np.float32), mask.astype(np.float32))
    assert count == c_ref
    assert abs(float(loss.item()) - l_ref) <= 1e-3 * abs(l_ref)
    dW_ref, db_ref = mlp_ref.backward(Ws, acts, dz_ref)
    for l in range(3):
        for got, ref, what in ((model.gW[l], dW_ref[l], "W"), (model.gb[l], db_ref[l], "b")):
            got = got.cpu().numpy()
            rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
            assert rel < 2e-2, (what, l, rel)


def _small_dem_task():
    return learned.DemTask(seed=5, n_det=216, n_mech=800, max_size=64, W=3)


def _check_learned(task, sess, det, lg, o, scores, per_column, budget, maxw):
    """The learned decode of every shot against learned_ref, and after the
    budget against the oracle's plain result (its count plus the pass's)."""
    r, bits, _ = task.decode(sess, det, lg, scores=torch.from_numpy(scores).cuda(), per_column=per_column,
                             budget=budget, max_weight=maxw)
    bits = bits.cpu().numpy()
    det_np = det.cpu().numpy()
    hits = 0
    for i in range(det.shape[0]):
        s = det_np[i]
        if not s.any():
            assert r["found"][i] == 0 and r["total_tested_combinations"][i] == 0
            continue
        hit, tried, tam = learned_ref.learned_pass(task.H, s, scores[i], per_column, maxw, budget, task.max_size)
        if hit is not None:
            hits += 1
            assert r["found"][i] == 1 and r["iteration"][i] == 1 and r["tam"][i] == tam, i
            assert r["total_tested_combinations"][i] == tried, i
            assert sorted(np.nonzero(bits[i])[0].tolist()) == hit, i
        else:
            for f in ("found", "found_weight", "degenerate", "tam", "iteration"):
                assert r[f][i] == o[f][i], (i, f)
            assert r["total_tested_combinations"][i] == o["total_tested_combinations"][i] + tried, i
    return hits


def test_learned_pass_matches_cpu_statement(oracle):
    task = _small_dem_task()
    det, lg = task.sample(300, seed=4)
    rng = np.random.default_rng(2)
    sess = task.session()
    rc, o = T.run_cln(oracle, G.KIND_PROB_CLN, task.H, task.L, det.cpu().numpy(), lg.cpu().numpy(), cfg=task.cfg,
                      file=task.phase_file)[:2]
    assert rc == 0
    for per_column, budget, maxw in ((True, 50, 4), (True, 20_000, 4), (False, 2000, 4), (True, 3000, 8)):
        width = task.m if per_column else task.max_size
        scores = rng.standard_normal((det.shape[0], width)).astype(np.float32)
        scores[0, :] = 0.0                                          # ties keep list order
        _check_learned(task, sess, det, lg, o, scores, per_column, budget, maxw)
    sess.close()


@pytest.fixture(scope="module")
def dem_full_learned(oracle):
    """The learned-order benchmark's DEM (216 x 3000, max_error_weight 4,
    max_size_active_positions 64), 200 shots, and the oracle's plain
    grand_decode_prob_cln results on them."""
    task = learned.DemTask()
    det, lg = task.sample(200, seed=6)
    rc, o = T.run_cln(oracle, G.KIND_PROB_CLN, task.H, task.L, det.cpu().numpy(), lg.cpu().numpy(), cfg=task.cfg,
                      file=task.phase_file)[:2]
    assert rc == 0
    return task, det, lg, o


@pytest.mark.parametrize("per_column", [True, False])
def test_learned_pass_dem216x3000_w4(dem_full_learned, per_column):
    """Both heads (a score per mechanism, a score per active-list position) with
    random scores, budgets 1e3 and 1e5, patterns up to weight 8."""
    task, det, lg, o = dem_full_learned
    assert (task.n, task.m, task.W, task.max_size) == (216, 3000, 4, 64)
    rng = np.random.default_rng(7 if per_column else 8)
    sess = task.session()
    width = task.m if per_column else task.max_size
    scores = rng.standard_normal((det.shape[0], width)).astype(np.float32)
    for budget in (1000, 100_000):
        hits = _check_learned(task, sess, det, lg, o, scores, per_column, budget, 8)
        assert hits >= 20
    sess.close()


def test_budget_zero_is_plain():
    task = _small_dem_task()
    det, lg = task.sample(2000, seed=8)
    sess = task.session()
    r0, b0, l0 = task.decode(sess, det, lg)
    scores = torch.randn(det.shape[0], task.m, device="cuda")
    r1, b1, l1 = task.decode(sess, det, lg, scores=scores, per_column=True, budget=0)
    for f in G.PARITY_FIELDS:
        assert (r0[f] == r1[f]).all(), f
    assert torch.equal(b0, b1) and torch.equal(l0, l1)
    sess.close()


def test_training_gain_over_plain():
    """Trained on plain-decoder pairs, the learned order tests fewer patterns
    per shot than plain decoding, at a logical error rate no higher."""
    task = _small_dem_task()
    sess = task.session()
    det, lg = task.sample(60_000, seed=21)
    r, bits, _ = task.decode(sess, det, lg)
    y, mask = task.labels(det, bits, r["found"], "column")
    model = learned.MLP((task.n, 256, 256, task.m), seed=1)
    hist = learned.train(task, model, det, y, mask, epochs=3, batch=2048)
    assert hist[-1] < hist[0]
    det_e, lg_e = task.sample(20_000, seed=22)
    r_pl, _, lb_pl = task.decode(sess, det_e, lg_e)
    r_le, _, lb_le = task.decode(sess, det_e, lg_e, scores=model.predict(det_e), per_column=True, budget=100_000)
    mean_pl = r_pl["total_tested_combinations"].mean()
    mean_le = r_le["total_tested_combinations"].mean()
    ler_pl = learned.logical_error_rate(r_pl, lb_pl, lg_e, det_e)
    ler_le = learned.logical_error_rate(r_le, lb_le, lg_e, det_e)
    print(f"mean patterns plain {mean_pl:.1f} learned {mean_le:.1f}; LER plain {ler_pl:.4f} learned {ler_le:.4f}")
    assert mean_le < mean_pl
    assert ler_le <= ler_pl + 3 * np.sqrt(ler_pl * (1 - ler_pl) / det_e.shape[0])
    sess.close()


def test_session_decode_is_ordered_after_default_stream_work():
    """Inputs written by torch kernels on the default stream and decoded at once
    (no synchronization in between) give the results of decoding them after a
    full synchronization: a NULL stream means the default stream."""
    task = _small_dem_task()
    sess = task.session()
    det, lg = task.sample(400_000, seed=31)          # queued on the default stream, not waited for
    r1, b1, l1 = task.decode(sess, det, lg)
    torch.cuda.synchronize()
    r2, b2, l2 = task.decode(sess, det, lg)
    for f in G.PARITY_FIELDS:
        assert (r1[f] == r2[f]).all(), f
    assert torch.equal(b1, b2) and torch.equal(l1, l2)
    sess.close()
