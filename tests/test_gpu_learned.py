"""Learned start on the GPU: MLP kernels against the oracle, sketch features,
learned start factors and the HOSVD fallback of dt_hooi_dense."""
import numpy as np
import pytest

import dtucker as dt
from oracle import dense as D
from oracle import learned as Lr

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def random_params(m, rng, scale=0.3):
    p = rng.standard_normal(m.nparam).astype(np.float32) * scale
    return p


@pytest.mark.parametrize("din,hid,dout,rows", [(32, 64, 16, 200), (8, 96, 5, 70), (32, 512, 16, 384)])
def test_mlp_forward_and_gradient_vs_oracle(ctx, din, hid, dout, rows):
    """Gradients of the hand-written backward kernels against the oracle's
    (FP64 accumulation, the same bf16 rounding of every product operand).
    FP32 accumulation on the device and rare bf16 rounding-boundary flips of
    FP32 intermediates bound the difference: 1e-2 relative (norm)."""
    print("mlp", din, hid, dout, rows, flush=True)
    rng = np.random.default_rng(din + hid + rows)
    m = dt.MLP(ctx, din, hid, dout, seed=3)
    try:
        p = random_params(m, rng)
        m.set_params(p)
        assert np.array_equal(m.get_params(), p)
        X = rng.standard_normal((rows, din)).astype(np.float32)
        T = rng.standard_normal((rows, dout)).astype(np.float32)
        dX, dT = dt.upload_rows(ctx, X), dt.upload_rows(ctx, T)
        Z = dt.download_rows(m.forward(dX, rows))
        Zr, _ = Lr.forward(p, X, din, hid, dout)
        assert rel(Z, Zr) < 1e-3
        loss, g = m.loss_grad(dX, dT, rows, want_grad=True)
        lr_, gr = Lr.loss_grad(p, X, T, din, hid, dout)
        assert abs(loss - lr_) / lr_ < 1e-3
        assert rel(g, gr) < 1e-2
        # per block (W1, b1, W2, b2, W3, b3)
        o = 0
        for shape in ((din, hid), (hid,), (hid, hid), (hid,), (hid, dout), (dout,)):
            n = int(np.prod(shape))
            assert rel(g[o:o + n], gr[o:o + n]) < 2e-2, shape
            o += n
    finally:
        m.close()


def test_mlp_initial_state_is_the_residual_identity(ctx):
    """Zero last layer: z = x[:, :d_out] exactly."""
    m = dt.MLP(ctx, 32, 64, 16, seed=1)
    try:
        X = np.random.default_rng(0).standard_normal((100, 32)).astype(np.float32)
        Z = dt.download_rows(m.forward(dt.upload_rows(ctx, X), 100))
        assert np.array_equal(Z, X[:, :16])
    finally:
        m.close()


def test_mlp_adam_fits_a_linear_map(ctx):
    """Training with the device kernels drives the loss of a learnable
    target (a fixed linear map of x) down by two orders of magnitude."""
    rng = np.random.default_rng(5)
    din, hid, dout, rows = 16, 128, 4, 1024
    W = rng.standard_normal((din, dout)) * 0.3
    X = rng.standard_normal((rows, din)).astype(np.float32)
    T = (X[:, :dout] + X @ W).astype(np.float32)
    m = dt.MLP(ctx, din, hid, dout, seed=2)
    try:
        dX, dT = dt.upload_rows(ctx, X), dt.upload_rows(ctx, T)
        l0 = m.loss_grad(dX, dT, rows)
        for _ in range(300):
            m.loss_grad(dX, dT, rows)
            m.adam_step(3e-3)
        l1 = m.loss_grad(dX, dT, rows)
        print("loss", l0, "->", l1, flush=True)
        assert l1 < 1e-2 * l0
    finally:
        m.close()


@pytest.mark.parametrize("shape,n,S", [((32, 24, 20), 0, 8), ((32, 24, 20), 1, 8), ((20, 16, 64), 2, 12),
                                       ((64, 64, 64), 1, 32)])
def test_sketch_features_vs_oracle(ctx, shape, n, S):
    print("sketch", shape, n, S, flush=True)
    rng = np.random.default_rng(sum(shape) + n)
    G = rng.standard_normal((4, 4, 4))
    Us = [np.linalg.qr(rng.standard_normal((I, 4)))[0] for I in shape]
    A = (D.reconstruct(G, Us) + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    X = ctx.sketch_features(ctx.to_device(A), n, S, seed=77)
    Xr = Lr.sketch_features(A.astype(np.float64), n, S, 77)
    # the leading columns (well separated singular values) agree to FP32 accuracy
    assert rel(X[:, :4], Xr[:, :4]) < 1e-4
    # all columns: same spanned space and the same column norms
    assert D.sin_max_angle(np.linalg.qr(X)[0], np.linalg.qr(Xr)[0]) < 1e-3
    assert np.allclose(np.linalg.norm(X, axis=0), np.linalg.norm(Xr, axis=0), rtol=1e-4)


@pytest.mark.parametrize("shape,n,q,dtype", [((64, 48, 40), 0, 4, dt.DT_F32), ((64, 48, 40), 1, 4, dt.DT_F32),
                                             ((128, 64, 40), 2, 5, dt.DT_F32), ((128, 64, 40), 2, 5, dt.DT_BF16),
                                             ((16, 12, 10, 8), 1, 3, dt.DT_F32)])
def test_sketch_features_kron_vs_oracle(ctx, shape, n, q, dtype):
    """Kronecker sketch through dt_ttmc (the fused leading pair for n >= 2)."""
    print("kron sketch", shape, n, q, dtype, flush=True)
    rng = np.random.default_rng(sum(shape) + 7 * n)
    r = 3
    G = 50.0 * rng.standard_normal((r,) * len(shape))  # signal well above the noise
    Us = [np.linalg.qr(rng.standard_normal((I, r)))[0] for I in shape]
    A = (D.reconstruct(G, Us) + 0.005 * rng.standard_normal(shape)).astype(np.float32)
    if dtype == dt.DT_BF16:
        Ab = dt.bf16_from_f32(A)
        dA = ctx.array(shape, dt.DT_BF16).upload(Ab)
        A = dt.f32_from_bf16(Ab)
    else:
        dA = ctx.to_device(A)
    X = ctx.sketch_features_kron(dA, n, q, seed=11)
    Xr = Lr.sketch_features_kron(A.astype(np.float64), n, q, 11)
    tol = 1e-4 if dtype == dt.DT_F32 else 1e-2  # bf16 rounds the factor of the first product
    assert rel(X[:, :r], Xr[:, :r]) < tol
    assert D.sin_max_angle(np.linalg.qr(X[:, :r])[0], Us[n]) < 0.1


def test_learned_start_with_untrained_model_is_the_sketch_range(ctx):
    """A fresh model (zero last layer) gives QR of the first R feature columns:
    the randomized range finder's leading subspace."""
    shape, ranks = (48, 40, 32), (6, 6, 6)
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=31)
    m = dt.MLP(ctx, 16, 64, 6, seed=0)
    try:
        Us = [u.download() for u in ctx.learned_start(m, g["A"], ranks, seed=9)]
        A = g["A"].download().astype(np.float64)
        for n in range(3):
            U = Us[n]
            assert np.allclose(U.T @ U, np.eye(6), atol=1e-12)
            Xr = Lr.sketch_features(A, n, 16, 9)
            assert D.sin_max_angle(U, np.linalg.qr(Xr[:, :6])[0]) < 1e-4
    finally:
        m.close()


@pytest.mark.parametrize("shape,q,r,dtype", [((48, 40, 32), 4, 5, dt.DT_BF16), ((16, 12, 10, 8), 3, 3, dt.DT_F32)])
def test_learned_start_kron_shares_the_sketch_prefix_exactly(ctx, shape, q, r, dtype):
    """The batched learned start contracts A over the modes above 1 once for the
    Kronecker sketches of modes 0 and 1; with a fresh model its start is QR of
    the first R feature columns.  The per-mode features (their own full chain,
    then the same Gram, Jacobi and feature kernels) must give that subspace to
    the rounding of the FP64 QR: a sketch formed differently would differ at
    FP32 rounding, about 1e-7."""
    print("kron prefix", shape, q, r, dtype, flush=True)
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal(shape).astype(np.float32)
    if dtype == dt.DT_BF16:
        dA = ctx.array(shape, dt.DT_BF16).upload(dt.bf16_from_f32(A))
    else:
        dA = ctx.to_device(A)
    S = q ** (len(shape) - 1)
    m = dt.MLP(ctx, S, 64, r, seed=0)
    try:
        Us = [u.download() for u in ctx.learned_start(m, dA, (r,) * len(shape), seed=5,
                                                      sketch=dt.SKETCH_KRONECKER)]
        for n in range(len(shape)):
            X = ctx.sketch_features_kron(dA, n, q, seed=5)
            assert D.sin_max_angle(Us[n], np.linalg.qr(X[:, :r].astype(np.float64))[0]) < 1e-10, n
    finally:
        m.close()


def test_learned_init_runs_hooi_and_reports_first_sweeps(ctx):
    """DT_INIT_LEARNED with DT_FLAG_LEARNED_CHECK: both first sweeps are
    reported and the better start is kept; the result matches the
    HOSVD-started HOOI's fit."""
    shape, ranks = (64, 64, 64), (8, 8, 8)
    g = ctx.gen_tucker(shape, ranks, 1e-2, seed=41, decay=0.8)
    m = dt.MLP(ctx, 16, 64, 8, seed=0)
    try:
        out = ctx.hooi_dense(g["A"], ranks, max_iter=5, init=dt.INIT_LEARNED, model=m.handle,
                             flags=dt.FLAG_LEARNED_CHECK | dt.FLAG_FP64_FIT, seed=3)
        ref = ctx.hooi_dense(g["A"], ranks, max_iter=5, init=dt.INIT_HOSVD, flags=dt.FLAG_FP64_FIT)
        print("learned first", out["learned_first_res"], "hosvd first", out["hosvd_first_res"],
              "used", out["init_used"], flush=True)
        assert out["hosvd_first_res"] > 0 and out["learned_first_res"] > 0
        want = dt.INIT_LEARNED if out["learned_first_res"] <= out["hosvd_first_res"] else dt.INIT_HOSVD
        assert out["init_used"] == want
        assert abs(out["residual"] - ref["residual"]) < 1e-6
    finally:
        m.close()


def test_learned_init_falls_back_to_hosvd_for_a_bad_model(ctx):
    """A model whose output ignores the data (huge random weights) starts
    far from the subspace; with one inner ALS step per mode its first sweep
    stays behind the HOSVD start's, and with the check the HOSVD start is taken."""
    shape, ranks = (64, 64, 64), (8, 8, 8)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=43, decay=0.8)
    m = dt.MLP(ctx, 16, 64, 8, seed=0)
    try:
        rng = np.random.default_rng(0)
        m.set_params(rng.standard_normal(m.nparam).astype(np.float32) * 5.0)
        out = ctx.hooi_dense(g["A"], ranks, max_iter=3, init=dt.INIT_LEARNED, model=m.handle,
                             flags=dt.FLAG_LEARNED_CHECK | dt.FLAG_FP64_FIT, seed=3, inner_iter=1)
        print("bad model: learned first", out["learned_first_res"], "hosvd", out["hosvd_first_res"],
              flush=True)
        assert out["learned_first_res"] > out["hosvd_first_res"]
        assert out["init_used"] == dt.INIT_HOSVD
    finally:
        m.close()


def test_learned_start_acceptance_rule(ctx):
    """DT_FLAG_LEARNED_ACCEPT (no HOSVD computed unless the test fails): from
    the sketch-range start of an untrained network HOOI has converged after
    one sweep (the second sweep moves the fit by less than 1e-4) and the start
    is kept; from the start of a network with huge random weights and one
    inner ALS step per mode it has not, and the HOSVD start is taken instead --
    whose run then matches the HOSVD-started HOOI."""
    shape, ranks = (64, 64, 64), (8, 8, 8)
    g = ctx.gen_tucker(shape, ranks, 1e-3, seed=45, decay=0.8)
    m = dt.MLP(ctx, 16, 64, 8, seed=0)
    try:
        good = ctx.hooi_dense(g["A"], ranks, max_iter=2, init=dt.INIT_LEARNED, model=m.handle,
                              flags=dt.FLAG_LEARNED_ACCEPT | dt.FLAG_FP64_FIT, seed=3)
        assert good["init_used"] == dt.INIT_LEARNED
        rng = np.random.default_rng(0)
        m.set_params(rng.standard_normal(m.nparam).astype(np.float32) * 5.0)
        # noise 0.1, slow decay: from that start the second sweep still moves the fit by ~4e-4
        g2 = ctx.gen_tucker(shape, ranks, 1e-1, seed=45, decay=0.95)
        bad = ctx.hooi_dense(g2["A"], ranks, max_iter=2, init=dt.INIT_LEARNED, model=m.handle,
                             flags=dt.FLAG_LEARNED_ACCEPT | dt.FLAG_FP64_FIT, seed=3, inner_iter=1)
        ref = ctx.hooi_dense(g2["A"], ranks, max_iter=2, init=dt.INIT_HOSVD, flags=dt.FLAG_FP64_FIT, inner_iter=1)
        print("accept: good", good["init_used"], "bad", bad["init_used"], bad["residuals"], ref["residuals"],
              flush=True)
        assert bad["init_used"] == dt.INIT_HOSVD
        assert abs(bad["residual"] - ref["residual"]) < 1e-9
    finally:
        m.close()


def test_learned_init_without_model_is_einval(ctx):
    dA = ctx.to_device(np.ones((8, 8, 8), np.float32))
    with pytest.raises(dt.DTError) as e:
        ctx.hooi_dense(dA, (2, 2, 2), max_iter=1, init=dt.INIT_LEARNED)
    assert e.value.code == dt.DT_EINVAL


def test_trained_model_128cube_against_hosvd_and_sketch_only(ctx):
    """What the committed network gains over the method without it, on 8
    held-out tensors of the learned-init workload (128^3, ranks 16, decay
    U(0.7, 0.95), noise log-uniform [1e-3, 1e-1]).  Training (tensors
    0..3519, the library's own kernels) lowers the squared loss but never the
    validation tensors' (3520..3583) subspace error, so the early stop keeps
    the untrained network: the learned start IS the sketch range, and the
    network gains nothing in start fit, first-sweep fit or sweeps.  Every
    start reaches the reference fit within 1e-4 in one sweep; the sketch start
    does so in less wall time than HOSVD (the sketch replaces the three Gram +
    eigen-solves)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "diana-tucker_amd", "python"))
    import learned as LN
    cfg = LN.FAMILY_128
    path = os.path.join(LN.MODEL_DIR, "learned_tucker128_r16.npz")
    m = LN.load_model(ctx, path)
    try:
        with np.load(path, allow_pickle=False) as f:
            import json
            hist = json.loads(str(f["meta"]))["history"]
        kept = hist[-1]
        vals = [h["val_subspace"] for h in hist if "epoch" in h]
        assert kept["val_subspace"] == min(vals)
        print("kept epoch", kept["kept_epoch"], "validation subspace errors", vals, flush=True)
        ids = list(range(cfg["n_train"], cfg["n_train"] + 8))
        # warm-up (first launches of every kernel)
        LN.evaluate(ctx, m, ids[:1], cfg, max_sweeps=3, ref_sweeps=5, log=lambda *a: None)
        rows = LN.evaluate(ctx, m, ids, cfg, max_sweeps=6, ref_sweeps=20, log=lambda *a: None)
        s = LN.summarize(rows)
        print("summary", s, flush=True)
        for name in ("hosvd", "sketch", "learned"):
            assert s[name]["reached"] == len(ids)
        assert s["learned"]["mean_sweeps"] <= s["hosvd"]["mean_sweeps"]
        assert s["learned"]["mean_ms"] < 0.8 * s["hosvd"]["mean_ms"]
        gain0 = s["learned"]["mean_start_gap"] - s["sketch"]["mean_start_gap"]
        gain1 = s["learned"]["mean_first_gap"] - s["sketch"]["mean_first_gap"]
        print("network gain over the sketch alone: start fit", gain0, "first-sweep fit", gain1, flush=True)
        # the kept network is never worse than the sketch range on the
        # validation tensors; on held-out tensors neither fit may fall behind
        assert gain0 >= -1e-9 and gain1 >= -1e-9
    finally:
        m.close()


def test_training_keeps_the_epoch_with_the_best_validation_subspace(ctx):
    """train() with validation rows returns the parameters of the epoch whose
    starts are closest to the targets' subspaces, epoch 0 (the untrained
    network) included, and its subspace_error is that epoch's."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "diana-tucker_amd", "python"))
    import learned as LN
    cfg = dict(LN.FAMILY_128, shape=(32, 32, 32), ranks=(4, 4, 4), S=8, hidden=32, n_train=6, n_total=8)
    X, T = LN.make_examples(ctx, range(5), cfg, sweeps=4, log=lambda *a: None)
    Xv, Tv = LN.make_examples(ctx, range(5, 7), cfg, sweeps=4, log=lambda *a: None)
    m, hist = LN.train(ctx, X, T, cfg, epochs=3, batch=64, lr=3e-3, Xv=Xv, Tv=Tv, log=lambda *a: None)
    try:
        errs = [h["val_subspace"] for h in hist if "epoch" in h]
        assert len(errs) == 4
        assert hist[-1]["kept_epoch"] == int(np.argmin(errs))
        dXv = dt.upload_rows(ctx, Xv)
        assert abs(LN.subspace_error(m, dXv, Tv, cfg) - min(errs)) < 1e-6
        dXv.free()
    finally:
        m.close()
