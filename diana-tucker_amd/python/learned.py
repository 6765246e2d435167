"""Training data, training and evaluation of the learned HOOI start.

Everything runs on the GPU through libdiana_hip.so: tensors come from the
device Tucker generator, targets from the plain HOOI's converged factors, the
network is trained by the library's own MLP forward/backward/Adam kernels.

Workload of the learned-init configuration: 4096 seeded dense 3-way FP32
tensors of 128x128x128 with ranks (16,16,16), core spectra decaying at a rate
drawn U(0.7, 0.95), noise relative norm log-uniform in [1e-3, 1e-1]; tensors
0..3583 train, 3584..4095 evaluate.  Per mode the features are the Gaussian
sketch X_(n) Ω (Ω: 16384 x 32) in its own singular frame
(dt_sketch_features); the target for a mode with converged factor U* is the
projection of the sketch's leading directions onto span(U*):
T = U* U*ᵀ X[:, :R].  This fixes a basis of span(U*) in the input's own
frame, so a row-wise network can be trained with a squared loss.

    python learned.py train [--config tucker128_r16|tucker256_r64] [--n-train N] [--epochs E] [--model PATH]
    python learned.py eval  [--config tucker128_r16|tucker256_r64] [--n-eval N] [--model PATH] [--report PATH] [--rows PATH]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import dtucker as dt  # noqa: E402

MODEL_DIR = os.path.join(os.path.dirname(HERE), "models")

FAMILY_128 = {"name": "tucker128_r16", "shape": (128, 128, 128), "ranks": (16, 16, 16), "S": 32, "hidden": 512,
      "base_seed": 20000905, "n_train": 3584, "n_total": 4096, "decay": (0.7, 0.95), "noise": (1e-3, 1e-1),
      "sketch": dt.SKETCH_GAUSSIAN}
# The 1024^3 rank-64 mixed-precision workload's family (core decay 0.93, noise
# 1e-3) at 256^3: the network is row-wise with features normalised by
# sqrt(I_n), so it is trained on smaller tensors of the same model.  Its
# sketch is the Kronecker one (q = 9 per mode, S = 81): at 1024^3 a dense
# Gaussian Ω would have 2^20 x S entries per mode.
FAMILY_256_R64 = {"name": "tucker256_r64", "shape": (256, 256, 256), "ranks": (64, 64, 64), "S": 81, "q": 9, "hidden": 512,
      "base_seed": 30000905, "n_train": 192, "n_total": 224, "decay": (0.93, 0.93), "noise": (1e-3, 1e-3),
      "sketch": dt.SKETCH_KRONECKER}
CONFIGS = {"tucker128_r16": FAMILY_128, "tucker256_r64": FAMILY_256_R64}


def tensor_params(t, cfg):
    """(decay, noise) of tensor t: decay ~ U(lo, hi), noise log-uniform in [lo, hi]."""
    rng = np.random.default_rng([cfg["base_seed"], t])
    d0, d1 = cfg["decay"]
    n0, n1 = cfg["noise"]
    return float(rng.uniform(d0, d1)), float(10.0 ** rng.uniform(math.log10(n0), math.log10(n1)))


def generate(ctx, t, cfg, **kw):
    decay, noise = tensor_params(t, cfg)
    return ctx.gen_tucker(cfg["shape"], cfg["ranks"], noise, seed=cfg["base_seed"] + t, decay=decay, **kw)


def features(ctx, A, cfg, seed):
    if cfg["sketch"] == dt.SKETCH_KRONECKER:
        return [ctx.sketch_features_kron(A, n, cfg["q"], seed) for n in range(len(A.shape))]
    return [ctx.sketch_features(A, n, cfg["S"], seed) for n in range(len(A.shape))]


def learned_flags(cfg):
    return dt.FLAG_SKETCH_KRON if cfg["sketch"] == dt.SKETCH_KRONECKER else 0


def make_examples(ctx, ids, cfg, sweeps=12, log=print):
    """Features and targets of tensors `ids`: X (rows x S), T (rows x R)."""
    R = cfg["ranks"][0]
    Xs, Ts = [], []
    t0 = time.time()
    for k, t in enumerate(ids):
        g = generate(ctx, t, cfg)
        out = ctx.hooi_dense(g["A"], cfg["ranks"], max_iter=sweeps, init=dt.INIT_HOSVD, tol=1e-9)
        X = features(ctx, g["A"], cfg, cfg["base_seed"] + t)
        for n in range(len(cfg["shape"])):
            U = out["U"][n]
            Xs.append(X[n].astype(np.float32))
            Ts.append((U @ (U.T @ X[n][:, :R].astype(np.float64))).astype(np.float32))
        g["A"].free()
        if (k + 1) % 64 == 0:
            log(f"[learned] examples {k + 1}/{len(ids)} ({time.time() - t0:.1f} s)")
    return np.concatenate(Xs), np.concatenate(Ts)


def train(ctx, X, T, cfg, epochs=8, batch=16384, lr=2e-4, seed=0, Xv=None, Tv=None, log=print):
    """Adam on the device; cosine learning-rate decay; returns (model, history).

    With validation rows (Xv, Tv) the parameters of the epoch whose starts
    are closest to the targets' subspaces on the validation tensors are kept
    (subspace_error; epoch 0, the untrained network whose start is the sketch
    range itself, included), so the returned model is never worse there than
    the method without it.  The squared loss alone is no such measure: a
    per-column rescaling lowers it and leaves every start subspace as it was."""
    R = cfg["ranks"][0]
    m = dt.MLP(ctx, cfg["S"], cfg["hidden"], R, seed=seed)
    rows = X.shape[0]
    nb = max(1, rows // batch)
    total = epochs * nb
    rng = np.random.default_rng(seed)
    hist = []
    dXv = dt.upload_rows(ctx, Xv) if Xv is not None else None
    dTv = dt.upload_rows(ctx, Tv) if Tv is not None else None
    best = None
    if dXv is not None:
        hist.append({"epoch": 0, "val": m.loss_grad(dXv, dTv, Xv.shape[0]),
                     "val_subspace": subspace_error(m, dXv, Tv, cfg)})
        best = (hist[-1]["val_subspace"], 0, m.get_params())
        log(f"[learned] epoch 0 val {hist[-1]['val']:.7g} subspace {hist[-1]['val_subspace']:.7g}")
    step = 0
    for ep in range(epochs):
        perm = rng.permutation(rows)
        dX = dt.upload_rows(ctx, X[perm[:nb * batch]])
        dT = dt.upload_rows(ctx, T[perm[:nb * batch]])
        tl = 0.0
        for b in range(nb):
            xb = _RowsView(dX, b * batch, cfg["S"])
            tb = _RowsView(dT, b * batch, R)
            tl += m.loss_grad(xb, tb, batch)
            step += 1
            m.adam_step(lr * 0.5 * (1.0 + math.cos(math.pi * step / total)) + 1e-6)
        dX.free()
        dT.free()
        h = {"epoch": ep + 1, "train": tl / nb}
        if dXv is not None:
            h["val"] = m.loss_grad(dXv, dTv, Xv.shape[0])
            h["val_subspace"] = subspace_error(m, dXv, Tv, cfg)
            if h["val_subspace"] < best[0]:
                best = (h["val_subspace"], ep + 1, m.get_params())
        hist.append(h)
        log(f"[learned] epoch {ep + 1}: train {h['train']:.7g}" +
            (f" val {h['val']:.7g} subspace {h['val_subspace']:.7g}" if "val" in h else ""))
    if best is not None:
        m.set_params(best[2])
        hist.append({"kept_epoch": best[1], "val_subspace": best[0]})
        log(f"[learned] kept epoch {best[1]} (subspace {best[0]:.7g})")
    if dXv is not None:
        dXv.free()
        dTv.free()
    return m, hist


def subspace_error(m, dX, T, cfg):
    """Mean over the (tensor, mode) blocks of the rows (I_n rows each, in the
    order make_examples writes them) of sum_k sin^2(theta_k), the principal
    angles between the span of the network's output block and the span of
    its target block (= span(U*))."""
    rows, R = T.shape
    Z = m.forward(dX, rows)
    Zh = dt.download_rows(Z).astype(np.float64)
    Z.free()
    sizes = list(cfg["shape"])
    out, r0, b = [], 0, 0
    while r0 < rows:
        I = sizes[b % len(sizes)]
        qz = np.linalg.qr(Zh[r0:r0 + I])[0]
        qt = np.linalg.qr(T[r0:r0 + I].astype(np.float64))[0]
        out.append(R - float(np.sum((qz.T @ qt) ** 2)))
        r0 += I
        b += 1
    return float(np.mean(out))


class _RowsView:
    """Row offset into a row-major device buffer (pointer + offset)."""

    def __init__(self, base, row0, width):
        self.ptr = base.ptr + row0 * width * 4


def save_model(m, path, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=m.get_params(), d_in=m.d_in, hidden=m.hidden, d_out=m.d_out,
                        meta=json.dumps(meta))


def load_model(ctx, path):
    with np.load(path, allow_pickle=False) as f:
        m = dt.MLP(ctx, int(f["d_in"]), int(f["hidden"]), int(f["d_out"]))
        m.set_params(f["params"])
    return m


def sweeps_to_fit(ctx, A, cfg, init, model=None, max_sweeps=10, target=None, seed=0):
    """Residual of the start itself (core of the start factors, FP64
    accumulated), per-sweep residuals from `init`, and the first sweep whose
    fit is within 1e-4 of the target fit (or None)."""
    s = ctx.hooi_session(A, cfg["ranks"], init=init, model=model.handle if model else None, seed=seed,
                         flags=dt.FLAG_FP64_FIT | learned_flags(cfg))
    start = s.fit(fp64=True)
    res = [s.sweep() for _ in range(max_sweeps)]
    s.close()
    hit = None
    if target is not None:
        for k, r in enumerate(res):
            if (1.0 - r) >= target - 1e-4:
                hit = k + 1
                break
    return start, res, hit


def timed_run(ctx, A, cfg, init, sweeps, model=None, seed=0):
    """Wall time (ms) of start + `sweeps` plain sweeps, host clock around synchronised calls."""
    ctx.sync()
    t0 = time.perf_counter()
    s = ctx.hooi_session(A, cfg["ranks"], init=init, model=model.handle if model else None, seed=seed,
                         flags=learned_flags(cfg))
    for _ in range(sweeps):
        s.sweep()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1e3
    s.close()
    return ms


def evaluate(ctx, model, ids, cfg, max_sweeps=10, ref_sweeps=25, log=print):
    """HOSVD start vs the untrained (sketch-only) start vs the learned start on
    tensors `ids`: sweeps and wall time to reach within 1e-4 of the reference
    fit (the converged fit of a 25-sweep FP32 HOOI with FP64-accumulated core)."""
    R = cfg["ranks"][0]
    plain = dt.MLP(ctx, cfg["S"], cfg["hidden"], R, seed=0)  # zero last layer: sketch-only start
    rows = []
    for k, t in enumerate(ids):
        g = generate(ctx, t, cfg)
        A = g["A"]
        ref = ctx.hooi_dense(A, cfg["ranks"], max_iter=ref_sweeps, init=dt.INIT_HOSVD, flags=dt.FLAG_FP64_FIT,
                             want_factors=False)
        target = 1.0 - ref["residual"]
        row = {"t": int(t), "target_fit": target}
        for name, init, mdl in (("hosvd", dt.INIT_HOSVD, None), ("sketch", dt.INIT_LEARNED, plain),
                                ("learned", dt.INIT_LEARNED, model)):
            start, res, hit = sweeps_to_fit(ctx, A, cfg, init, mdl, max_sweeps, target, seed=cfg["base_seed"] + t)
            ms = timed_run(ctx, A, cfg, init, hit if hit else max_sweeps, mdl, seed=cfg["base_seed"] + t)
            row[name] = {"start_res": start, "gap0": (1.0 - start) - target, "first_res": res[0],
                         "gap1": (1.0 - res[0]) - target, "sweeps": hit, "ms": ms}
        rows.append(row)
        A.free()
        if (k + 1) % 32 == 0:
            log(f"[learned] evaluated {k + 1}/{len(ids)}")
    plain.close()
    return rows


def summarize(rows):
    out = {}
    for name in ("hosvd", "sketch", "learned"):
        sw = [r[name]["sweeps"] for r in rows]
        ok = [s for s in sw if s is not None]
        out[name] = {"reached": len(ok), "of": len(rows),
                     "mean_sweeps": float(np.mean(ok)) if ok else None,
                     "mean_ms": float(np.mean([r[name]["ms"] for r in rows])),
                     "median_ms": float(np.median([r[name]["ms"] for r in rows])),
                     "mean_start_gap": float(np.mean([r[name]["gap0"] for r in rows])),
                     "mean_first_gap": float(np.mean([r[name]["gap1"] for r in rows]))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["train", "eval"])
    ap.add_argument("--config", choices=sorted(CONFIGS), default="tucker128_r16")
    ap.add_argument("--n-train", type=int, default=None)
    ap.add_argument("--n-eval", type=int, default=None)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--model", default=None)
    ap.add_argument("--report", default=None, help="report path (default: models/learned_<config>_report.json)")
    ap.add_argument("--rows", default=None, help="also write every evaluated tensor's row here")
    args = ap.parse_args()
    cfg = dict(CONFIGS[args.config])
    if args.n_train is None:
        args.n_train = cfg["n_train"]
    if args.n_eval is None:
        args.n_eval = cfg["n_total"] - cfg["n_train"]
    if args.model is None:
        args.model = os.path.join(MODEL_DIR, f"learned_{cfg['name']}.npz")
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    with dt.Context(0) as ctx:
        if args.what == "train":
            # validation tensors are the last of the training ids, so the
            # held-out ids (n_train and up) stay unseen by the early stop
            n_val = min(64, args.n_train // 6)
            tr = list(range(args.n_train - n_val))
            va = list(range(args.n_train - n_val, args.n_train))
            t0 = time.time()
            X, T = make_examples(ctx, tr, cfg, log=log)
            Xv, Tv = make_examples(ctx, va, cfg, log=log)
            t1 = time.time()
            m, hist = train(ctx, X, T, cfg, epochs=args.epochs, lr=args.lr, Xv=Xv, Tv=Tv, log=log)
            t2 = time.time()
            meta = {"cfg": {k: list(v) if isinstance(v, tuple) else v for k, v in cfg.items()},
                    "n_train": args.n_train, "n_val": n_val, "epochs": args.epochs, "lr": args.lr, "history": hist,
                    "data_s": t1 - t0, "train_s": t2 - t1}
            save_model(m, args.model, meta)
            print(json.dumps(meta))
            m.close()
        else:
            m = load_model(ctx, args.model)
            ids = list(range(cfg["n_train"], cfg["n_train"] + args.n_eval))
            evaluate(ctx, m, ids[:1], cfg, max_sweeps=2, ref_sweeps=2, log=lambda *a: None)  # first launches
            rows = evaluate(ctx, m, ids, cfg, log=log)
            summ = summarize(rows)
            print(json.dumps({"summary": summ, "n_eval": len(ids)}))
            # the model's report: its training record, the held-out summary
            # and the commands that made both
            with np.load(args.model, allow_pickle=False) as f:
                meta = json.loads(str(f["meta"]))
            meta.update({"eval_summary": summ, "n_eval": len(ids),
                         "command_train": f"python diana-tucker_amd/python/learned.py train --config {cfg['name']}",
                         "command_eval": f"python diana-tucker_amd/python/learned.py eval --config {cfg['name']}"})
            report = args.report or os.path.join(MODEL_DIR, f"learned_{cfg['name']}_report.json")
            with open(report, "w") as f:
                json.dump(meta, f, indent=1)
            if args.rows:
                with open(args.rows, "w") as f:
                    json.dump(rows, f)
            m.close()


if __name__ == "__main__":
    main()
