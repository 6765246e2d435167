#!/usr/bin/env python3
"""Benchmark of the HOOI-ALS hot path on one MI355X.

Headline workload (printed as the JSON `value`): dense HOOI on a 4-way FP32
tensor of size 256x256x256x256 (17.2 GB) from a seeded Tucker model (core
16^4, orthonormal 256x16 factors, noise 1e-2), ranks (16,16,16,16), HOSVD
start, dimension-tree schedule.  One step = one HOOI sweep over all modes
including the core and the residual.  Inputs are generated on the device and
are resident in HBM before the timed region.

`python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]`
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "diana-tucker_amd", "python"))

HBM_PEAK_GBS = 8000.0
# the two full-tensor passes of a tree-schedule sweep at 256^4: phase 1 reads A
# and writes A x_2 U2^T x_3 U3^T in one fused pass (ttm2_trail_f32_stream);
# phase 2 reads A and writes A x_0 U0^T x_1 U1^T in one fused pass
# (ttm2_lead_f32_stream).  ttm_mk_f32_stream is phase 1's single product when
# the trailing pair does not fit.
FULL_PASS_KERNELS = ("ttm2_trail_f32_stream", "ttm2_lead_f32_stream", "ttm_mk_f32_stream")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tree_contractions(shape, ranks):
    """The full list of (input elements, K, J) contractions of one sweep of the
    dimension-tree schedule, mirroring sweep_once() in dt_hooi.hip."""
    N = len(shape)
    hm = N // 2
    out = []

    def chain(sh, modes):
        sh = list(sh)
        for m in modes:
            out.append((math.prod(sh), sh[m], ranks[m]))
            sh[m] = ranks[m]
        return sh

    # phase 1: modes N-2 and N-1 in the fused pass (stage 1 over N-2), then the rest
    TL = chain(shape, [N - 2, N - 1] + list(range(N - 3, hm - 1, -1)))
    for n in range(hm):
        chain(TL, [i for i in range(hm - 1, -1, -1) if i != n])
    TR = chain(shape, list(range(hm - 1, -1, -1)))
    last = None
    for n in range(hm, N):
        last = chain(TR, [i for i in range(N - 1, hm - 1, -1) if i != n])
    chain(last, [N - 1])
    return out


def dense_256x4(ctx, dt, steps, warmup, shape=(256,) * 4, ranks=(16,) * 4, noise=1e-2,
             seed=20000905):
    log(f"[bench] generating {shape} FP32 Tucker model (ranks {ranks}, noise {noise})")
    g = ctx.gen_tucker(shape, ranks, noise, seed=seed)
    ctx.sync()
    log("[bench] HOSVD start factors")
    s = ctx.hooi_session(g["A"], ranks, init=dt.INIT_HOSVD, schedule=dt.SCHED_TREE)
    for i in range(warmup):
        r = s.sweep()
        log(f"[bench] warmup sweep {i}: residual {r:.9f}")
    # live HIP-event timing of the dominant kernel only (other launches are
    # not bracketed, so the timed sweeps carry no per-launch event overhead)
    ctx.timing_filter(",".join(FULL_PASS_KERNELS))
    ctx.timing(True)
    ctx.timing_reset()
    ctx.sync()
    t0 = time.perf_counter()
    res = []
    for i in range(steps):
        res.append(s.sweep())
    ctx.sync()
    t1 = time.perf_counter()
    ktimes = ctx.timing_list()
    ctx.timing(False)
    ctx.timing_filter("")
    # per-kernel breakdown from two extra, separately timed sweeps
    ctx.timing(True)
    ctx.timing_reset()
    for i in range(2):
        s.sweep()
    breakdown = {k: v[1] / 2 for k, v in sorted(ctx.timing_list().items())}
    ctx.timing(False)
    out = s.result(want_factors=False)
    s.close()
    wall = t1 - t0
    log(f"[bench] {steps} sweeps in {wall:.3f} s, residual {res[-1]:.9f}")
    # algorithmic work of one sweep
    con = tree_contractions(shape, ranks)
    flops = sum(2 * e * J for e, K, J in con)
    total = math.prod(shape)
    # algorithmic bytes per launch: read A once, write the contracted result
    kbytes = {"ttm_mk_f32_stream": total * 4 + total // shape[-1] * ranks[-1] * 4,
              "ttm2_trail_f32_stream": total * 4 + total // (shape[-2] * shape[-1]) * ranks[-2] * ranks[-1] * 4,
              "ttm2_lead_f32_stream": total * 4 + total // (shape[0] * shape[1]) * ranks[0] * ranks[1] * 4}
    return {"wall": wall, "res": res, "ktimes": ktimes, "breakdown": breakdown, "flops_per_sweep": flops,
            "kbytes": kbytes, "init_ms": out["init_ms"], "A": g["A"], "total": total}


def _median_ms(fn, reps=3):
    vals = sorted(fn() for _ in range(reps))
    return vals[len(vals) // 2]


def mixed_1024x3(ctx, dt, shape=(1024,) * 3, ranks=(64,) * 3, noise=1e-3, decay=0.93, seed=20000905 + 2,
             max_sweeps=4):
    """Time-to-fit on the 1024^3 rank-64 workload (core 64^3 with mode spectra
    decaying as 0.93^k, noise 1e-3): wall time, including the start, until
    the fit is within 1e-4 of the reference fit (FP32 HOOI from the HOSVD
    start with FP64-accumulated fit, 3 sweeps).  The sweep count each path
    needs is found in a run that evaluates the fit in FP64 after every sweep;
    the timed runs then do exactly that many sweeps:
      plain FP32  : HOSVD start on the FP32 tensor + FP32 sweeps (their own residual);
      mixed+learned: learned start (Kronecker sketch of the BF16 tensor -> MLP
                    -> QR) + BF16 sweeps + the acceptance sweep (the HOSVD
                    start is taken when it moves the fit by 1e-4 or more) + one
                    fit of the factors on the FP32 copy;
      checked     : the same with the HOSVD comparison instead of the
                    acceptance sweep (the HOSVD start is computed every time).
    The network's early stop kept its untrained state (no epoch brought the
    validation starts closer to the converged subspaces), so its start is the
    sketch range: the speedup is that of the sketch start and the BF16 sweeps,
    which the line says ("speedup_source")."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "diana-tucker_amd", "python"))
    import learned as LN
    log(f"[bench] mixed precision: generating {shape} BF16 + FP32 (ranks {ranks})")
    g = ctx.gen_tucker(shape, ranks, noise, seed=seed, decay=decay, bf16=True, want_factors=True)
    A32, A16 = g["A"], g["A_bf16"]
    ref = ctx.hooi_dense(A32, ranks, max_iter=3, init=dt.INIT_HOSVD, flags=dt.FLAG_FP64_FIT)
    target = 1.0 - ref["residual"]
    model_path = os.path.join(LN.MODEL_DIR, "learned_tucker256_r64.npz")
    model = LN.load_model(ctx, model_path)
    with np.load(model_path, allow_pickle=False) as f:
        kept_epoch = json.loads(str(f["meta"]))["history"][-1].get("kept_epoch")
    kron = dt.FLAG_SKETCH_KRON

    def sweeps_needed(A, init, **kw):
        s = ctx.hooi_session(A, ranks, init=init, **kw)
        res = []
        for k in range(max_sweeps):
            res.append(s.sweep())
            if 1.0 - res[-1] >= target - 1e-4:
                break
        out = s.result(want_factors=True)
        s.close()
        return len(res), res, out["U"]

    k_plain, res_plain, U_plain = sweeps_needed(A32, dt.INIT_HOSVD, flags=dt.FLAG_FP64_FIT)
    k_mixed, res_mixed, U_mixed = sweeps_needed(A16, dt.INIT_LEARNED, model=model.handle, A_f32=A32,
                                                flags=dt.FLAG_FP64_FIT | kron, seed=seed)

    def run_plain():
        ctx.sync()
        t0 = time.perf_counter()
        s = ctx.hooi_session(A32, ranks, init=dt.INIT_HOSVD)
        for _ in range(k_plain):
            s.sweep()
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3
        s.close()
        return ms

    fit_mixed_reported = []
    fit_mixed_fp64 = []
    accepted = []

    def run_mixed(check=True):
        """learned start, its BF16 sweeps, the acceptance test of
        DT_FLAG_LEARNED_ACCEPT (one more sweep must move the fit by less than
        1e-4: HOOI from the learned start has converged, else the HOSVD start
        is taken) and the fit on the FP32 copy -- all inside the timed region"""
        ctx.sync()
        t0 = time.perf_counter()
        s = ctx.hooi_session(A16, ranks, init=dt.INIT_LEARNED, model=model.handle, A_f32=A32, flags=kron,
                             seed=seed)
        r = [s.sweep() for _ in range(k_mixed)]
        r.append(s.sweep())
        ok = abs(r[-1] - r[-2]) < dt.LEARNED_ACCEPT_FIT
        accepted.append(ok)
        if not ok:
            s.close()
            s = ctx.hooi_session(A16, ranks, init=dt.INIT_HOSVD, A_f32=A32)
            for _ in range(k_mixed):
                s.sweep()
        fit_mixed_reported.append(1.0 - s.fit(fp64=False))
        ctx.sync()
        ms = (time.perf_counter() - t0) * 1e3
        # outside the timed region: the same factors' fit with FP64 accumulation
        if check:
            fit_mixed_fp64.append(1.0 - s.fit(fp64=True))
        s.close()
        return ms

    def run_mixed_hosvd_checked():
        """the same with the HOSVD comparison of DT_FLAG_LEARNED_CHECK: the
        learned start is kept only when its first sweep fits at least as well
        as the first sweep from the HOSVD start (that start is computed too)"""
        ctx.sync()
        t0 = time.perf_counter()
        ctx.hooi_dense(A16, ranks, max_iter=k_mixed, init=dt.INIT_LEARNED, model=model.handle, A_f32=A32,
                       flags=kron | dt.FLAG_LEARNED_CHECK, seed=seed, want_factors=False)
        ctx.sync()
        return (time.perf_counter() - t0) * 1e3

    run_plain()
    run_mixed()  # first launches of every kernel
    plain_ms = _median_ms(run_plain)
    mixed_ms = _median_ms(run_mixed)
    run_mixed_hosvd_checked()
    checked_ms = _median_ms(run_mixed_hosvd_checked)
    # breakdown of one mixed run and of one plain run
    ctx.timing(True)
    ctx.timing_reset()
    run_mixed(check=False)
    kt = {k: round(v[1], 3) for k, v in sorted(ctx.timing_list().items(), key=lambda kv: -kv[1][1])[:12]}
    ctx.timing_reset()
    run_plain()
    kt_plain = {k: round(v[1], 3) for k, v in sorted(ctx.timing_list().items(), key=lambda kv: -kv[1][1])[:12]}
    ctx.timing(False)
    model.close()

    def angle(Ua, Ub):
        """sine of the largest principal angle between span(Ua) and span(Ub)"""
        return float(np.linalg.svd(Ua - Ub @ (Ub.T @ Ua), compute_uv=False)[0])

    Ut = [u.download() for u in g["U"]]
    out = {"workload": "dense HOOI 1024^3, core 64^3 decaying 0.93^k, noise 1e-3, ranks (64,64,64)",
           "reference_fit": target,
           "closed_form_fit": 1.0 - noise * math.sqrt(1 - np.prod(ranks) / np.prod(shape)) / math.sqrt(1 + noise ** 2),
           "plain_fp32": {"sweeps": k_plain, "ms": plain_ms, "fit": 1.0 - res_plain[-1], "kernel_ms": kt_plain,
                          "max_sin_angle_vs_reference": max(angle(a, b) for a, b in zip(U_plain, ref["U"])),
                          "max_sin_angle_vs_true": max(angle(a, b) for a, b in zip(U_plain, Ut))},
           "mixed_learned": {"sweeps_bf16": k_mixed, "ms": mixed_ms, "fit": fit_mixed_fp64[-1],
                             "fit_after_first_sweeps": 1.0 - res_mixed[-1],
                             "fit_reported_fp32": fit_mixed_reported[-1],
                             "fit_error_vs_fp64": fit_mixed_reported[-1] - fit_mixed_fp64[-1],
                             "learned_start_accepted": all(accepted),
                             "ms_with_hosvd_check": checked_ms,
                             "max_sin_angle_vs_reference": max(angle(a, b) for a, b in zip(U_mixed, ref["U"])),
                             "max_sin_angle_vs_true": max(angle(a, b) for a, b in zip(U_mixed, Ut)),
                             "kernel_ms": kt},
           "speedup": plain_ms / mixed_ms,
           "speedup_with_hosvd_check": plain_ms / checked_ms,
           "network_kept_epoch": kept_epoch,
           "speedup_source": "sketch start + BF16 sweeps (the network kept its untrained state)" if kept_epoch == 0
           else "sketch start + trained network + BF16 sweeps"}
    for a in (A32, A16):
        a.free()
    log(f"[bench] time-to-fit plain FP32 {plain_ms:.1f} ms ({k_plain} sweeps), mixed+learned "
        f"{mixed_ms:.1f} ms ({k_mixed} BF16 sweeps): {plain_ms / mixed_ms:.2f}x")
    return out


def sparse_zipf(ctx, dt, dims=(1 << 20, 1 << 18, 1 << 16), ndraws=100_000_000, seed=20000905, sweeps=10):
    """Sparse HOOI on the 2^20 x 2^18 x 2^16 COO tensor from 1e8 Zipf(1.1)
    draws per mode (duplicates merged), ranks (16,16,16), 10 sweeps.  The
    TTMc kernel of each mode is timed with HIP events inside the sweeps; its
    algorithmic bytes are nnz (4N + 4) + I_n prod_{m!=n} R_m 4 + sum_{m!=n} I_m R_m 4
    (read the mode-sorted nonzeros, write Y_(n), read the factors once).  Beside
    that rate, each mode carries the HBM traffic the hardware counters measured
    for the kernel (profiles/kernel_counters_sparse_zipf.txt: the factor rows
    that miss L2 are fetched again) over this run's time."""
    log(f"[bench] sparse: generating {dims}, {ndraws:.0e} Zipf(1.1) draws")
    ctx.sync()
    t0 = time.perf_counter()
    idx, val, nnz = dt.sp_gen_zipf(ctx, dims, ndraws, 1.1, seed)
    ctx.sync()
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sp = dt.SpTensor(ctx, dims, idx, val, nnz)
    ctx.sync()
    build_s = time.perf_counter() - t0
    for a in idx + [val]:
        a.free()
    log(f"[bench] sparse: nnz {nnz} (gen {gen_s:.2f} s, sorted copies {build_s:.2f} s)")
    sp.hooi(max_iter=1, seed=seed, want_factors=False)  # first launches
    ctx.timing_filter("sp_ttmc_m")
    ctx.timing(True)
    ctx.timing_reset()
    out = sp.hooi(max_iter=sweeps, seed=seed, want_factors=False)
    kt = ctx.timing_list()
    ctx.timing(False)
    ctx.timing_filter("")
    ctx.timing(True)
    ctx.timing_reset()
    sp.hooi(max_iter=1, seed=seed, want_factors=False)
    breakdown = {k: round(v[1], 3) for k, v in sorted(ctx.timing_list().items(), key=lambda kv: -kv[1][1])[:10]}
    ctx.timing(False)
    sp.close()
    modes = {}
    traffic = measured_sparse_traffic()
    for n in range(3):
        cnt, ms = kt.get(f"sp_ttmc_m{n}", (0, 0.0))
        if not cnt:
            continue
        others = [m for m in range(3) if m != n]
        nbytes = nnz * 16 + dims[n] * 256 * 4 + sum(dims[m] * 16 * 4 for m in others)
        avg = ms / cnt
        modes[f"mode{n}"] = {"launches": cnt, "avg_ms": avg, "bytes": nbytes,
                             "GB_per_s": nbytes / (avg * 1e-3) / 1e9,
                             "frac_of_8TBs": nbytes / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if traffic:
            # the measured HBM traffic (counters) over this run's time
            modes[f"mode{n}"]["hbm_traffic_bytes"] = traffic[n]
            modes[f"mode{n}"]["traffic_frac_of_8TBs"] = traffic[n] / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS
    return {"workload": "sparse HOOI 2^20 x 2^18 x 2^16, 1e8 Zipf(1.1) draws per mode merged, ranks (16,16,16)",
            "nnz": nnz, "sweeps": sweeps, "ms_per_sweep": out["sweeps_ms"] / sweeps,
            "residual": out["residual"], "residuals": out["residuals"], "ttmc": modes,
            "kernel_ms_one_sweep": breakdown, "generate_s": gen_s, "sorted_copies_s": build_s}


def copy_calibration(ctx, dt, nbytes=1 << 30):
    """A plain FP32 copy of known size (the `convert` kernel reads and writes
    nbytes each); profiler counter runs use it to scale FETCH_SIZE/WRITE_SIZE."""
    n = nbytes // 4
    src = ctx.array((n,), dt.DT_F32)
    dst = ctx.array((n,), dt.DT_F32)
    ctx.zero(src)
    ctx.convert(dst, src)
    ctx.timing_filter("convert")
    ctx.timing(True)
    ctx.timing_reset()
    ctx.convert(dst, src)
    ctx.sync()
    cnt, ms = ctx.timing_list().get("convert", (0, 0.0))
    ctx.timing(False)
    ctx.timing_filter("")
    src.free()
    dst.free()
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
            "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9 if ms else None}


def measured_traffic(name):
    """HBM bytes per launch of `name` from the committed counter summary
    profiles/hbm_traffic_dense_256x4.txt (lines `kernel fetch_bytes write_bytes`), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_dense_256x4.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == name:
                    return float(parts[1]) + float(parts[2])
    except OSError:
        return None
    return None


def measured_sparse_traffic():
    """HBM bytes (read + write) per launch of the sparse TTMc of each mode from
    the committed counter summary profiles/kernel_counters_sparse_zipf.txt
    (its `HBM read bytes (GB)` and `HBM write bytes (GB)` rows), or None."""
    path = os.path.join(ROOT, "profiles", "kernel_counters_sparse_zipf.txt")
    rows = {}
    try:
        with open(path) as f:
            for line in f:
                for key in ("HBM read bytes (GB)", "HBM write bytes (GB)"):
                    if line.startswith(key):
                        rows[key] = [float(x) * 1e9 for x in line[len(key):].split()]
    except (OSError, ValueError):
        return None
    if len(rows) != 2 or any(len(v) != 3 for v in rows.values()):
        return None
    return [r + w for r, w in zip(rows["HBM read bytes (GB)"], rows["HBM write bytes (GB)"])]


CPU_HOOI = os.path.join(ROOT, "oracle", "bin", "cpu_hooi")


def cpu_baseline_dense(shape=(256,) * 4, ranks=(16,) * 4):
    """The C++/OpenMP FP64 restatement of the reference's HOOI_ALS schedule
    (oracle/cpu_hooi.cpp: A.copy(), unfold copy -> blocked DGEMM -> fold per
    product, 5-step ALS_) on the host cores at the workload's full size: its
    two kinds of chain are timed as the sweep runs them and the sweep is
    composed as chain 0 + N x chain 1 (a whole FP64 sweep at 256^4 would take
    minutes).  The tensor is N(0,1): the cost does not depend on the values.
    Also reports the host, the thread count and the restatement's own DGEMM
    GFLOP/s at 4096^3."""
    import subprocess
    if not os.path.exists(CPU_HOOI):
        log(f"[bench] {CPU_HOOI} missing (make -C oracle): no CPU baseline")
        return None
    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    cmd = [CPU_HOOI, "pieces", "--shape", ",".join(map(str, shape)), "--ranks", ",".join(map(str, ranks)),
           "--dgemm", "4096"]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True, env=env).stdout
    r = json.loads(out.strip().splitlines()[-1])
    sweep_s = r["sweep_s_composed"]
    return {"value": 1.0 / sweep_s, "unit": "iters/s", "cores": r["threads"], "kind": "port",
            "host": r["host"], "nproc": r["nproc"], "dgemm_gflops_4096": r["dgemm_gflops"],
            "sample": f"C++/OpenMP FP64 restatement of the reference schedule at {shape} ranks {ranks}: "
                      f"chains 0 and 1 of a sweep timed at full size ({r['chain_s'][0]:.2f} s, {r['chain_s'][1]:.2f} s), "
                      f"sweep = chain 0 + {len(shape)} x chain 1 = {sweep_s:.2f} s on {r['threads']} threads",
            "chain1_parts_s": r["parts_chain1_s"], "command": " ".join(["oracle/bin/cpu_hooi"] + cmd[1:])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-sparse", action="store_true")
    ap.add_argument("--section", choices=["all", "dense", "mixed", "sparse"], default="all",
                    help="profiler runs: 'mixed' or 'sparse' runs that section alone and prints its own JSON; "
                         "'dense' runs the dense section alone (no CPU baseline)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build keeps one rank's block on one "
              "MI355X (no collectives); --gpus > 1 is not supported.", file=sys.stderr)
        sys.exit(2)
    import dtucker as dt
    if args.section in ("mixed", "sparse"):
        with dt.Context(0) as ctx:
            out = mixed_1024x3(ctx, dt) if args.section == "mixed" else sparse_zipf(ctx, dt)
        print(json.dumps({"section": args.section, **out}), flush=True)
        return
    if args.section == "dense":
        args.no_cpu = args.no_mixed = args.no_sparse = True
    with dt.Context(0) as ctx:
        r = dense_256x4(ctx, dt, args.steps, args.warmup)
        del r["A"]
        calib = copy_calibration(ctx, dt)
        log(f"[bench] copy calibration: {calib['GB_per_s']:.1f} GB/s")
        c3 = None if args.no_mixed else mixed_1024x3(ctx, dt)
        c4 = None if args.no_sparse else sparse_zipf(ctx, dt)
    steps = args.steps
    value = steps / r["wall"]
    passes = {}
    for k in FULL_PASS_KERNELS:
        cnt, ms = r["ktimes"].get(k, (0, 0.0))
        if cnt:
            avg = ms / cnt
            passes[k] = {"launches": cnt, "avg_launch_ms": avg, "bytes_per_launch": r["kbytes"][k],
                         "GB_per_s": r["kbytes"][k] / (avg * 1e-3) / 1e9}
    # the dominant kernel: the full-pass kernel with the largest total time
    dom_name = max(passes, key=lambda k: passes[k]["avg_launch_ms"] * passes[k]["launches"]) if passes else None
    dom = passes.get(dom_name, {})
    achieved = dom.get("GB_per_s")
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (C++/OpenMP restatement of the reference schedule) ...")
        cpu = cpu_baseline_dense()
        if cpu:
            log(f"[bench] CPU baseline {cpu['value']:.4g} iters/s")
    line = {
        "metric": "HOOI iters/s (dense 256^4 FP32, ranks 16^4)",
        "value": value,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Tucker model generated on the device)",
        "config": {"workload": "dense HOOI 256x256x256x256 FP32, core 16^4, noise 1e-2, "
                               "ranks (16,16,16,16), HOSVD init, dimension-tree schedule "
                               "(2 full-tensor passes per sweep: modes 2 and 3 fused, then modes 0 and 1 fused)",
                   "residual_after": r["res"][-1],
                   "tflops_per_s": r["flops_per_sweep"] * value / 1e12,
                   "init_ms": r["init_ms"],
                   "kernel_ms_per_sweep": r["breakdown"],
                   "full_passes": passes,
                   "copy_calibration": calib,
                   "time_to_fit_mixed_learned_vs_fp32": c3,
                   "sparse_ttmc": c4},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": measured_traffic(dom_name) if dom_name else None,
                     "kernel": dom_name, "bytes_per_launch": dom.get("bytes_per_launch"),
                     "avg_launch_ms": dom.get("avg_launch_ms"), "launches": dom.get("launches")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
