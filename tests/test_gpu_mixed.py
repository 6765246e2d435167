"""Mixed precision (BF16 tensor, FP32 accumulation, FP64 eigen/QR/fit) on the
GPU through the C ABI: the full 1024^3 rank-64 workload against its closed
form and true factors, and a reduced BF16 HOOI against the FP64 oracle run
on the same bf16-rounded tensor from the same start."""
import math
import os
import sys

import numpy as np
import pytest

import dtucker as dt
from oracle import dense as D

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "diana-tucker_amd", "python"))


def sin_max(Ua, Ub):
    return float(np.linalg.svd(Ua - Ub @ (Ub.T @ Ua), compute_uv=False)[0])


def test_mixed_learned_1024cube_rank64_full_size(ctx):
    """1024^3 stored in BF16 with its FP32 copy, core 64^3 with mode spectra
    decaying as 0.93^k, noise 1e-3, ranks (64, 64, 64): the learned start
    (Kronecker sketch of the BF16 tensor -> MLP -> QR) and BF16 sweeps until
    the FP64 fit is within 1e-4 of the closed form
        fit* = 1 - eps sqrt(1 - prod R / prod I) / sqrt(1 + eps^2)
    (noise orthogonal to the model's subspace up to O(eps/sqrt(I))).
    Checks: the FP64-accumulated fit on the FP32 copy within 1e-4 of fit*; the
    largest principal-angle sine to the true factors at the noise level
    (first-order perturbation ~ eps ||X|| / sigma_64(X_(n)), observed ~1.1e-3;
    bound 5e-3); the FP32-copy fit (FP32 core with the factor-copy
    correction) within 1e-4 of the FP64 fit -- one FP32 ulp of ||G||^2
    (relative 2^-24 = 6e-8) moves sqrt(1 - ||G||^2/||A||^2) by
    6e-8 / (2 * 1e-3) = 3e-5 at this residual, and the core's three FP32
    contractions (the first over K = 1024) leave ||G||^2 a few ulps off:
    -9e-6 after the first sweep and +5e-5 after the second have been
    measured, so the bound is about three ulps."""
    import learned as LN
    shape, ranks, eps, decay, seed = (1024,) * 3, (64,) * 3, 1e-3, 0.93, 20000905 + 2
    print("C3 full size: generating", shape, flush=True)
    g = ctx.gen_tucker(shape, ranks, eps, seed=seed, decay=decay, bf16=True, want_factors=True)
    A32, A16 = g["A"], g["A_bf16"]
    fit_star = 1.0 - eps * math.sqrt(1 - np.prod(ranks) / np.prod(shape)) / math.sqrt(1 + eps ** 2)
    model = LN.load_model(ctx, os.path.join(LN.MODEL_DIR, "learned_tucker256_r64.npz"))
    s = ctx.hooi_session(A16, ranks, init=dt.INIT_LEARNED, model=model.handle, A_f32=A32,
                         flags=dt.FLAG_SKETCH_KRON, seed=seed)
    fit64 = None
    for k in range(3):
        s.sweep()
        fit64 = 1.0 - s.fit(fp64=True)
        print("C3 sweep", k + 1, "fit64", fit64, "closed form", fit_star, flush=True)
        if abs(fit64 - fit_star) < 1e-4:
            break
    fit32 = 1.0 - s.fit(fp64=False)
    out = s.result(want_factors=True)
    s.close()
    model.close()
    Ut = [u.download() for u in g["U"]]
    angles = [sin_max(a, b) for a, b in zip(out["U"], Ut)]
    print("C3 fit64", fit64, "fit32", fit32, "closed form", fit_star, "angles", angles, flush=True)
    assert abs(fit64 - fit_star) < 1e-4
    assert max(angles) < 5e-3
    assert abs(fit32 - fit64) < 1e-4
    for a in (A32, A16):
        a.free()


def test_bf16_hooi_matches_oracle_on_rounded_tensor(ctx):
    """BF16 HOOI (256 x 192 x 160, ranks 16, noise 1e-3, 3 sweeps; modes 0 and
    1 take the fused BF16 pair) against oracle/dense.py:hooi_als on the same
    bf16-rounded tensor in FP64 from the same (oracle HOSVD) start.  What
    differs: FP32 accumulation (~1e-6 relative in each Y_(n)) and the bf16
    rounding of the factor in the product that reads the tensor (|delta| <=
    2^-9 per entry).  A factor perturbation inside span(U_n) leaves span(Y_(n))
    unchanged; the part outside it only meets the noise, so its effect on the
    next factor is ~ 2^-9 sqrt(R) eps ||A|| / sigma_R(X_(n)) ~ 2e-5, and the fit,
    stationary in the factors, moves by its square (~4e-10).  Bounds: relative
    fit 1e-8, principal-angle sine 1e-4 (observed 1e-11 and 2e-7 from the
    converged HOSVD start)."""
    shape, ranks, eps = (256, 192, 160), (16, 16, 16), 1e-3
    g = ctx.gen_tucker(shape, ranks, eps, seed=77, bf16=True)
    A16 = g["A_bf16"]
    Ab = dt.f32_from_bf16(A16.download()).astype(np.float64)
    U0 = D.hosvd_factors(Ab, ranks)
    _, Ur, res_ref = D.hooi_als(Ab, ranks, 3, U0)
    out = ctx.hooi_dense(A16, ranks, max_iter=3, init=dt.INIT_GIVEN, U_init=[ctx.to_device(u) for u in U0],
                         flags=dt.FLAG_FP64_FIT)
    fit_gpu, fit_ref = 1 - out["residuals"][-1], 1 - res_ref[-1]
    angles = [sin_max(out["U"][n], Ur[n]) for n in range(3)]
    print("bf16 vs oracle fit", fit_gpu, fit_ref, "angles", angles, flush=True)
    assert abs(fit_gpu - fit_ref) / fit_ref < 1e-8
    assert max(angles) < 1e-4
