"""FP64 numpy restatement of the reference's dense HOOI-ALS path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Conventions (reference): tensors are column-major, first index fastest
(template/tensor_mpi.tpp:49-53); here a tensor is a numpy array of its
natural shape whose flat Fortran-order view is the reference buffer.
"""
from __future__ import annotations

import numpy as np

from .glibc_rand import GlibcRand, randn_fill
from . import philox


# ---------------------------------------------------------------------------
# Layout maps (template/operator/operator_cpu.tpp)
def tenmat(A: np.ndarray, n: int) -> np.ndarray:
    """A_(n): I_n x prod(others), columns ordered lower modes fastest
    (Operator::tenmat, operator_cpu.tpp:56-72)."""
    return np.moveaxis(A, n, 0).reshape(A.shape[n], -1, order="F")


def tenmatt(A: np.ndarray, n: int) -> np.ndarray:
    """A_(n)^T (Operator::tenmatt, operator_cpu.tpp:83-103)."""
    return tenmat(A, n).T


def mattten(B: np.ndarray, shape, n: int) -> np.ndarray:
    """Inverse of tenmatt: B is prod(others) x I_n (Operator::mattten,
    operator_cpu.tpp:115-135)."""
    shape = list(shape)
    rest = shape[:n] + shape[n + 1:]
    T = B.T.reshape([shape[n]] + rest, order="F")
    return np.moveaxis(T, 0, n)


def tenmat_loops(flat, shape, n):
    """Literal loop transcription of Operator::tenmat (operator_cpu.tpp:56-72)
    on a flat column-major buffer; used only on tiny shapes to pin tenmat()."""
    size = int(np.prod(shape))
    br = shape[n]
    bc = int(np.prod(shape[:n])) if n > 0 else 1
    B = [0.0] * size
    for idx in range(0, size, br * bc):
        for i in range(bc):
            for j in range(br):
                B[idx + i * br + j] = flat[idx + j * bc + i]
    return B


def tenmatt_loops(flat, shape, n):
    """Literal loop transcription of Operator::tenmatt (operator_cpu.tpp:83-103)."""
    size = int(np.prod(shape))
    br = shape[n]
    bc = int(np.prod(shape[:n])) if n > 0 else 1
    col = size // br
    B = [0.0] * size
    for idx in range(0, size, br * bc):
        for j in range(br):
            for i in range(bc):
                B[j * col + idx // br + i] = flat[idx + j * bc + i]
    return B


# ---------------------------------------------------------------------------
# Function layer (template/function/tensor.tpp, matrix.tpp)
def ttm(A: np.ndarray, M: np.ndarray, n: int) -> np.ndarray:
    """Function::ttm (tensor.tpp:272-335) at one rank: tenmatt, then
    matmulNT(C, B, M, remain, J, I_n) = B M^T, then mattten."""
    B = tenmatt(A, n)
    Cm = B @ M.T
    shape = list(A.shape)
    shape[n] = M.shape[0]
    return mattten(Cm, shape, n)


def gram(A: np.ndarray, n: int) -> np.ndarray:
    """Function::gram (tensor.tpp:18-120) at one rank: A_(n) A_(n)^T."""
    X = tenmat(A, n)
    return X @ X.T


def ttt_except(A: np.ndarray, B: np.ndarray, n: int) -> np.ndarray:
    """Function::ttt_except (tensor.tpp:132-259) at one rank: A_(n) B_(n)^T."""
    return tenmat(A, n) @ tenmat(B, n).T


def reduced_qr(A: np.ndarray):
    """Function::reduced_QR -> Operator::QR (matrix.tpp:85-100,
    operator_double.cpp:131-155): Householder QR (LAPACK dgeqrf/dorgqr)."""
    assert A.shape[0] >= A.shape[1]
    return np.linalg.qr(A, mode="reduced")


def inverse(A: np.ndarray) -> np.ndarray:
    """Function::inverse -> dgetrf + dgetri (operator_double.cpp:84-103)."""
    return np.linalg.inv(A)


def fnorm(A: np.ndarray) -> float:
    """Function::fnorm (tensor.tpp:467-480): sqrt of the FP64 sum of squares."""
    return float(np.sqrt(np.sum(np.asarray(A, dtype=np.float64) ** 2)))


# ---------------------------------------------------------------------------
# Algorithm layer (template/algorithm/tucker/hooi_als.tpp)
def als_(Y: np.ndarray, n: int, L_initial: np.ndarray, max_iter: int = 5) -> np.ndarray:
    """Algorithm::Tucker::ALS_ (hooi_als.tpp:7-34), step by step."""
    L = L_initial.copy()
    for it in range(max_iter):
        if it == 0:
            LG_inv = L
        else:
            G = L.T @ L                                  # matmulTN(L, L)
            LG_inv = L @ inverse(G)                      # matmulNN(L, G_inv)
        G_invLt = LG_inv.T                               # transpose
        G_invLtY = ttm(Y, G_invLt, n)
        YYtLG_inv = ttt_except(Y, G_invLtY, n)
        G_R = LG_inv.T @ YYtLG_inv                       # matmulTN
        L = YYtLG_inv @ inverse(G_R)                     # matmulNN(., G_R_inv)
    q, _ = reduced_qr(L)
    return q


def reference_random_factors(shape, ranks, seed: int):
    """Start factors of HOOI_ALS (hooi_als.tpp:76-81) after srand(seed):
    for each mode, U_rand.randn() in buffer order, then reduced_QR.
    Returns (raw Gaussian matrices, orthonormal factors)."""
    g = GlibcRand(seed)
    raws, Us = [], []
    for I, R in zip(shape, ranks):
        raw = np.array(randn_fill(g, I * R), dtype=np.float64).reshape(I, R, order="F")
        raws.append(raw)
        Us.append(reduced_qr(raw)[0])
    return raws, Us


def hosvd_factors(A: np.ndarray, ranks):
    """Leading R_n eigenvectors of A_(n) A_(n)^T (no reference counterpart)."""
    Us = []
    for n, R in enumerate(ranks):
        w, V = np.linalg.eigh(gram(A, n))
        Us.append(V[:, ::-1][:, :R].copy())
    return Us


def core(A: np.ndarray, Us) -> np.ndarray:
    G = A
    for n, U in enumerate(Us):
        G = ttm(G, U.T, n)
    return G


def hooi_als(A: np.ndarray, ranks, max_iter: int, U_init, inner_iter: int = 5):
    """Algorithm::Tucker::HOOI_ALS (hooi_als.tpp:64-123) from given start
    factors.  Returns (G, U, residual per sweep) with the reference's printed
    residual sqrt(1 - ||G||^2/||A||^2) (hooi_als.tpp:110-114)."""
    A = np.asarray(A, dtype=np.float64)
    N = A.ndim
    U = [u.copy() for u in U_init]
    A_norm = fnorm(A)
    res = []
    for _ in range(max_iter):
        for n in range(N):
            Y = A.copy()
            for i in range(N):
                if i == n:
                    continue
                Y = ttm(Y, U[i].T, i)
            U[n] = als_(Y, n, U[n], inner_iter)
        G = core(A, U)
        G_norm = fnorm(G)
        res.append(float(np.sqrt(max(0.0, 1 - (G_norm * G_norm) / (A_norm * A_norm)))))
    G = core(A, U)
    return G, U, res


# ---------------------------------------------------------------------------
# Metrics
def sin_max_angle(Ua: np.ndarray, Ub: np.ndarray) -> float:
    """sin of the largest principal angle between span(Ua) and span(Ub):
    ||(I - Ub Ub^T) Ua||_2 for orthonormal Ua, Ub (no arccos near 1)."""
    Qa, _ = np.linalg.qr(Ua)
    Qb, _ = np.linalg.qr(Ub)
    D = Qa - Qb @ (Qb.T @ Qa)
    return float(np.linalg.norm(D, 2))


def reconstruct(G: np.ndarray, Us) -> np.ndarray:
    X = G
    for n, U in enumerate(Us):
        X = ttm(X, U, n)
    return X


def direct_residual(A, G, Us) -> float:
    """||A - G x_1 U_1 ... x_N U_N|| / ||A|| in FP64."""
    A = np.asarray(A, dtype=np.float64)
    return float(np.linalg.norm((A - reconstruct(G, Us)).ravel()) / np.linalg.norm(A.ravel()))


# ---------------------------------------------------------------------------
# Synthetic Tucker model (mirrors dt_gen_tucker, diana-tucker_amd/csrc/dt_gen.hip)
def orth_posdiag(Z: np.ndarray) -> np.ndarray:
    """Q of the QR of Z with a positive R diagonal (what Cholesky-QR returns)."""
    Q, R = np.linalg.qr(Z)
    s = np.sign(np.diag(R))
    s[s == 0] = 1
    return Q * s


def tucker_model(shape, ranks, noise, seed, decay=0.0):
    N = len(shape)
    Us = [orth_posdiag(philox.normal(I * R, seed, 2000 + n).reshape(I, R, order="F"))
          for n, (I, R) in enumerate(zip(shape, ranks))]
    G = philox.normal(int(np.prod(ranks)), seed, 1000).reshape(ranks, order="F")
    if decay > 0:
        for n in range(N):
            d = decay ** np.arange(ranks[n], dtype=np.float64)
            sh = [1] * N
            sh[n] = ranks[n]
            G = G * d.reshape(sh)
    X = reconstruct(G, Us).astype(np.float32)
    normX2 = float(np.sum(X.astype(np.float64) ** 2))
    A = X.astype(np.float64)
    if noise > 0:
        E = philox.normal(int(np.prod(shape)), seed, 3000).reshape(shape, order="F")
        scale = noise * np.sqrt(normX2 / float(np.sum(E * E)))
        A = A + scale * E
    return A.astype(np.float32), Us, G
