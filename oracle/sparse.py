"""Sparse path restated in numpy -- TEST INFRASTRUCTURE ONLY.

* Zipf COO generator with duplicate merging (no reference counterpart; the
  reference reads .tns files, src/read_data.cpp:36-129): P(k) ~ (k+1)^-s on
  [0, I_n), cumulative table in FP64 (sequential sums, each entry divided by
  the total), index = first k with cdf[k] >= u, u = Philox uniform stream
  7000 + n; values Philox uniform stream 7100 in FP32; duplicates merged in
  increasing composite-key order (mode 0 lowest), value = FP64 sum in draw
  order rounded to FP32.
* Stable sort by one mode: the SpTensor constructor's counting sort
  (include/sp_tensor.hpp:117-147) == argsort(kind="stable").
* TTMc: Y_(n)[i, r_a + R_a r_b] = sum over nonzeros of slice i of
  val * U_a[i_a, r_a] * U_b[i_b, r_b] (ttmc_mTR / ttmc_mTL with
  add_outer_product_all, template/function/sp_tensor.tpp:236-417, as intended:
  the Kronecker row of a slice reset per slice and paired with row i of U_n).
* ttmc_mTR / ttmc_mTL (sp_tensor.tpp:236-365, as intended): Y_(n)ᵀ U_n and
  (Y_(n) L)ᵀ.
* Duplicate coordinates merged at their first occurrence (FP64 sums in input
  order), the rest in input order.
* Sparse HOOI (sp_hooi_als.tpp:48-118): per sweep and mode n,
  U_n <- orth(Y_(n) Y_(n)ᵀ U_n) with orth = QR with positive R diagonal (the
  span and basis modified Gram-Schmidt (orth2) gives); core = Y_(N-1)ᵀ U_{N-1}.
"""
import numpy as np

from . import dense as D
from . import philox


def zipf_cdf(dim, s):
    p = np.power(np.arange(1, dim + 1, dtype=np.float64), -float(s))
    c = np.cumsum(p)
    return c / c[-1]


def bits_for(dim):
    b = 0
    while b < 62 and (1 << b) < dim:
        b += 1
    return max(b, 1)


def zipf_draws(dims, ndraws, s, seed):
    idx = []
    for n, I in enumerate(dims):
        u = philox.uniform(ndraws, seed, 7000 + n)
        idx.append(np.searchsorted(zipf_cdf(I, s), u, side="left").astype(np.int64))
    vals = philox.uniform(ndraws, seed, 7100).astype(np.float32)
    return idx, vals


def merge_duplicates(dims, idx, vals):
    key = np.zeros(len(vals), dtype=np.uint64)
    shift = 0
    for n, I in enumerate(dims):
        key |= idx[n].astype(np.uint64) << np.uint64(shift)
        shift += bits_for(I)
    order = np.argsort(key, kind="stable")
    ks = key[order]
    heads = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    ends = np.r_[heads[1:], len(ks)]
    v = vals[order].astype(np.float64)
    out = np.empty(len(heads), dtype=np.float32)
    for j, (h, e) in enumerate(zip(heads, ends)):
        out[j] = np.float32(np.cumsum(v[h:e])[-1])   # sequential FP64 sum
    uidx = [idx[n][order[heads]].astype(np.int32) for n in range(len(dims))]
    return uidx, out


def gen_zipf(dims, ndraws, s, seed):
    idx, vals = zipf_draws(dims, ndraws, s, seed)
    return merge_duplicates(dims, idx, vals)


def sort_mode(keys, dim):
    perm = np.argsort(np.asarray(keys), kind="stable").astype(np.int32)
    rowptr = np.r_[0, np.cumsum(np.bincount(np.asarray(keys), minlength=dim))].astype(np.int32)
    return perm, rowptr


def densify(dims, idx, vals):
    A = np.zeros(dims, dtype=np.float64)
    np.add.at(A, tuple(np.asarray(i) for i in idx), np.asarray(vals, dtype=np.float64))
    return A


def ttmc(dims, idx, vals, Us, n):
    """Y_(n) (I_n x prod R_{-n}) with the other modes' entry index r_a + R_a r_b ..."""
    others = [m for m in range(len(dims)) if m != n]
    rows = np.asarray(idx[n])
    K = np.asarray(vals, dtype=np.float64)[:, None]
    for m in others:  # Kronecker row, first other mode fastest
        Um = np.asarray(Us[m], dtype=np.float64)[np.asarray(idx[m])]
        K = (K[:, :, None] * Um[:, None, :]).reshape(len(rows), -1, order="F")
    Y = np.zeros((dims[n], K.shape[1]))
    np.add.at(Y, rows, K)
    return Y


def ttmc_mtr(dims, idx, vals, Us, n):
    """Function::ttmc_mTR (sp_tensor.tpp:236-286) as intended: L = Y_(n)ᵀ U_n,
    K x R_n with K in the Kronecker order of ttmc."""
    return ttmc(dims, idx, vals, Us, n).T @ np.asarray(Us[n], dtype=np.float64)


def ttmc_mtl(dims, idx, vals, Us, n, L):
    """Function::ttmc_mTL (sp_tensor.tpp:288-365) as intended: (Y_(n) L)ᵀ,
    R_n x I_n."""
    return (ttmc(dims, idx, vals, Us, n) @ np.asarray(L, dtype=np.float64)).T


def merge_first_occurrence(idx, vals):
    """Equal coordinates merged: one entry per coordinate, at its first
    occurrence, value = FP64 sum of its entries in input order; the survivors
    keep their input order (ingest step after read_data, read_data.cpp:80-127)."""
    coords = np.stack([np.asarray(i, dtype=np.int64) for i in idx], axis=1)
    _, first, inv = np.unique(coords, axis=0, return_index=True, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    sums = np.zeros(len(first))
    v = np.asarray(vals, dtype=np.float64)
    for k in range(len(v)):  # sequential: input order
        sums[inv[k]] += v[k]
    order = np.argsort(first, kind="stable")
    keep = first[order]
    return [np.asarray(i)[keep] for i in idx], sums[order]


def hooi(dims, idx, vals, U0, sweeps, lnorms=None):
    """Sparse HOOI from factors U0 (orthonormal columns); residual per sweep.
    When lnorms is a list, ||L||_F of every mode and sweep is appended to it
    (L = Y_(n)ᵀ U_n before the update: what Sp_TTMC_ALS_ prints as ||G||_F,
    sp_hooi_als.tpp:51-53)."""
    Us = [np.asarray(u, dtype=np.float64) for u in U0]
    normA2 = float(np.sum(np.asarray(vals, dtype=np.float64) ** 2))
    res = []
    N = len(dims)
    for _ in range(sweeps):
        for n in range(N):
            Y = ttmc(dims, idx, vals, Us, n)
            L = Y.T @ Us[n]
            if lnorms is not None:
                lnorms.append(float(np.linalg.norm(L)))
            Us[n] = D.orth_posdiag(Y @ L)
        Y = ttmc(dims, idx, vals, Us, N - 1)
        G = Y.T @ Us[N - 1]
        res.append(np.sqrt(max(0.0, 1.0 - float(np.sum(G * G)) / normA2)))
    return Us, G, res
