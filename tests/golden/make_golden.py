"""Writes the golden fixtures of tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

The fixtures come from NumPy/SciPy, never from the oracle or the product:

* mttkrp_*.npz  -- numpy.einsum MTTKRP of tiny 3- and 4-way tensors, laid out
                   as PLANC's k x I_n (common/tensor.hpp:246-327).
* nnls.npz      -- scipy.optimize.nnls (Lawson-Hanson) solutions of random
                   normal-equation problems, the reference's BPP solver target
                   (nnls/bppnnls.hpp).
* r2.npz        -- brute force over the four active sets of the two-variable
                   NNLS solved by distnmf/distr2.hpp:35-61.
* nmf_traj.npz  -- three MU and HALS iterations of a 6 x 5, k = 2 problem
                   computed with an explicit NumPy loop over the reference
                   update order (nmf/mu.hpp:51-89, nmf/hals.hpp:50-116).
* csr_subset.npz -- scipy.sparse csr_matrix[:, cols] of a small random matrix
                   (hiernmf/node.hpp:46-64 column selection).
"""
import os

import numpy as np
import scipy.sparse as sp
from scipy.optimize import nnls

OUT = os.path.dirname(os.path.abspath(__file__))
EPS16 = 1e-17


def mttkrp_einsum(X, F):
    N = X.ndim
    subs = "abcdefgh"[:N]
    out = []
    for n in range(N):
        others = [i for i in range(N) if i != n]
        expr = subs + "," + ",".join(subs[o] + "r" for o in others) + "->r" + subs[n]
        out.append(np.einsum(expr, X, *[F[o] for o in others]))
    return out


def make_mttkrp(rng):
    for name, dims, k in (("mttkrp_3way", (3, 4, 5), 3), ("mttkrp_4way", (2, 3, 4, 5), 3)):
        X = rng.random(dims)
        F = [rng.random((d, k)) for d in dims]
        M = mttkrp_einsum(X, F)
        arrs = {"X": X}
        for i, f in enumerate(F):
            arrs[f"F{i}"] = f
        for i, m in enumerate(M):
            arrs[f"M{i}"] = m
        np.savez(os.path.join(OUT, name + ".npz"), **arrs)


def make_nnls(rng):
    k, nrhs = 8, 12
    Wt = rng.random((40, k))
    B = rng.standard_normal((40, nrhs)) + 0.5
    AtA = Wt.T @ Wt
    AtB = Wt.T @ B
    X = np.zeros((k, nrhs))
    for j in range(nrhs):
        X[:, j], _ = nnls(Wt, B[:, j])
    np.savez(os.path.join(OUT, "nnls.npz"), AtA=AtA, AtB=AtB, X=X)


def r2_bruteforce(G, r):
    best, bestx = np.inf, None
    for act in ((1, 1), (1, 0), (0, 1), (0, 0)):
        idx = [i for i in range(2) if act[i]]
        x = np.zeros(2)
        if idx:
            x[idx] = np.linalg.solve(G[np.ix_(idx, idx)], r[idx])
        if np.any(x < 0):
            continue
        f = 0.5 * x @ G @ x - r @ x
        if f < best - 1e-15:
            best, bestx = f, x
    return bestx


def make_r2(rng):
    Wt = rng.random((50, 2))
    G = Wt.T @ Wt
    # A >= 0 and W >= 0 make the right-hand side nonnegative, the only case
    # the rank-2 rule is defined for (it does not clamp).
    R = Wt.T @ (rng.random((50, 40)) ** 4 * rng.random((1, 40)) * 3)
    R[:, ::3] *= np.array([[1.0], [0.02]])
    R[:, 1::5] *= np.array([[0.03], [1.0]])
    X = np.stack([r2_bruteforce(G, R[:, j]) for j in range(R.shape[1])], axis=1)
    np.savez(os.path.join(OUT, "r2.npz"), G=G, R=R, X=X)


def nmf_numpy(algo, A, W, H, iters):
    W = W.copy()
    H = H.copy()
    normA = np.linalg.norm(A)
    rel = []
    k = W.shape[1]

    def normalize_by_w(W, H):
        n2 = np.sqrt((W ** 2).sum(0))
        for i in range(k):
            if n2[i] > 0:
                W[:, i] /= n2[i]
                H[:, i] *= n2[i]

    if algo == 1:
        normalize_by_w(W, H)
    for _ in range(iters):
        if algo == 0:
            AtW = A.T @ W
            WtW = W.T @ W
            H = H * AtW / (H @ WtW + EPS16)
            AH = A @ H
            HtH = H.T @ H
            W = W * AH / (W @ HtH + EPS16)
        else:
            WtA = W.T @ A
            WtW = W.T @ W
            for x in range(k):
                hx = H[:, x] + (WtA[x, :] - H @ WtW[:, x])
                hx = np.where(hx < EPS16, EPS16, hx)
                if np.linalg.norm(hx) != 0:
                    H[:, x] = hx
            AH = A @ H
            HtH = H.T @ H
            for x in range(k):
                wx = W[:, x] * HtH[x, x] + (AH[:, x] - W @ HtH[:, x])
                wx = np.where(wx < EPS16, EPS16, wx)
                nc = np.linalg.norm(wx)
                if nc != 0:
                    W[:, x] = wx / nc
            normalize_by_w(W, H)
        rel.append(np.linalg.norm(A - W @ H.T) / normA)
    if algo == 0:
        normalize_by_w(W, H)
    return W, H, np.array(rel)


def make_nmf(rng):
    A = rng.random((6, 5))
    W0 = rng.random((6, 2))
    H0 = rng.random((5, 2))
    arrs = {"A": A, "W0": W0, "H0": H0}
    for algo, name in ((0, "mu"), (1, "hals")):
        W, H, rel = nmf_numpy(algo, A, W0, H0, 3)
        arrs[f"{name}_W"] = W
        arrs[f"{name}_H"] = H
        arrs[f"{name}_rel"] = rel
    np.savez(os.path.join(OUT, "nmf_traj.npz"), **arrs)


def make_csr(rng):
    A = sp.random(30, 40, density=0.15, format="csr", random_state=np.random.RandomState(3))
    A.data = np.exp(rng.standard_normal(A.nnz)).astype(np.float32)
    cols = np.sort(rng.choice(40, 17, replace=False)).astype(np.int64)
    S = A[:, cols].tocsr()
    S.sort_indices()
    np.savez(os.path.join(OUT, "csr_subset.npz"), indptr=A.indptr.astype(np.int64),
             indices=A.indices.astype(np.int32), data=A.data, shape=np.array(A.shape),
             cols=cols, sub_indptr=S.indptr.astype(np.int64),
             sub_indices=S.indices.astype(np.int32), sub_data=S.data)


def main():
    rng = np.random.default_rng(20240611)
    make_mttkrp(rng)
    make_nnls(rng)
    make_r2(rng)
    make_nmf(rng)
    make_csr(rng)


if __name__ == "__main__":
    main()
