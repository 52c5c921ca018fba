"""CPU tests that pin the FP64 oracle (oracle/planc_oracle.cpp) against the
committed NumPy/SciPy fixtures, closed forms and invariants."""
import os

import numpy as np
import pytest
from scipy.optimize import nnls

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_gemm_matches_numpy(ta, tb):
    rng = np.random.default_rng(1)
    M, N, K = 37, 19, 5000  # long K exercises the split-K path
    A = rng.random((M, K) if ta == "N" else (K, M))
    B = rng.random((K, N) if tb == "N" else (N, K))
    ref = (A if ta == "N" else A.T) @ (B if tb == "N" else B.T)
    np.testing.assert_allclose(O.gemm(ta, tb, A, B), ref, rtol=1e-12, atol=1e-10)


@pytest.mark.parametrize("name", ["mttkrp_3way", "mttkrp_4way"])
def test_mttkrp_matches_einsum_fixture(name):
    g = load(name + ".npz")
    X = np.asfortranarray(g["X"])
    N = X.ndim
    F = [g[f"F{i}"] for i in range(N)]
    for n in range(N):
        ref = g[f"M{n}"]
        np.testing.assert_allclose(O.mttkrp(X, F, n, method=0), ref, rtol=1e-13, atol=1e-13)
        for split in range(N - 1):
            np.testing.assert_allclose(O.mttkrp(X, F, n, method=1, split=split), ref, rtol=1e-13,
                                       atol=1e-13)


def test_mttkrp_rank1_closed_form():
    # X = a o b o c  ->  M_0 = a * ((b^T B) o (c^T C))  (k x I_0 layout: transposed)
    rng = np.random.default_rng(2)
    a, b, c = rng.random(4), rng.random(5), rng.random(6)
    X = np.asfortranarray(np.einsum("i,j,k->ijk", a, b, c))
    F = [rng.random((4, 3)), rng.random((5, 3)), rng.random((6, 3))]
    ref = np.outer((b @ F[1]) * (c @ F[2]), a)
    np.testing.assert_allclose(O.mttkrp(X, F, 0), ref, rtol=1e-13)
    np.testing.assert_allclose(O.mttkrp(X, F, 0, method=1), ref, rtol=1e-13)


def test_krp_row_order_lowest_mode_fastest():
    rng = np.random.default_rng(3)
    F = [rng.random((d, 2)) for d in (2, 3, 4)]
    K = O.krp_leave_out_one(F, 0)
    for i1 in range(3):
        for i2 in range(4):
            np.testing.assert_allclose(K[i1 + 3 * i2], F[1][i1] * F[2][i2], rtol=1e-15)
    K = O.krp_leave_out_one(F, 1)
    for i0 in range(2):
        for i2 in range(4):
            np.testing.assert_allclose(K[i0 + 2 * i2], F[0][i0] * F[2][i2], rtol=1e-15)


def test_bppnnls_matches_scipy_fixture():
    g = load("nnls.npz")
    X, it = O.bppnnls(g["AtA"], g["AtB"])
    assert it > 0
    np.testing.assert_allclose(X, g["X"], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("k,nrhs", [(1, 5), (4, 1), (12, 300), (32, 2500)])
def test_bppnnls_kkt_and_scipy(k, nrhs):
    rng = np.random.default_rng(k * 1000 + nrhs)
    Wt = rng.random((3 * k + 5, k))
    B = rng.standard_normal((3 * k + 5, nrhs)) + 0.3
    AtA, AtB = Wt.T @ Wt, Wt.T @ B
    X, it = O.bppnnls(AtA, AtB)
    assert it >= 0
    Y = AtA @ X - AtB
    assert X.min() >= 0
    assert Y.min() >= -1e-9 * np.abs(AtB).max()
    assert np.abs((X * Y).sum(0)).max() < 1e-8 * np.abs(AtB).max() ** 2
    for j in range(min(nrhs, 20)):
        x, _ = nnls(Wt, B[:, j])
        np.testing.assert_allclose(X[:, j], x, rtol=1e-8, atol=1e-9)


def test_bppnnls_reports_nonconvergence_for_indefinite_gram():
    # An indefinite "Gram" makes the pivoting cycle; the reference exits with
    # EXIT_FAILURE (nnls/bppnnls.hpp:185-188); the oracle returns -1.
    AtA = np.array([[1.0, 2.0], [2.0, 1.0]])
    AtB = np.array([[1.0], [1.0]])
    X, it = O.bppnnls(AtA, AtB)
    assert it == -1 or (X.min() >= 0)


def test_r2_update_matches_bruteforce_fixture():
    g = load("r2.npz")
    np.testing.assert_allclose(O.r2_update(g["G"], g["R"]), g["X"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("algo,name", [(0, "mu"), (1, "hals")])
def test_nmf_matches_numpy_trajectory_fixture(algo, name):
    g = load("nmf_traj.npz")
    W, H, rel, obj, rc = O.nmf(algo, g["A"], g["W0"], g["H0"], 3)
    assert rc == 0
    np.testing.assert_allclose(W, g[f"{name}_W"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(H, g[f"{name}_H"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(rel, g[f"{name}_rel"], rtol=1e-9)


def _lowrank(rng, m, n, k, noise=0.01):
    A = rng.random((m, k)) @ rng.random((k, n))
    return A + noise * np.abs(rng.standard_normal((m, n)))


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_nmf_error_decreases_and_matches_direct_norm(algo):
    rng = np.random.default_rng(10 + algo)
    A = _lowrank(rng, 60, 45, 5)
    W0, H0 = rng.random((60, 5)), rng.random((45, 5))
    W, H, rel, obj, rc = O.nmf(algo, A, W0, H0, 12)
    assert rc == 0
    # testcases.md:3 -- the objective error decreases at every iteration.
    assert np.all(np.diff(rel) <= 1e-12)
    # the printed relative error equals the directly computed one
    np.testing.assert_allclose(rel[-1], np.linalg.norm(A - W @ H.T) / np.linalg.norm(A),
                               rtol=1e-8)
    # the final normalize_by_W leaves unit-norm columns of W
    np.testing.assert_allclose(np.linalg.norm(W, axis=0), 1.0, rtol=1e-12)


def test_mu_fixed_point_at_exact_factorisation():
    rng = np.random.default_rng(5)
    W0, H0 = rng.random((20, 3)), rng.random((15, 3))
    A = W0 @ H0.T
    W, H, rel, obj, rc = O.nmf(0, A, W0, H0, 2)
    np.testing.assert_allclose(W @ H.T, A, rtol=1e-12)
    # fit^2 = ||A||^2 - 2 tr + tr cancels to rounding; common/nmf.hpp:462
    # prints ||A|| (relative error 1) when that rounding is negative.
    assert rel[-1] < 1e-6 or rel[-1] == 1.0


def test_bpp_nmf_halfsteps_satisfy_kkt():
    rng = np.random.default_rng(6)
    A = _lowrank(rng, 40, 30, 4)
    W0, H0 = rng.random((40, 4)), rng.random((30, 4))
    W, H, rel, obj, rc = O.nmf(2, A, W0, H0, 1)
    # After one iteration H solves min ||A - W H^T|| s.t. H >= 0 for this W
    # (normalize_by_W rescales W and H by the same diagonal, which keeps KKT).
    G = W.T @ W
    Y = H @ G - A.T @ W
    assert H.min() >= 0
    assert Y.min() > -1e-9
    assert np.abs((H * Y).sum()) < 1e-8


def test_hals_column_step_is_coordinate_minimiser():
    rng = np.random.default_rng(7)
    A = _lowrank(rng, 25, 20, 3)
    W0, H0 = rng.random((25, 3)), rng.random((20, 3))
    # one iteration; then perturbing any single column of H away from the
    # iterate cannot lower the objective by more than rounding: the last
    # W sweep uses H, and H's last column step was exact w.r.t. the W before
    # that sweep -- check the H half-step directly with W fixed instead.
    W, H, rel, obj, rc = O.nmf(1, A, W0, H0, 1)
    W1 = W / np.linalg.norm(W, axis=0)
    # rerun one H sweep from (W1, H) in numpy-free form: oracle with W fixed is
    # not exposed, so check the optimality of the final column of H against W1
    # scaled back: column k-1 of H minimises ||A - W H^T|| given the others.
    Hs = H * np.linalg.norm(W, axis=0)
    x = Hs.shape[1] - 1
    R = A - W1 @ Hs.T + np.outer(W1[:, x], Hs[:, x])

    def f(h):
        return np.linalg.norm(R - np.outer(W1[:, x], h)) ** 2

    best = np.maximum(R.T @ W1[:, x], 0)  # exact minimiser for a unit column
    assert f(best) <= f(best + 1e-3 * rng.standard_normal(best.shape).clip(0)) + 1e-12


@pytest.mark.parametrize("algo", [0, 1, 2])
@pytest.mark.parametrize("dimtree", [0, 1])
def test_ntf_error_decreases_and_matches_reconstruction(algo, dimtree):
    rng = np.random.default_rng(20 + algo)
    dims, k = (6, 7, 8), 3
    F0 = [rng.random((d, k)) for d in dims]
    X = np.einsum("ir,jr,kr->ijk", *F0) + 0.01 * np.abs(rng.standard_normal(dims))
    X = np.asfortranarray(X)
    init = [rng.random((d, k)) for d in dims]
    F, lam, rel, rc = O.ntf(algo, X, init, 15, dimtree=dimtree)
    assert rc == 0
    if algo != 0:
        assert np.all(np.diff(rel) <= 1e-10)
    Xh = np.einsum("r,ir,jr,kr->ijk", lam, *F)
    np.testing.assert_allclose(rel[-1], np.linalg.norm(X - Xh) / np.linalg.norm(X), rtol=1e-9)
    # every factor column is unit norm after normalize(mode), except columns
    # that BPP drove to zero, which normalize skips (common/ncpfactors.hpp:349)
    for f in F:
        nr = np.linalg.norm(f, axis=0)
        assert np.all((np.abs(nr - 1.0) < 1e-12) | (nr == 0.0))


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_ntf_dimtree_equals_direct(algo):
    # distntf/testcases.md:3-67 records identical trajectories with and without --dimtree.
    rng = np.random.default_rng(30)
    dims, k = (5, 6, 4, 3), 2
    X = np.asfortranarray(rng.random(dims))
    init = [rng.random((d, k)) for d in dims]
    _, _, r0, _ = O.ntf(algo, X, init, 6, dimtree=0)
    _, _, r1, _ = O.ntf(algo, X, init, 6, dimtree=1)
    np.testing.assert_allclose(r0, r1, rtol=1e-10)


def test_cp_loss_gradient_central_differences():
    rng = np.random.default_rng(40)
    dims, k = (4, 3, 5), 2
    X = np.asfortranarray(rng.random(dims))
    F = [rng.random((d, k)) for d in dims]
    loss, G = O.cp_loss(X, F, grads=True)
    ref = 0.5 * np.linalg.norm(X - np.einsum("ir,jr,kr->ijk", *F)) ** 2
    np.testing.assert_allclose(loss, ref, rtol=1e-12)
    h = 1e-6
    for n in range(3):
        for (i, c) in [(0, 0), (dims[n] - 1, 1), (1, 1)]:
            Fp = [f.copy() for f in F]
            Fm = [f.copy() for f in F]
            Fp[n][i, c] += h
            Fm[n][i, c] -= h
            fd = (O.cp_loss(X, Fp) - O.cp_loss(X, Fm)) / (2 * h)
            np.testing.assert_allclose(G[n][i, c], fd, rtol=1e-6, atol=1e-8)


def _small_csr(m, n, density, seed):
    import scipy.sparse as sp
    A = sp.random(m, n, density=density, format="csr", random_state=np.random.RandomState(seed))
    A.data = np.exp(np.random.default_rng(seed).standard_normal(A.nnz))
    A.sort_indices()
    return A


def test_r2_nmf_csr_matches_dense_numpy_loop():
    """DistAUNMF order (H given W, then W given H) with DistR2's rule, run
    densely in NumPy with the brute-force-pinned r2_update; every half-step
    is a KKT point of its two-variable NNLS."""
    A = _small_csr(40, 30, 0.2, 1)
    cols = np.array([1, 2, 5, 7, 8, 11, 12, 20, 21, 29])
    rng = np.random.default_rng(2)
    W0, H0 = rng.random((40, 2)), rng.random((len(cols), 2))
    W, H = O.r2_nmf_csr(40, 30, A.indptr, A.indices, A.data, cols, W0, H0, 5)
    Ad = A.toarray()[:, cols]
    Wn, Hn = W0.copy(), H0.copy()
    for _ in range(5):
        Hn = O.r2_update(Wn.T @ Wn, Wn.T @ Ad).T
        Wn = O.r2_update(Hn.T @ Hn, (Ad @ Hn).T).T
    np.testing.assert_allclose(H, Hn, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(W, Wn, rtol=1e-10, atol=1e-12)
    # the last W step: each row solves min ||a_i - H w|| over w >= 0 (KKT)
    G = H.T @ H
    Y = W @ G - Ad @ H
    assert W.min() >= 0 and Y.min() > -1e-9 and np.abs((W * Y).sum()) < 1e-8


def test_powiter_csr_matches_svd():
    A = _small_csr(60, 50, 0.15, 3)
    cols = np.arange(0, 50, 2)
    Ad = A.toarray()[:, cols]
    q0 = np.random.default_rng(4).standard_normal(len(cols))
    s = O.powiter_csr(60, 50, A.indptr, A.indices, A.data, cols, q0, 300)
    sv = np.linalg.svd(Ad, compute_uv=False)
    np.testing.assert_allclose(s, sv[0] ** 2, rtol=1e-8)
