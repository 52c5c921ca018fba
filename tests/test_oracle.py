"""CPU tests that pin the FP64 oracle (oracle/planc_oracle.cpp) against the
committed NumPy/SciPy fixtures, closed forms and invariants."""
import os

import numpy as np
import pytest
from scipy.optimize import nnls

from oracle import oracle as O
from planc_amd import datagen as D

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


def _hals_case(seed):
    rng = np.random.default_rng(seed)
    m, n, k = 25, 20, 4
    A = _lowrank(rng, m, n, 2) + 0.05 * rng.random((m, n))
    # initial factors far from A's structure, so that some unconstrained
    # coordinate minimisers are negative and the 1e-17 clamp is exercised
    W0 = rng.random((m, k)) ** 4
    H0 = 3.0 * rng.random((n, k))
    return A, W0, H0


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_hals_h_sweep_columns_are_clamped_coordinate_minimisers(seed):
    """nmf/hals.hpp:67-75: after the constructor's normalize_by_W, column c of
    H is the minimiser of ||A - W H^T||^2 over H(:, c) >= 1e-17 given the
    columns before it (already swept) and after it (not yet).  Checked
    through the gradient of the explicit residual, not the Gram form."""
    A, W0, H0 = _hals_case(seed)
    k = W0.shape[1]
    W, H, rel, obj, rc = O.nmf(1, A, W0, H0, 1)
    nrm = np.linalg.norm(W0, axis=0)
    W1, H1 = W0 / nrm, H0 * nrm  # HALSNMF constructor (nmf/hals.hpp:46)
    # The iteration ends with the W sweep (unit columns) and normalize_by_W,
    # whose norms are 1 to rounding, so H is the H sweep's result.
    scale = np.abs(A).sum()
    clamped = 0
    for c in range(k):
        Hc = np.concatenate([H[:, :c + 1], H1[:, c + 1:]], axis=1)
        g = -(A - W1 @ Hc.T).T @ W1[:, c]
        at_floor = H[:, c] <= 1e-16
        clamped += int(at_floor.sum())
        assert np.all(np.abs(g[~at_floor]) <= 1e-11 * scale)
        assert np.all(g[at_floor] >= -1e-11 * scale)
        # the floor is 1e-17, not 0 (common/utils.hpp:59-64 with repl = 1e-17)
        np.testing.assert_allclose(H[at_floor, c], 1e-17, rtol=1e-12)
    if seed == 7:
        assert clamped > 0


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_hals_w_sweep_columns_are_normalised_coordinate_minimisers(seed):
    """nmf/hals.hpp:89-102: column c of W is the minimiser w* of
    ||A - W H^T||^2 over W(:, c) (others: swept columns before c, the
    constructor's W after c), scaled by HtH(c, c), floored at 1e-17 and
    normalised.  w* comes from the 1-D least squares on the residual."""
    A, W0, H0 = _hals_case(seed)
    k = W0.shape[1]
    W, H, rel, obj, rc = O.nmf(1, A, W0, H0, 1)
    W1 = W0 / np.linalg.norm(W0, axis=0)
    for c in range(k):
        others = [t for t in range(k) if t != c]
        Wc = np.concatenate([W[:, :c], W1[:, c:]], axis=1)
        R = A - Wc[:, others] @ H[:, others].T
        hc = H[:, c]
        wstar = R @ hc / (hc @ hc)
        x = np.maximum(wstar * (hc @ hc), 1e-17)
        np.testing.assert_allclose(W[:, c], x / np.linalg.norm(x), rtol=1e-9, atol=1e-14)


def _krp_brute(F, n):
    """Khatri-Rao product of the factors other than n, rows ordered with the
    lowest remaining mode fastest, by explicit multi-index loops."""
    import itertools
    others = [f for i, f in enumerate(F) if i != n]
    dims = [f.shape[0] for f in others]
    k = F[0].shape[1]
    K = np.zeros((int(np.prod(dims)), k))
    for idx in itertools.product(*[range(d) for d in reversed(dims)]):
        idx = idx[::-1]  # idx[0] is the lowest remaining mode
        row, mul = 0, 1
        for i, d in zip(idx, dims):
            row += i * mul
            mul *= d
        v = np.ones(k)
        for i, f in zip(idx, others):
            v = v * f[i]
        K[row] = v
    return K


def _unfold(X, n):
    """Mode-n unfolding whose columns follow _krp_brute's row order."""
    return np.moveaxis(X, n, 0).reshape(X.shape[n], -1, order="F")


def _ntf_mode_case(seed, n=1, k=4, dims=(6, 5, 7)):
    rng = np.random.default_rng(seed)
    F = [rng.random((d, k)) for d in dims]
    F = [f / np.linalg.norm(f, axis=0) for f in F]  # unit columns: S(i, i) = 1
    X = np.asfortranarray(np.einsum("ir,jr,kr->ijk", *[rng.random((d, k)) for d in dims])
                          + 0.1 * rng.random(dims))
    K = _krp_brute(F, n)
    Xn = _unfold(X, n)
    return rng, F, X, K, Xn


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_ntf_hals_mode_update_columns_are_clamped_coordinate_minimisers(seed):
    """ntf/ntfhals.hpp:13-29: every column of the sweep minimises
    0.5 ||X_(n) - H K^T||^2 over H(:, i) >= 1e-17 with the columns before i
    already updated and those after i not yet."""
    rng, F, X, K, Xn = _ntf_mode_case(seed)
    H0 = 5.0 * rng.random((Xn.shape[0], K.shape[1]))
    Hn, rc = O.ntf_mode_update(1, H0, K.T @ Xn.T, K.T @ K)
    assert rc == 0
    scale = np.abs(Xn).sum() * np.abs(K).max()
    clamped = 0
    for c in range(K.shape[1]):
        Hc = np.concatenate([Hn[:, :c + 1], H0[:, c + 1:]], axis=1)
        g = -(Xn - Hc @ K.T) @ K[:, c]
        at_floor = Hn[:, c] <= 1e-16
        clamped += int(at_floor.sum())
        assert np.all(np.abs(g[~at_floor]) <= 1e-12 * scale)
        assert np.all(g[at_floor] >= -1e-12 * scale)
        assert np.all(Hn[at_floor, c] == 1e-17)
    assert clamped > 0


@pytest.mark.parametrize("seed", [14, 15])
def test_ntf_bpp_mode_update_satisfies_kkt(seed):
    """ntf/ntfanlsbpp.hpp:15-57: the factor solves min 0.5||X_(n) - H K^T||^2
    over H >= 0 exactly: H >= 0, gradient >= 0, complementary slackness."""
    rng, F, X, K, Xn = _ntf_mode_case(seed, k=5)
    H0 = rng.random((Xn.shape[0], K.shape[1]))
    Hn, rc = O.ntf_mode_update(2, H0, K.T @ Xn.T, K.T @ K)
    assert rc == 0
    G = -(Xn - Hn @ K.T) @ K
    scale = np.abs(Xn).sum() * np.abs(K).max()
    assert Hn.min() >= 0
    assert G.min() >= -1e-12 * scale
    assert np.abs(Hn * G).sum() <= 1e-12 * scale * np.abs(Hn).sum()
    assert (Hn == 0).any() or G.min() > -1e-12 * scale


def test_ntf_mu_exact_factorisation_shows_eps_1e6():
    """ntf/ntfmu.hpp:11-16: H <- H .* M^T ./ (H S + 1e-6).  At an exact
    factorisation X_(n) = H* K^T the numerator equals H* S, so the step
    shrinks each entry by exactly 1e-6 / (H*S + 1e-6); with H* of order 1e-4
    that is a relative change near 1e-3, which the 1e-17 of the NMF rule
    would not produce."""
    rng = np.random.default_rng(16)
    dims, k, n = (6, 5, 7), 3, 1
    F = [rng.random((d, k)) for d in dims]
    F = [f / np.linalg.norm(f, axis=0) for f in F]
    Hstar = 1e-4 * (0.5 + rng.random((dims[n], k)))
    F[n] = Hstar
    X = np.asfortranarray(np.einsum("ir,jr,kr->ijk", *F))
    K = _krp_brute(F, n)
    Xn = _unfold(X, n)
    Hn, rc = O.ntf_mode_update(0, Hstar, K.T @ Xn.T, K.T @ K)
    HS = Hstar @ (K.T @ K)
    shrink = (Hstar - Hn) / Hstar
    np.testing.assert_allclose(shrink, 1e-6 / (HS + 1e-6), rtol=1e-6)


@pytest.mark.parametrize("seed", [17, 18, 19])
def test_ntf_mu_mode_update_does_not_increase_objective(seed):
    rng, F, X, K, Xn = _ntf_mode_case(seed)
    H0 = rng.random((Xn.shape[0], K.shape[1]))
    Hn, rc = O.ntf_mode_update(0, H0, K.T @ Xn.T, K.T @ K)

    def f(H):
        return 0.5 * np.linalg.norm(Xn - H @ K.T) ** 2
    assert f(Hn) <= f(H0) * (1 + 1e-14)
    assert f(Hn) < f(H0)


@pytest.mark.parametrize("algo", [0, 1, 2])
def test_ntf_iteration_equals_mode_updates_on_explicit_krp(algo):
    """ntf/auntf.hpp:128-177 wiring: per mode S = K^T K (the Hadamard of the
    other Grams), M = K^T X_(n)^T with the brute-force Khatri-Rao product,
    the (pinned) mode update, then normalize(mode) replacing lambda
    (common/ncpfactors.hpp:346-351); the constructor's normalize() first."""
    rng = np.random.default_rng(21 + algo)
    dims, k = (5, 4, 6), 3
    X = np.asfortranarray(np.einsum("ir,jr,kr->ijk", *[rng.random((d, k)) for d in dims])
                          + 0.02 * rng.random(dims))
    init = [rng.random((d, k)) for d in dims]
    Fo, lam_o, rel_o, rc = O.ntf(algo, X, init, 2, dimtree=0)
    assert rc == 0
    F = [f / np.linalg.norm(f, axis=0) for f in init]
    lam = np.prod([np.linalg.norm(f, axis=0) for f in init], axis=0)
    for it in range(2):
        for n in range(3):
            K = _krp_brute(F, n)
            Hn, rc = O.ntf_mode_update(algo, F[n], K.T @ _unfold(X, n).T, K.T @ K)
            assert rc == 0
            lam = np.linalg.norm(Hn, axis=0)
            F[n] = Hn / np.where(lam > 0, lam, 1.0)
        Xh = np.einsum("r,ir,jr,kr->ijk", lam, *F)
        np.testing.assert_allclose(rel_o[it], np.linalg.norm(X - Xh) / np.linalg.norm(X),
                                   rtol=1e-9)
    for a, b in zip(Fo, F):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(lam_o, lam, rtol=1e-9)


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


# ---- sparse build ---------------------------------------------------------------
def _sparse_case(m, n, density, stream):
    import scipy.sparse as sp
    rng = np.random.default_rng(stream)
    A = sp.random(m, n, density=density, format="csr", random_state=rng,
                  data_rvs=lambda s: rng.random(s) + 0.1)
    A.sort_indices()
    return A


@pytest.mark.parametrize("algo", [0, 1])
def test_oracle_nmf_sparse_equals_dense_loop(algo):
    """MU / HALS of the sparse build run exactly the dense loop on the
    densified matrix."""
    A = _sparse_case(60, 40, 0.2, 1)
    W0, H0 = D.init_factors([(60, 4), (40, 4)], stream=31)
    Ws, Hs, rs, rc = O.nmf_sparse(algo, 60, 40, A.indptr, A.indices, A.data, W0, H0, 6)
    Wd, Hd, rd, _, rc2 = O.nmf(algo, A.toarray(), W0, H0, 6)
    assert rc == 0 and rc2 == 0
    assert np.array_equal(rs, rd) and np.array_equal(Ws, Wd) and np.array_equal(Hs, Hd)


def test_oracle_nmf_sparse_bpp_starts_normalised():
    """ANLS-BPP of the sparse build equals the dense loop started from
    normalize_by_W'd factors (nmf/bppnmf.hpp:195-201 via nmf/hals.hpp:46).
    The first W solve absorbs any column scaling of H, so this start changes
    rounding only; the check is that the two loops agree to rounding."""
    A = _sparse_case(60, 40, 0.2, 2)
    W0, H0 = D.init_factors([(60, 4), (40, 4)], stream=32)
    nrm = np.linalg.norm(W0, axis=0)
    Wn, Hn = W0 / nrm, H0 * nrm
    Ws, Hs, rs, rc = O.nmf_sparse(2, 60, 40, A.indptr, A.indices, A.data, W0, H0, 5)
    Wd, Hd, rd, _, _ = O.nmf(2, A.toarray(), Wn, Hn, 5)
    assert rc == 0
    assert np.allclose(rs, rd, rtol=1e-12, atol=1e-14)
    assert np.allclose(Ws, Wd, rtol=1e-10, atol=1e-13)


def test_oracle_sparse_fit_error_matches_dense_norm():
    """sqrt(nnzsse + ||Rw Rh^T||^2 - nnzwh) equals ||A - W H^T||_F computed
    densely (the identity common/utils.hpp:225-232 relies on)."""
    for (m, n, k, dens, st) in ((50, 30, 3, 0.1, 3), (7, 90, 5, 0.5, 4), (120, 80, 1, 0.02, 5)):
        A = _sparse_case(m, n, dens, st)
        W = D.uniform(D.INIT_SEED, 600 + st, (m, k))
        H = D.uniform(D.INIT_SEED, 700 + st, (n, k))
        C = A.tocsc()
        C.sort_indices()
        e = O.sparse_fit_error(m, n, C.indptr, C.indices, C.data, W, H)
        ref = np.linalg.norm(A.toarray() - W @ H.T)
        assert abs(e - ref) <= 1e-10 * ref


# ---- AO-ADMM and GNSym ------------------------------------------------------------
def test_admm_solve_converges_to_nnls():
    """Repeated AO-ADMM solves with the dual carried over converge to the
    NNLS solution of G x = r (scipy nnls on the Cholesky factor), for every
    column; a wrong sign or a dropped rho term does not."""
    rng = np.random.default_rng(5)
    k, n = 6, 9
    B = rng.random((40, k))
    G = B.T @ B
    Y = rng.random((40, n)) - 0.3
    R = B.T @ Y
    X = np.zeros((n, k))
    D = np.zeros((n, k))
    for _ in range(400):
        X, D, rc = O.admm_solve(G, R, X, D)
        assert rc == 0
    for j in range(n):
        ref, _ = nnls(B, Y[:, j])
        assert np.allclose(X[j], ref, rtol=1e-6, atol=1e-8), (j, X[j], ref)


def _admm_np(G0, R, X, D):
    k = G0.shape[0]
    rho = np.trace(G0) / k
    rho = rho if rho > 0 else 0.01
    G = G0 + rho * np.eye(k)
    for _ in range(5):
        Xp = X.copy()
        T = np.linalg.solve(G, R + rho * (X + D).T)
        X = np.maximum(T.T - D, 0.0)
        D = D + X - T.T
        if (np.linalg.norm(X - T.T) < 0.01 * np.linalg.norm(X) and
                np.linalg.norm(X - Xp) < 0.01 * np.linalg.norm(D)):
            break
    return X, D


def test_aoadmm_nmf_first_iterations_match_numpy():
    """Two AO-ADMM NMF iterations (nmf/aoadmm.hpp) against a NumPy
    restatement: constructor normalize_by_W, H then W, duals carried over."""
    A = D.lowrank_matrix(30, 20, 3, 0.01, stream=41)
    W0, H0 = D.init_factors([(30, 3), (20, 3)], stream=41)
    Wr, Hr, rel, _, rc = O.nmf(4, A, W0, H0, 2)
    assert rc == 0
    nrm = np.linalg.norm(W0, axis=0)
    W, H = W0 / nrm, H0 * nrm
    U, V = np.zeros_like(W), np.zeros_like(H)
    for it in range(2):
        H, V = _admm_np(W.T @ W, W.T @ A, H, V)
        W, U = _admm_np(H.T @ H, (A @ H).T, W, U)
        err = np.linalg.norm(A - W @ H.T) / np.linalg.norm(A)
        assert abs(rel[it] - err) < 1e-10
    assert np.allclose(Wr, W, rtol=1e-9, atol=1e-12) and np.allclose(Hr, H, rtol=1e-9, atol=1e-12)


def test_aoadmm_ntf_dimtree_equals_direct_and_error_is_direct():
    X, _ = D.cp_tensor((8, 7, 6), 3, 0.01, stream=42)
    F0 = D.init_factors([(8, 3), (7, 3), (6, 3)], stream=42)
    Fa, la, ra, rc = O.ntf(4, X, F0, 4, dimtree=1)
    Fb, lb, rb, rc2 = O.ntf(4, X, F0, 4, dimtree=0)
    assert rc == 0 and rc2 == 0
    assert np.allclose(ra, rb, rtol=1e-10)
    Xh = np.einsum("az,bz,cz->abc", Fa[0] * la, Fa[1], Fa[2])
    assert abs(ra[-1] - np.linalg.norm(X - Xh) / np.linalg.norm(X)) < 1e-10
    assert np.all(ra < 1.0)


def test_gnsym_exact_factorisation_is_a_fixed_point():
    Hs = D.uniform(D.DATA_SEED, 43, (25, 3))
    A = Hs @ Hs.T
    W0 = np.zeros((25, 3))
    _, H, rel, obj, rc = O.nmf(7, A, W0, Hs, 3)
    assert rc == 0
    assert np.array_equal(H, Hs)
    # fit^2 = ||A||^2 - 2 tr(A H H^T) + tr((H^T H)^2) cancels to rounding; a
    # non-positive value prints as relative error 1 (printObjective's guard,
    # nmf/gnsym.hpp:166), so the fit itself is checked.
    assert np.all(np.abs(obj) < 1e-10 * np.linalg.norm(A) ** 2)


def test_gnsym_first_iteration_matches_numpy_and_error_is_direct():
    rng = np.random.default_rng(44)
    Hs = rng.random((40, 4))
    A = Hs @ Hs.T + 0.01 * rng.random((40, 40))
    A = 0.5 * (A + A.T)
    H0 = rng.random((40, 4))
    _, H1, rel, _, rc = O.nmf(7, A, np.zeros((40, 4)), H0, 1)
    assert rc == 0
    # NumPy restatement of one GNSYMNMF iteration (nmf/gnsym.hpp:178-260).
    k = 4
    H = H0.copy()
    HtH = H.T @ H
    grad = -2 * (H.T @ A - HtH @ H.T)
    P, Dt = grad.copy(), np.zeros_like(grad)
    rs = (grad * grad).sum()
    if rs > 1e-4:
        for _ in range(k):
            Y = 2 * (HtH @ P + (P @ H) @ H.T)
            al = rs / (P * Y).sum()
            Dt += al * P
            grad -= al * Y
            rn = (grad * grad).sum()
            if rn < 1e-4:
                break
            P = grad + (rn / rs) * P
            rs = rn
    H = np.maximum(H - Dt.T, 0)
    assert np.allclose(H1, H, rtol=1e-9, atol=1e-12)
    assert abs(rel[0] - np.linalg.norm(A - H @ H.T) / np.linalg.norm(A)) < 1e-9
    _, _, rel20, _, _ = O.nmf(7, A, np.zeros((40, 4)), H0, 20)
    assert rel20[-1] < rel20[0]


# ---- NTF-NES -------------------------------------------------------------------
def test_sym_eig_minmax_matches_numpy():
    rng = np.random.default_rng(70)
    for k in (1, 2, 7, 33, 64):
        B = rng.random((k + 5, k))
        S = B.T @ B
        mn, mx = O.sym_eig_minmax(S)
        ev = np.linalg.eigvalsh(S)
        assert abs(mn - ev[0]) <= 1e-12 * ev[-1] and abs(mx - ev[-1]) <= 1e-12 * ev[-1]


def test_nes_mode_update_solves_the_proximal_nnls():
    """The Nesterov inner loop stops once |grad o acc| <= 1e-2 and grad >=
    -1e-2 (ntf/ntfnes.hpp:56-66): its result is the solution of the proximal
    NNLS min 0.5 x^T (S + lambda I) x - (m + lambda h*)^T x, x >= 0, to that
    tolerance; the BPP oracle (pinned against scipy) solves the same problem."""
    rng = np.random.default_rng(71)
    k, I = 5, 30
    F = [rng.random((20, k)), rng.random((25, k))]
    F = [f / np.linalg.norm(f, axis=0) for f in F]
    S = (F[0].T @ F[0]) * (F[1].T @ F[1])
    M = rng.random((k, I)) * 2
    Hprox = rng.random((I, k))
    H, rc = O.nes_mode_update(S, M, Hprox)
    assert rc == 0
    ev = np.linalg.eigvalsh(S)
    q = ev[-1] / ev[0]
    lam = 10 * ev[0] if q > 1e6 else (ev[0] if q > 1e3 else ev[0] / 10)
    G = S + lam * np.eye(k)
    R = M + lam * Hprox.T
    X, it = O.bppnnls(G, R)
    assert it >= 0
    # (the stop test is taken at the extrapolated point, one step past the
    # returned iterate, so the returned factor meets it only approximately)
    grad = G @ H.T - R
    assert np.all(H >= 0)
    assert np.abs(grad * H.T).max() <= 5e-2 and grad.min() >= -5e-2
    assert np.linalg.norm(H.T - X) / np.linalg.norm(X) < 0.05


def test_nes_ntf_error_is_direct_and_dimtree_equals_direct():
    X, _ = D.cp_tensor((9, 8, 7), 3, 0.01, stream=72)
    F0 = D.init_factors([(9, 3), (8, 3), (7, 3)], stream=72)
    Fa, la, ra, rc = O.ntf(5, X, F0, 12, dimtree=1)
    Fb, lb, rb, rc2 = O.ntf(5, X, F0, 12, dimtree=0)
    assert rc == 0 and rc2 == 0
    assert np.allclose(ra, rb, rtol=1e-9)
    Xh = np.einsum("az,bz,cz->abc", Fa[0] * la, Fa[1], Fa[2])
    # after iteration 6 the factors may be the accelerated ones (accepted
    # steps lower the error), so the returned factors reproduce an error no
    # larger than the last recorded one
    assert np.linalg.norm(X - Xh) / np.linalg.norm(X) <= ra[-1] + 1e-12
    assert ra[-1] < ra[0]
