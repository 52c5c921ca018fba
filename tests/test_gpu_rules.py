"""GPU tests of the further update rules against the FP64 oracle, through the
C ABI: AO-ADMM NMF (nmf/aoadmm.hpp) and NTF (ntf/ntfaoadmm.hpp), symmetric
Gauss-Newton NMF (nmf/gnsym.hpp), Nesterov NTF (ntf/ntfnes.hpp).
Trajectories are held to 1e-4 for AO-ADMM; GNSym to 1e-3 over 8 iterations
and NES to 1e-3: their stopping tests (the CG residual, the Nesterov
gradient test) compare absolute quantities, so an FP32 rounding can end an
inner loop one step earlier or later than in FP64, and the Gauss-Newton
trajectory is not monotone (the error rises in some iterations), which
amplifies such differences; the tests print both trajectories."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import planc_amd as P  # noqa: E402
from oracle import oracle as O  # noqa: E402
from planc_amd import datagen as D  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(x, ref):
    return np.linalg.norm(np.asarray(x, np.float64) - ref) / max(np.linalg.norm(ref), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = P.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("m,n,k,reg", [(500, 300, 10, False), (2000, 1000, 20, False),
                                       (400, 350, 16, True), (3000, 64, 64, False)])
def test_aoadmm_nmf_matches_oracle(ctx, m, n, k, reg):
    A = D.lowrank_matrix(m, n, k, 0.01, stream=51).astype(np.float32)
    W0, H0 = D.init_factors([(m, k), (n, k)], stream=51)
    W0, H0 = W0.astype(np.float32), H0.astype(np.float32)
    regW, regH = ((0.5, 0.1), (0.25, 0.0)) if reg else ((0.0, 0.0), (0.0, 0.0))
    Wr, Hr, rel_r, _, rc = O.nmf(4, A.astype(np.float64), W0.astype(np.float64), H0.astype(np.float64), 10,
                                 regW=regW, regH=regH)
    assert rc == 0
    dW, dH = P.colmajor(W0), P.colmajor(H0)
    rel = ctx.nmf_run(P.colmajor(A), dW, dH, P.AOADMM, 10, regW=regW, regH=regH)
    say(f"  aoadmm {m}x{n} k={k}: gpu {np.array2string(rel, precision=8)}")
    say(f"                  oracle {np.array2string(rel_r, precision=8)}")
    assert np.abs(rel - rel_r).max() < 1e-4
    assert rel_err(dW.cpu().numpy(), Wr) < 1e-3
    assert rel_err(dH.cpu().numpy(), Hr) < 1e-3


@pytest.mark.parametrize("dims,k", [((20, 18, 16), 5), ((40, 36, 32), 8), ((12, 10, 9, 8), 4)])
@pytest.mark.parametrize("dimtree", [True, False])
def test_aoadmm_ntf_matches_oracle(ctx, dims, k, dimtree):
    X, _ = D.cp_tensor(dims, k, 0.01, stream=52)
    X32 = X.astype(np.float32)
    F0 = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=52)]
    Fr, lam_r, rel_r, rc = O.ntf(4, X32.astype(np.float64), [f.astype(np.float64) for f in F0], 10,
                                 dimtree=1 if dimtree else 0)
    assert rc == 0
    dF = [P.colmajor(f) for f in F0]
    lam, rel = ctx.ntf_run(P.colmajor(X32), dF, P.AOADMM, 10, dimtree=dimtree)
    say(f"  ntf aoadmm {dims}: gpu {rel[-1]:.8f} oracle {rel_r[-1]:.8f} max diff {np.abs(rel - rel_r).max():.2e}")
    assert np.abs(rel - rel_r).max() < 1e-4
    for f, fr in zip(dF, Fr):
        assert rel_err(f.cpu().numpy(), fr) < 1e-3


@pytest.mark.parametrize("n,k", [(300, 8), (2000, 16), (1500, 64)])
def test_gnsym_matches_oracle(ctx, n, k):
    B = D.uniform(D.DATA_SEED, 16 * 53 + k, (n, k))
    A = B @ B.T
    A = (A + 0.01 * A.mean() * np.abs(D.normal(D.NOISE_SEED, 53, (n, n)))).astype(np.float32)
    A = 0.5 * (A + A.T)
    H0 = D.uniform(D.INIT_SEED, 16 * 53, (n, k))
    H0 = (2 * np.sqrt(A.astype(np.float64).mean() / k) * H0).astype(np.float32)
    _, Hr, rel_r, _, rc = O.nmf(7, A.astype(np.float64), np.zeros((n, k)), H0.astype(np.float64), 8)
    assert rc == 0
    dW, dH = P.colmajor(np.zeros((n, k), np.float32)), P.colmajor(H0)
    rel = ctx.nmf_run(P.colmajor(A), dW, dH, P.GNSYM, 8)
    say(f"  gnsym n={n} k={k}: gpu {np.array2string(rel, precision=8)}")
    say(f"                 oracle {np.array2string(rel_r, precision=8)}")
    assert np.abs(rel - rel_r).max() < 1e-3
    assert abs(rel[-1] - rel_r[-1]) < 1e-4
    assert rel_err(dH.cpu().numpy(), Hr) < 1e-2
    assert np.all(dW.cpu().numpy() == 0)  # W is not used
    with pytest.raises(P.PlancError):
        ctx.nmf_run(P.colmajor(A[:, :n - 1]), dW, P.colmajor(H0[:n - 1]), P.GNSYM, 1)


def test_sparse_aoadmm_and_gnsym(ctx):
    """The sparse-input loop runs the same rules (products from the CSR)."""
    sp = pytest.importorskip("scipy.sparse")
    rng = np.random.default_rng(54)
    m = n = 600
    A = sp.random(m, n, density=0.05, format="csr", random_state=rng, data_rvs=lambda s: rng.random(s) + 0.1)
    A = (0.5 * (A + A.T)).tocsr().astype(np.float32)
    A.sort_indices()
    ip = torch.from_numpy(A.indptr.astype(np.int64)).cuda()
    ix = torch.from_numpy(A.indices.astype(np.int32)).cuda()
    vv = torch.from_numpy(A.data).cuda()
    W0, H0 = D.init_factors([(m, 6), (n, 6)], stream=54)
    W0, H0 = W0.astype(np.float32), H0.astype(np.float32)
    for algo in (P.AOADMM, P.GNSYM):
        _, _, rel_r, rc = O.nmf_sparse(algo, m, n, A.indptr, A.indices, A.data.astype(np.float64),
                                       W0.astype(np.float64), H0.astype(np.float64), 6)
        assert rc == 0
        dW, dH = P.colmajor(W0), P.colmajor(H0)
        rel = ctx.nmf_sparse_run(m, n, ip, ix, vv, dW, dH, algo, 6)
        say(f"  sparse algo={algo}: gpu {rel[-1]:.8f} oracle {rel_r[-1]:.8f}")
        assert np.abs(rel - rel_r).max() < (1e-4 if algo == P.AOADMM else 1e-3)


@pytest.mark.parametrize("dims,k", [((20, 18, 16), 5), ((40, 36, 32), 8), ((12, 10, 9, 8), 4)])
@pytest.mark.parametrize("dimtree", [True, False])
def test_nes_ntf_matches_oracle(ctx, dims, k, dimtree):
    """NTFNES: the proximal Nesterov mode updates and, from iteration 6 on,
    accelerate() with its accept / reject decision, 14 iterations."""
    X, _ = D.cp_tensor(dims, k, 0.01, stream=55)
    X32 = X.astype(np.float32)
    F0 = [f.astype(np.float32) for f in D.init_factors([(d, k) for d in dims], stream=55)]
    Fr, lam_r, rel_r, rc = O.ntf(P.NESTEROV, X32.astype(np.float64), [f.astype(np.float64) for f in F0], 14,
                                 dimtree=1 if dimtree else 0)
    assert rc == 0
    dF = [P.colmajor(f) for f in F0]
    lam, rel = ctx.ntf_run(P.colmajor(X32), dF, P.NESTEROV, 14, dimtree=dimtree)
    say(f"  nes {dims}: gpu {np.array2string(rel, precision=7)}")
    say(f"       oracle {np.array2string(rel_r, precision=7)}")
    assert np.abs(rel - rel_r).max() < 1e-3
    assert abs(rel[-1] - rel_r[-1]) < 2e-4


def test_admm_update_entry(ctx):
    """planc_admm_update, the solve that AO-ADMM NMF and NTF call for every
    factor (nmf/aoadmm.hpp, ntf/ntfaoadmm.hpp:23-77): two calls in a row with
    the dual carried against the oracle's solve (1e-4); repeated calls
    converge to the NNLS solution of each column (the oracle's BPP, pinned to
    scipy's NNLS); a Gram whose shift is not positive definite, set from the
    host, returns PLANC_ENOCONV as the oracle's solve fails, with X and D
    untouched (the sweeps are skipped on the device)."""
    say("admm_update entry")
    rng = np.random.default_rng(77)
    k, n = 16, 3000
    B = rng.random((60, k))
    G = B.T @ B
    R = (B.T @ (rng.random((60, n)) - 0.2)).astype(np.float32)
    X0 = rng.random((n, k)).astype(np.float32)
    Xr, Dr = X0.astype(np.float64), np.zeros((n, k))
    dG, dR = P.colmajor(G), P.colmajor(R)
    dX, dD = P.colmajor(X0), P.colmajor(np.zeros((n, k), np.float32))
    for call in range(2):
        Xr, Dr, rc = O.admm_solve(G, R.astype(np.float64), Xr, Dr)
        assert rc == 0
        ctx.admm_update(dG, dR, dX, dD)
        ex, ed = rel_err(dX.cpu().numpy(), Xr), rel_err(dD.cpu().numpy(), Dr)
        say(f"  call {call}: ||dX||/||X|| {ex:.2e}, ||dD||/||D|| {ed:.2e}")
        assert ex < 1e-4 and ed < 1e-4
    # Fixed point: the NNLS solutions of G x = r_j.
    Xs, _ = O.bppnnls(G, R[:, :200].astype(np.float64))
    dR2 = P.colmajor(R[:, :200])
    dX2, dD2 = P.colmajor(np.zeros((200, k), np.float32)), P.colmajor(np.zeros((200, k), np.float32))
    for _ in range(300):
        ctx.admm_update(dG, dR2, dX2, dD2)
    es = rel_err(dX2.cpu().numpy(), Xs.T)
    say(f"  300 calls vs NNLS: {es:.2e}")
    assert es < 1e-4
    # Not positive definite: G = -I (rho falls back to 0.01, G + rho I = -0.99 I).
    Gbad = -np.eye(k)
    _, _, rc = O.admm_solve(Gbad, R[:, :10].astype(np.float64), np.zeros((10, k)), np.zeros((10, k)))
    assert rc == -1
    before = dX.cpu().numpy().copy(), dD.cpu().numpy().copy()
    with pytest.raises(P.PlancError) as e:
        ctx.admm_update(P.colmajor(Gbad), dR, dX, dD)
    assert e.value.rc == P._lib.PLANC_ENOCONV
    assert np.array_equal(dX.cpu().numpy(), before[0]) and np.array_equal(dD.cpu().numpy(), before[1])
    # The context stays usable after the error.
    ctx.admm_update(dG, dR, dX, dD)
