"""GPU tests of the sparse-input NMF (the reference's BUILD_SPARSE build):
the CSR transpose bit-exact against scipy.sparse, the sparse products against
NumPy, the sparse error helper (common/utils.hpp:190-241) against the oracle
and the dense norm, and the MU / HALS / ANLS-BPP loops against the FP64
oracle's sparse build (trajectories 1e-5, BPP 1e-4)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
sp = pytest.importorskip("scipy.sparse")

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


def _csr(m, n, density, seed, empty_rows=(), empty_cols=()):
    """Random CSR (float32 values in [0.1, 1.1)) with chosen empty rows and
    columns; sorted column ids."""
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="lil", random_state=rng,
                  data_rvs=lambda s: rng.random(s) + 0.1)
    for r in empty_rows:
        A[r, :] = 0
    for c in empty_cols:
        A[:, c] = 0
    A = A.tocsr().astype(np.float32)
    A.eliminate_zeros()
    A.sort_indices()
    return A


def _dev(A):
    return (torch.from_numpy(A.indptr.astype(np.int64)).cuda(),
            torch.from_numpy(A.indices.astype(np.int32)).cuda(),
            torch.from_numpy(A.data.astype(np.float32)).cuda())


@pytest.mark.parametrize("m,n,density", [(37, 53, 0.1), (1000, 700, 0.01), (5, 300000, 0.0003),
                                         (20000, 3, 0.3)])
def test_csr_transpose_bit_exact(ctx, m, n, density):
    A = _csr(m, n, density, m + n, empty_rows=(0,), empty_cols=(n - 1,))
    say(f"csr transpose {m}x{n} nnz={A.nnz}")
    tp, ti, tv = ctx.csr_transpose(m, n, *_dev(A))
    T = A.T.tocsr()
    T.sort_indices()
    assert np.array_equal(tp.cpu().numpy(), T.indptr.astype(np.int64))
    assert np.array_equal(ti.cpu().numpy(), T.indices.astype(np.int32))
    assert np.array_equal(tv.cpu().numpy(), T.data.astype(np.float32))


def test_csr_transpose_empty_matrix(ctx):
    A = sp.csr_matrix((6, 9), dtype=np.float32)
    ip, ix, vv = _dev(A)
    tp, _, _ = ctx.csr_transpose(6, 9, ip, torch.zeros(1, dtype=torch.int32, device="cuda"),
                                 torch.zeros(1, dtype=torch.float32, device="cuda"))
    assert np.array_equal(tp.cpu().numpy(), np.zeros(10, dtype=np.int64))


@pytest.mark.parametrize("k", [1, 7, 32, 64])
def test_spmm_matches_numpy(ctx, k):
    A = _csr(3000, 1700, 0.01, k, empty_rows=(5, 6), empty_cols=(2,))
    F = D.uniform(D.INIT_SEED, 800 + k, (1700, k)).astype(np.float32)
    out = ctx.spmm(3000, 1700, *_dev(A), P.colmajor(F)).cpu().numpy()
    ref = A.astype(np.float64) @ F.astype(np.float64)
    assert rel_err(out, ref) < 1e-6
    assert np.all(out[5] == 0) and np.all(out[6] == 0)


def test_sparse_fit_error_matches_oracle_and_dense(ctx):
    m, n, k = 800, 600, 9
    A = _csr(m, n, 0.05, 17)
    W = D.uniform(D.INIT_SEED, 901, (m, k)).astype(np.float32)
    H = D.uniform(D.INIT_SEED, 902, (n, k)).astype(np.float32)
    e = ctx.sparse_fit_error(m, n, *_dev(A), P.colmajor(W), P.colmajor(H))
    C = A.tocsc()
    C.sort_indices()
    e_or = O.sparse_fit_error(m, n, C.indptr, C.indices, C.data, W, H)
    ref = np.linalg.norm(A.toarray().astype(np.float64) - W.astype(np.float64) @ H.astype(np.float64).T)
    say(f"  sparse error gpu {e:.10e} oracle {e_or:.10e} dense {ref:.10e}")
    assert abs(e - e_or) <= 1e-9 * e_or
    assert abs(e - ref) <= 1e-9 * ref


def _csr_lowrank(m, n, k, density, seed, empty_rows=(), empty_cols=()):
    """A masked low-rank CSR: (W H)(i, j) at random positions (the sparse
    build's rand_lowrank form, nmf/nmf.cpp:103-117)."""
    M = _csr(m, n, density, seed, empty_rows, empty_cols)
    Wt = D.uniform(D.DATA_SEED, 900 + seed, (m, k))
    Ht = D.uniform(D.DATA_SEED, 901 + seed, (k, n))
    r = np.repeat(np.arange(m), np.diff(M.indptr))
    M.data = np.einsum("ek,ke->e", Wt[r], Ht[:, M.indices]).astype(np.float32)
    return M


@pytest.mark.parametrize("algo", [P.MU, P.HALS, P.ANLSBPP])
@pytest.mark.parametrize("m,n,k,density", [(500, 400, 8, 0.05), (20000, 8000, 16, 0.002)])
def test_nmf_sparse_matches_oracle(ctx, algo, m, n, k, density):
    """HALS runs on a masked low-rank matrix of density 0.5.  At density 0.2
    or on a random sparse matrix its 1e-17 clamps collapse columns at points
    so close that two FP64 evaluation orders already part by 2e-3 in relative
    error by the third iteration (the trajectory is chaotic, not an FP32
    effect); at density 0.5 an FP32 restatement stays within 4e-9 of the
    FP64 oracle over ten iterations."""
    if algo == P.HALS:
        A = _csr_lowrank(m // 4, n // 4, k, 0.5, 100, empty_rows=(1,), empty_cols=(3,))
        m, n = m // 4, n // 4
    else:
        A = _csr(m, n, density, 100 + algo, empty_rows=(1,), empty_cols=(3,))
    W0, H0 = D.init_factors([(m, k), (n, k)], stream=33)
    W0, H0 = W0.astype(np.float32), H0.astype(np.float32)
    iters = 10
    Wr, Hr, rel_r, rc = O.nmf_sparse(algo, m, n, A.indptr, A.indices, A.data.astype(np.float64),
                                     W0.astype(np.float64), H0.astype(np.float64), iters)
    assert rc == 0
    dW, dH = P.colmajor(W0), P.colmajor(H0)
    rel = ctx.nmf_sparse_run(m, n, *_dev(A), dW, dH, algo, iters)
    say(f"  sparse nmf algo={algo} {m}x{n}: gpu {rel[-1]:.10f} oracle {rel_r[-1]:.10f}")
    tol = 1e-4 if algo == P.ANLSBPP else 1e-5
    assert np.abs(rel - rel_r).max() < tol
    ftol = 1e-3 if algo == P.ANLSBPP else 1e-4
    assert rel_err(dW.cpu().numpy(), Wr) < ftol
    assert rel_err(dH.cpu().numpy(), Hr) < ftol
    assert np.all(np.diff(rel) <= 1e-7)


def test_nmf_sparse_rejects_16bit(ctx):
    A = _csr(50, 40, 0.1, 5)
    W0, H0 = D.init_factors([(50, 3), (40, 3)], stream=34)
    d = P.NMFDesc()
    d.algo, d.dtype, d.m, d.n, d.k, d.iters, d.compute_error = P.MU, P.BF16, 50, 40, 3, 1, 1
    ip, ix, vv = _dev(A)
    dW, dH = P.colmajor(W0.astype(np.float32)), P.colmajor(H0.astype(np.float32))
    import ctypes
    rc = ctx.lib.planc_nmf_sparse_run(ctx.ctx, ctypes.byref(d), ctypes.c_void_p(ip.data_ptr()),
                                      ctypes.c_void_p(ix.data_ptr()), ctypes.c_void_p(vv.data_ptr()),
                                      ctypes.c_void_p(dW.data_ptr()), ctypes.c_void_p(dH.data_ptr()), None)
    assert rc == P._lib.PLANC_EUNSUPPORTED
