"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so, the FP64 CPU restatement of PLANC's
shared-memory NMF/NTF path (oracle/planc_oracle.cpp).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (planc_amd) never does.

How each oracle function is pinned (tests/test_oracle.py):

* gemm                -- numpy matmul on random operands, all four op() pairs.
* bppnnls             -- scipy.optimize.nnls (Lawson-Hanson, unique solution for
                         an SPD Gram) on random problems, and the KKT conditions
                         x >= 0, y = AtA x - Atb >= 0, x.y = 0; committed fixtures.
* krp_leave_out_one   -- brute-force product over explicit multi-indices.
* mttkrp (direct and dimension tree)
                      -- numpy.einsum on tiny tensors (committed fixtures), the
                         closed form for rank-1 tensors, and dimtree == direct,
                         the invariant the reference records in
                         distntf/testcases.md:3-67.
* nmf MU / HALS / BPP -- NumPy step-by-step restatements on tiny problems
                         (committed trajectory fixtures), monotone objective
                         (testcases.md:3), the exact fixed point A = W H^T for
                         MU, KKT of each BPP half-step; every HALS column step
                         (H sweep and W sweep) is the clamped coordinate
                         minimiser of ||A - W H^T|| given the other columns,
                         checked through the explicit residual.
* ntf_mode_update     -- HALS: every column of the sweep is the clamped
                         coordinate minimiser of 0.5||X_(n) - H K^T||^2 given the
                         columns before it (new) and after it (old), checked
                         through the gradient of the explicit residual; BPP: KKT
                         of the same objective; MU: at an exact factorisation the
                         step shrinks H by exactly eps/(HS + eps) with eps = 1e-6
                         (ntf/ntfmu.hpp:13), and the objective does not increase.
* ntf HALS / MU / BPP -- one iteration equals mode updates on S and M built from
                         an explicit (brute-force) Khatri-Rao product, with the
                         column normalisation and lambda of normalize(mode);
                         monotone error; relative error checked against a direct
                         numpy reconstruction; dimtree == direct.
* cp_loss             -- central differences in FP64 (h = 1e-6) and the einsum
                         reconstruction.
* admm_solve / AO-ADMM NMF and NTF
                      -- repeated solves (the dual carried over) converge to the
                         NNLS solution (scipy.optimize.nnls, unique for an SPD
                         Gram); one AO-ADMM NMF iteration against a NumPy
                         restatement; NTF AO-ADMM dimtree == direct and its
                         error against a direct reconstruction.
* NTF-NES             -- Jacobi eigenvalues against numpy.linalg.eigvalsh; a mode
                         update against the BPP (pinned) solution of the same
                         proximal NNLS (S + lambda I) x = m + lambda h*, to the
                         reference's absolute stopping tolerance; the error
                         trajectory against a direct reconstruction; dimtree ==
                         direct; the accelerated run ends below the plain one's
                         start.
* GNSym               -- the exact factorisation A = H H^T is a fixed point
                         (zero gradient, no step, zero error); the fast error
                         formula against ||A - H H^T||_F; one iteration against a
                         NumPy restatement; decreasing error on a well-posed case.
* nmf_sparse          -- the dense loop on the densified matrix (identical
                         trajectories for MU / HALS), and for ANLS-BPP the dense
                         loop started from normalize_by_W'd factors (the sparse
                         build's start, nmf/bppnmf.hpp:195-201).
* sparse_fit_error    -- the direct dense ||A - W H^T||_F on small matrices.
* r2_update           -- brute force over the four active sets of the
                         two-variable NNLS.
* r2_nmf_csr          -- dense NumPy loop of the same DistAUNMF order on a tiny
                         node; every H/W step a KKT point of its 2-variable
                         NNLS (r2_update pinned above).
* powiter_csr         -- numpy.linalg.svd: sigma_max(A_node)^2 on a matrix
                         with a clear spectral gap.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.oracle_gemm.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_long, ctypes.c_long,
                                  ctypes.c_long, _dp, ctypes.c_long, _dp, ctypes.c_long, _dp,
                                  ctypes.c_long]
        L.oracle_bppnnls.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_long, _dp]
        L.oracle_bppnnls.restype = ctypes.c_int
        L.oracle_nmf.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_long, _dp, _dp,
                                 ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp]
        L.oracle_nmf.restype = ctypes.c_int
        L.oracle_krp_leave_out_one.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                                               ctypes.POINTER(_dp), ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_mttkrp.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), _dp,
                                    ctypes.POINTER(_dp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _dp]
        L.oracle_ntf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_long), _dp,
                                 ctypes.POINTER(_dp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _dp, _dp]
        L.oracle_ntf.restype = ctypes.c_int
        L.oracle_ntf_mode_update.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_int, _dp, _dp,
                                             _dp]
        L.oracle_ntf_mode_update.restype = ctypes.c_int
        L.oracle_cp_loss.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), _dp,
                                     ctypes.POINTER(_dp), ctypes.c_int, ctypes.POINTER(_dp)]
        L.oracle_cp_loss.restype = ctypes.c_double
        L.oracle_r2_update.argtypes = [_dp, _dp, ctypes.c_long, _dp]
        _lp = ctypes.POINTER(ctypes.c_long)
        _ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_r2_nmf_csr.argtypes = [ctypes.c_long, _lp, _ip, _dp, _lp, ctypes.c_long,
                                        ctypes.c_long, _dp, _dp, ctypes.c_int]
        L.oracle_powiter_csr.argtypes = [ctypes.c_long, _lp, _ip, _dp, _lp, ctypes.c_long,
                                         ctypes.c_long, _dp, ctypes.c_int]
        L.oracle_powiter_csr.restype = ctypes.c_double
        _lp = ctypes.POINTER(ctypes.c_long)
        _ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_nmf_sparse.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long, _lp, _ip, _dp, _dp,
                                        _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.oracle_nmf_sparse.restype = ctypes.c_int
        L.oracle_sparse_fit_error.argtypes = [ctypes.c_long, ctypes.c_long, _lp, _ip, _dp, _dp, _dp,
                                              ctypes.c_int]
        L.oracle_sparse_fit_error.restype = ctypes.c_double
        L.oracle_admm_solve.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_long, _dp, _dp]
        L.oracle_admm_solve.restype = ctypes.c_int
        L.oracle_sym_eig_minmax.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.oracle_nes_mode_update.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_long, _dp, _dp]
        L.oracle_nes_mode_update.restype = ctypes.c_int
        L.oracle_use_blas.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.oracle_use_blas.restype = ctypes.c_int
        L.oracle_last_loop_seconds.restype = ctypes.c_double
        _fp = ctypes.POINTER(ctypes.c_float)
        L.oracle_mu_f32.argtypes = [_fp, ctypes.c_long, ctypes.c_long, _fp, _fp, ctypes.c_int,
                                    ctypes.c_int, _dp]
        L.oracle_mu_f32.restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, order="F"):
    return np.require(np.asarray(a, dtype=np.float64), requirements=[order, "A", "W", "O"])


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return lib().oracle_get_threads()


# Host BLAS libraries tried, in order, for the CPU baseline (bench.py only).
BLAS_CANDIDATES = ("/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.2",
                   "libopenblas.so.0", "libopenblas.so")


def use_blas(nthreads, candidates=BLAS_CANDIDATES):
    """Routes the oracle's GEMMs through a host CBLAS (MKL's GNU-OpenMP layer
    when MKL is found).  Returns (path, has_sgemm) or (None, False) when no
    library loads, in which case the oracle's blocked GEMM stays in use."""
    import os
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    os.environ.setdefault("MKL_INTERFACE_LAYER", "LP64")
    for path in candidates:
        rc = lib().oracle_use_blas(path.encode(), int(nthreads))
        if rc >= 0:
            return path, rc == 1
    return None, False


def no_blas():
    lib().oracle_use_blas(None, 0)


def last_loop_seconds():
    """Seconds of the iteration loop of the last nmf() / mu_f32() call."""
    return lib().oracle_last_loop_seconds()


def mu_f32(A, W0, H0, iters):
    """FP32 MU loop through the loaded cblas_sgemm (CPU baseline only).
    Returns (W, H, relerr, rc); rc -3 when no sgemm is loaded."""
    A = np.require(A, dtype=np.float32, requirements=["F", "A"])
    W = np.require(W0, dtype=np.float32, requirements=["F", "A", "W", "O"]).copy(order="F")
    H = np.require(H0, dtype=np.float32, requirements=["F", "A", "W", "O"]).copy(order="F")
    m, n = A.shape
    k = W.shape[1]
    rel = np.zeros(iters)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = lib().oracle_mu_f32(A.ctypes.data_as(fp), m, n, W.ctypes.data_as(fp),
                             H.ctypes.data_as(fp), k, iters, _p(rel))
    return W, H, rel, rc


def gemm(ta, tb, A, B):
    A = _f64(A)
    B = _f64(B)
    M = A.shape[0] if ta == "N" else A.shape[1]
    K = A.shape[1] if ta == "N" else A.shape[0]
    N = B.shape[1] if tb == "N" else B.shape[0]
    C = np.zeros((M, N), order="F")
    lib().oracle_gemm(ta.encode(), tb.encode(), M, N, K, _p(A), A.shape[0], _p(B), B.shape[0],
                      _p(C), M)
    return C


def bppnnls(AtA, AtB):
    """Returns (X, iterations); iterations is -1 when BPP did not converge."""
    AtA = _f64(AtA)
    AtB = _f64(AtB)
    if AtB.ndim == 1:
        AtB = AtB.reshape(-1, 1, order="F")
    n, nrhs = AtB.shape
    X = np.zeros((n, nrhs), order="F")
    it = lib().oracle_bppnnls(_p(AtA), n, _p(AtB), nrhs, _p(X))
    return X, it


def nmf(algo, A, W0, H0, iters, regW=(0.0, 0.0), regH=(0.0, 0.0), symm_reg=-1.0):
    """algo 0 MU, 1 HALS, 2 ANLS-BPP. Returns (W, H, relerr, objective, rc)."""
    A = _f64(A)
    W = _f64(W0).copy(order="F")
    H = _f64(H0).copy(order="F")
    m, n = A.shape
    k = W.shape[1]
    rel = np.zeros(iters)
    obj = np.zeros(iters)
    rW = np.asarray(regW, dtype=np.float64)
    rH = np.asarray(regH, dtype=np.float64)
    rc = lib().oracle_nmf(algo, _p(A), m, n, _p(W), _p(H), k, iters, _p(rW), _p(rH),
                          float(symm_reg), _p(rel), _p(obj))
    return W, H, rel, obj, rc


def admm_solve(G, R, X, D):
    """One AO-ADMM solve (at most 5 sweeps): returns the new (X, D, rc)."""
    G = _f64(G)
    R = _f64(R)
    X = _f64(X).copy(order="F")
    D = _f64(D).copy(order="F")
    rc = lib().oracle_admm_solve(_p(G), G.shape[0], _p(R), R.shape[1], _p(X), _p(D))
    return X, D, rc


def sym_eig_minmax(S):
    S = _f64(S)
    mn, mx = ctypes.c_double(), ctypes.c_double()
    lib().oracle_sym_eig_minmax(_p(S), S.shape[0], ctypes.byref(mn), ctypes.byref(mx))
    return mn.value, mx.value


def nes_mode_update(S, M, Hprox):
    """NTFNES update of one mode from the proximal centre Hprox (I x k):
    returns (H, rc)."""
    S = _f64(S)
    M = _f64(M)
    Hp = _f64(Hprox)
    H = np.zeros_like(Hp, order="F")
    rc = lib().oracle_nes_mode_update(_p(S), S.shape[0], _p(M), Hp.shape[0], _p(Hp), _p(H))
    return H, rc


def _csr_arrays(indptr, indices, values):
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    vv = np.ascontiguousarray(values, dtype=np.float64)
    return ip, ix, vv


def nmf_sparse(algo, m, n, indptr, indices, values, W0, H0, iters, regW=(0.0, 0.0),
               regH=(0.0, 0.0)):
    """Sparse build of the NMF loop, A as CSR.  Returns (W, H, relerr, rc)."""
    ip, ix, vv = _csr_arrays(indptr, indices, values)
    W = _f64(W0).copy(order="F")
    H = _f64(H0).copy(order="F")
    k = W.shape[1]
    rel = np.zeros(iters)
    rW = np.asarray(regW, dtype=np.float64)
    rH = np.asarray(regH, dtype=np.float64)
    rc = lib().oracle_nmf_sparse(algo, m, n, ip.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                 ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(vv), _p(W), _p(H),
                                 k, iters, _p(rW), _p(rH), _p(rel))
    return W, H, rel, rc


def sparse_fit_error(m, n, col_ptrs, row_indices, values, W, H):
    """computeObjectiveError(A, W, H) of common/utils.hpp:190-241, A as CSC."""
    cp, ri, vv = _csr_arrays(col_ptrs, row_indices, values)
    W = _f64(W)
    H = _f64(H)
    return lib().oracle_sparse_fit_error(m, n, cp.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                         ri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(vv),
                                         _p(W), _p(H), W.shape[1])


def _factor_ptrs(factors):
    fs = [_f64(f) for f in factors]
    arr = (_dp * len(fs))(*[_p(f) for f in fs])
    return fs, arr


def krp_leave_out_one(factors, n):
    fs, arr = _factor_ptrs(factors)
    dims = (ctypes.c_long * len(fs))(*[f.shape[0] for f in fs])
    k = fs[0].shape[1]
    rows = int(np.prod([f.shape[0] for i, f in enumerate(fs) if i != n]))
    out = np.zeros((rows, k), order="F")
    lib().oracle_krp_leave_out_one(len(fs), dims, arr, k, n, _p(out))
    return out


def mttkrp(X, factors, n, method=0, split=None):
    """MTTKRP of mode n as PLANC lays it out: k x I_n.  X is mode-0-fastest
    (pass a numpy array in Fortran order)."""
    Xf = _f64(X)
    fs, arr = _factor_ptrs(factors)
    N = len(fs)
    dims = (ctypes.c_long * N)(*Xf.shape)
    k = fs[0].shape[1]
    out = np.zeros((k, Xf.shape[n]), order="F")
    if split is None:
        split = N // 2
    lib().oracle_mttkrp(N, dims, _p(Xf), arr, k, n, method, split, _p(out))
    return out


def ntf(algo, X, factors0, iters, dimtree=1):
    """algo 0 MU, 1 HALS, 2 BPP. Returns (factors, lambda, relerr, rc)."""
    Xf = _f64(X)
    fs = [_f64(f).copy(order="F") for f in factors0]
    arr = (_dp * len(fs))(*[_p(f) for f in fs])
    N = len(fs)
    dims = (ctypes.c_long * N)(*Xf.shape)
    k = fs[0].shape[1]
    lam = np.zeros(k)
    rel = np.zeros(iters)
    rc = lib().oracle_ntf(algo, N, dims, _p(Xf), arr, k, iters, dimtree, _p(lam), _p(rel))
    return fs, lam, rel, rc


def ntf_mode_update(algo, H, M, S):
    """One NTF mode update (algo 0 MU, 1 HALS, 2 BPP) of the factor H (I x k)
    given M (k x I, the MTTKRP) and S (k x k).  Returns (H_new, rc)."""
    Hn = _f64(H).copy(order="F")
    Mf = _f64(M)
    Sf = _f64(S)
    In, k = Hn.shape
    rc = lib().oracle_ntf_mode_update(algo, In, k, _p(Hn), _p(Mf), _p(Sf))
    return Hn, rc


def cp_loss(X, factors, grads=False):
    Xf = _f64(X)
    fs, arr = _factor_ptrs(factors)
    N = len(fs)
    dims = (ctypes.c_long * N)(*Xf.shape)
    k = fs[0].shape[1]
    if grads:
        gs = [np.zeros(f.shape, order="F") for f in fs]
        garr = (_dp * N)(*[_p(g) for g in gs])
        loss = lib().oracle_cp_loss(N, dims, _p(Xf), arr, k, garr)
        return loss, gs
    return lib().oracle_cp_loss(N, dims, _p(Xf), arr, k, None)


def r2_update(left, right):
    left = _f64(left)
    right = _f64(right)
    out = np.zeros_like(right, order="F")
    lib().oracle_r2_update(_p(left), _p(right), right.shape[1], _p(out))
    return out


def _csr_args(indptr, indices, data):
    ip = np.ascontiguousarray(indptr, dtype=np.int64)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    dv = np.ascontiguousarray(data, dtype=np.float64)
    return ip, ix, dv


def r2_nmf_csr(m, n, indptr, indices, data, cols, W0, H0, iters):
    """Rank-2 NMF of the node A(:, cols) from W0 (m x 2), H0 (ncols x 2)."""
    ip, ix, dv = _csr_args(indptr, indices, data)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    W = _f64(W0).copy(order="F")
    H = _f64(H0).copy(order="F")
    lp = ctypes.POINTER(ctypes.c_long)
    lib().oracle_r2_nmf_csr(m, ip.ctypes.data_as(lp), ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            _p(dv), c.ctypes.data_as(lp), len(c), n, _p(W), _p(H), iters)
    return W, H


def powiter_csr(m, n, indptr, indices, data, cols, q0, iters):
    ip, ix, dv = _csr_args(indptr, indices, data)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    q = np.ascontiguousarray(q0, dtype=np.float64)
    lp = ctypes.POINTER(ctypes.c_long)
    return lib().oracle_powiter_csr(m, ip.ctypes.data_as(lp),
                                    ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(dv),
                                    c.ctypes.data_as(lp), len(c), n, _p(q), iters)
