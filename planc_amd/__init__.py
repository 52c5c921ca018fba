"""planc_amd -- MI355X (gfx950) implementation of PLANC's shared-memory NMF/NTF
hot path behind a C ABI (include/planc_hip.h, libplanc_hip.so).

This module is the Python face of that ABI, used by the tests and bench.py;
the C++ face is planc_amd/host (NMF<T>, AUNTF, NCPFactors, ParseCommandLine).
Matrices follow Armadillo's column-major layout: a torch tensor of shape
(m, n) with strides (1, ld).  Tensors are stored with mode 0 fastest
(strides (1, I0, I0*I1, ...)).  PyTorch only provides device memory and the
stream; every computation is a HIP kernel of libplanc_hip.so.
"""
import atexit
import ctypes
import sys
import weakref

import numpy as np

from . import _lib
from ._lib import (ANLSBPP, AOADMM, BF16, F16, F32, GNSYM, HALS, MU, NESTEROV, R2, HierDesc,  # noqa: F401
                   HierNode, NMFDesc, NTFDesc)

__all__ = ["Context", "PlancError", "colmajor", "F32", "BF16", "F16", "MU", "HALS", "ANLSBPP"]


class PlancError(RuntimeError):
    def __init__(self, name, rc, msg):
        super().__init__(f"{name} failed with status {rc}: {msg}")
        self.rc = rc


def _torch():
    import torch
    return torch


def colmajor(a, device="cuda", dtype=None):
    """numpy array -> torch device tensor with Fortran (column-major) strides."""
    torch = _torch()
    arr = np.asfortranarray(a if dtype is None else np.asarray(a, dtype=dtype))
    t = torch.from_numpy(arr)
    return t.to(device)


def empty_colmajor(shape, dtype=None, device="cuda"):
    torch = _torch()
    dtype = dtype or torch.float32
    rev = tuple(reversed(shape))
    t = torch.empty(rev, dtype=dtype, device=device)
    return t.permute(*reversed(range(len(shape))))


def _ld(t):
    """Leading dimension of a column-major 2-D tensor (element (i, j) at
    i + ld*j); the stride of a dimension of extent 1 is irrelevant."""
    if t.dim() != 2:
        raise ValueError("expected a 2-D tensor")
    m, n = t.shape
    if m > 1 and t.stride(0) != 1:
        raise ValueError("expected a column-major 2-D tensor (stride(0) == 1)")
    ld = t.stride(1) if n > 1 else m
    return max(ld, m, 1)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _is_fortran_tensor(t):
    exp = 1
    for d in range(t.dim()):
        if t.shape[d] > 1 and t.stride(d) != exp:
            return False
        exp *= t.shape[d]
    return True


# Contexts still open when the interpreter exits are destroyed by an atexit
# hook, while the HIP runtime and torch's streams are still alive.  It is
# registered on the first Context (after torch has initialised HIP, so it
# runs before torch's own exit handlers).  A context collected during
# interpreter finalisation is not destroyed: its stream, graphs and scratch
# go with the process.
_LIVE = weakref.WeakSet()
_ATEXIT_REGISTERED = False


def close_all():
    """Destroys every open Context (called at interpreter exit)."""
    for c in list(_LIVE):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """One planc_ctx (device, stream, scratch).  Calls are ordered on torch's
    current stream of that device."""

    def __init__(self, device=0, use_torch_stream=True):
        global _ATEXIT_REGISTERED
        self.lib = _lib.load()
        self.device = device
        p = ctypes.c_void_p()
        rc = self.lib.planc_create(device, ctypes.byref(p))
        if rc != 0:
            raise PlancError("planc_create", rc, "no HIP device")
        self.ctx = p
        if use_torch_stream:
            torch = _torch()
            s = torch.cuda.current_stream(device).cuda_stream
            self._check(self.lib.planc_set_stream(self.ctx, ctypes.c_void_p(s)), "planc_set_stream")
        _LIVE.add(self)
        if not _ATEXIT_REGISTERED:
            atexit.register(close_all)
            _ATEXIT_REGISTERED = True

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.planc_destroy(self.ctx)
            self.ctx = None
        _LIVE.discard(self)

    def __del__(self, _finalizing=sys.is_finalizing):
        # (the bound default keeps the check usable after the sys module has
        # been torn down at interpreter exit)
        try:
            if _finalizing():
                return
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, name):
        if rc != 0:
            msg = self.lib.planc_last_error(self.ctx)
            raise PlancError(name, rc, msg.decode() if msg else "")

    def sync(self):
        self._check(self.lib.planc_sync(self.ctx), "planc_sync")

    # ---- synthetic inputs on the device ----------------------------------
    def gen_uniform(self, shape, seed, stream):
        """Column-major FP32 U[0,1) tensor (same stream as datagen.uniform)."""
        out = empty_colmajor(shape, device=f"cuda:{self.device}")
        n = int(np.prod(shape))
        self._check(self.lib.planc_gen_uniform(self.ctx, _ptr(out), n, seed, stream),
                    "planc_gen_uniform")
        return out

    def gen_lowrank(self, W0, H0, sigma, seed, stream):
        m, r = W0.shape
        n = H0.shape[1]
        A = empty_colmajor((m, n), device=W0.device)
        self._check(self.lib.planc_gen_lowrank(self.ctx, m, n, r, _ptr(W0), _ptr(H0), float(sigma),
                                               seed, stream, _ptr(A), m), "planc_gen_lowrank")
        return A

    def gen_cp_tensor(self, factors, sigma, seed, stream):
        N = len(factors)
        dims = [f.shape[0] for f in factors]
        X = empty_colmajor(tuple(dims), device=factors[0].device)
        cd = (ctypes.c_long * N)(*dims)
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        self._check(self.lib.planc_gen_cp_tensor(self.ctx, N, cd, arr, factors[0].shape[1],
                                                 float(sigma), seed, stream, _ptr(X)),
                    "planc_gen_cp_tensor")
        return X

    # ---- kernel-level ----------------------------------------------------
    def convert(self, A, dtype, scaled=True):
        """16-bit copy of a column-major FP32 matrix; returns (A16, colscale)."""
        torch = _torch()
        m, n = A.shape
        A16 = empty_colmajor((m, n), dtype=torch.int16, device=A.device)
        cs = torch.empty(n, dtype=torch.float32, device=A.device) if scaled else None
        self._check(self.lib.planc_convert(self.ctx, dtype, _ptr(A), m, n, _ld(A), _ptr(A16), m,
                                           _ptr(cs)), "planc_convert")
        return A16, cs

    def gemm_nn(self, A, B, dtype=F32, colscale=None):
        """C = A (L x K) * B (K x k); A may be a 16-bit copy from convert()."""
        torch = _torch()
        L, K = A.shape
        k = B.shape[1]
        C = empty_colmajor((L, k), device=A.device)
        self._check(self.lib.planc_gemm_nn(self.ctx, dtype, _ptr(A), _ld(A), L, K, _ptr(colscale),
                                           _ptr(B), _ld(B), k, _ptr(C), L), "planc_gemm_nn")
        return C

    def gemm_tn(self, A, B, dtype=F32, colscale=None):
        """C (k x L) = B^T (k x K) * A (K x L)."""
        K, L = A.shape
        k = B.shape[1]
        C = empty_colmajor((k, L), device=A.device)
        self._check(self.lib.planc_gemm_tn(self.ctx, dtype, _ptr(A), _ld(A), K, L, _ptr(colscale),
                                           _ptr(B), _ld(B), k, _ptr(C), k), "planc_gemm_tn")
        return C

    def read_bandwidth(self, buf, reps=20):
        """Device read bandwidth (GB/s) over the bytes of a device tensor
        (16-byte aligned, at least 16 MiB): the achievable ceiling that
        roofline fractions are quoted against besides the 8 TB/s nameplate."""
        out = ctypes.c_double(0.0)
        nbytes = buf.numel() * buf.element_size()
        self._check(self.lib.planc_read_bandwidth(self.ctx, _ptr(buf), nbytes, int(reps), ctypes.byref(out)),
                    "planc_read_bandwidth")
        return out.value

    def gram(self, X):
        torch = _torch()
        rows, k = X.shape
        G = empty_colmajor((k, k), dtype=torch.float64, device=X.device)
        self._check(self.lib.planc_gram(self.ctx, _ptr(X), rows, k, _ld(X), _ptr(G)), "planc_gram")
        return G

    def mu_update(self, H, N, G, eps, n_rs=None, n_cs=None):
        n, k = H.shape
        if n_rs is None:
            n_rs, n_cs = 1, _ld(N)
        self._check(self.lib.planc_mu_update(self.ctx, _ptr(H), n, k, _ld(H), _ptr(N), n_rs, n_cs,
                                             _ptr(G), float(eps)), "planc_mu_update")

    def hals_update_rows(self, H, N, G, n_rs=None, n_cs=None):
        n, k = H.shape
        if n_rs is None:
            n_rs, n_cs = 1, _ld(N)
        self._check(self.lib.planc_hals_update_rows(self.ctx, _ptr(H), n, k, _ld(H), _ptr(N), n_rs,
                                                    n_cs, _ptr(G)), "planc_hals_update_rows")

    def hals_update_cols(self, W, AH, G):
        m, k = W.shape
        self._check(self.lib.planc_hals_update_cols(self.ctx, _ptr(W), m, k, _ld(W), _ptr(AH),
                                                    _ld(AH), _ptr(G)), "planc_hals_update_cols")

    def normalize_by_w(self, W, H):
        m, k = W.shape
        n = H.shape[0]
        self._check(self.lib.planc_normalize_by_w(self.ctx, _ptr(W), m, _ld(W), _ptr(H), n, _ld(H),
                                                  k), "planc_normalize_by_w")

    def normalize_cols(self, A):
        torch = _torch()
        rows, k = A.shape
        lam = torch.empty(k, dtype=torch.float64, device=A.device)
        self._check(self.lib.planc_normalize_cols(self.ctx, _ptr(A), rows, k, _ld(A), _ptr(lam)),
                    "planc_normalize_cols")
        return lam

    def hadamard_grams(self, grams, skip):
        torch = _torch()
        k = grams[0].shape[0]
        out = empty_colmajor((k, k), dtype=torch.float64, device=grams[0].device)
        arr = (ctypes.c_void_p * len(grams))(*[g.data_ptr() for g in grams])
        self._check(self.lib.planc_hadamard_grams(self.ctx, arr, len(grams), skip, k, _ptr(out)),
                    "planc_hadamard_grams")
        return out

    def mttkrp(self, X, factors, mode, dtype=F32):
        """MTTKRP of `mode` (k x I_mode); X has mode-0-fastest strides."""
        N = len(factors)
        if not _is_fortran_tensor(X) and dtype == F32:
            raise ValueError("X must be stored mode-0 fastest")
        dims = (ctypes.c_long * N)(*[f.shape[0] for f in factors])
        k = factors[0].shape[1]
        out = empty_colmajor((k, factors[mode].shape[0]), device=factors[0].device)
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        self._check(self.lib.planc_mttkrp(self.ctx, dtype, _ptr(X), N, dims, arr, k, mode, _ptr(out),
                                          k), "planc_mttkrp")
        return out

    def mttkrp_partial(self, X, factors, split, side, dtype=F32, out=None):
        """Dimension-tree partial MTTKRP: side 0 -> L x k right contraction,
        side 1 -> R x k left contraction (split s: left = modes 0..s)."""
        N = len(factors)
        dims = [f.shape[0] for f in factors]
        k = factors[0].shape[1]
        L = int(np.prod(dims[:split + 1]))
        R = int(np.prod(dims[split + 1:]))
        rows = L if side == 0 else R
        if out is None:
            out = empty_colmajor((rows, k), device=factors[0].device)
        cd = (ctypes.c_long * N)(*dims)
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        self._check(self.lib.planc_mttkrp_partial(self.ctx, dtype, _ptr(X), N, cd, arr, k, split,
                                                  side, _ptr(out), rows), "planc_mttkrp_partial")
        return out

    def multi_ttv(self, Pm, factors, target):
        """Contract projection Pm (E x k) over the block modes except target."""
        nb = len(factors)
        k = factors[0].shape[1]
        out = empty_colmajor((k, factors[target].shape[0]), device=Pm.device)
        cd = (ctypes.c_long * nb)(*[f.shape[0] for f in factors])
        arr = (ctypes.c_void_p * nb)(*[f.data_ptr() for f in factors])
        self._check(self.lib.planc_multi_ttv(self.ctx, _ptr(Pm), _ld(Pm), nb, cd, arr, target, k,
                                             _ptr(out), k), "planc_multi_ttv")
        return out

    def bpp(self, G, B):
        """Batched FP64 BPP NNLS: X (k x nrhs) with X(:,j) = argmin over x >= 0
        given G = A^T A (k x k FP64) and B = A^T [a_j] (k x nrhs FP32)."""
        k, nrhs = B.shape
        X = empty_colmajor((k, nrhs), device=B.device)
        self._check(self.lib.planc_bpp(self.ctx, _ptr(G), k, _ptr(B), _ld(B), 1, nrhs, _ptr(X), k,
                                       1), "planc_bpp")
        return X

    def admm_update(self, G, R, X, D):
        """One AO-ADMM solve in place: X (n x k FP32) and its dual D (n x k)
        given the regularised Gram G (k x k FP64) and the right-hand side R
        (k x n FP32); raises PlancError (rc PLANC_ENOCONV) when G + rho I is
        not positive definite, leaving X and D as they were."""
        k, n = R.shape
        if G.shape != (k, k) or X.shape != (n, k) or D.shape != (n, k):
            raise ValueError("admm_update: G must be k x k, R k x n, X and D n x k")
        self._check(self.lib.planc_admm_update(self.ctx, _ptr(G), k, _ptr(R), _ld(R), 1, _ptr(X),
                                               _ld(X), _ptr(D), _ld(D), n), "planc_admm_update")
        return X, D

    def cp_loss(self, X, factors):
        """0.5 ||X - [[A_0..A_{N-1}]]||^2 (FP64 float), fused on the device."""
        N = len(factors)
        cd = (ctypes.c_long * N)(*[f.shape[0] for f in factors])
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        out = ctypes.c_double()
        self._check(self.lib.planc_cp_loss_fwd(self.ctx, _ptr(X), N, cd, arr, factors[0].shape[1],
                                               ctypes.byref(out)), "planc_cp_loss_fwd")
        return out.value

    def cp_loss_noalloc(self, X, factors, out):
        """Same loss through planc_cp_loss_fwd_noalloc: written to out (a
        device float64 tensor with one element), stream-ordered; the library
        allocates nothing for it."""
        N = len(factors)
        cd = (ctypes.c_long * N)(*[f.shape[0] for f in factors])
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        self._check(self.lib.planc_cp_loss_fwd_noalloc(self.ctx, _ptr(X), N, cd, arr,
                                                       factors[0].shape[1], _ptr(out)),
                    "planc_cp_loss_fwd_noalloc")
        return out

    def cp_loss_grad(self, X, factors):
        """[A_n S_n - MTTKRP_n^T] for every mode (I_n x k, column-major)."""
        N = len(factors)
        grads = [empty_colmajor(tuple(f.shape), device=f.device) for f in factors]
        cd = (ctypes.c_long * N)(*[f.shape[0] for f in factors])
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        garr = (ctypes.c_void_p * N)(*[g.data_ptr() for g in grads])
        self._check(self.lib.planc_cp_loss_bwd(self.ctx, _ptr(X), N, cd, arr, factors[0].shape[1],
                                               garr), "planc_cp_loss_bwd")
        return grads

    def r2_update(self, G, R):
        """Rank-2 NNLS update of DistR2 (distnmf/distr2.hpp:35-61): G a 2 x 2
        FP64 device Gram, R a 2 x n FP32 device right-hand side (column-major);
        returns X (2 x n FP32)."""
        torch = _torch()
        assert G.dtype == torch.float64 and G.shape == (2, 2) and R.shape[0] == 2
        n = R.shape[1]
        X = empty_colmajor((2, n), device=R.device)
        self._check(self.lib.planc_r2_update(self.ctx, _ptr(G), _ptr(R), _ld(R), 1, n, _ptr(X), _ld(X), 1),
                    "planc_r2_update")
        return X

    def rel_error_gram(self, normA2, trace, WtW, HtH):
        """Relative NMF error from the trace identity (device FP64 scalars and
        k x k Grams); returns it as a host float."""
        torch = _torch()
        k = WtW.shape[0]
        out = torch.empty(1, dtype=torch.float64, device=WtW.device)
        self._check(self.lib.planc_rel_error_gram(self.ctx, _ptr(normA2), _ptr(trace), _ptr(WtW), _ptr(HtH), k,
                                                  _ptr(out)), "planc_rel_error_gram")
        return float(out.cpu()[0])

    def csr_col_subset(self, indptr, indices, data, n, cols):
        """Device CSR column subset A(:, cols) (cols sorted, unique).  indptr,
        indices, data are device tensors (int64, int32, float32)."""
        torch = _torch()
        m = indptr.numel() - 1
        c = np.ascontiguousarray(cols, dtype=np.int64)
        out_ptr = torch.empty(m + 1, dtype=torch.int64, device=indptr.device)
        cnt = ctypes.c_long()
        lp = ctypes.POINTER(ctypes.c_long)
        self._check(self.lib.planc_csr_col_subset(self.ctx, m, n, _ptr(indptr), _ptr(indices),
                                                  _ptr(data), c.ctypes.data_as(lp), len(c),
                                                  _ptr(out_ptr), None, None, ctypes.byref(cnt)),
                    "planc_csr_col_subset")
        oi = torch.empty(max(cnt.value, 1), dtype=torch.int32, device=indptr.device)
        ov = torch.empty(max(cnt.value, 1), dtype=torch.float32, device=indptr.device)
        self._check(self.lib.planc_csr_col_subset(self.ctx, m, n, _ptr(indptr), _ptr(indices),
                                                  _ptr(data), c.ctypes.data_as(lp), len(c),
                                                  _ptr(out_ptr), _ptr(oi), _ptr(ov),
                                                  ctypes.byref(cnt)), "planc_csr_col_subset")
        return out_ptr, oi[:cnt.value], ov[:cnt.value]

    def hiernmf_run(self, m, n, indptr, indices, data, depth, iters, initseed, want_H=False,
                    want_W=False, nodes=0):
        """Hierarchical rank-2 NMF (host CSR input): the complete tree of
        `depth` levels (nodes <= 0) or the best-first frontier with a budget
        of `nodes` (depth capping the levels).  Returns (nodes: list of
        dicts in the reported order, cols (int64 array), H or None, W or
        None; W holds 2*m doubles per reported node, in that order)."""
        d = HierDesc(m, n, depth, iters, initseed, nodes)
        cap = max(3, 2 * nodes - 1) if nodes > 0 else (1 << depth) - 1
        table = (HierNode * cap)()
        nn = ctypes.c_int()
        ip = np.ascontiguousarray(indptr, dtype=np.int64)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        dv = np.ascontiguousarray(data, dtype=np.float32)
        cols = np.zeros(max(n * depth, 1), dtype=np.int64)
        H = np.zeros(2 * max(n * depth, 1)) if want_H else None
        W = np.zeros(2 * m * cap) if want_W else None
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(self.lib.planc_hiernmf_run(
            self.ctx, ctypes.byref(d), ip.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
            ix.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            dv.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), table, ctypes.byref(nn),
            cols.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
            H.ctypes.data_as(dp) if H is not None else None,
            W.ctypes.data_as(dp) if W is not None else None), "planc_hiernmf_run")
        out = []
        for i in range(nn.value):
            x = table[i]
            out.append({"index": x.index, "level": x.level, "ncols": x.ncols,
                        "col_offset": x.col_offset, "sigma": x.sigma, "score": x.score,
                        "lvalid": x.lvalid, "rvalid": x.rvalid, "split": x.split,
                        "parent_side": x.parent_side, "slot": i})
        return out, cols, H, W

    # ---- whole loops -------------------------------------------------------
    def transpose16(self, A16):
        """Transposed copy (n x m, column-major) of a 16-bit m x n matrix."""
        torch = _torch()
        m, n = A16.shape
        out = empty_colmajor((n, m), dtype=A16.dtype, device=A16.device)
        self._check(self.lib.planc_transpose16(self.ctx, _ptr(A16), m, n, _ld(A16), _ptr(out), n),
                    "planc_transpose16")
        return out

    def nmf_run(self, A, W, H, algo, iters, dtype=F32, regW=(0.0, 0.0), regH=(0.0, 0.0),
                compute_error=True, A16=None, colscale=None, symm_reg=-1.0, A16T=None,
                normA2=0.0):
        """computeNMF on the device; W (m x k) and H (n x k) are updated in
        place.  Returns the per-iteration relative error (numpy).  symm_reg > 0
        turns on ANLS-BPP's symmetric regularisation (square A).  A16T (from
        transpose16(A16)) and normA2 (||A||_F^2) may be passed in once made,
        as the reference's constructors form At and normA once."""
        m, n = A.shape
        k = W.shape[1]
        d = NMFDesc()
        d.algo, d.dtype, d.m, d.n, d.k, d.iters = algo, dtype, m, n, k, iters
        d.regW[0], d.regW[1] = regW
        d.regH[0], d.regH[1] = regH
        d.compute_error = 1 if compute_error else 0
        d.symm_reg = float(symm_reg)
        d.A16T = A16T.data_ptr() if A16T is not None else None
        d.normA2 = float(normA2)
        if _ld(W) != m or _ld(H) != n:
            raise ValueError("W and H must be tight column-major (ld = m, n)")
        rel = np.zeros(max(iters, 1))
        self._check(self.lib.planc_nmf_run(self.ctx, ctypes.byref(d), _ptr(A), _ld(A), _ptr(A16),
                                           _ptr(colscale), _ptr(W), _ptr(H),
                                           rel.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "planc_nmf_run")
        return rel[:iters]

    # ---- sparse input (CSR: indptr int64, indices int32, values float32) ----
    def csr_transpose(self, m, n, indptr, indices, values):
        """Device CSR of A^T (n x m); returns (indptr, indices, values)."""
        torch = _torch()
        nnz = int(indices.numel())
        tp = torch.empty(n + 1, dtype=torch.int64, device=indptr.device)
        ti = torch.empty(max(nnz, 1), dtype=torch.int32, device=indptr.device)
        tv = torch.empty(max(nnz, 1), dtype=torch.float32, device=indptr.device)
        self._check(self.lib.planc_csr_transpose(self.ctx, m, n, _ptr(indptr), _ptr(indices), _ptr(values),
                                                 _ptr(tp), _ptr(ti), _ptr(tv)), "planc_csr_transpose")
        return tp, ti[:nnz], tv[:nnz]

    def spmm(self, rows, cols, indptr, indices, values, F):
        """A (rows x cols, CSR) * F (cols x k, column-major) -> rows x k."""
        k = F.shape[1]
        out = empty_colmajor((rows, k), device=F.device)
        self._check(self.lib.planc_spmm(self.ctx, rows, cols, _ptr(indptr), _ptr(indices), _ptr(values),
                                        _ptr(F), _ld(F), k, _ptr(out), rows), "planc_spmm")
        return out

    def sparse_fit_error(self, m, n, indptr, indices, values, W, H):
        """computeObjectiveError(A, W, H) of common/utils.hpp:190-241."""
        err = ctypes.c_double()
        self._check(self.lib.planc_sparse_fit_error(self.ctx, m, n, _ptr(indptr), _ptr(indices),
                                                    _ptr(values), _ptr(W), _ld(W), _ptr(H), _ld(H),
                                                    W.shape[1], ctypes.byref(err)),
                    "planc_sparse_fit_error")
        return err.value

    def nmf_sparse_run(self, m, n, indptr, indices, values, W, H, algo, iters, regW=(0.0, 0.0),
                       regH=(0.0, 0.0), compute_error=True, symm_reg=-1.0):
        """computeNMF of the sparse build on the device; W (m x k) and H
        (n x k) updated in place.  Returns the per-iteration relative error."""
        k = W.shape[1]
        d = NMFDesc()
        d.algo, d.dtype, d.m, d.n, d.k, d.iters = algo, F32, m, n, k, iters
        d.regW[0], d.regW[1] = regW
        d.regH[0], d.regH[1] = regH
        d.compute_error = 1 if compute_error else 0
        d.symm_reg = float(symm_reg)
        if _ld(W) != m or _ld(H) != n:
            raise ValueError("W and H must be tight column-major (ld = m, n)")
        rel = np.zeros(max(iters, 1))
        self._check(self.lib.planc_nmf_sparse_run(self.ctx, ctypes.byref(d), _ptr(indptr), _ptr(indices),
                                                  _ptr(values), _ptr(W), _ptr(H),
                                                  rel.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "planc_nmf_sparse_run")
        return rel[:iters]

    def ntf_tree_split(self, dims):
        """The dimension tree's split s (left block modes 0..s)."""
        d = (ctypes.c_long * len(dims))(*dims)
        s = self.lib.planc_ntf_tree_split(len(dims), d)
        if s < 0:
            raise ValueError(f"bad tensor dimensions {dims}")
        return s

    def ntf_copies16(self, X, dtype):
        """Unscaled 16-bit copy of the tensor X (modes 0-fastest in memory,
        any torch shape) and its dimension-tree transpose, the pair
        ntf_run(X16=, X16T=) takes: made once, as the reference forms At once
        for NMF."""
        torch = _torch()
        dims = [int(x) for x in X.shape]
        s = self.ntf_tree_split(dims)
        L = int(np.prod(dims[:s + 1]))
        R = int(np.prod(dims[s + 1:]))
        flatX = X.permute(*reversed(range(X.dim()))).reshape(-1)[:, None]
        X16, _ = self.convert(flatX, dtype, scaled=False)
        X16T = torch.empty(L * R, dtype=torch.int16, device=X.device)
        self._check(self.lib.planc_transpose16(self.ctx, _ptr(X16), L, R, L, _ptr(X16T), R),
                    "planc_transpose16")
        return X16, X16T

    def ntf_run(self, X, factors, algo, iters, dtype=F32, dimtree=True, compute_error=True,
                X16=None, X16T=None, normX2=0.0):
        """computeNTF on the device; factors (I_i x k, tight column-major) are
        updated in place.  Returns (lambda, relerr).  X16 / X16T: a caller's
        unscaled 16-bit copy and its dimension-tree transpose (ntf_copies16);
        normX2 (> 0): ||X||_F^2 computed once by the caller."""
        torch = _torch()
        N = len(factors)
        d = NTFDesc()
        d.algo, d.dtype, d.nmodes = algo, dtype, N
        for i, f in enumerate(factors):
            d.dims[i] = f.shape[0]
            if _ld(f) != f.shape[0]:
                raise ValueError("factors must be tight column-major")
        d.k = factors[0].shape[1]
        d.iters, d.dimtree, d.compute_error = iters, 1 if dimtree else 0, 1 if compute_error else 0
        d.X16T = X16T.data_ptr() if X16T is not None else None
        d.normX2 = float(normX2)
        lam = torch.empty(d.k, dtype=torch.float64, device=factors[0].device)
        arr = (ctypes.c_void_p * N)(*[f.data_ptr() for f in factors])
        rel = np.zeros(max(iters, 1))
        self._check(self.lib.planc_ntf_run(self.ctx, ctypes.byref(d), _ptr(X), _ptr(X16), arr,
                                           _ptr(lam),
                                           rel.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "planc_ntf_run")
        return lam, rel[:iters]
