"""torch.autograd.Function for the CP reconstruction loss
0.5 ||X - [[A_0..A_{N-1}]]||^2, backed by libplanc_hip.so: the forward is the
fused reconstruct + residual kernel (planc_cp_loss_fwd), the backward the
MTTKRP-based gradient A_n S_n - MTTKRP_n^T (planc_cp_loss_bwd) with the
Khatri-Rao products recomputed on the fly.

X must be FP32 with mode 0 fastest; factors FP32 I_n x k column-major
(planc_amd.colmajor / empty_colmajor).
"""
import torch

from . import Context

_CTX = {}


def context(device=0):
    if device not in _CTX:
        _CTX[device] = Context(device)
    return _CTX[device]


def close():
    """Destroys the contexts this module opened (also done at interpreter
    exit by planc_amd.close_all)."""
    while _CTX:
        _, c = _CTX.popitem()
        c.close()


class CPLoss(torch.autograd.Function):
    @staticmethod
    def forward(fctx, X, *factors):
        c = context(X.device.index or 0)
        loss = c.cp_loss(X, list(factors))
        fctx.save_for_backward(X, *factors)
        return torch.tensor(loss, dtype=torch.float64, device=X.device)

    @staticmethod
    def backward(fctx, grad_out):
        X, *factors = fctx.saved_tensors
        c = context(X.device.index or 0)
        grads = c.cp_loss_grad(X, factors)
        g = grad_out.to(torch.float32)
        return (None, *[gr * g for gr in grads])


def cp_loss(X, factors):
    """Differentiable CP loss (a 0-d FP64 tensor)."""
    return CPLoss.apply(X, *factors)
