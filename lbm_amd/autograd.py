"""torch.autograd face of the differentiation C ABI, and the coarse-to-fine
porosity map of the inverse problem.

`LbmObjective.apply(alpha, tape)` runs lbm_adj_forward on a fine porosity
field (torch tensor, (zl, yl, xl), float64) and returns J as a 0-d tensor;
its backward calls lbm_adj_backward. A porosity on the domain's GPU goes
through lbm_adj_forward_device / lbm_adj_backward_device and never visits
the host. `upsample_periodic` is the trilinear,
periodic coarse -> fine map as a torch expression, so torch differentiates
the composition (gradient with respect to the coarse grid) by itself.
"""
from __future__ import annotations

import numpy as np


def trilinear_matrix(n_coarse: int, n_fine: int) -> np.ndarray:
    """(n_fine, n_coarse) 1-D periodic linear interpolation: fine cell i sits
    at coarse coordinate (i + 1/2) n_coarse / n_fine - 1/2."""
    m = np.zeros((n_fine, n_coarse))
    for i in range(n_fine):
        xc = (i + 0.5) * n_coarse / n_fine - 0.5
        i0 = int(np.floor(xc))
        t = xc - i0
        m[i, i0 % n_coarse] += 1.0 - t
        m[i, (i0 + 1) % n_coarse] += t
    return m


def upsample_np(coarse: np.ndarray, fine_shape) -> np.ndarray:
    """Trilinear periodic upsampling of a (zc, yc, xc) array to fine_shape (z, y, x)."""
    mz = trilinear_matrix(coarse.shape[0], fine_shape[0])
    my = trilinear_matrix(coarse.shape[1], fine_shape[1])
    mx = trilinear_matrix(coarse.shape[2], fine_shape[2])
    return np.einsum("Zz,Yy,Xx,zyx->ZYX", mz, my, mx, coarse, optimize=True)


def upsample_np_adjoint(fine_grad: np.ndarray, coarse_shape) -> np.ndarray:
    """Transpose of upsample_np."""
    mz = trilinear_matrix(coarse_shape[0], fine_grad.shape[0])
    my = trilinear_matrix(coarse_shape[1], fine_grad.shape[1])
    mx = trilinear_matrix(coarse_shape[2], fine_grad.shape[2])
    return np.einsum("Zz,Yy,Xx,ZYX->zyx", mz, my, mx, fine_grad, optimize=True)


def _torch():
    import torch
    return torch


def upsample_periodic(coarse, fine_shape):
    """torch version of upsample_np (differentiable by torch)."""
    torch = _torch()
    mats = [torch.as_tensor(trilinear_matrix(coarse.shape[k], fine_shape[k]), dtype=coarse.dtype,
                            device=coarse.device) for k in range(3)]
    t = torch.tensordot(mats[0], coarse, dims=([1], [0]))            # Z y x
    t = torch.tensordot(mats[1], t, dims=([1], [1])).transpose(0, 1)  # Z Y x
    t = torch.tensordot(t, mats[2], dims=([2], [1]))                  # Z Y X
    return t


def make_objective_function():
    torch = _torch()

    class LbmObjective(torch.autograd.Function):
        """J(alpha) through lbm_adj_forward / lbm_adj_backward."""

        @staticmethod
        def forward(ctx, alpha, tape):
            ctx.tape = tape
            ctx.shape = alpha.shape
            ctx.dtype, ctx.device = alpha.dtype, alpha.device
            # a porosity already on the domain's GPU stays there
            ctx.on_device = alpha.is_cuda and alpha.device.index == tape.dom.device
            if ctx.on_device:
                j = tape.forward_device(alpha.detach().to(torch.float64).contiguous())
            else:
                j = tape.forward(alpha.detach().cpu().double().numpy())
            return torch.tensor(j, dtype=alpha.dtype, device=alpha.device)

        @staticmethod
        def backward(ctx, grad_out):
            if ctx.on_device:
                g, _ = ctx.tape.backward_device(float(grad_out))
                return g.to(ctx.dtype).reshape(ctx.shape), None
            g, _ = ctx.tape.backward(float(grad_out))
            return torch.as_tensor(g, dtype=ctx.dtype, device=ctx.device).reshape(ctx.shape), None

    return LbmObjective
