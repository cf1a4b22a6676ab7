"""torch.autograd.Function over the adjoint pass (mi_adjoint).

``sigma(params, ctx, run)`` returns sigma (mb) for params = (alphaS(MZ), renorm_fac, gluon a,
gluon b) as a differentiable FP64 tensor; its backward uses the derivatives the device computed
in the same pass, on the same points (common random numbers), so autograd sees the exact
derivative of the Monte-Carlo estimate.
"""
from __future__ import annotations

import copy

import torch

from . import _abi as A
from .core import Context


def _run_with(run: A.MiRun, params) -> A.MiRun:
    r = copy.copy(run)
    r.phys = copy.copy(run.phys)
    a_s, xi, a_g, b_g = (float(x) for x in params)
    r.phys.alphaS_value = a_s
    r.phys.renorm_fac = xi
    r.phys.pdf_g[0] = a_g
    r.phys.pdf_g[1] = b_g
    return r


class SigmaAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params: torch.Tensor, mctx: Context, run: A.MiRun):
        g = mctx.adjoint(_run_with(run, params.detach().cpu().tolist()))
        grad = torch.tensor([g.d_alphaS_MZ, g.d_renorm_fac, g.d_pdf_g_a, g.d_pdf_g_b], dtype=torch.float64,
                            device=params.device)
        fctx.save_for_backward(grad)
        fctx.errors = (g.sigma_err, g.d_alphaS_MZ_err, g.d_renorm_fac_err, g.d_pdf_g_a_err, g.d_pdf_g_b_err)
        return torch.tensor(g.sigma, dtype=torch.float64, device=params.device)

    @staticmethod
    def backward(fctx, grad_out):
        (grad,) = fctx.saved_tensors
        return grad_out * grad, None, None


def sigma(params: torch.Tensor, mctx: Context, run: A.MiRun) -> torch.Tensor:
    """sigma(params) with d sigma / d params from the adjoint pass."""
    return SigmaAdjoint.apply(params, mctx, run)
