"""Pins of the torch flow reference (CPU): log|det| equals the log-determinant of the autograd
Jacobian, the inverse undoes the forward, and the identity initialisation is the identity."""
import torch

import os, sys
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import flow_ref as F  # noqa: E402


def _params(seed, scale):
    g = torch.Generator().manual_seed(seed)
    return (F.identity_params(seed) + scale * torch.randn(F.L * F.NPL, generator=g)).double()


def test_identity_initialisation():
    y = torch.rand(500, 3, dtype=torch.float64)
    z, ld = F.forward(y, F.identity_params(3).double())
    # the float32-stored derivative bias sets the knot slopes to 1 within 6e-8
    assert (z - y).abs().max() < 1e-9 and ld.abs().max() < 1e-7


def test_logdet_is_log_jacobian():
    """In FP64 without the BF16 rounding (whose straight-through gradients would blur the full
    Jacobian), log|det| of the autograd Jacobian equals the summed spline log-derivatives."""
    p = _params(1, 0.1)
    y = torch.rand(6, 3, dtype=torch.float64)
    F.BF16 = False
    try:
        for i in range(6):
            J = torch.autograd.functional.jacobian(lambda t: F.forward(t.view(1, 3), p)[0].view(3), y[i])
            assert abs(torch.logdet(J).item() - F.forward(y[i:i + 1], p)[1].item()) < 1e-10
    finally:
        F.BF16 = True


def test_inverse_round_trip():
    p = _params(2, 0.2)
    y = torch.rand(2000, 3, dtype=torch.float64)
    z, ld = F.forward(y, p)
    yb, ldi = F.inverse(z, p)
    assert (yb - y).abs().max() < 1e-9 and (ld + ldi).abs().max() < 1e-8
    assert torch.all((z >= 0) & (z <= 1))
