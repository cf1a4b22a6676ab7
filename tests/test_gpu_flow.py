"""GPU flow kernels (MFMA BF16 MLP + FP32 splines) against the plain PyTorch FP32 reference of
the same architecture (tests/flow_ref.py, which rounds weights and 64-wide activations to BF16
where the device does), and training / integration checks."""
import math
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import flow_ref as F  # noqa: E402

pytestmark = pytest.mark.gpu


def _noisy(seed, scale):
    g = torch.Generator().manual_seed(seed)
    return F.identity_params(seed) + scale * torch.randn(F.L * F.NPL, generator=g)


def test_identity_init(ctx):
    ctx.flow_init(5)
    y = torch.rand(4096, 3, device="cuda")
    z, ld = ctx.flow_forward(y)
    assert (z - y).abs().max().item() < 2e-6 and ld.abs().max().item() < 2e-5
    yb, ldi = ctx.flow_inverse(z)
    assert (yb - y).abs().max().item() < 2e-6


def test_forward_inverse_vs_reference(ctx):
    p = _noisy(3, 0.15)
    ctx.flow_set_params(p.numpy())
    assert np.array_equal(ctx.flow_get_params(), p.numpy())
    y = torch.rand(8192, 3, generator=torch.Generator().manual_seed(1))
    z, ld = ctx.flow_forward(y.cuda())
    zr, ldr = F.forward(y, p)
    dz = (z.cpu() - zr).abs()
    dl = (ld.cpu() - ldr).abs()
    assert dz.max() < 2e-3 and dz.mean() < 2e-5, (dz.max(), dz.mean())
    assert dl.max() < 2e-2 and dl.mean() < 2e-4, (dl.max(), dl.mean())
    yb, ldi = ctx.flow_inverse(z)
    assert (yb - y.cuda()).abs().max().item() < 1e-4
    # The BF16 hidden activations make the MLP piecewise constant at the 1e-3 level: where the
    # inverse sees conditioners that differ from the forward's by the round-trip error (~1e-5), a
    # hidden unit can round the other way.  The torch reference with the same rounding shows the
    # same kind of isolated outliers (max 7e-4, 99.9% quantile 5e-5 on these inputs).
    e = (ldi + ld).abs().cpu()
    q999 = torch.quantile(e, 0.999).item()
    assert e.mean().item() < 1e-5 and q999 < 2e-4 and e.max().item() < 1e-2, (e.mean().item(), q999, e.max().item())
    yr, ldir = F.inverse(zr, p)
    assert (yb.cpu() - yr).abs().mean() < 2e-5


def test_backward_vs_autograd(ctx):
    """Gradient of sum_i g_i log q(y_i): BF16 operands in the backward GEMMs, FP32 accumulation;
    agreement with FP32 autograd of the reference to ~1e-2 relative (L2) per layer."""
    p = _noisy(4, 0.1)
    ctx.flow_set_params(p.numpy())
    gen = torch.Generator().manual_seed(2)
    n = 4096
    y = torch.rand(n, 3, generator=gen)
    g = torch.randn(n, generator=gen) / n
    grad = ctx.flow_backward(y.cuda(), g.cuda()).cpu()
    pr = p.clone().requires_grad_(True)
    _, ld = F.forward(y, pr)
    (gr,) = torch.autograd.grad((g * ld).sum(), pr)
    for l in range(F.L):
        a, b = grad[l * F.NPL:(l + 1) * F.NPL], gr[l * F.NPL:(l + 1) * F.NPL]
        rel = (a - b).norm() / b.norm()
        cos = torch.dot(a, b) / (a.norm() * b.norm())
        assert rel < 3e-2 and cos > 0.999, (l, rel.item(), cos.item())


def test_training_and_unbiased_integration(ctx, O):
    phys = mi.physics_default(1)
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 8, 1.5)
    vegas = ctx.integrate(mi.make_run(phys, 1 << 22, key=77, sampler=mi.MI_SAMPLER_VEGAS))
    ctx.flow_init(11)
    flow0 = ctx.integrate(mi.make_run(phys, 1 << 22, key=77, sampler=mi.MI_SAMPLER_FLOW))
    # the identity flow reproduces VEGAS on the same events
    assert flow0.sigmaGen() == pytest.approx(vegas.sigmaGen(), rel=1e-5)
    info = ctx.flow_train(mi.make_run(phys, 1 << 20, key=43), 60, 1 << 16, 2e-3)
    assert info.loss_last < info.loss_first, (info.loss_first, info.loss_last)
    est = ctx.integrate(mi.make_run(phys, 1 << 22, key=78, sampler=mi.MI_SAMPLER_FLOW))
    q, _ = O.had_quadrature(O.hadcfg_default(), 12, 12, 24)
    assert abs(est.sigmaGen() - q) < 3.5 * est.sigmaErr(), (est.sigmaGen(), est.sigmaErr(), q)
    print(f"flow: loss {info.loss_first:.4f} -> {info.loss_last:.4f}, err VEGAS {vegas.sigmaErr():.3e} "
          f"flow {est.sigmaErr():.3e}, train {info.kernel_ms:.1f} ms")
