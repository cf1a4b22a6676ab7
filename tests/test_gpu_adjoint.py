"""GPU adjoint (mi_adjoint) against the oracle, finite differences with common random numbers,
the alphaS^2 scaling, and the torch.autograd harness."""
import math
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi
from cupythia_amd import adjoint

pytestmark = pytest.mark.gpu


def oracle_adj(O, cfg, n, key=42, ctr0=0):
    import ctypes as C
    out = (C.c_double * 9)()
    O.lib().orc_adjoint(C.byref(cfg), key, ctr0, n, out)
    return np.array(list(out))


@pytest.mark.parametrize("prec,tol", [(mi.MI_PREC_FP64, 1e-7), (mi.MI_PREC_FP32, 2e-5)])
def test_adjoint_vs_oracle(ctx, O, prec, tol):
    n = 1 << 16
    g = ctx.adjoint(mi.make_run(mi.physics_default(1), n, precision=prec, key=42))
    r = oracle_adj(O, O.hadcfg_default(), n)
    got = np.array([g.sigma, g.d_alphaS_MZ, g.d_renorm_fac, g.d_pdf_g_a, g.d_pdf_g_b]) * n
    rel = np.abs(got - r[:5]) / np.abs(r[:5])
    assert rel.max() < tol, rel


@pytest.mark.parametrize("mask", [0x3FF, 0x3C0, 0x100, 0x200])
@pytest.mark.parametrize("prec,tol", [(mi.MI_PREC_FP64, 1e-7), (mi.MI_PREC_FP32, 3e-5)])
def test_adjoint_heavy_vs_oracle(ctx, O, mask, prec, tol):
    """The heavy-quark families (codes 121..124, pT > 5 GeV): every family's coupling at its own
    scale pT^2 + m^2, the gluon shape through the g g -> Q Qbar luminosity."""
    n = 1 << 16
    g = ctx.adjoint(mi.make_run(mi.physics_default(1, proc_mask=mask, pTHatMin=5.0), n, precision=prec, key=42))
    r = oracle_adj(O, O.hadcfg_default(procMask=mask, pTHatMin=5.0), n)
    got = np.array([g.sigma, g.d_alphaS_MZ, g.d_renorm_fac, g.d_pdf_g_a, g.d_pdf_g_b]) * n
    k = 3 if mask == 0x200 else 5
    rel = np.abs(got[:k] - r[:k]) / np.abs(r[:k])
    assert rel.max() < tol, rel
    if mask == 0x200:                      # q qbar -> b bbar alone: no gluon in the initial state
        assert r[3] == 0.0 and r[4] == 0.0
        assert g.d_pdf_g_a == 0.0 and g.d_pdf_g_b == 0.0


@pytest.fixture(scope="module")
def grid_file(tmp_path_factory, O):
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import pdfgrid_util as PG
    path = str(tmp_path_factory.mktemp("adjgrid") / "toyq.dat")
    PG.write_lhagrid1(path, PG.nnpdf_like_x_grid(50, 50, 1e-9), [PG.q_grid(1.0, 1.0e4, 50)], PG.toy_fn(O, Qdep=0.05))
    return path


@pytest.mark.parametrize("mask", [0x3F, 0x3FF])
@pytest.mark.parametrize("prec,tol", [(mi.MI_PREC_FP64, 1e-7), (mi.MI_PREC_FP32, 3e-5)])
def test_adjoint_grid_vs_oracle(ctx, O, grid_file, mask, prec, tol):
    """Parton densities from a PDF grid: the coupling derivatives as with the toy PDFs, the toy
    gluon's shape parameters then enter nothing (derivatives exactly 0)."""
    ctx.pdf_load(grid_file)
    n = 1 << 16
    phys = mi.physics_default(1, proc_mask=mask, pTHatMin=5.0, pdf_source=mi.MI_PDF_GRID, fact_fac=1.3)
    g = ctx.adjoint(mi.make_run(phys, n, precision=prec, key=42))
    r = oracle_adj(O, O.hadcfg_default(procMask=mask, pTHatMin=5.0, grid=O.LhaGrid(grid_file), factFac=1.3), n)
    got = np.array([g.sigma, g.d_alphaS_MZ, g.d_renorm_fac]) * n
    rel = np.abs(got - r[:3]) / np.abs(r[:3])
    assert rel.max() < tol, rel
    assert r[3] == 0.0 and r[4] == 0.0
    assert g.d_pdf_g_a == 0.0 and g.d_pdf_g_b == 0.0 and g.d_pdf_g_a_err == 0.0


def test_alphas_squared_scaling(ctx):
    phys = mi.physics_default(1, alphaS_order=0, alphaS_value=0.13)
    g = ctx.adjoint(mi.make_run(phys, 1 << 22, key=3))
    assert abs(g.d_alphaS_MZ * 0.13 / g.sigma - 2.0) < 1e-6
    assert g.d_renorm_fac == 0.0


def _sigma64(ctx, n, key, **over):
    e = ctx.integrate(mi.make_run(mi.physics_default(1, **over), n, precision=mi.MI_PREC_FP64, key=key))
    return e.sigmaGen()


@pytest.mark.parametrize("mask", [0x1, 0x1F])
def test_c4_workload_vs_finite_differences(ctx, mask):
    """2^26 points, FP32 forward + FP64 accumulation: the adjoint equals central differences
    (h = 1e-4, same events, FP64 forward) to 1e-4, far inside the statistical error."""
    n = 1 << 26
    base = dict(proc_mask=mask)
    g = ctx.adjoint(mi.make_run(mi.physics_default(1, **base), n, key=42))
    h = 1e-4
    fd = {
        "as": (_sigma64(ctx, n, 42, alphaS_value=0.118 + h, **base) - _sigma64(ctx, n, 42, alphaS_value=0.118 - h, **base)) / (2 * h),
        "xi": (_sigma64(ctx, n, 42, renorm_fac=1 + h, **base) - _sigma64(ctx, n, 42, renorm_fac=1 - h, **base)) / (2 * h),
        "a": (_sigma64(ctx, n, 42, pdf_g=(-0.1 + h, 5.0), **base) - _sigma64(ctx, n, 42, pdf_g=(-0.1 - h, 5.0), **base)) / (2 * h),
        "b": (_sigma64(ctx, n, 42, pdf_g=(-0.1, 5.0 + h), **base) - _sigma64(ctx, n, 42, pdf_g=(-0.1, 5.0 - h), **base)) / (2 * h),
    }
    pairs = [(g.d_alphaS_MZ, fd["as"], g.d_alphaS_MZ_err), (g.d_renorm_fac, fd["xi"], g.d_renorm_fac_err),
             (g.d_pdf_g_a, fd["a"], g.d_pdf_g_a_err), (g.d_pdf_g_b, fd["b"], g.d_pdf_g_b_err)]
    for adj, f, err in pairs:
        assert abs(adj - f) < 1e-4 * abs(f), (adj, f, err)
        assert err < 0.05 * abs(adj)
    print(f"C4 mask {mask:#x}: sigma {g.sigma:.6e}, d/dalphaS {g.d_alphaS_MZ:.6e}, d/dxi {g.d_renorm_fac:.6e}, "
          f"d/da {g.d_pdf_g_a:.6e}, d/db {g.d_pdf_g_b:.6e}, {g.kernel_ms:.2f} ms")


def test_autograd_harness(ctx):
    run = mi.make_run(mi.physics_default(1), 1 << 22, key=9)
    p = torch.tensor([0.118, 1.0, -0.1, 5.0], dtype=torch.float64, requires_grad=True)
    s = adjoint.sigma(p, ctx, run)
    (gr,) = torch.autograd.grad(s, p)
    g = ctx.adjoint(run)
    assert float(s) == g.sigma
    assert torch.equal(gr, torch.tensor([g.d_alphaS_MZ, g.d_renorm_fac, g.d_pdf_g_a, g.d_pdf_g_b], dtype=torch.float64))
    # central differences through the same Function's forward (common random numbers)
    h = 1e-4
    for q in range(4):
        pp, pm = p.detach().clone(), p.detach().clone()
        pp[q] += h
        pm[q] -= h
        fd = (float(adjoint.sigma(pp, ctx, run)) - float(adjoint.sigma(pm, ctx, run))) / (2 * h)
        assert abs(fd - float(gr[q])) < 2e-3 * abs(float(gr[q])), (q, fd, float(gr[q]))
