"""Scale and PDF variations of events (mi_reweight) against re-running the same events at the
varied scales or with the varied PDFs (FP64 points, 1e-10): every process code with the toy PDFs
and with grid PDFs, the heavy pairs (scale pT^2 + m^2), and PDF variations from a second grid in
slot 1 (as the members of an LHAPDF set would be); for mu_R alone it equals mi_reweight_scales."""
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pdfgrid_util as PG  # noqa: E402

pytestmark = pytest.mark.gpu

N = 1 << 12


@pytest.fixture(scope="module")
def grids(tmp_path_factory, O):
    d = tmp_path_factory.mktemp("rw")
    central = str(d / "central.dat")
    member = str(d / "member.dat")
    xs, qs = PG.nnpdf_like_x_grid(40, 40, 1e-8), PG.q_grid(1.0, 1.0e4, 30)
    PG.write_lhagrid1(central, xs, [qs], PG.toy_fn(O, Qdep=0.05))
    PG.write_lhagrid1(member, xs, [qs], PG.toy_fn(O, Qdep=0.12))
    return central, member


def _events(c, phys, key):
    w, kin = c.points(mi.make_run(phys, N, precision=mi.MI_PREC_FP64, key=key))
    cols = [kin[:, j].contiguous() for j in range(4)]
    return w, cols


def _check(rw, ref, w0):
    m = (ref > 1e-250) & (w0 > 1e-250)
    assert int(m.sum()) > N // 2
    rel = ((rw[m] - ref[m]).abs() / ref[m]).max().item()
    assert rel < 1e-10, rel


CODES = [(0, 111), (1, 112), (2, 113), (3, 114), (4, 115), (5, 116), (6, 121), (7, 122), (8, 123), (9, 124)]


@pytest.mark.parametrize("bit,code", CODES)
def test_reweight_toy_pdfs_matches_reruns(ctx, bit, code):
    base = dict(proc_mask=1 << bit, pTHatMin=10.0, renorm_fac=1.3)
    phys = mi.physics_default(1, **base)
    w, (sH, tH, x1, x2) = _events(ctx, phys, 5 + bit)
    pid = torch.full((N,), code, dtype=torch.int32, device="cuda")
    w0 = w[:, 0].contiguous()
    variations = [(1.3, 1.0, 0), (0.5, 1.0, 0), (2.0, 3.0, 0)]
    rw = ctx.reweight(phys, sH, tH, x1, x2, pid, w0, variations)
    for k, (r, _, _) in enumerate(variations):
        wf, _ = ctx.points(mi.make_run(mi.physics_default(1, **{**base, "renorm_fac": r}), N,
                                       precision=mi.MI_PREC_FP64, key=5 + bit))
        _check(rw[:, k], wf[:, 0], w0)
    # mu_R alone: the same as mi_reweight_scales
    rs = ctx.reweight_scales(phys, sH, tH, w0, [r for (r, _, _) in variations], pid=pid)
    m = w0 > 0
    assert ((rs[m] - rw[m]).abs() <= 1e-12 * rs[m].abs()).all()


@pytest.mark.parametrize("bit,code", CODES)
def test_reweight_grid_scales_and_members_match_reruns(ctx, grids, bit, code):
    central, member = grids
    ctx.pdf_load(central)
    ctx.pdf_load(member, slot=1)
    base = dict(proc_mask=1 << bit, pTHatMin=10.0, pdf_source=mi.MI_PDF_GRID, renorm_fac=1.0, fact_fac=0.8)
    phys = mi.physics_default(1, **base)
    w, (sH, tH, x1, x2) = _events(ctx, phys, 11 + bit)
    pid = torch.full((N,), code, dtype=torch.int32, device="cuda")
    w0 = w[:, 0].contiguous()
    variations = [(1.0, 0.8, 0), (0.5, 2.0, 0), (2.0, 0.5, 0), (1.0, 0.8, 1), (0.5, 0.5, 1)]
    rw = ctx.reweight(phys, sH, tH, x1, x2, pid, w0, variations)
    with mi.Context(0) as c2:
        c2.pdf_load(member)
        for k, (r, f, slot) in enumerate(variations):
            c = ctx if slot == 0 else c2
            wf, _ = c.points(mi.make_run(mi.physics_default(1, **{**base, "renorm_fac": r, "fact_fac": f}), N,
                                         precision=mi.MI_PREC_FP64, key=11 + bit))
            _check(rw[:, k], wf[:, 0], w0)
    # the variations move the weights (the grids have a Q dependence, the member another one)
    m = w0 > 0
    for k in (1, 2, 3):
        assert ((rw[m, k] - w0[m]).abs() > 1e-6 * w0[m]).any()


def test_reweight_unweighted_events_and_errors(ctx, grids):
    """On compacted events: the variation of the run's own scales returns the event weights;
    errors for bad slots, unloaded slots, and PDF slots with the toy PDFs."""
    central, member = grids
    ctx.pdf_load(central)
    ctx.pdf_load(member, slot=2)
    phys = mi.physics_default(1, proc_mask=0x3FF, pdf_source=mi.MI_PDF_GRID)
    ev = ctx.unweight_compact(mi.make_run(phys, 1 << 18, key=4), wmax=2.0)
    assert ev.n > 100
    rw = ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(1.0, 1.0, 0), (1.0, 1.0, 2)])
    assert torch.allclose(rw[:, 0], ev.weight, rtol=1e-13, atol=0)
    assert not torch.equal(rw[:, 1], ev.weight)
    with pytest.raises(mi.MIError):
        ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(1.0, 1.0, 9)])
    with pytest.raises(mi.MIError) as e:
        ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(1.0, 1.0, 5)])
    assert e.value.status == mi._abi.MI_ERR_STATE
    with pytest.raises(mi.MIError):
        ctx.reweight(mi.physics_default(1), ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(1.0, 1.0, 1)])
    with pytest.raises(mi.MIError):
        ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(0.0, 1.0, 0)])
