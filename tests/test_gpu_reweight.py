"""Scale and PDF variations of events.

mi_reweight_events (events regenerated from their counters, the reference's pinned-seed
reweighting): every variation bit for bit equal to a re-run at those scales / PDFs and within
1e-12 of the long-double oracle, for every code, toy and grid PDFs, PDF members from a second grid
in slot 1 (as the members of an LHAPDF set would be), and unweighted events of VEGAS runs.  The
per-point FP64 weights with running alphaS at renorm_fac 0.5 and 2 and grid PDFs at fact_fac != 1
against the oracle at 1e-12.

mi_reweight / mi_reweight_scales (events given by their stored sHat, tHat, x1, x2): against
re-runs and the oracle at 1e-10, the precision the stored (sHat, tHat) carry next to the pT cut."""
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


# ------------------------------------------------------------------------------ against the oracle
# Running alphaS at mu_R = renorm_fac * pT (renorm_fac 0.5 and 2) and the grid PDFs at mu_F =
# fact_fac * pT (fact_fac 0.7, 1.6): the FP64 per-point weights of every code against the
# long-double oracle (Pythia's forms: StandardModel.cc:188-215 alphaS, SigmaProcess.cc:846-925
# scales, SigmaQCD.cc sigmaHat, PartonDistributions.cc:648-745 grid).  Bound 1e-12 for the toy and
# the grid PDFs alike (measured maximum 5e-14).
SCALE_SETS = [(0.5, 0.7), (2.0, 1.6)]
GRID_TOL = 1e-12


def _phys_cfg(O, bit, r, f, grid=None, **kw):
    base = dict(proc_mask=1 << bit, pTHatMin=10.0, renorm_fac=r, fact_fac=f, **kw)
    if grid is not None:
        base["pdf_source"] = mi.MI_PDF_GRID
    phys = mi.physics_default(1, **base)
    cfg = O.hadcfg_default(procMask=1 << bit, pTHatMin=10.0, renormFac=r, factFac=f,
                           **({"grid": grid} if grid is not None else {}))
    return phys, cfg


def _rel(dev, ref, floor=1e-250):
    m = ref > floor
    assert int(m.sum()) > 0
    assert np.all(np.abs(dev[~m]) < 1e-240)
    return float(np.max(np.abs(dev[m] - ref[m]) / ref[m]))


@pytest.mark.parametrize("bit,code", CODES)
@pytest.mark.parametrize("pdf", ["toy", "grid"])
def test_points_running_alphas_vs_oracle(ctx, O, grids, bit, code, pdf):
    g = None
    if pdf == "grid":
        ctx.pdf_load(grids[0])
        g = O.LhaGrid(grids[0])
    for r, f in SCALE_SETS:
        phys, cfg = _phys_cfg(O, bit, r, f, g)
        w, _ = ctx.points(mi.make_run(phys, N, precision=mi.MI_PREC_FP64, key=21 + bit, counter0=3))
        wld, _ = O.had_points_ld(cfg, 21 + bit, 3, N)
        rel = _rel(w.cpu().numpy()[:, 1 + bit], wld[:, 1 + bit])
        print(f"code {code} {pdf} xiR {r} xiF {f}: max rel {rel:.2e}")
        assert rel < (1e-12 if g is None else GRID_TOL), (code, pdf, r, f, rel)


VARS = [(1.0, 0.8, 0), (0.5, 2.0, 0), (2.0, 0.5, 0), (1.0, 0.8, 1), (0.5, 0.5, 1)]


@pytest.mark.parametrize("bit,code", CODES)
def test_reweight_events_equals_reruns_and_oracle(ctx, O, grids, bit, code):
    """mi_reweight_events regenerates each event from its counter (the reference's pinned-seed
    reweighting): each variation's weight is bit for bit the per-point weight of a run at those
    scales / PDFs (cuPythia/pipeline/VALIDATION.md:19: one pass == pinned re-runs, max|diff| = 0),
    and within 1e-12 of the long-double oracle; sparse, unordered counters."""
    central, member = grids
    ctx.pdf_load(central)
    ctx.pdf_load(member, slot=1)
    gc, gm = O.LhaGrid(central), O.LhaGrid(member)
    phys = mi.physics_default(1, proc_mask=1 << bit, pTHatMin=10.0, pdf_source=mi.MI_PDF_GRID, renorm_fac=1.0,
                              fact_fac=0.8)
    key, c0, stride = 31 + bit, 7, 3
    run = mi.make_run(phys, N, precision=mi.MI_PREC_FP64, key=key)
    idx = (torch.arange(N, dtype=torch.int64, device="cuda") * stride + c0).flip(0).contiguous()
    pid = torch.full((N,), code, dtype=torch.int32, device="cuda")
    rw = ctx.reweight_events(run, idx, pid, VARS)
    order = idx.cpu().numpy() - c0
    with mi.Context(0) as c2:
        c2.pdf_load(member)
        for k, (r, f, slot) in enumerate(VARS):
            c = ctx if slot == 0 else c2
            pv = mi.physics_default(1, proc_mask=1 << bit, pTHatMin=10.0, pdf_source=mi.MI_PDF_GRID, renorm_fac=r,
                                    fact_fac=f)
            wf, _ = c.points(mi.make_run(pv, N * stride, precision=mi.MI_PREC_FP64, key=key, counter0=c0),
                             with_kin=False)
            rerun = wf[order, 1 + bit]
            assert torch.equal(rw[:, k], rerun), (code, k)
            cfg = O.hadcfg_default(procMask=1 << bit, pTHatMin=10.0, renormFac=r, factFac=f,
                                   grid=gc if slot == 0 else gm)
            wld, _ = O.had_points_ld(cfg, key, c0, N * stride)
            rel = _rel(rw[:, k].cpu().numpy(), wld[order, 1 + bit])
            assert rel < GRID_TOL, (code, k, rel)
    # toy PDFs: the same at 1e-12, and the ratio form w * W_k / W_0 on the run's own weights
    physt = mi.physics_default(1, proc_mask=1 << bit, pTHatMin=10.0, renorm_fac=1.3)
    runt = mi.make_run(physt, N, precision=mi.MI_PREC_FP64, key=key)
    tv = [(1.3, 1.0, 0), (0.5, 1.0, 0), (2.0, 1.0, 0)]
    w0 = ctx.reweight_events(runt, idx, pid, tv[:1])[:, 0].contiguous()
    rt = ctx.reweight_events(runt, idx, pid, tv, w=w0)
    assert torch.equal(rt[:, 0], w0)
    for k, (r, f, _) in enumerate(tv):
        cfg = O.hadcfg_default(procMask=1 << bit, pTHatMin=10.0, renormFac=r)
        wld, _ = O.had_points_ld(cfg, key, c0, N * stride)
        rel = _rel(rt[:, k].cpu().numpy(), wld[order, 1 + bit])
        assert rel < 2e-12, (code, k, rel)


@pytest.mark.parametrize("prec", [mi.MI_PREC_FP32, mi.MI_PREC_FP64])
def test_reweight_events_of_unweighted_vegas_events(ctx, O, prec):
    """Unweighted events of a VEGAS run, all ten codes: w * W_k / W_0 regenerated through the
    context's grid against the oracle's ratio at the same points (FP64 at the same uniforms, also for
    an FP32 run).  For an FP64 run also the stored-kinematics paths (mi_reweight, mi_reweight_scales)
    against the same oracle at their bound 1e-10 (uHat = -(sHat + tHat) from the stored values cancels
    next to the pT cut; the FP32 run's stored kinematics are the FP32 point's)."""
    phys = mi.physics_default(1, proc_mask=0x3FF, pTHatMin=10.0)
    ctx.vegas_init(3, 16)
    ctx.vegas_adapt(mi.make_run(phys, 1 << 16, sampler=mi.MI_SAMPLER_VEGAS, key=3), 3)
    edges = ctx.vegas_get_grid()
    n = 1 << 17
    run = mi.make_run(phys, n, sampler=mi.MI_SAMPLER_VEGAS, precision=prec, key=77)
    ev = ctx.unweight_compact(run, wmax=0.0)
    assert ev.n > 1000
    assert len(set(ev.pid.cpu().tolist())) >= 5      # gg and qg dominate; the heavy codes are rare
    var = [(1.0, 1.0, 0), (0.5, 1.0, 0), (2.0, 1.0, 0)]
    rw = ctx.reweight_events(run, ev.index, ev.pid, var, w=ev.weight)
    idx = ev.index.cpu().numpy()
    slot = np.array([O.CODES.index(c) for c in ev.pid.cpu().tolist()])
    w = ev.weight.cpu().numpy()
    W = []
    for r, _, _ in var:
        cfg = O.hadcfg_default(procMask=0x3FF, pTHatMin=10.0, renormFac=r)
        wld, _ = O.had_points_ld(cfg, 77, 0, n, edges=edges)
        W.append(wld[idx, 1 + slot])
    for k in range(len(var)):
        assert _rel(rw[:, k].cpu().numpy(), w * (W[k] / W[0])) < 2e-12, k
    if prec == mi.MI_PREC_FP64:
        old = ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, var).cpu().numpy()
        sc = ctx.reweight_scales(phys, ev.sHat, ev.tHat, ev.weight, [r for r, _, _ in var], pid=ev.pid).cpu().numpy()
        for k in range(len(var)):
            ref = w * (W[k] / W[0])
            assert _rel(old[:, k], ref) < 1e-10 and _rel(sc[:, k], ref) < 1e-10, k
