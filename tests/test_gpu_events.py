"""Device event record (mi_build_events) against the oracle: integer fields (flavours, status,
colour tags, mothers, daughters) bit for bit and four-momenta to 1e-12 of the event energy, on
compacted events of the 10-process library (toy PDFs, a PDF grid) and of the partonic gg -> gg
run; invariants at 2^26 candidates; the LHE writer's agreement with the record; the record
stride; errors."""
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

INT_FIELDS = ("pdg", "status", "col", "acol", "mo1", "mo2", "d1", "d2")


def sample(ctx, run):
    """Compacted events of run with wMax = 20 <w> (the exact maximum accepts few at pT > 5 GeV)."""
    est = ctx.integrate(run)
    return ctx.unweight_compact(run, wmax=20 * est.sigmaGen())


def compare(O, cfg, rec, ev, key, partonic=False, sHat=0.0):
    n = ev.n
    ref = O.build_events(cfg, key, ev.index.cpu().numpy().astype(np.uint64), ev.pid.cpu().numpy(),
                         ev.sHat.cpu().numpy(), ev.tHat.cpu().numpy(),
                         None if partonic else ev.x1.cpu().numpy(), None if partonic else ev.x2.cpu().numpy(),
                         partonic=partonic, sHat_partonic=sHat)
    assert np.array_equal(rec.n_part.cpu().numpy(), ref["n_part"])
    for f in INT_FIELDS:
        got = getattr(rec, f).cpu().numpy()[:, :4]
        bad = np.nonzero(np.any(got != ref[f], axis=1))[0]
        assert bad.size == 0, (f, bad[:5], got[bad[:2]], ref[f][bad[:2]])
    p = rec.p4().cpu().numpy()[:, :4]
    etot = ref["p"][:, 0, 3] + ref["p"][:, 1, 3]
    err = np.max(np.abs(p - ref["p"]), axis=(1, 2)) / etot
    assert err.max() < 1e-12, err.max()
    assert np.array_equal(rec.m.cpu().numpy()[:, :4], ref["m"])
    assert np.allclose(rec.scale.cpu().numpy(), ref["scale"], rtol=1e-14, atol=0)
    return n


@pytest.mark.parametrize("mask,prec", [(0x3FF, mi.MI_PREC_FP32), (0x3FF, mi.MI_PREC_FP64), (0x3F, mi.MI_PREC_FP32)])
def test_record_vs_oracle_toy(ctx, O, mask, prec):
    phys = mi.physics_default(1, proc_mask=mask, pTHatMin=5.0)
    ev = sample(ctx, mi.make_run(phys, 1 << 20, key=33, precision=prec))
    assert ev.n > 1000

    rec = ctx.build_events(phys, ev, key=77)
    compare(O, O.hadcfg_default(procMask=mask, pTHatMin=5.0), rec, ev, 77)
    if mask == 0x3FF:                      # the heavy codes are rare here: a heavy-only sample too
        evh = sample(ctx, mi.make_run(mi.physics_default(1, proc_mask=0x3C0, pTHatMin=5.0), 1 << 20, key=34))
        codes = set(evh.pid.cpu().tolist())
        assert {121, 123} <= codes and len(codes) >= 3, codes
        compare(O, O.hadcfg_default(procMask=0x3C0, pTHatMin=5.0),
                ctx.build_events(mi.physics_default(1, proc_mask=0x3C0, pTHatMin=5.0), evh, key=78), evh, 78)


def test_record_vs_oracle_grid(ctx, O, tmp_path):
    import pdfgrid_util as PG
    path = str(tmp_path / "toyq.dat")
    PG.write_lhagrid1(path, PG.nnpdf_like_x_grid(50, 50, 1e-9), [PG.q_grid(1.0, 1.0e4, 50)], PG.toy_fn(O, Qdep=0.05))
    ctx.pdf_load(path)
    phys = mi.physics_default(1, proc_mask=0x3FF, pTHatMin=5.0, pdf_source=mi.MI_PDF_GRID, fact_fac=1.3)
    ev = sample(ctx, mi.make_run(phys, 1 << 20, key=35))
    rec = ctx.build_events(phys, ev, key=5)
    cfg = O.hadcfg_default(procMask=0x3FF, pTHatMin=5.0, grid=O.LhaGrid(path), factFac=1.3)
    compare(O, cfg, rec, ev, 5)


def test_record_vs_oracle_partonic(ctx, O):
    phys = mi.physics_default(0)
    ev = ctx.unweight_compact(mi.make_run(phys, 1 << 18, key=42, precision=mi.MI_PREC_FP64))
    rec = ctx.build_events(phys, ev, key=42)
    compare(O, O.hadcfg_default(), rec, ev, 42, partonic=True, sHat=phys.sHat)
    e = rec.e.cpu().numpy()
    assert np.all(e == 0.5 * np.sqrt(phys.sHat))                  # 50 GeV per parton at sqrt(s) = 100 GeV
    # the reference's 2 -> 2 four-momenta at sqrt(s) = 100 GeV (cuPythia/12_phase_space/README.md:8-16):
    # max |p_cons| 3.6e-15, max |p^2 - m^2| 1.7e-12 GeV^2; the record does at least as well
    p = rec.p4().cpu().numpy()
    cons = np.abs(p[:, 2] + p[:, 3] - p[:, 0] - p[:, 1]).max()
    msq = np.abs(p[:, :, 3] ** 2 - (p[:, :, :3] ** 2).sum(axis=2)).max()
    assert cons <= 3.6e-15 and msq <= 1.7e-12, (cons, msq)


def test_invariants_at_2p26(ctx):
    """The unweighting sample of 2^26 candidates through the adapted 3-D VEGAS grid: conservation to
    FP64 rounding (the stable boost), on-shell masses, (p1 + p2)^2 = x1 x2 S, (p1 - p3)^2 = the
    record's tHat at x1 x2 S, colour tags 501..504."""
    phys = mi.physics_default(1)
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(mi.make_run(phys, 1 << 24, key=42), 12, 1.5)
    ev = ctx.unweight_compact(mi.make_run(phys, 1 << 26, key=42, sampler=mi.MI_SAMPLER_VEGAS))
    assert ev.n > 500000
    rec = ctx.build_events(phys, ev, key=42)
    p = rec.p4()
    etot = p[:, 0, 3] + p[:, 1, 3]
    bal = (p[:, 2] + p[:, 3] - p[:, 0] - p[:, 1]).abs().amax(dim=1) / etot
    assert float(bal.max()) < 4e-15, float(bal.max())

    def mdot(a, b):
        return a[..., 3] * b[..., 3] - a[..., 0] * b[..., 0] - a[..., 1] * b[..., 1] - a[..., 2] * b[..., 2]

    assert float((mdot(p, p).abs() / etot[:, None] ** 2).max()) < 1e-14   # massless library
    s_lab = ev.x1 * ev.x2 * phys.eCM ** 2
    assert float(((2 * mdot(p[:, 0], p[:, 1]) - s_lab).abs() / s_lab).max()) < 1e-13      # massless: 2 p1.p2
    # the record's FP32 kinematics: sHat = x1 x2 S to FP32 precision; the final state is built at
    # x1 x2 S with the record's angle, so (p1 - p3)^2 = tHat x1 x2 S / sHat, checked against the
    # rounding scale of the lab-frame product, (E1 + E2)^2 (a soft parton 1 against a hard,
    # boosted parton 3 makes the squares of q = p1 - p3 that large)
    assert float(((s_lab - ev.sHat).abs() / ev.sHat).max()) < 1e-5
    q = p[:, 0] - p[:, 2]
    assert float(((mdot(q, q) - ev.tHat * s_lab / ev.sHat).abs() / etot ** 2).max()) < 1e-14
    assert int(rec.n_part.min()) == 4 and int(rec.n_part.max()) == 4
    col = torch.cat([rec.col[:, :4], rec.acol[:, :4]], 1)
    assert int(col[col > 0].min()) == 501 and int(col.max()) <= 504


def test_lhe_matches_record(ctx, tmp_path):
    phys = mi.physics_default(1, proc_mask=0x3FF, pTHatMin=5.0)
    ev = sample(ctx, mi.make_run(phys, 1 << 18, key=9))
    rec = ctx.build_events(phys, ev, key=12)
    path = str(tmp_path / "ev.lhe")
    ctx.write_lhe(path, phys, ev, key=12)
    lines = open(path).read().split("\n")
    parts, k = [], 0
    while k < len(lines):
        if lines[k] == "<event>":
            parts.append([lines[k + 2 + j].split() for j in range(4)])
            k += 6
        else:
            k += 1
    assert len(parts) == ev.n
    pdg, col, acol = rec.pdg.cpu().numpy(), rec.col.cpu().numpy(), rec.acol.cpu().numpy()
    p = rec.p4().cpu().numpy()
    for i in range(0, ev.n, max(1, ev.n // 2000)):
        for j in range(4):
            f = parts[i][j]
            assert int(f[0]) == pdg[i, j] and int(f[4]) == col[i, j] and int(f[5]) == acol[i, j]
            if j >= 2:
                v = np.array([float(x) for x in f[6:10]])
                assert np.allclose(v, p[i, j], rtol=1e-9, atol=1e-9 * p[i, j, 3])


def test_stride_and_errors(ctx):
    phys = mi.physics_default(1)
    ev = ctx.unweight_compact(mi.make_run(phys, 1 << 16, key=3))
    r4 = ctx.build_events(phys, ev, key=1)
    r8 = ctx.build_events(phys, ev, key=1, max_part=8)
    assert torch.equal(r8.px[:, :4], r4.px) and torch.equal(r8.col[:, :4], r4.col)
    assert int(r8.pdg[:, 4:].abs().max()) == 0 and float(r8.e[:, 4:].abs().max()) == 0.0
    # a stride that is not a multiple of 4: the per-element stores, the same record
    r5 = ctx.build_events(phys, ev, key=1, max_part=5)
    for f in ("px", "py", "pz", "e", "m", "pdg", "status", "col", "acol", "mo1", "mo2", "d1", "d2"):
        assert torch.equal(getattr(r5, f)[:, :4], getattr(r4, f)), f
    assert int(r5.pdg[:, 4].abs().max()) == 0 and torch.equal(r5.scale, r4.scale)
    with pytest.raises(mi.MIError):
        ctx.build_events(phys, ev, key=1, max_part=3)
    # an event with a code outside the library gets no partons; the others are unchanged
    bad = mi.Events(ev.sHat, ev.tHat, ev.x1, ev.x2, ev.weight, ev.pid.clone(), ev.index, ev.n, ev.wmax)
    bad.pid[3] = 999
    rb = ctx.build_events(phys, bad, key=1)
    nb = rb.n_part.cpu().numpy()
    assert nb[3] == 0 and np.all(np.delete(nb, 3) == 4)
    keep = torch.ones(ev.n, dtype=torch.bool, device=rb.px.device)
    keep[3] = False
    assert torch.equal(rb.px[keep], r4.px[keep]) and torch.equal(rb.pdg[keep], r4.pdg[keep])
    assert int(rb.pdg[3].abs().max()) == 0 and float(rb.e[3].abs().max()) == 0.0
    with mi.Context(0) as fresh:
        with pytest.raises(mi.MIError) as e:
            fresh.build_events(mi.physics_default(1, pdf_source=mi.MI_PDF_GRID), ev, key=1)
        assert e.value.status == mi._abi.MI_ERR_STATE
