"""PDF grids on the device (mi_pdf_load / mi_pdf_xf and the integrand with pdf_source =
MI_PDF_GRID) against the oracle's LHAGrid1 restatement, on grid files written by the test
(the shapes of the toy PDFs with a Q dependence on an NNPDF-like node set, and a grid with two Q
subgrids, the second of 3 nodes)."""
import math
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pdfgrid_util as PG  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grid_files(tmp_path_factory, O):
    d = tmp_path_factory.mktemp("grids")
    main = str(d / "toyq.dat")
    PG.write_lhagrid1(main, PG.nnpdf_like_x_grid(50, 50, 1e-9), [PG.q_grid(1.0, 1.0e4, 50)], PG.toy_fn(O, Qdep=0.05))
    sub = str(d / "sub.dat")
    PG.write_lhagrid1(sub, PG.nnpdf_like_x_grid(30, 20, 1e-7), [PG.q_grid(1.0, 100.0, 12), np.array([100.0, 1e3, 1e4])],
                      PG.toy_fn(O, Qdep=0.08))
    return main, sub


def _xq_points(n, seed):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(math.log(1e-10), 0.0, n))
    q = np.exp(rng.uniform(math.log(0.5), math.log(2e4), n))
    x[:8] = [1e-9, 1e-12, 0.1, 0.5, 0.999999, 1.0, 3e-9, 1e-3]      # nodes, below xMin, x = 1
    q[8:12] = [1.0, 1e4, 0.3, 5e4]                                   # q borders, outside
    return x, q * q


def test_pdf_xf_vs_oracle(ctx, O, grid_files):
    for path in grid_files:
        ctx.pdf_load(path)
        g = O.LhaGrid(path)
        x, Q2 = _xq_points(4096, 3)
        dev = ctx.pdf_xf(torch.tensor(x), torch.tensor(Q2)).cpu().numpy()
        ref = np.array([g.xf(float(a), float(b)) for a, b in zip(x, Q2)])
        ref[:, 11] = 0.0
        scale = np.maximum(np.abs(ref), 1e-300)
        err = np.abs(dev - ref)
        assert np.all(err <= 1e-12 * scale + 1e-300), (np.max(err / scale), path)


def test_grid_points_fp64_vs_long_double(ctx, O, grid_files):
    main, _ = grid_files
    ctx.pdf_load(main)
    g = O.LhaGrid(main)
    for ff in (1.0, 0.6):
        phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=ff, proc_mask=0x3F)
        cfg = O.hadcfg_default(grid=g, factFac=ff, procMask=0x3F)
        n = 1 << 16
        w, kin = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=42, counter0=99))
        w = w.cpu().numpy()
        wld, _ = O.had_points_ld(cfg, 42, 99, n)
        m = wld > 1e-250
        rel = np.abs(w[m] - wld[m]) / wld[m]
        assert rel.max() < 1e-11, rel.max()
        assert np.all(w[~m] < 1e-240)


def test_grid_points_fp32_bit_exact(ctx, O, grid_files):
    for path in grid_files:
        ctx.pdf_load(path)
        g = O.LhaGrid(path)
        phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=1.3)
        cfg = O.hadcfg_default(grid=g, factFac=1.3)
        n = 1 << 17
        w, kin = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP32, key=42, counter0=5))
        w32, k32 = O.had_points32(cfg, 42, 5, n)
        assert np.array_equal(w.cpu().numpy().astype(np.float32), w32), path
        assert np.array_equal(kin.cpu().numpy().astype(np.float32), k32)


def test_grid_integral_vs_quadrature(ctx, O, grid_files):
    main, _ = grid_files
    ctx.pdf_load(main)
    g = O.LhaGrid(main)
    q, qp = O.had_quadrature(O.hadcfg_default(grid=g), 12, 12, 24)
    phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID)
    e = ctx.integrate(mi.make_run(phys, 1 << 24, key=42))
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    for i, code in enumerate((111, 112, 113, 114, 115)):
        assert abs(e.sigmaGen(code) - qp[i]) < 3.0 * e.sigmaErr(code) + 1e-12, code
    # the toy run differs (the grid carries a Q dependence)
    et = ctx.integrate(mi.make_run(mi.physics_default(1), 1 << 24, key=42))
    assert abs(et.sigmaGen() - e.sigmaGen()) > 5 * e.sigmaErr()


def test_grid_flow_sampler(ctx, O, grid_files):
    """The flow sampler with grid PDFs: a short training, then 2^22 points through the flow,
    sigma within 3 sigma_stat of the grid quadrature; the same events with the hardware
    transcendentals (MI_PREC_FP32_HW) within 1e-5 of the FP32 estimate."""
    main, _ = grid_files
    ctx.pdf_load(main)
    g = O.LhaGrid(main)
    q, _ = O.had_quadrature(O.hadcfg_default(grid=g), 12, 12, 24)
    phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID)
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 6, 1.5)
    ctx.flow_init(9)
    info = ctx.flow_train(mi.make_run(phys, 1 << 20, key=43), 20, 1 << 16, 2e-3)
    assert math.isfinite(info.loss_last) and info.loss_last < info.loss_first
    e = ctx.integrate(mi.make_run(phys, 1 << 22, key=79, sampler=mi.MI_SAMPLER_FLOW))
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    eh = ctx.integrate(mi.make_run(phys, 1 << 22, key=79, sampler=mi.MI_SAMPLER_FLOW, precision=mi.MI_PREC_FP32_HW))
    assert eh.sigmaGen() == pytest.approx(e.sigmaGen(), rel=1e-5)


@pytest.mark.parametrize("which,n,sampler,mask", [
    (0, 1 << 19, mi.MI_SAMPLER_FLAT, 0x1F),
    (1, (1 << 19) + 4096, mi.MI_SAMPLER_FLAT, 0x3F),        # an odd count of 256-wide eval blocks
    (0, 1 << 19, mi.MI_SAMPLER_VEGAS, 0x1F),
    (1, 1 << 18, mi.MI_SAMPLER_VEGAS, 0x3FF),
])
def test_grid_unweight_bit_exact(ctx, O, grid_files, which, n, sampler, mask):
    """Compacted event indices and codes of FP32 grid-PDF unweighting (the evaluation reads the
    grid's reachable window from LDS) equal the oracle's FP32 mirror bit for bit."""
    path = grid_files[which]
    ctx.pdf_load(path)
    g = O.LhaGrid(path)
    phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, proc_mask=mask)
    cfg = O.hadcfg_default(grid=g, procMask=mask)
    edges = None
    if sampler == mi.MI_SAMPLER_VEGAS:
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(mi.make_run(phys, 1 << 18, key=7), 4, 1.5)
        edges = ctx.vegas_get_grid()
    ev = ctx.unweight_compact(mi.make_run(phys, n, key=42, counter0=1000, sampler=sampler))
    cnt, flags, pid, w = O.unweight32(cfg, 42, 1000, n, ev.wmax, edges)
    assert ev.wmax == float(w.max()) and ev.n == cnt and cnt > 0
    idx = np.nonzero(flags)[0]
    assert np.array_equal(ev.index.cpu().numpy(), idx + 1000)
    assert np.array_equal(ev.pid.cpu().numpy(), pid[idx])
    est = ev.estimate.raw
    assert est.wMax[0] == float(w.max()) and est.nNonZero == int(np.count_nonzero(w > 0))


@pytest.mark.parametrize("sampler", [mi.MI_SAMPLER_FLAT, mi.MI_SAMPLER_VEGAS])
@pytest.mark.parametrize("which,ptmin", [(0, 20.0), (1, 20.0), (0, 60.0)])
def test_grid_integration_window_vs_points(ctx, grid_files, sampler, which, ptmin):
    """FP32 integration with grid PDFs stages the reachable rows and columns of the grid in LDS;
    its per-code maxima and non-zero count equal those of the same points' weights from the
    gather path (mi_points) exactly, and its sums agree to the FP64 summation order (1e-13), which
    a single mis-read node among 2^20 points would exceed."""
    ctx.pdf_load(grid_files[which])
    phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=1.3, pTHatMin=ptmin)
    if sampler == mi.MI_SAMPLER_VEGAS:
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(mi.make_run(phys, 1 << 18, key=3), 3, 1.5)
    run = mi.make_run(phys, 1 << 20, key=17, counter0=5, sampler=sampler)
    est = ctx.integrate(run)
    w, _ = ctx.points(run, with_kin=False)
    w = w.cpu().numpy()
    assert est.raw.nNonZero == int(np.count_nonzero(w[:, 0] > 0))
    for c in range(6):
        assert est.raw.wMax[c] == w[:, c].max(), c
        s, s2 = w[:, c].sum(), (w[:, c] ** 2).sum()
        assert abs(est.raw.sumW[c] - s) <= 1e-13 * s, (c, est.raw.sumW[c], s)
        assert abs(est.raw.sumW2[c] - s2) <= 1e-13 * s2, c


def test_grid_errors(ctx, O, grid_files):
    with mi.Context(0) as fresh:
        with pytest.raises(mi.MIError) as e:
            fresh.integrate(mi.make_run(mi.physics_default(1, pdf_source=mi.MI_PDF_GRID), 1024))
        assert e.value.status == mi._abi.MI_ERR_STATE
        with pytest.raises(mi.MIError):
            fresh.pdf_load("/nonexistent/file.dat")
    ctx.pdf_load(grid_files[0])
    with pytest.raises(mi.MIError):
        ctx.adjoint(mi.make_run(mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=0.0), 1 << 16))
    with pytest.raises(mi.MIError):
        ctx.integrate(mi.make_run(mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=0.0), 1024))
