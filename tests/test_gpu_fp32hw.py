"""MI_PREC_FP32_HW: the FP32 integrand with the hardware exp2 / log2 / reciprocal / square root
(v_exp_f32, v_log_f32, v_rcp_f32, v_sqrt_f32) instead of this library's own bit-reproducible
forms.  Per-point weights against the long-double Pythia-form oracle within 1e-5 relative wherever
the weight is representable (> 1e-30 mb) -- the bound the bit-exact FP32 path is held to --
for the five-process and the ten-process libraries and the flat and VEGAS maps; with grid PDFs
the bit-exact FP32 path's grid criterion (99.9% within 1e-5, all within 5e-5, of the weights above
1e-8 of the mean);
integration within 3 sigma_stat of the quadrature; on the same events the sums agree with the FP64
path to 2e-6; unweighting of an FP32_HW run uses the bit-exact FP32 math (indices equal the
oracle's)."""
import os
import sys

import numpy as np
import pytest

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pdfgrid_util as PG  # noqa: E402

pytestmark = pytest.mark.gpu
HW = mi.MI_PREC_FP32_HW


def _rel(w, wld, floor=1e-30):
    m = wld > floor
    assert int(m.sum()) > 1000
    r = np.abs(w[m] - wld[m]) / wld[m]
    return float(r.max()), float(np.median(r))


@pytest.mark.parametrize("mask", [0x1F, 0x3FF])
@pytest.mark.parametrize("vegas", [False, True])
def test_points_vs_long_double_oracle(ctx, O, mask, vegas):
    phys = mi.physics_default(1, proc_mask=mask)
    cfg = O.hadcfg_default(procMask=mask)
    n = 1 << 18
    edges = None
    sampler = mi.MI_SAMPLER_FLAT
    if vegas:
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 4, 1.5)
        edges = ctx.vegas_get_grid()
        sampler = mi.MI_SAMPLER_VEGAS
    w, _ = ctx.points(mi.make_run(phys, n, precision=HW, sampler=sampler, key=42, counter0=777), with_kin=False)
    wld, _ = O.had_points_ld(cfg, 42, 777, n, edges=edges)
    mx, med = _rel(w.cpu().numpy(), wld)
    print(f"FP32_HW mask {mask:#x} vegas {vegas}: max rel {mx:.2e}, median {med:.2e}")
    assert mx < 1e-5 and med < 1e-6


def test_points_grid_pdf_vs_oracle(ctx, O, tmp_path):
    """Grid PDFs: (1) on the configuration where the bit-exact FP32 path's grid criterion is pinned
    (tests/test_oracle_pdfgrid.py::test_fp32_grid_mirror_accuracy: 99.9% of the weights above 1e-8
    of the mean within 1e-5, all within 5e-5, median below 3e-6), the same criterion; (2) on a
    finer grid and 2^17 points, no worse than the bit-exact FP32 path on the same points (the
    cubic interpolation of the steep large-x tail cancels in FP32 with either math)."""
    xs, qs = PG.nnpdf_like_x_grid(40, 40, 1e-7), PG.q_grid(1.0, 1.0e4, 30)
    path = str(tmp_path / "toyq.dat")
    PG.write_lhagrid1(path, xs, [qs], PG.toy_fn(O, Qdep=0.05))
    ctx.pdf_load(path)
    g = O.LhaGrid(path)
    phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=0.8)
    cfg = O.hadcfg_default(grid=g, factFac=0.8)
    n = 1 << 14
    w, _ = ctx.points(mi.make_run(phys, n, precision=HW, key=42, counter0=0), with_kin=False)
    w = w.cpu().numpy()
    wld, _ = O.had_points_ld(cfg, 42, 0, n)
    for col in range(1, 7):
        m = wld[:, col] > 1e-30
        if not m.any():
            continue
        rel = np.abs(w[m, col] - wld[m, col]) / wld[m, col]
        big = wld[m, col] > 1e-8 * wld[m, col].mean()
        print(f"FP32_HW grid col {col}: max rel (big) {rel[big].max():.2e}, median {np.median(rel):.2e}")
        assert rel[big].max() < 5e-5, (col, rel[big].max())
        assert np.quantile(rel[big], 0.999) < 1e-5, col
        assert np.median(rel) < 3e-6
    path2 = str(tmp_path / "toyq2.dat")
    PG.write_lhagrid1(path2, PG.nnpdf_like_x_grid(50, 50, 1e-9), [PG.q_grid(1.0, 1.0e4, 50)], PG.toy_fn(O, Qdep=0.05))
    ctx.pdf_load(path2)
    cfg2 = O.hadcfg_default(grid=O.LhaGrid(path2), factFac=1.3)
    phys2 = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID, fact_fac=1.3)
    n = 1 << 17
    wh = ctx.points(mi.make_run(phys2, n, precision=HW, key=42, counter0=5), with_kin=False)[0].cpu().numpy()
    wo = ctx.points(mi.make_run(phys2, n, precision=mi.MI_PREC_FP32, key=42, counter0=5), with_kin=False)[0].cpu().numpy()
    wld, _ = O.had_points_ld(cfg2, 42, 5, n)
    for col in range(1, 7):
        m = wld[:, col] > 1e-30
        if not m.any():
            continue
        big = wld[m, col] > 1e-8 * wld[m, col].mean()
        rh = (np.abs(wh[m, col] - wld[m, col]) / wld[m, col])[big]
        ro = (np.abs(wo[m, col] - wld[m, col]) / wld[m, col])[big]
        print(f"grid 50x50 col {col}: FP32_HW max {rh.max():.2e} median {np.median(rh):.2e}; "
              f"FP32 max {ro.max():.2e} median {np.median(ro):.2e}")
        assert rh.max() < 2.0 * ro.max() + 1e-5, col
        assert np.median(rh) < 3e-6


def test_integral_and_same_event_sums(ctx, O):
    phys = mi.physics_default(1)
    q, qp = O.had_quadrature(O.hadcfg_default(), 12, 12, 24)
    n = 1 << 24
    e = ctx.integrate(mi.make_run(phys, n, precision=HW, key=42))
    e2 = ctx.integrate(mi.make_run(phys, n, precision=HW, key=42))
    assert e.sumW() == e2.sumW()
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    for i, code in enumerate((111, 112, 113, 114, 115)):
        assert abs(e.sigmaGen(code) - qp[i]) < 3.0 * e.sigmaErr(code) + 1e-12, code
    assert e.sigmaGen(116) == 0.0
    e64 = ctx.integrate(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=42))
    assert abs(e64.sumW() - e.sumW()) < 2e-6 * e64.sumW()
    big = ctx.integrate(mi.make_run(phys, 1 << 28, precision=HW, key=42))
    assert abs(big.sigmaGen() - q) < 3.0 * big.sigmaErr(), (big.sigmaGen(), big.sigmaErr(), q)


def test_unweighting_of_an_fp32_hw_run_is_bit_exact(ctx, O):
    phys = mi.physics_default(1)
    n = 1 << 18
    ev = ctx.unweight_compact(mi.make_run(phys, n, precision=HW, key=9), wmax=2.0)
    cnt, flags, pid, _ = O.unweight32(O.hadcfg_default(), 9, 0, n, 2.0)
    assert ev.n == cnt
    assert np.array_equal(ev.index.cpu().numpy(), np.nonzero(flags)[0])


@pytest.mark.parametrize("pt,xi", [(1.0, 1.0), (3.0, 0.5), (1.2, 2.0)])
def test_points_across_the_flavour_thresholds(ctx, O, pt, xi):
    """Runs whose scales cross m_c^2 and m_b^2 (pTmin 1-3 GeV, renorm_fac 0.5-2): alphaS takes
    its full flavour chain (nf 3, 4, 5), not the one-test form of runs above m_b^2.  The
    bit-exact FP32 path equals the host mirror, FP32_HW stays within 1e-5 of the long-double
    oracle, for the ten-process library."""
    phys = mi.physics_default(1, pTHatMin=pt, renorm_fac=xi, proc_mask=0x3FF)
    cfg = O.hadcfg_default(pTHatMin=pt, renormFac=xi, procMask=0x3FF)
    n = 1 << 17
    w32, _ = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP32, key=11, counter0=3), with_kin=False)
    o32, _ = O.had_points32(cfg, 11, 3, n)
    assert np.array_equal(w32.cpu().numpy().astype(np.float32), o32)
    wh, _ = ctx.points(mi.make_run(phys, n, precision=HW, key=11, counter0=3), with_kin=False)
    wld, _ = O.had_points_ld(cfg, 11, 3, n)
    mx, med = _rel(wh.cpu().numpy(), wld)
    print(f"FP32_HW pT {pt} xi {xi}: max rel {mx:.2e}, median {med:.2e}")
    assert mx < 1e-5 and med < 1e-6
