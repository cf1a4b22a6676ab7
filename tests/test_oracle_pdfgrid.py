"""The oracle's LHAGrid1 reader and interpolation (Pythia's LHAGrid1, PartonDistributions.cc:
478-745), pinned by: the reference's published gluon value of the default NNPDF2.3 LO set,
that set's momentum and valence sum rules, exactness on grids of cubic polynomials in
(ln x, ln q), node exactness, the freezing rules, Q subgrids, and the hadronic cross section
of the toy PDFs tabulated as a grid."""
import math
import os

import numpy as np
import pytest

import pdfgrid_util as PG

NNPDF = os.path.join("/root/reference", "pythia8317", "share", "Pythia8", "pdfdata", "NNPDF23_lo_as_0130_qed_0000.dat")
need_nnpdf = pytest.mark.skipif(not os.path.exists(NNPDF), reason="NNPDF2.3 LO data file not available here")


@need_nnpdf
def test_nnpdf_gluon_known_answer(O):
    """cuPythia/pipeline/PRECISION.md:92: Pythia 8.317's default proton PDF gives
    xf_g(x = 0.01, Q = 100 GeV) = 7.9801."""
    g = O.LhaGrid(NNPDF)
    assert (g.g.nx, g.g.nq, g.g.nqSub) == (100, 50, 1)
    assert round(g.xf(0.01, 1e4)[0], 4) == 7.9801


@need_nnpdf
def test_nnpdf_sum_rules(O):
    """Momentum sum (partons + photon) and valence numbers u - ubar = 2, d - dbar = 1,
    s - sbar = c - cbar = 0 by quadrature in ln x of the interpolated grid, at Q = 2, 10, 100 GeV
    (below the top threshold, whose columns Pythia drops).  The LO set's momentum sum is 1.005
    and, as evolution conserves it, the same at every Q to 1e-5."""
    g = O.LhaGrid(NNPDF)
    lx = np.linspace(math.log(1e-9), 0.0, 20001)
    xs = np.exp(lx)
    dlx = lx[1] - lx[0]
    wts = np.full(len(xs), dlx); wts[0] = wts[-1] = 0.5 * dlx            # trapezoid in ln x
    moms = []
    for Q in (2.0, 10.0, 100.0):
        vals = np.array([g.xf(float(xx), Q * Q) for xx in xs])
        moms.append(float(np.sum(wts * xs * vals.sum(axis=1))))       # int sum xf dx
        num = lambda a, b: float(np.sum(wts * (vals[:, a] - vals[:, b])))   # int (f_a - f_b) dx
        assert abs(num(2, 7) - 2.0) < 1e-2, num(2, 7)                  # u - ubar
        assert abs(num(1, 6) - 1.0) < 1e-2, num(1, 6)                  # d - dbar
        assert abs(num(3, 8)) < 1e-3 and abs(num(4, 9)) < 1e-3        # s - sbar, c - cbar
    assert abs(moms[0] - 1.0) < 1e-2
    assert max(moms) - min(moms) < 1e-4, moms


def test_cubic_exactness_nodes_and_freezing(O, tmp_path):
    xs = PG.nnpdf_like_x_grid(20, 20, 1e-6)
    qs = PG.q_grid(1.0, 1e4, 12)
    fn = PG.cubic_fn(3)
    path = str(tmp_path / "cubic.dat")
    PG.write_lhagrid1(path, xs, [qs], fn)
    g = O.LhaGrid(path)
    rng = np.random.default_rng(1)
    for _ in range(300):
        x = math.exp(rng.uniform(math.log(1.01e-6), math.log(0.999)))
        q = math.exp(rng.uniform(0.01, math.log(9.9e3)))
        v = g.xf(x, q * q)
        for i in PG.IDS:
            ref = fn(i, x, q)
            assert abs(v[PG.SLOT[i]] - ref) < 1e-9 * (1 + abs(ref)), (i, x, q, v[PG.SLOT[i]], ref)
    # nodes: exact for any values (q = qMin and q = qMax are the frozen border rows)
    for ix in (5, 19, 38):
        for iq in (0, 3, 11):
            v = g.xf(float(xs[ix]), float(qs[iq]) ** 2)
            for i in PG.IDS:
                ref = fn(i, xs[ix], qs[iq])
                assert abs(v[PG.SLOT[i]] - ref) < 1e-9 * (1 + abs(ref))
    # q frozen outside the grid, x frozen below xMin, 0 at x = 1 (x >= xMax)
    x = 3e-3
    assert g.xf(x, 0.25) == g.xf(x, 1.0) and g.xf(x, 4e8) == g.xf(x, 1e8)
    lo1, lo2 = g.xf(1e-8, 100.0), g.xf(1e-7, 100.0)
    assert lo1 == lo2
    at0 = np.array(g.xf(float(xs[0]) * (1 + 1e-12), 100.0))
    assert np.allclose(np.array(lo1), at0, rtol=1e-9, atol=1e-9)
    assert g.xf(1.0, 100.0) == [0.0] * 12


def test_q_subgrids_and_linear_subgrid(O, tmp_path):
    """Two subgrids sharing a border (averaged for the search, kept for the weights), the
    second of only 3 points (linear in ln q): a function linear in ln q is reproduced
    everywhere, and a cubic one inside the 4+ point subgrid."""
    xs = PG.nnpdf_like_x_grid(10, 10, 1e-5)
    q1 = PG.q_grid(1.0, 100.0, 8)
    q2 = np.array([100.0, 1000.0, 1.0e4])
    lin = lambda i, x, q: (1 + abs(i)) * (2.0 + math.log(q)) * (3.0 + 0.2 * math.log(x) + 0.01 * math.log(x) ** 2)
    path = str(tmp_path / "sub.dat")
    PG.write_lhagrid1(path, xs, [q1, q2], lin)
    g = O.LhaGrid(path)
    assert (g.g.nq, g.g.nqSub, g.g.nqSum[0], g.g.nqSum[1]) == (11, 2, 8, 11)
    assert g.g.qDiv[0] == pytest.approx(100.0)
    rng = np.random.default_rng(2)
    for _ in range(200):
        x = math.exp(rng.uniform(math.log(2e-5), math.log(0.5)))
        q = math.exp(rng.uniform(0.05, math.log(9.5e3)))
        v = g.xf(x, q * q)
        # quadratic in ln x and linear in ln q: reproduced exactly in both subgrids
        ref = lin(21, x, q)
        assert abs(v[0] - ref) < 1e-10 * abs(ref), (x, q, v[0], ref)
        # the ratio of two flavours isolates the q dependence exactly
        assert v[1] / v[0] == pytest.approx(2.0 / 22.0, rel=1e-12)


def test_toy_pdfs_as_grid_reproduce_toy_sigma(O, tmp_path):
    """The 5-process sigma with the toy PDFs tabulated on a fine grid (no Q dependence) equals
    the toy-PDF sigma: the flavour slots, the clipping and the flux sums are wired alike."""
    xs = np.exp(np.linspace(math.log(1e-6), 0.0, 400))
    xs[-1] = 1.0
    qs = PG.q_grid(1.0, 1.0e4, 6)
    path = str(tmp_path / "toy.dat")
    PG.write_lhagrid1(path, xs, [qs], PG.toy_fn(O))
    g = O.LhaGrid(path)
    cfg_t = O.hadcfg_default()
    cfg_g = O.hadcfg_default(grid=g)
    q_t, p_t = O.had_quadrature(cfg_t, 8, 8, 16)
    q_g, p_g = O.had_quadrature(cfg_g, 8, 8, 16)
    assert abs(q_g - q_t) < 2e-4 * q_t, (q_g, q_t)
    for a, b in zip(p_g, p_t):
        assert abs(a - b) < 3e-4 * q_t


def test_grid_point_fp64_vs_long_double(O, tmp_path):
    xs = PG.nnpdf_like_x_grid(40, 40, 1e-7)
    qs = PG.q_grid(1.0, 1.0e4, 30)
    path = str(tmp_path / "toyq.dat")
    PG.write_lhagrid1(path, xs, [qs], PG.toy_fn(O, Qdep=0.05))
    g = O.LhaGrid(path)
    cfg = O.hadcfg_default(grid=g, factFac=1.3)
    wd, _ = O.had_points_ld(cfg, 42, 0, 4096)
    rng = np.random.default_rng(4)
    for _ in range(200):
        u = rng.integers(0, 2 ** 24, 3) / 2.0 ** 24
        t64, w64, _ = O.had_point(cfg, u)
        import ctypes as C
        uu = (C.c_double * 3)(*u); w = (C.c_double * O.NPROC)(); k = O.OrcKin()
        tld = O.lib().orc_had_point_ld(C.byref(cfg), uu, w, C.byref(k), 0)
        if tld > 0:
            assert abs(t64 - tld) < 1e-9 * tld
    assert np.all(wd >= 0) and wd[:, 0].mean() > 0


def test_fp32_grid_mirror_accuracy(O, tmp_path):
    """The FP32 mirror of the device's grid evaluation against the long-double Pythia-form oracle
    (same grid, same points): median 1.2e-6 (the FP32 resolution of ln x and ln q in the Lagrange
    weights), 99.9% of the weights above 1e-8 of the mean within 1e-5 and all of them within 5e-5;
    the cubic interpolation of the steep large-x tail cancels to larger relative errors only on
    weights ~1e-20 of the mean."""
    xs = PG.nnpdf_like_x_grid(40, 40, 1e-7)
    qs = PG.q_grid(1.0, 1.0e4, 30)
    path = str(tmp_path / "toyq.dat")
    PG.write_lhagrid1(path, xs, [qs], PG.toy_fn(O, Qdep=0.05))
    g = O.LhaGrid(path)
    cfg = O.hadcfg_default(grid=g, factFac=0.8)
    n = 1 << 14
    wld, _ = O.had_points_ld(cfg, 42, 0, n)
    w32, _ = O.had_points32(cfg, 42, 0, n)
    for col in range(1, 7):
        m = wld[:, col] > 1e-30
        if not m.any():
            continue
        rel = np.abs(w32[m, col].astype(np.float64) - wld[m, col]) / wld[m, col]
        big = wld[m, col] > 1e-8 * wld[m, col].mean()
        assert rel[big].max() < 5e-5, (col, rel[big].max())
        assert np.quantile(rel[big], 0.999) < 1e-5, col
        assert np.median(rel) < 3e-6
