"""The oracle's heavy-quark pair processes 121..124 (Pythia's Sigma2gg2QQbar, Sigma2qqbar2QQbar,
SigmaQCD.cc, with m3 = m4 = m_c or m_b) and their massive phase space (PhaseSpace.cc:1166-1172,
1256-1337, 1453-1568; SigmaProcess.cc:846-925), pinned by: the textbook total partonic cross
sections (Combridge 1979; Gluck, Owens, Reya 1978) integrated from Pythia's d sigma / d t, the
massless limit against codes 112 and 116 per flavour, the massive map's Jacobian by quadrature
against the Monte Carlo of the Pythia-form map, and the FP32 mirror of the device arithmetic."""
import math

import numpy as np
import pytest


def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


@pytest.mark.parametrize("code", [121, 122, 123, 124])
@pytest.mark.parametrize("sH,m", [(40.0, 1.5), (2.0e3, 1.5), (120.0, 4.8), (5.0e4, 4.8), (9.3, 1.5)])
def test_massive_sigma_integrates_to_textbook_total(O, code, sH, m):
    L = O.lib()
    m2 = m * m
    beta = math.sqrt(1 - 4 * m2 / sH)
    tlo, thi = m2 - 0.5 * sH * (1 + beta), m2 - 0.5 * sH * (1 - beta)
    # Gauss-Legendre on panels of cos(theta) (the integrand peaks at the t ends when beta -> 1)
    x, w = _gl(40)
    edges = np.concatenate([[-1.0], -1 + np.logspace(-8, 0, 30)[:-1], [0.0], 1 - np.logspace(0, -8, 30)[1:], [1.0]])
    tot = 0.0
    for c0, c1 in zip(edges[:-1], edges[1:]):
        cs = 0.5 * (c0 + c1) + 0.5 * (c1 - c0) * x
        for ci, wi in zip(cs, w):
            t = m2 - 0.5 * sH * (1 - beta * ci)
            u = 2 * m2 - sH - t
            tot += 0.5 * (c1 - c0) * wi * (0.5 * sH * beta) * L.orc_sigma_hat_massive(code, sH, t, u, m2, 0.2)
    ref = L.orc_sigma_tot_massive_textbook(code, sH, m2, 0.2)
    assert (tlo, thi) < (0.0, 0.0)
    assert tot == pytest.approx(ref, rel=1e-10), (tot, ref)


def test_textbook_threshold_and_high_energy_limits(O):
    """The textbook totals themselves: zero at threshold, and at s >> m^2 the massless
    integrated forms (qqbar: 8 pi alpS^2 / (27 s); gg: pi alpS^2/(3 s) (ln(s/m^2) - 7/4))."""
    L = O.lib()
    for code in (121, 122):
        assert L.orc_sigma_tot_massive_textbook(code, 4 * 2.25, 2.25, 0.3) == pytest.approx(0.0, abs=1e-15)
    s, m2, a = 1.0e8, 1.0, 0.2
    assert L.orc_sigma_tot_massive_textbook(122, s, m2, a) == pytest.approx(8 * math.pi * a * a / (27 * s), rel=1e-7)
    assert L.orc_sigma_tot_massive_textbook(121, s, m2, a) == pytest.approx(
        math.pi * a * a / (3 * s) * (math.log(s / m2) - 1.75), rel=1e-6)


def test_massless_limit_equals_light_codes(O):
    """m -> 0: gg -> QQbar is gg -> qqbar (112) and qqbar -> QQbar is qqbar -> q'qbar' (116),
    each for one outgoing flavour."""
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(200):
        sH = math.exp(rng.uniform(math.log(50.0), math.log(1e6)))
        t = -sH * rng.uniform(0.01, 0.99)
        u = -sH - t
        for heavy, light in ((121, 112), (123, 112), (122, 116), (124, 116)):
            a = L.orc_sigma_hat_massive(heavy, sH, t, u, 0.0, 0.15)
            b = L.orc_sigma_hat_pythia(light, sH, t, u, 0.15, 2, -2, 1)
            assert a == pytest.approx(b, rel=1e-13)


def test_massive_sigma_symmetry_and_positivity(O):
    """t <-> u symmetry of both codes and positivity over the physical region."""
    L = O.lib()
    rng = np.random.default_rng(5)
    for _ in range(300):
        m2 = rng.choice([2.25, 23.04])
        sH = 4 * m2 * math.exp(rng.uniform(0.001, 8.0))
        beta = math.sqrt(1 - 4 * m2 / sH)
        c = rng.uniform(-1, 1)
        t = m2 - 0.5 * sH * (1 - beta * c)
        u = 2 * m2 - sH - t
        for code in (121, 122):
            v = L.orc_sigma_hat_massive(code, sH, t, u, m2, 0.2)
            assert v > 0
            assert v == pytest.approx(L.orc_sigma_hat_massive(code, sH, u, t, m2, 0.2), rel=1e-12)


def _heavy_cfg(O, **kw):
    kw.setdefault("procMask", 0x3C0)
    kw.setdefault("pTHatMin", 5.0)
    return O.hadcfg_default(**kw)


@pytest.mark.parametrize("mask", [0x3C0, 0x3FF, 0x300])
def test_heavy_quadrature_vs_mc(O, mask):
    """The Pythia-form massive map (tauMin from mT, p2Abs, the z split) as a Monte Carlo over
    2^20 points against the quadrature in (tau, y, ln(a/b)) of the same sigma: per code within
    4 sigma, and the massive codes' totals are a few per mille of the jet cross section."""
    cfg = _heavy_cfg(O, procMask=mask)
    q, qp = O.had_quadrature(cfg, 12, 12, 24)
    w, _ = O.had_points_ld(cfg, 7, 0, 1 << 20)
    mean = w.mean(axis=0)
    err = w.std(axis=0) / math.sqrt(len(w))
    assert abs(mean[0] - q) < 4 * err[0], (mean[0], q, err[0])
    for i, code in enumerate(O.CODES):
        if mask >> i & 1:
            assert qp[i] > 0
            assert abs(mean[1 + i] - qp[i]) < 4 * err[1 + i] + 1e-12 * q, (code, mean[1 + i], qp[i], err[1 + i])
        else:
            assert qp[i] == 0 and np.all(w[:, 1 + i] == 0)


def test_heavy_masses_move_the_cross_section(O):
    """Heavier quarks raise tauMin and the scale: sigma(b bbar) < sigma(c cbar) at the same
    pTmin, and a lighter charm moves the c cbar codes only."""
    cfg = _heavy_cfg(O)
    _, qp = O.had_quadrature(cfg, 10, 10, 20)
    assert qp[8] < qp[6] and qp[9] < qp[7]
    cfg2 = _heavy_cfg(O)
    cfg2.mQ[0] = 1.2
    _, qp2 = O.had_quadrature(cfg2, 10, 10, 20)
    assert qp2[6] > qp[6] and qp2[7] > qp[7]
    assert qp2[8] == pytest.approx(qp[8], rel=1e-12) and qp2[9] == pytest.approx(qp[9], rel=1e-12)


@pytest.mark.parametrize("order", [0, 1])
def test_fp32_heavy_mirror_accuracy(O, order):
    """The FP32 device arithmetic (the massless map with pTmin^2 -> mT^2, Q^2 = sH a b,
    mu = r m^2 / (4 mT^2)) against the long-double Pythia-form massive map at the same points:
    every code's weight above the FP32 normal range (1e-30; the x -> 1 tail underflows) to 2e-5,
    median 3e-6; tHat = m^2 - sH a to 1e-5 relative."""
    cfg = _heavy_cfg(O, procMask=0x3C0, alphaS_order=order)
    n = 1 << 15
    wld, kld = O.had_points_ld(cfg, 42, 0, n)
    w32, k32 = O.had_points32(cfg, 42, 0, n)
    for col in range(7, 11):
        m = wld[:, col] > 1e-30
        assert m.mean() > 0.99
        rel = np.abs(w32[m, col].astype(np.float64) - wld[m, col]) / wld[m, col]
        assert rel.max() < 2e-5, (col, rel.max())
        assert np.median(rel) < 3e-6
    assert np.allclose(k32[:, 1], kld[:, 1], rtol=1e-5, atol=0)
    assert np.allclose(k32[:, 0], kld[:, 0], rtol=1e-5, atol=0)


def test_fp32_heavy_family_alone_matches_combined(O):
    """A family's weights do not depend on which other families are selected (each family has
    its own map at the same uniforms)."""
    n = 1 << 12
    wa, _ = O.had_points32(_heavy_cfg(O, procMask=0x3FF), 42, 0, n)
    wc, _ = O.had_points32(_heavy_cfg(O, procMask=0x0C0), 42, 0, n)
    wb, kb = O.had_points32(_heavy_cfg(O, procMask=0x300), 42, 0, n)
    assert np.array_equal(wa[:, 7:9], wc[:, 7:9]) and np.array_equal(wa[:, 9:11], wb[:, 9:11])
    assert np.all(wc[:, 9:] == 0) and np.all(wc[:, 1:7] == 0)
    # the b bbar family's tHat carries its mass: tHat - m_b^2 < 0, sHat > 4 mT^2
    assert np.all(kb[:, 1] < 23.04) and np.all(kb[:, 0] >= np.float32(4 * (23.04 + 25.0)) * (1 - 1e-6))
