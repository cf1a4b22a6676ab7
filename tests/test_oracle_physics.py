"""Pins of the oracle's physics (CPU): sigmaHat, alphaS, toy PDFs, the (tau, y, z) map, the
hadronic integrand, the FP32 device-arithmetic mirror and the quadratures.

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign, a swapped t/u,
a factor 2 in a Jacobian) fails it: independent textbook forms, closed forms, sum rules,
published reference numbers and brute-force quadrature."""
import math

import numpy as np
import pytest

CODES = [111, 112, 113, 114, 115, 116]


def _random_kin(rng, n):
    s = 10 ** rng.uniform(2, 8, n)
    c = rng.uniform(-0.999, 0.999, n)
    t = -0.5 * s * (1 - c)
    u = -0.5 * s * (1 + c)
    return s, t, u


def test_sigma_hat_pythia_vs_textbook(O):
    """Pythia's literal brackets (SigmaQCD.cc:115-378) == ESW/Combridge textbook forms < 1e-12."""
    L = O.lib()
    rng = np.random.default_rng(7)
    s, t, u = _random_kin(rng, 4000)
    pairs = [(21, 21), (1, 2), (2, 2), (2, -2), (-1, 3), (3, -3)]
    worst = 0.0
    for code in CODES:
        for (a, b) in pairs:
            for i in range(len(s)):
                py = L.orc_sigma_hat_pythia(code, s[i], t[i], u[i], 0.118, a, b, 3)
                tb = L.orc_bracket_textbook(code, s[i], t[i], u[i], a, b, 3) * math.pi * 0.118 ** 2 / s[i] ** 2
                worst = max(worst, abs(py - tb) / abs(tb))
    assert worst < 1e-12, worst


def test_sigma_hat_symmetries(O):
    """gg->gg, gg->qqbar, qqbar->gg and identical-quark scattering are t<->u symmetric;
    qg->qg and q q' (different flavours) are not."""
    L = O.lib()
    s, t, u = 1e4, -2e3, -8e3
    for code, ids in [(111, (21, 21)), (112, (21, 21)), (115, (1, -1)), (114, (2, 2)), (116, (1, -1))]:
        a = L.orc_sigma_hat_pythia(code, s, t, u, 0.118, *ids, 3)
        b = L.orc_sigma_hat_pythia(code, s, u, t, 0.118, *ids, 3)
        assert abs(a - b) <= 1e-14 * abs(a)
    for code, ids in [(113, (1, 21)), (114, (1, 2))]:
        a = L.orc_sigma_hat_pythia(code, s, t, u, 0.118, *ids, 3)
        b = L.orc_sigma_hat_pythia(code, s, u, t, 0.118, *ids, 3)
        assert abs(a - b) > 1e-3 * abs(a)


def test_gg2gg_closed_form_and_reference_simpson(O):
    """Simpson(2e6) = 1.323768e-4 mb at conv 0.3893793721 (cuPythia/03_fused_resident_me/README.md:13)
    and equals the closed-form integral of the gg->gg bracket."""
    L = O.lib()
    simp = L.orc_c1_simpson(2000000, 1e4, 0.118, 0.9, O.CONV_CUPYTHIA)
    closed = L.orc_c1_closed(1e4, 0.118, 0.9, O.CONV_CUPYTHIA)
    assert abs(simp - closed) < 1e-13 * closed
    assert round(simp, 10) == pytest.approx(1.323768e-4, abs=5e-11)


def test_reference_cpumc_kat(O):
    """cpuMC of fused_gg2gg.cu:53-62 at seed 0xABCDEF01, 2e8 trials: 1.324063e-4 mb
    (cuPythia/03_fused_resident_me/README.md:15)."""
    sig = O.lib().orc_c1_cpumc_compat(0xABCDEF01, 200000000, 1e4, 0.118, 0.9, O.CONV_CUPYTHIA)
    assert abs(sig - 1.324063e-4) < 5e-11


def test_process_library_reference_numbers(O):
    """sigma_cut per process at sqrt(sHat)=100 GeV, |cos|<0.8, alphaS=0.118 in pb
    (cuPythia/08_qcd_process_library/README.md:14-21): qq'->qq' is the t-channel bracket of
    different flavours, gg->qqbar per flavour (nQuarkNew = 1)."""
    L = O.lib()
    s, cmax, conv = 1e4, 0.8, O.CONV_CUPYTHIA * 1e9
    ref = {111: (5.99e4, (21, 21), 1), 113: (2.61e4, (1, 21), 1), 114: (1.07e4, (1, 2), 1),
           115: (1.72e3, (1, -1), 1), 112: (4.83e2, (21, 21), 1)}
    N = 20000
    h = 2 * cmax / N
    for code, (val, ids, nqn) in ref.items():
        acc = 0.0
        for i in range(N + 1):
            c = -cmax + i * h
            t, u = -0.5 * s * (1 - c), -0.5 * s * (1 + c)
            f = L.orc_sigma_hat_pythia(code, s, t, u, 0.118, ids[0], ids[1], nqn)
            acc += (1 if i in (0, N) else (4 if i % 2 else 2)) * f
        sig = acc * h / 3 * (s / 2) * conv
        assert abs(sig - val) / val < 5e-3, (code, sig, val)


def test_alphas_pythia_semantics(O):
    """AlphaStrong order 1 (StandardModel.cc:58-62,207-215): alphaS(MZ^2) = 0.118 exactly,
    continuous at mb and mt, 1-loop running 12 pi / (b0 ln(Q^2/Lambda^2)) with b0 = 33 - 2 nf."""
    import ctypes as C
    L = O.lib()
    a = O.OrcAlphas()
    L.orc_alphas_init(C.byref(a), 0.118, 1, 6)
    assert abs(L.orc_alphas_eval(C.byref(a), 91.188 ** 2) - 0.118) < 1e-14
    for m2 in (a.mb2, a.mt2, a.mc2):
        lo, hi = L.orc_alphas_eval(C.byref(a), m2 * (1 - 1e-12)), L.orc_alphas_eval(C.byref(a), m2 * (1 + 1e-12))
        assert abs(lo - hi) < 1e-9
    q2 = 1000.0
    lam5 = 91.188 * math.exp(-6 * math.pi / (23 * 0.118))
    assert abs(L.orc_alphas_eval(C.byref(a), q2) - 12 * math.pi / (23 * math.log(q2 / lam5 ** 2))) < 1e-14
    assert L.orc_alphas_eval(C.byref(a), 1e-6) == L.orc_alphas_eval(C.byref(a), a.scale2Min)
    a0 = O.OrcAlphas()
    L.orc_alphas_init(C.byref(a0), 0.13, 0, 6)
    assert L.orc_alphas_eval(C.byref(a0), 1e5) == 0.13


def test_toy_pdf_sum_rules(O):
    """Number sum rules (u - ubar = 2, d - dbar = 1, s - sbar = 0) and the momentum sum rule."""
    from scipy import integrate
    cfg = O.hadcfg_default()
    L = O.lib()
    import ctypes as C
    pdf = C.byref(cfg.pdf)

    def num(q):
        f = lambda x: (L.orc_toypdf_xf(pdf, q, x) - L.orc_toypdf_xf(pdf, -q, x)) / x
        return integrate.quad(f, 0, 1, limit=200)[0]

    assert abs(num(2) - 2) < 1e-6 and abs(num(1) - 1) < 1e-6 and abs(num(3)) < 1e-12
    mom = integrate.quad(lambda x: sum(L.orc_toypdf_xf(pdf, q, x) for q in list(range(-5, 0)) + list(range(1, 6)) + [21]),
                         0, 1, limit=400)[0]
    assert abs(mom - 1.0) < 1e-6
    assert L.orc_toypdf_xf(pdf, 21, 0.1) == pytest.approx(cfg.pdf.Ag * 0.1 ** -0.1 * 0.9 ** 5, rel=1e-14)


def test_phase_space_invariants(O):
    cfg = O.hadcfg_default()
    rng = np.random.default_rng(3)
    S = cfg.eCM ** 2
    for _ in range(2000):
        u = rng.random(3)
        tot, w, k = O.had_point(cfg, u)
        assert 0 < k.x1 < 1 and 0 < k.x2 < 1
        assert abs(k.sH - k.x1 * k.x2 * S) < 1e-12 * k.sH
        assert abs(k.sH + k.tH + k.uH) < 1e-9 * k.sH
        assert abs(k.pT2 - k.tH * k.uH / k.sH) < 1e-12 * k.pT2
        assert k.pT2 >= cfg.pTHatMin ** 2 * (1 - 1e-9)


def test_phase_space_jacobian(O):
    """E_u[wtTau wtY wtZ / sHat] = int dtau/tau * 2 yMax * zMax (the t-range length is sHat*zMax):
    a Jacobian missing a factor or a wrong range fails this."""
    from scipy import integrate
    cfg = O.hadcfg_default()
    S = cfg.eCM ** 2
    tmin = 4 * cfg.pTHatMin ** 2 / S
    exact = integrate.quad(lambda v: (-v) * math.sqrt(max(0.0, 1 - tmin / math.exp(v))), math.log(tmin), 0,
                           limit=400)[0]
    n = 1 << 14
    g = (np.arange(n) + 0.5) / n
    rng = np.random.default_rng(11)
    vals = []
    for i in range(n):
        u = [g[i], rng.random(), rng.random()]
        _, _, k = O.had_point(cfg, u)
        vals.append(k.jac / k.sH)
    est = float(np.mean(vals))
    assert abs(est - exact) / exact < 2e-3, (est, exact)


def test_long_double_and_literal_forms(O):
    """FP64 Pythia arithmetic agrees with its 80-bit evaluation to 1e-9 everywhere (it loses digits
    by subtraction next to the pT cut and at x -> 1) and to 1e-11 away from the z edges; the store2Kin literal
    uHat = -(sH + tH) differs from the stable value only by that cancellation."""
    cfg = O.hadcfg_default()
    rng = np.random.default_rng(5)
    worst, worst_mid = 0.0, 0.0
    for _ in range(3000):
        u = rng.integers(0, 2 ** 24, 3) / 2.0 ** 24
        t64, _, _ = O.had_point(cfg, u)
        import ctypes as C
        uu = (C.c_double * 3)(*u); w = (C.c_double * O.NPROC)(); k = O.OrcKin()
        tld = O.lib().orc_had_point_ld(C.byref(cfg), uu, w, C.byref(k), 0)
        if tld > 0:
            r = abs(t64 - tld) / tld
            worst = max(worst, r)
            if 0.05 < u[2] < 0.95 and 0.05 < u[1] < 0.95 and u[0] > 0.01:
                worst_mid = max(worst_mid, r)
    assert worst < 1e-9 and worst_mid < 1e-11, (worst, worst_mid)


def test_fp32_mirror_accuracy(O):
    """The FP32 device arithmetic (own exp/log, stable a = -t/s, b = -u/s) stays within 1e-5
    relative of the long-double Pythia-form oracle at every point whose weight is representable
    (> 1e-30 mb; 2.5e-7 median)."""
    cfg = O.hadcfg_default()
    n = 1 << 15
    wld, kld = O.had_points_ld(cfg, 42, 0, n)
    w32, k32 = O.had_points32(cfg, 42, 0, n)
    nz = wld[:, 0] > 0
    for col in range(6):
        # FP32 underflows below ~1e-38 mb; such points carry no weight (sigma ~ 0.1 mb)
        m = wld[:, col] > 1e-30
        rel = np.abs(w32[m, col].astype(np.float64) - wld[m, col]) / wld[m, col]
        assert rel.max() < 1e-5, (col, rel.max())
        assert np.all(np.abs(w32[~m, col] - wld[~m, col]) < 1e-30)
    rel = np.abs(w32[nz, 0] - wld[nz, 0]) / wld[nz, 0]
    assert np.median(rel) < 1e-6
    assert np.all((w32[:, 6] == 0))                 # 116 is not in the 5-process mask
    assert np.allclose(k32[:, 0], kld[:, 0], rtol=1e-5)


def test_log_table_in_device_source_matches_oracle(O):
    """The device log's table is a literal in mi_device.h; the oracle computes it from its
    defining formula.  Both must agree bit for bit for the FP32 mirror to be exact."""
    import ctypes as C
    import os
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "cupythia_amd", "csrc", "mi_device.h")).read()
    body = src[src.index("kLogTab[128] = {"):]
    body = body[:body.index("};")]
    vals = [float.fromhex(t[:-1]) for t in re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+f", body)]
    assert len(vals) == 256
    G, L = (C.c_float * 128)(), (C.c_float * 128)()
    O.lib().orc32_logtab(G, L)
    assert np.array_equal(np.array(vals[0::2], np.float32), np.array(G[:], np.float32))
    assert np.array_equal(np.array(vals[1::2], np.float32), np.array(L[:], np.float32))
    # every entry reduces its mantissa interval to |m G_j - 1| <= 2^-8
    for j in range(128):
        for m in (1 + j / 128, 1 + (j + 1) / 128):
            assert abs(m * G[j] - 1) <= 2.0 ** -8


def test_own_exp_log_accuracy(O):
    L = O.lib()
    xs = np.linspace(-80, 80, 20001, dtype=np.float32)
    e = np.array([L.orc_expf(float(x)) for x in xs])
    ref = np.exp(xs.astype(np.float64))
    assert np.max(np.abs(e - ref) / ref) < 4e-7
    ys = np.float32(10.0) ** np.linspace(-30, 30, 20001).astype(np.float32)
    lg = np.array([L.orc_logf(float(y)) for y in ys])
    refl = np.log(ys.astype(np.float64))
    assert np.max(np.abs(lg - refl)) < 1e-5 and np.max(np.abs(lg - refl) / np.maximum(np.abs(refl), 1)) < 4e-7
    # within 2^-8 of 1 (log(1 - x) for small x) the error is absolute, ~ulp(2^-8)/2: what matters
    # there, since the value enters b log(1 - x) + a log x + log A of size O(1)
    zs = (1.0 - np.float32(2.0) ** -np.linspace(8, 24, 4001)).astype(np.float32)
    zs = np.concatenate([zs, (1.0 + np.float32(2.0) ** -np.linspace(8, 23, 4001)).astype(np.float32)])
    lz = np.array([L.orc_logf(float(z)) for z in zs])
    refz = np.log(zs.astype(np.float64))
    assert np.max(np.abs(lz - refz)) < 5e-10
    # exp: the Cody-Waite + fitted polynomial keeps ~1 ulp over the whole range used
    xs2 = np.linspace(-86, 0, 40001, dtype=np.float32)
    e2 = np.array([L.orc_expf(float(x)) for x in xs2])
    assert np.max(np.abs(e2 - np.exp(xs2.astype(np.float64))) / np.exp(xs2.astype(np.float64))) < 2.5e-7


def test_hadronic_quadrature_converges_and_matches_mc(O):
    cfg = O.hadcfg_default()
    q1, p1 = O.had_quadrature(cfg, 8, 8, 16)
    q2, p2 = O.had_quadrature(cfg, 12, 12, 24)
    assert abs(q1 - q2) < 1e-7 * q2
    n = 1 << 17
    w, _ = O.had_points_ld(cfg, 42, 0, n)
    m, err = w[:, 0].mean(), w[:, 0].std() / math.sqrt(n)
    assert abs(m - q2) < 4 * err, (m, err, q2)
    for col in range(5):
        mc, e = w[:, 1 + col].mean(), w[:, 1 + col].std() / math.sqrt(n)
        assert abs(mc - p2[col]) < 4 * e + 1e-12
