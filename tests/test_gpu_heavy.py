"""Heavy-quark pair processes 121..124 (gg -> c cbar, q qbar -> c cbar, gg -> b bbar,
q qbar -> b bbar; Pythia's Sigma2gg2QQbar / Sigma2qqbar2QQbar with the massive phase space) on
the device against the oracle: per-point FP64 weights against the long-double Pythia-form massive
map, FP32 weights and kinematics bit for bit, sigma against the quadrature, unweighting decisions
bit for bit with each event's kinematics from its own process's family, the batched massive
sigmaHat, scale reweighting with the massive scale, and LHEF output with massive final states."""
import math
import os
import sys

import numpy as np
import pytest
import torch

import cupythia_amd as mi

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

PT = 5.0          # a low pT cut, where the quark masses matter


def phys_h(mask, **kw):
    return mi.physics_default(1, proc_mask=mask, pTHatMin=PT, **kw)


def cfg_h(O, mask, **kw):
    return O.hadcfg_default(procMask=mask, pTHatMin=PT, **kw)


@pytest.mark.parametrize("mask", [0x3C0, 0x3FF, 0x0C0, 0x200])
def test_heavy_points_fp64_vs_long_double(ctx, O, mask):
    n = 1 << 15
    w, kin = ctx.points(mi.make_run(phys_h(mask), n, precision=mi.MI_PREC_FP64, key=42, counter0=11))
    w, kin = w.cpu().numpy(), kin.cpu().numpy()
    wld, kld = O.had_points_ld(cfg_h(O, mask), 42, 11, n)
    assert w.shape == wld.shape == (n, 11)
    m = wld > 1e-250
    rel = np.abs(w[m] - wld[m]) / wld[m]
    assert rel.max() < 1e-11, rel.max()
    assert np.all(w[~m] < 1e-240)
    for i in range(10):
        if not (mask >> i) & 1:
            assert np.all(w[:, 1 + i] == 0)
    # kinematics of the first selected family: sHat, and tHat = m^2 - sHat a
    assert np.max(np.abs(kin[:, 0] - kld[:, 0]) / kld[:, 0]) < 1e-12
    assert np.max(np.abs(kin[:, 1] - kld[:, 1]) / np.abs(kld[:, 1])) < 1e-9


@pytest.mark.parametrize("mask,order", [(0x3C0, 1), (0x3FF, 1), (0x300, 0)])
def test_heavy_points_fp32_bit_exact(ctx, O, mask, order):
    n = 1 << 17
    w, kin = ctx.points(mi.make_run(phys_h(mask, alphaS_order=order), n, key=42, counter0=5))
    w32, k32 = O.had_points32(cfg_h(O, mask, alphaS_order=order), 42, 5, n)
    assert np.array_equal(w.cpu().numpy().astype(np.float32), w32)
    assert np.array_equal(kin.cpu().numpy().astype(np.float32), k32)


def test_heavy_points_fp32_vegas_map_bit_exact(ctx, O):
    """Through an adapted VEGAS grid (the grid's FP32 tables, complement column) as well."""
    mask = 0x3FF
    ctx.vegas_init(3, 32)
    ctx.vegas_adapt(mi.make_run(phys_h(mask), 1 << 18, key=3), 3, 1.5)
    edges = ctx.vegas_get_grid()
    n = 1 << 16
    w, kin = ctx.points(mi.make_run(phys_h(mask), n, key=42, counter0=9, sampler=mi.MI_SAMPLER_VEGAS))
    w32, k32 = O.had_points32(cfg_h(O, mask), 42, 9, n, edges=edges)
    assert np.array_equal(w.cpu().numpy().astype(np.float32), w32)
    assert np.array_equal(kin.cpu().numpy().astype(np.float32), k32)


@pytest.fixture(scope="module")
def quad_h(O):
    return O.had_quadrature(cfg_h(O, 0x3FF), 12, 12, 24)


def test_heavy_integral_vs_quadrature(ctx, O, quad_h):
    q, qp = quad_h
    e = ctx.integrate(mi.make_run(phys_h(0x3FF), 1 << 24, key=42))
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    for i, code in enumerate(mi.PROC_CODES):
        assert abs(e.sigmaGen(code) - qp[i]) < 3.0 * e.sigmaErr(code) + 1e-12 * q, (code, e.sigmaGen(code), qp[i])
        assert e.raw.code[1 + i] == code
    # the massive families alone give the same per-code sums (own maps, same uniforms)
    eh = ctx.integrate(mi.make_run(phys_h(0x3C0), 1 << 24, key=42))
    for code in (121, 122, 123, 124):
        assert eh.sumW(code) == e.sumW(code) and eh.wMax(code) == e.wMax(code)
    assert eh.sigmaGen(111) == 0.0
    # FP64 on the same points agrees to the FP32 accuracy
    e64 = ctx.integrate(mi.make_run(phys_h(0x3C0), 1 << 22, precision=mi.MI_PREC_FP64, key=42))
    e32 = ctx.integrate(mi.make_run(phys_h(0x3C0), 1 << 22, key=42))
    assert abs(e64.sumW() - e32.sumW()) < 2e-6 * e64.sumW()


@pytest.mark.parametrize("exact_max", [True, False])
def test_heavy_unweight_bit_exact_and_event_kinematics(ctx, O, exact_max):
    """Exact maximum, and a low wMax (20 x the mean weight; heavier events keep their weight) so
    that the heavy codes, 0.3% of sigma at pT > 5 GeV, show up among the accepted events."""
    mask = 0x3FF
    n = 1 << 19 if exact_max else 1 << 20
    wm = 0.0 if exact_max else 20 * 9.15
    ev = ctx.unweight_compact(mi.make_run(phys_h(mask), n, key=42, counter0=1000), wmax=wm)
    cnt, flags, pid, w = O.unweight32(cfg_h(O, mask), 42, 1000, n, ev.wmax, None)
    assert ev.n == cnt and cnt > 0
    assert ev.wmax == (float(w.max()) if exact_max else wm)
    idx = np.nonzero(flags)[0]
    assert np.array_equal(ev.index.cpu().numpy(), idx + 1000)
    got_pid = ev.pid.cpu().numpy()
    assert np.array_equal(got_pid, pid[idx])
    if not exact_max:
        seen = set(np.unique(got_pid).tolist())
        assert {121, 123} <= seen and seen & {122, 124}, seen
    # each event's (sHat, tHat, x1, x2) is its own family's: compare with that family alone
    sH, tH = ev.sHat.cpu().numpy(), ev.tHat.cpu().numpy()
    x1 = ev.x1.cpu().numpy()
    for fam_mask, codes in ((0x3F, range(111, 117)), (0xC0, (121, 122)), (0x300, (123, 124))):
        _, kf = O.had_points32(cfg_h(O, fam_mask), 42, 1000, n)
        sel = np.isin(got_pid, list(codes))
        j = idx[sel]
        assert np.array_equal(sH[sel].astype(np.float32), kf[j, 0])
        assert np.array_equal(tH[sel].astype(np.float32), kf[j, 1])
        assert np.array_equal(x1[sel].astype(np.float32), kf[j, 2])


def test_heavy_sigma_hat_vs_pythia_massive(ctx, O):
    """mi_sigma_hat for 121..124 (mass from (s + t + u) / 2) == Kfactor * CONVERT2MB * Pythia's
    massive sigmaKin within 1e-11 per point; without uH the call is rejected."""
    L = O.lib()
    rng = np.random.default_rng(7)
    n = 4096
    for code, m in ((121, 1.5), (122, 1.5), (123, 4.8), (124, 4.8)):
        m2 = m * m
        s = 4 * m2 * np.exp(rng.uniform(0.01, 9.0, n))
        beta = np.sqrt(1 - 4 * m2 / s)
        c = rng.uniform(-0.999, 0.999, n)
        t = m2 - 0.5 * s * (1 - beta * c)
        u = 2 * m2 - s - t
        a = rng.uniform(0.1, 0.3, n)
        dev = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
        got = ctx.sigma_hat(code, dev(s), dev(t), dev(a), uH=dev(u), kfactor=1.5).cpu().numpy()
        ref = np.array([1.5 * O.CONVERT2MB * L.orc_sigma_hat_massive(code, s[k], t[k], u[k], m2, a[k])
                        for k in range(n)])
        assert np.max(np.abs(got - ref) / ref) < 1e-11, code
        with pytest.raises(mi.MIError):
            ctx.sigma_hat(code, dev(s), dev(t), dev(a))
    sig = mi.Sigma2gg2bbbar(ctx)
    assert (sig.code(), sig.inFlux(), sig.idNew()) == (123, "gg", 5)
    with pytest.raises(ValueError):
        sig.sigmaHatWrap(torch.ones(2, dtype=torch.float64, device="cuda"),
                         -torch.ones(2, dtype=torch.float64, device="cuda"),
                         torch.ones(2, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("mask,code", [(0x0C0, 121), (0x300, 124)])
def test_heavy_scale_reweighting_matches_reruns(ctx, mask, code):
    """With the event codes, mi_reweight_scales uses the massive scale pT^2 + m^2 and reproduces
    re-runs at each factor (FP64, 1e-10); as massless events (no pid) it would not."""
    phys = phys_h(mask)
    n = 1 << 14
    w0, kin = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=5))
    pid = torch.full((n,), code, dtype=torch.int32, device="cuda")
    facs = [1.0, 0.5, 2.0]
    sH, tH, w = kin[:, 0].contiguous(), kin[:, 1].contiguous(), w0[:, 0].contiguous()
    rw = ctx.reweight_scales(phys, sH, tH, w, facs, pid=pid)
    for k, f in enumerate(facs):
        wf, _ = ctx.points(mi.make_run(phys_h(mask, renorm_fac=f), n, precision=mi.MI_PREC_FP64, key=5))
        m = wf[:, 0] > 0
        rel = ((rw[m, k] - wf[m, 0]).abs() / wf[m, 0]).max().item()
        assert rel < 1e-10, (f, rel)
    rw0 = ctx.reweight_scales(phys, sH, tH, w, facs)
    assert ((rw0[:, 1] - rw[:, 1]).abs() > 1e-6 * rw[:, 1].abs()).any()


def test_heavy_lhe(ctx, tmp_path):
    """LHEF of c cbar / b bbar events: Pythia codes in the init block, final states 4 -4 / 5 -5
    (or the charge-conjugate order for qbar q), on-shell masses, four-momentum conservation,
    incoming g g or q qbar, valid colour flows, event scale = renorm_fac sqrt(pT^2 + m^2)."""
    from test_gpu_lhe import colour_ok, parse
    phys = phys_h(0x3C0)
    ev = ctx.unweight_compact(mi.make_run(phys, 1 << 18, key=21, counter0=7))
    assert ev.n > 100
    path = str(tmp_path / "hq.lhe")
    ctx.write_lhe(path, phys, ev, key=21)
    beam, procs, events = parse(path)
    assert {int(p[3]) for p in procs} == {121, 122, 123, 124}
    assert len(events) == ev.n
    pid = ev.pid.cpu().numpy()
    sH, tH = ev.sHat.cpu().numpy(), ev.tHat.cpu().numpy()
    for n, (head, parts) in enumerate(events):
        code = int(head[1])
        assert code == pid[n]
        mQ, idQ = (1.5, 4) if code in (121, 122) else (4.8, 5)
        ids = [int(q[0]) for q in parts]
        assert sorted(ids[2:]) == [-idQ, idQ]
        if code in (121, 123):
            assert ids[0] == ids[1] == 21
        else:
            assert ids[0] == -ids[1] and ids[0] != 21 and abs(ids[0]) <= 5
            assert (ids[2] > 0) == (ids[0] > 0)
        p = [[float(x) for x in q[6:11]] for q in parts]
        tot = [p[2][j] + p[3][j] - p[0][j] - p[1][j] for j in range(4)]
        scale = max(abs(x) for x in p[0][:4] + p[1][:4])
        assert max(abs(x) for x in tot) < 1e-9 * scale
        for q in p[2:]:
            m2 = q[3] ** 2 - q[0] ** 2 - q[1] ** 2 - q[2] ** 2
            assert q[4] == pytest.approx(mQ, rel=1e-12)
            assert m2 == pytest.approx(mQ * mQ, abs=1e-8 * q[3] ** 2)
        uH = 2 * mQ * mQ - sH[n] - tH[n]
        Q2 = (mQ * mQ - tH[n]) * (mQ * mQ - uH) / sH[n]
        assert float(head[3]) == pytest.approx(math.sqrt(Q2), rel=1e-9)
        assert colour_ok(parts), parts


def test_heavy_flow_sampler(ctx, quad_h):
    """The flow sampler's weight kernel with the heavy families: after a short training on the
    10-process integrand, sigma per code agrees with the quadrature."""
    q, qp = quad_h
    phys = phys_h(0x3FF)
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 6, 1.5)
    ctx.flow_init(5)
    info = ctx.flow_train(mi.make_run(phys, 1 << 20, key=43), 10, 1 << 16, 2e-3)
    assert math.isfinite(info.loss_last)
    est = ctx.integrate(mi.make_run(phys, 1 << 22, key=78, sampler=mi.MI_SAMPLER_FLOW))
    assert abs(est.sigmaGen() - q) < 3.0 * est.sigmaErr(), (est.sigmaGen(), est.sigmaErr(), q)
    for i, code in enumerate(mi.PROC_CODES):
        assert abs(est.sigmaGen(code) - qp[i]) < 3.0 * est.sigmaErr(code) + 1e-12 * q, code


def test_heavy_errors(ctx):
    with pytest.raises(mi.MIError):
        ctx.integrate(mi.make_run(phys_h(0x0C0, mass_Q=(0.0, 4.8)), 1024))
    with pytest.raises(mi.MIError):
        ctx.integrate(mi.make_run(mi.physics_default(1, proc_mask=0x400), 1024))
    with pytest.raises(mi.MIError):                       # 2 pTmin < eCM < 2 mT of the b pair
        ctx.integrate(mi.make_run(mi.physics_default(1, proc_mask=0x300, eCM=12.0, pTHatMin=5.0), 1024))
    with pytest.raises(mi.MIError):                       # the adjoint checks the heavy masses as well
        ctx.adjoint(mi.make_run(phys_h(0x0C0, mass_Q=(0.0, 4.8)), 1 << 16))
    # the light codes of the same physics are unaffected by the heavy masses
    a = ctx.integrate(mi.make_run(phys_h(0x3F, mass_Q=(1.2, 4.5)), 1 << 16, key=3))
    b = ctx.integrate(mi.make_run(phys_h(0x3F), 1 << 16, key=3))
    assert a.sumW() == b.sumW()
