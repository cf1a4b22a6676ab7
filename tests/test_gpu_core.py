"""GPU parity of the integration path against the CPU oracle, through the C ABI.

Bars: RNG words and FP32 per-point weights bit-exact; FP64 per-point weights within 1e-12
relative of the long-double Pythia-form oracle; sigma within 3 sigma_stat of the oracle's
deterministic quadrature; sums bit-identical run to run."""
import math

import numpy as np
import pytest
import torch

import cupythia_amd as mi

pytestmark = pytest.mark.gpu


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def test_philox_words_bit_exact(ctx, O):
    n = 1 << 16
    for stream in (0, 1, 2):
        assert np.array_equal(u32(ctx.philox_fill(42, 0, n, stream)), O.philox_fill(42, 0, n, stream))
    # counters crossing 2^32 and a full 64-bit key
    c0 = (1 << 32) - 1000
    assert np.array_equal(u32(ctx.philox_fill(0x0123456789abcdef, c0, 4096)),
                          O.philox_fill(0x0123456789abcdef, c0, 4096))
    # published vectors through the device path
    assert list(u32(ctx.philox_fill(42, 0, 2))) == [0x9ceaf053, 0x77f5493b, 0x12bf50ad, 0x5742b3d7,
                                                    0xfcdb2127, 0x53ba6cfd, 0x838f5a6e, 0x744e06fb]


def test_sigma_hat_fp64_vs_pythia(ctx, O):
    """mi_sigma_hat == Kfactor * CONVERT2MB * Pythia sigmaKin/sigmaHat within 1e-12 per point."""
    rng = np.random.default_rng(1)
    n = 3000
    s = 10 ** rng.uniform(2, 8, n)
    c = rng.uniform(-0.995, 0.995, n)
    t, u = -0.5 * s * (1 - c), -0.5 * s * (1 + c)
    a = rng.uniform(0.08, 0.3, n)
    dev = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")
    L = O.lib()
    for code in (111, 112, 113, 114, 115, 116):
        for ids in ((1, 2), (2, 2), (2, -2), (21, 21)):
            i1 = torch.full((n,), ids[0], dtype=torch.int32, device="cuda")
            i2 = torch.full((n,), ids[1], dtype=torch.int32, device="cuda")
            for with_u in (True, False):
                got = ctx.sigma_hat(code, dev(s), dev(t), dev(a), uH=dev(u) if with_u else None, id1=i1, id2=i2,
                                    nQuarkNew=3, kfactor=1.5).cpu().numpy()
                uu = u if with_u else -(s + t)
                ref = np.array([1.5 * O.CONVERT2MB * L.orc_sigma_hat_pythia(code, s[k], t[k], uu[k], a[k], ids[0],
                                                                           ids[1], 3) for k in range(n)])
                rel = np.abs(got - ref) / np.abs(ref)
                assert rel.max() < 1e-12, (code, ids, with_u, rel.max())


def test_partonic_points_fp64(ctx, O):
    phys = mi.physics_default(0)
    n = 1 << 16
    w, kin = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=42))
    w = w.cpu().numpy()
    words = O.philox_fill(42, 0, n)
    L = O.lib()
    ref = np.empty(n)
    for i in range(n):
        uu = L.orc_u53(int(words[4 * i]), int(words[4 * i + 1]))
        ref[i] = L.orc_c1_weight(1e4, 0.118, 0.9, 0.9 * (2 * uu - 1), O.CONVERT2MB)
    assert np.max(np.abs(w[:, 0] - ref) / ref) < 1e-12
    assert np.array_equal(w[:, 0], w[:, 1])


def test_partonic_integral(ctx, O):
    """2^24 events: sigma vs Simpson / closed form within 3 sigma; == the oracle's Philox MC of
    the same events up to summation order; bit-identical on a rerun."""
    phys = mi.physics_default(0)
    n = 1 << 24
    run = mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=42)
    e1 = ctx.integrate(run)
    e2 = ctx.integrate(run)
    assert e1.sumW() == e2.sumW() and e1.sumW2() == e2.sumW2()
    L = O.lib()
    closed = L.orc_c1_closed(1e4, 0.118, 0.9, O.CONVERT2MB)
    simp = L.orc_c1_simpson(1000000, 1e4, 0.118, 0.9, O.CONVERT2MB)
    assert abs(closed - simp) < 1e-12 * closed
    assert abs(e1.sigmaGen() - closed) < 3.0 * e1.sigmaErr()
    import ctypes as C
    sums = (C.c_double * 2)()
    L.orc_c1_mc_philox(42, 0, n, 1e4, 0.118, 0.9, O.CONVERT2MB, 8, sums)
    assert abs(e1.sumW() - sums[0]) < 1e-12 * sums[0]
    assert abs(e1.sumW2() - sums[1]) < 1e-12 * sums[1]
    assert e1.nTried() == n and e1.nAccepted() == n and e1.raw.nNonZero == n
    assert e1.sigmaGen(111) == e1.sigmaGen(0)


def test_hadronic_points_fp64_vs_long_double_oracle(ctx, O):
    cfg = O.hadcfg_default()
    n = 1 << 16
    for c0 in (0, 123456789):
        w, kin = ctx.points(mi.make_run(mi.physics_default(1), n, precision=mi.MI_PREC_FP64, key=42, counter0=c0))
        w, kin = w.cpu().numpy(), kin.cpu().numpy()
        wld, kld = O.had_points_ld(cfg, 42, c0, n)
        m = wld > 0
        rel = np.abs(w[m] - wld[m]) / wld[m]
        assert rel.max() < 1e-12, rel.max()
        assert np.array_equal(w == 0, wld == 0)
        assert np.max(np.abs(kin[:, 0] - kld[:, 0]) / kld[:, 0]) < 1e-13


def test_hadronic_points_fp32_bit_exact(ctx, O):
    cfg = O.hadcfg_default()
    n = 1 << 18
    w, kin = ctx.points(mi.make_run(mi.physics_default(1), n, precision=mi.MI_PREC_FP32, key=42, counter0=777))
    w32, k32 = O.had_points32(cfg, 42, 777, n)
    assert np.array_equal(w.cpu().numpy().astype(np.float32), w32)
    assert np.array_equal(kin.cpu().numpy().astype(np.float32), k32)


def test_hadronic_options_fp64(ctx, O):
    """Non-default physics through both paths: 6 processes, nQuarkIn 4, fixed alphaS, renorm_fac 2."""
    import ctypes as C
    phys = mi.physics_default(1, proc_mask=0x3F, nQuarkIn=4, renorm_fac=2.0, alphaS_order=0, alphaS_value=0.13,
                              kfactor=1.3, pdf_g=(-0.2, 6.0))
    cfg = O.hadcfg_default(procMask=0x3F, nQuarkIn=4, renormFac=2.0, alphaS_order=0, alphaS_value=0.13,
                           kfactor=1.3, pdf_g=(-0.2, 6.0))
    n = 1 << 14
    w, _ = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=7))
    wld, _ = O.had_points_ld(cfg, 7, 0, n)
    w = w.cpu().numpy()
    m = wld > 0
    assert np.max(np.abs(w[m] - wld[m]) / wld[m]) < 1e-12
    w32, _ = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP32, key=7))
    o32, _ = O.had_points32(cfg, 7, 0, n)
    assert np.array_equal(w32.cpu().numpy().astype(np.float32), o32)


@pytest.fixture(scope="module")
def quad(O):
    return O.had_quadrature(O.hadcfg_default(), 12, 12, 24)


def test_hadronic_integral_fp32(ctx, O, quad):
    q, qp = quad
    run = mi.make_run(mi.physics_default(1), 1 << 24, key=42)
    e = ctx.integrate(run)
    e2 = ctx.integrate(run)
    assert e.sumW() == e2.sumW() and e.sumW2() == e2.sumW2()
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    for i, code in enumerate((111, 112, 113, 114, 115)):
        assert abs(e.sigmaGen(code) - qp[i]) < 3.0 * e.sigmaErr(code) + 1e-12, code
    assert e.sigmaGen(116) == 0.0
    tot = sum(e.sigmaGen(c) for c in (111, 112, 113, 114, 115))
    assert abs(tot - e.sigmaGen()) < 1e-9 * tot
    # FP64 on the same events: the sums agree to the FP32 accuracy
    e64 = ctx.integrate(mi.make_run(mi.physics_default(1), 1 << 24, precision=mi.MI_PREC_FP64, key=42))
    assert abs(e64.sumW() - e.sumW()) < 2e-6 * e64.sumW()
    # Pythia's sigmaDelta with nSel = nAcc = nTry
    n = e.nTried()
    avg = e.sumW() / n
    assert e.sigmaErr() == pytest.approx(math.sqrt((e.sumW2() / n - avg * avg) / n), rel=1e-12)


def test_hadronic_integral_full_size(ctx, O, quad):
    """The 2^28-point workload: sigma within 3 sigma_stat of the quadrature (size-independent)."""
    q, _ = quad
    e = ctx.integrate(mi.make_run(mi.physics_default(1), 1 << 28, key=42))
    assert abs(e.sigmaGen() - q) < 3.0 * e.sigmaErr(), (e.sigmaGen(), e.sigmaErr(), q)
    assert e.raw.nNonZero > 0.99 * (1 << 28)
    print(f"2^28 FP32: sigma {e.sigmaGen():.8e} +- {e.sigmaErr():.2e}, quad {q:.8e}, {e.kernel_ms:.3f} ms")


def test_counter_partition_is_additive(ctx):
    """Disjoint counter ranges reproduce the single range: Philox events depend on the index only."""
    phys = mi.physics_default(1)
    full = ctx.points(mi.make_run(phys, 3000, key=9, counter0=100), with_kin=False)[0]
    a = ctx.points(mi.make_run(phys, 1000, key=9, counter0=100), with_kin=False)[0]
    b = ctx.points(mi.make_run(phys, 2000, key=9, counter0=1100), with_kin=False)[0]
    assert torch.equal(full, torch.cat([a, b]))


def test_error_paths(ctx):
    bad = mi.physics_default(1, proc_mask=0)
    with pytest.raises(mi.MIError) as ei:
        ctx.integrate(mi.make_run(bad, 1000))
    assert ei.value.status == -1
    ctx2 = mi.Context(0)
    with pytest.raises(mi.MIError) as ei:
        ctx2.integrate(mi.make_run(mi.physics_default(1), 1000, sampler=mi.MI_SAMPLER_VEGAS))
    assert ei.value.status == -3
    with pytest.raises(mi.MIError) as ei:
        ctx2.vegas_init(3, 1000)
    assert ei.value.status == -4
    ctx2.close()


def test_scale_reweighting_matches_reruns(ctx):
    """One pass of mi_reweight_scales == re-running the events at each renormalisation-scale factor
    (cuPythia/pipeline/VALIDATION.md:19 reports max|diff| = 0 for its analogue); here FP64 to 1e-10
    (the reweighting recomputes uHat = -(sHat + tHat))."""
    phys = mi.physics_default(1)
    n = 1 << 14
    w0, kin = ctx.points(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=5))
    facs = [1.0, 0.5, 2.0]
    rw = ctx.reweight_scales(phys, kin[:, 0].contiguous(), kin[:, 1].contiguous(), w0[:, 0].contiguous(), facs)
    for k, f in enumerate(facs):
        wf, _ = ctx.points(mi.make_run(mi.physics_default(1, renorm_fac=f), n, precision=mi.MI_PREC_FP64, key=5))
        m = wf[:, 0] > 0
        rel = ((rw[m, k] - wf[m, 0]).abs() / wf[m, 0]).max().item()
        assert rel < 1e-10, (f, rel)
