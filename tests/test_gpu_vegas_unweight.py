"""GPU parity of VEGAS adaptation and unweighting + compaction against the CPU oracle."""
import math

import numpy as np
import pytest
import torch

import cupythia_amd as mi

pytestmark = pytest.mark.gpu


def test_vegas_3d_matches_oracle(ctx, O):
    """Same events, same algorithm: grids agree to 1e-9 after 6 iterations (the per-bin sums are
    added in a different order on the device), per-iteration sums to 1e-9."""
    cfg = O.hadcfg_default()
    n, it, nb = 1 << 17, 6, 64
    ctx.vegas_init(3, nb)
    info = ctx.vegas_adapt(mi.make_run(mi.physics_default(1), n, key=42, counter0=5), it, 1.5)
    g = ctx.vegas_get_grid()
    eo, sums = O.vegas_adapt32(cfg, nb, 1.5, it, 42, 5, n)
    assert np.max(np.abs(g - eo)) < 1e-9
    assert info.last_sigma == pytest.approx(sums[-1, 0] / n, rel=1e-9)
    assert np.all(np.diff(g, axis=1) > 0) and np.all(g[:, 0] == 0) and np.all(g[:, -1] == 1)


def test_vegas_deterministic_and_alpha0(ctx):
    run = mi.make_run(mi.physics_default(1), 1 << 18, key=3)
    grids = []
    for _ in range(2):
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(run, 3, 1.5)
        grids.append(ctx.vegas_get_grid())
    assert np.array_equal(grids[0], grids[1])
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(run, 2, 0.0)
    assert np.allclose(ctx.vegas_get_grid(), np.tile(np.linspace(0, 1, 65), (3, 1)), atol=1e-14)


@pytest.mark.parametrize("nb", [64, 50, 256])
def test_vegas_refine_kernel_matches_oracle_refine(ctx, O, nb):
    """One adaptation step from a non-uniform grid equals the oracle's refinement of the oracle's
    histogram of the same events: 64 bins (one wave per dimension), 50 (a partial wave) and 256
    (the maximum, four waves: the block scan crosses waves)."""
    cfg = O.hadcfg_default()
    rng = np.random.default_rng(2)
    e0 = np.sort(rng.random((3, nb - 1)), axis=1)
    e0 = np.concatenate([np.zeros((3, 1)), e0, np.ones((3, 1))], axis=1)
    ctx.vegas_set_grid(e0)
    ctx.vegas_adapt(mi.make_run(mi.physics_default(1), 1 << 16, key=11), 1, 1.5)
    g = ctx.vegas_get_grid()
    import ctypes as C
    e = e0.copy()
    sums = np.zeros(2)
    O.lib().orc32_vegas_adapt(C.byref(cfg), nb, 1.5, 1, 11, 0, 1 << 16, O._dp(e), O._dp(sums))
    assert np.max(np.abs(g - e)) < 1e-10
    assert np.all(np.diff(g, axis=1) > 0) and np.all(g[:, 0] == 0) and np.all(g[:, -1] == 1)


def test_vegas_c3_workload(ctx, O):
    """12 x 2^24 with 64 bins / dim, alpha 1.5: the combined sigma agrees with the quadrature
    within 3 sigma and the weight spread shrinks (eta rises, error per point falls)."""
    q, _ = O.had_quadrature(O.hadcfg_default(), 12, 12, 24)
    flat = ctx.integrate(mi.make_run(mi.physics_default(1), 1 << 24, key=42, counter0=1 << 40))
    ctx.vegas_init(3, 64)
    info = ctx.vegas_adapt(mi.make_run(mi.physics_default(1), 1 << 24, key=42), 12, 1.5)
    assert abs(info.sigma - q) < 3.0 * info.sigma_err, (info.sigma, info.sigma_err, q)
    assert abs(info.last_sigma - q) < 3.0 * info.last_err
    assert info.last_err < 0.2 * flat.sigmaErr()
    assert info.last_eta > 5 * (flat.sigmaGen() / flat.wMax())
    ev = ctx.integrate(mi.make_run(mi.physics_default(1), 1 << 24, key=42, sampler=mi.MI_SAMPLER_VEGAS,
                                   counter0=1 << 30))
    assert abs(ev.sigmaGen() - q) < 3.0 * ev.sigmaErr()
    print(f"VEGAS: sigma {info.sigma:.8e} +- {info.sigma_err:.2e} chi2/dof {info.chi2_dof:.2f}, "
          f"eta {info.last_eta:.3f} (flat {flat.sigmaGen() / flat.wMax():.4f}), {info.kernel_ms:.2f} ms")


def test_partonic_vegas_1d_matches_oracle(ctx, O):
    phys = mi.physics_default(0)
    ctx.vegas_init(1, 128)
    info = ctx.vegas_adapt(mi.make_run(phys, 1 << 16, precision=mi.MI_PREC_FP64, key=42), 5, 1.5)
    g = ctx.vegas_get_grid()
    eo, sums = O.c1_vegas_adapt(1e4, 0.118, 0.9, O.CONVERT2MB, 128, 1.5, 5, 42, 0, 1 << 16)
    assert np.max(np.abs(g - eo)) < 1e-10
    closed = O.lib().orc_c1_closed(1e4, 0.118, 0.9, O.CONVERT2MB)
    assert abs(info.sigma - closed) < 3.0 * info.sigma_err


def test_unweight_reference_kat(ctx, O):
    """Reference kernel 10 (seed 0xC0B0, 2^24 candidates, |cos| < 0.9, wMax scan of 2e6 points):
    1,680,778 accepted (cuPythia/10_cub_compaction/README.md:12); indices equal the oracle's."""
    L = O.lib()
    n = 1 << 24
    wmax_ref = L.orc_c1_wmax_scan(2000000, 1e4, 0.118, 0.9)
    wmax_mb = wmax_ref * O.CONVERT2MB * (2 * 0.9) * (1e4 / 2)
    run = mi.make_run(mi.physics_default(0), n, precision=mi.MI_PREC_FP64, rng=mi.MI_RNG_SPLITMIX_COMPAT,
                      key=0xC0B0)
    ev = ctx.unweight_compact(run, wmax=wmax_mb)
    assert ev.n == 1680778
    flags = np.zeros(n, dtype=np.uint8)
    import ctypes as C
    L.orc_c1_unweight_compat(0xC0B0, n, 1e4, 0.118, 0.9, wmax_ref, flags.ctypes.data_as(C.POINTER(C.c_uint8)), None)
    assert np.array_equal(ev.index.cpu().numpy(), np.nonzero(flags)[0])
    assert torch.all(ev.pid == 111)


@pytest.mark.parametrize("vegas", [False, True])
def test_unweight_hadronic_bit_exact(ctx, O, vegas):
    cfg = O.hadcfg_default()
    n = 1 << 20
    edges = None
    sampler = mi.MI_SAMPLER_FLAT
    if vegas:
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(mi.make_run(mi.physics_default(1), 1 << 18, key=1), 4, 1.5)
        edges = ctx.vegas_get_grid()
        sampler = mi.MI_SAMPLER_VEGAS
    run = mi.make_run(mi.physics_default(1), n, key=42, counter0=1000, sampler=sampler)
    ev = ctx.unweight_compact(run)          # exact maximum over the candidates
    cnt, flags, pid, w = O.unweight32(cfg, 42, 1000, n, ev.wmax, edges)
    assert ev.wmax == float(w.max())
    assert ev.n == cnt and cnt > 0
    idx = np.nonzero(flags)[0]
    assert np.array_equal(ev.index.cpu().numpy(), idx + 1000)
    assert np.array_equal(ev.pid.cpu().numpy(), pid[idx])
    assert np.array_equal(ev.weight.cpu().numpy(), np.maximum(w[idx].astype(np.float64), ev.wmax))
    w32, k32 = O.had_points32(cfg, 42, 1000, n, edges)
    assert np.array_equal(ev.sHat.cpu().numpy().astype(np.float32), k32[idx, 0])
    assert np.array_equal(ev.x2.cpu().numpy().astype(np.float32), k32[idx, 3])
    est = ev.estimate
    assert est.nAccepted() == cnt and est.nTried() == n
    # unweighted sigma = eta * wMax agrees with the weighted estimate
    sig_unw = cnt / n * ev.wmax
    assert abs(sig_unw - est.sigmaGen()) < 5 * est.sigmaGen() * math.sqrt(1.0 / cnt)


def test_unweight_capacity_and_given_wmax(ctx, O):
    cfg = O.hadcfg_default()
    n = 1 << 18
    run = mi.make_run(mi.physics_default(1), n, key=5)
    full = ctx.unweight_compact(run, wmax=2e-3)
    cnt, flags, _, _ = O.unweight32(cfg, 5, 0, n, 2e-3)
    assert full.n == cnt
    with pytest.raises(mi.MIError) as ei:
        ctx.unweight_compact(run, wmax=2e-3, capacity=cnt // 2)
    assert ei.value.status == -6


def test_unweight_c3_workload(ctx, O):
    """2^26 candidates through the adapted 3-D grid: count == sum of flags, order preserved."""
    ctx.vegas_init(3, 64)
    ctx.vegas_adapt(mi.make_run(mi.physics_default(1), 1 << 24, key=42), 12, 1.5)
    run = mi.make_run(mi.physics_default(1), 1 << 26, key=42, counter0=1 << 32, sampler=mi.MI_SAMPLER_VEGAS)
    ev = ctx.unweight_compact(run)
    idx = ev.index
    assert ev.n > 0 and bool(torch.all(idx[1:] > idx[:-1]))
    assert ev.estimate.nAccepted() == ev.n
    sig_unw = ev.n / (1 << 26) * ev.wmax
    assert abs(sig_unw - ev.estimate.sigmaGen()) < 5 * ev.estimate.sigmaGen() / math.sqrt(ev.n)
    print(f"C3 unweighting: {ev.n} of 2^26 accepted (eta {ev.n / (1 << 26):.4f}), {ev.estimate.kernel_ms:.2f} ms")


@pytest.mark.parametrize("exact_max", [False, True])
def test_unweight_chunked_equals_split_calls(ctx, exact_max):
    """Above 2^27 candidates the weights are stored chunk by chunk: the events of one call over
    2^27 + 2^22 candidates equal, array for array, those of two calls over [0, 2^27) and
    [2^27, 2^27 + 2^22) at the same wMax; with the exact maximum, wMax is the maximum of the two
    calls' maxima and the weighted estimate folds both chunks."""
    phys = mi.physics_default(1)
    n1, n2 = 1 << 27, 1 << 22
    wm = 0.0 if exact_max else 2.7
    cap = 1 << 23
    full = ctx.unweight_compact(mi.make_run(phys, n1 + n2, key=8, counter0=3), wmax=wm, capacity=cap)
    w_used = full.wmax
    a = ctx.unweight_compact(mi.make_run(phys, n1, key=8, counter0=3), wmax=w_used, capacity=cap)
    b = ctx.unweight_compact(mi.make_run(phys, n2, key=8, counter0=3 + n1), wmax=w_used, capacity=cap)
    if exact_max:
        ea = ctx.unweight_compact(mi.make_run(phys, n1, key=8, counter0=3), capacity=1 << 12, wmax=1e30)
        eb = ctx.unweight_compact(mi.make_run(phys, n2, key=8, counter0=3 + n1), capacity=1 << 12, wmax=1e30)
        assert w_used == max(ea.estimate.wMax(), eb.estimate.wMax())
        assert full.estimate.sumW() == pytest.approx(ea.estimate.sumW() + eb.estimate.sumW(), rel=1e-13)
    assert full.n == a.n + b.n and b.n > 0
    for name in ("index", "sHat", "tHat", "x1", "x2", "weight", "pid"):
        assert torch.equal(getattr(full, name), torch.cat([getattr(a, name), getattr(b, name)])), name
    assert full.estimate.nTried() == n1 + n2 and full.estimate.nAccepted() == full.n


def test_unweight_fp64_matches_long_double_decisions(ctx, O):
    """FP64 unweighting (stored FP64 weights): the accepted indices are the candidates with
    u * wMax < w for the long-double Pythia-form weights w and the accept uniforms of Philox
    stream 1, except where w and u * wMax agree to 1e-10 (the FP64 vs long-double tolerance)."""
    phys = mi.physics_default(1, proc_mask=0x3FF, pTHatMin=10.0)
    n, wmax = 1 << 16, 5.0
    ev = ctx.unweight_compact(mi.make_run(phys, n, precision=mi.MI_PREC_FP64, key=31, counter0=77), wmax=wmax)
    wld, _ = O.had_points_ld(O.hadcfg_default(procMask=0x3FF, pTHatMin=10.0), 31, 77, n)
    words = O.philox_fill(31, 77, n, 1).reshape(n, 4).astype(np.uint64)
    u = (((words[:, 1] << np.uint64(32)) | words[:, 0]) >> np.uint64(11)).astype(np.float64) / 2.0 ** 53
    acc = u * wmax < wld[:, 0]
    got = np.zeros(n, dtype=bool)
    got[ev.index.cpu().numpy() - 77] = True
    close = np.abs(u * wmax - wld[:, 0]) <= 1e-10 * np.maximum(wld[:, 0], 1e-300)
    assert np.array_equal(got[~close], acc[~close])
    assert ev.n == int(got.sum()) and ev.n > 100
    # weights carried: max(w, wMax) of the FP64 weights
    w_ev = ev.weight.cpu().numpy()
    ref = np.maximum(wld[ev.index.cpu().numpy() - 77, 0], wmax)
    assert np.allclose(w_ev, ref, rtol=1e-11, atol=0)
