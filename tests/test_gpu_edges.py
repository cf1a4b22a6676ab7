"""Edge cases through the C ABI: run sizes outside 1..2^40 are refused (n_points = 0 -> MI_ERR_ARG,
above 2^40 -> MI_ERR_RANGE); ragged sizes (1, 3, 63, 257, 100,003 points: partial waves, blocks
and steps) give the per-code maxima and non-zero counts of the same points' weights exactly and
their sums to the FP64 summation order; unweighting of ragged candidate counts matches the
oracle's accept decisions bit for bit; empty event arrays pass through the event record, the
reweighting and the LHE writer."""
import numpy as np
import pytest
import torch

import cupythia_amd as mi
from cupythia_amd import _abi

pytestmark = pytest.mark.gpu


def test_run_size_limits(ctx):
    phys = mi.physics_default(1)
    for n, status in ((0, _abi.MI_ERR_ARG), ((1 << 40) + 1, _abi.MI_ERR_RANGE)):
        with pytest.raises(mi.MIError) as e:
            ctx.integrate(mi.make_run(phys, n))
        assert e.value.status == status
        with pytest.raises(mi.MIError) as e:
            ctx.unweight_compact(mi.make_run(phys, n), capacity=16)
        assert e.value.status == status


@pytest.mark.parametrize("sampler", [mi.MI_SAMPLER_FLAT, mi.MI_SAMPLER_VEGAS])
@pytest.mark.parametrize("n", [1, 3, 63, 257, 100003])
def test_ragged_integration_matches_points(ctx, n, sampler):
    phys = mi.physics_default(1)
    if sampler == mi.MI_SAMPLER_VEGAS:
        ctx.vegas_init(3, 64)
        ctx.vegas_adapt(mi.make_run(phys, 1 << 16, key=2), 2, 1.5)
    run = mi.make_run(phys, n, key=8, counter0=12345, sampler=sampler)
    est = ctx.integrate(run)
    w, _ = ctx.points(run, with_kin=False)
    w = w.cpu().numpy()
    assert est.raw.nTry == n and est.raw.nNonZero == int(np.count_nonzero(w[:, 0] > 0))
    for c in range(6):
        assert est.raw.wMax[c] == w[:, c].max(), c
        s = w[:, c].sum()
        assert abs(est.raw.sumW[c] - s) <= 1e-13 * s, (c, est.raw.sumW[c], s)


@pytest.mark.parametrize("n", [1, 257, 12345])
def test_ragged_unweighting_bit_exact(ctx, O, n):
    phys = mi.physics_default(1)
    ev = ctx.unweight_compact(mi.make_run(phys, n, key=9, counter0=77))
    cnt, flags, pid, w = O.unweight32(O.hadcfg_default(), 9, 77, n, ev.wmax, None)
    assert ev.wmax == float(w.max()) and ev.n == cnt
    idx = np.nonzero(flags)[0]
    assert np.array_equal(ev.index.cpu().numpy(), idx + 77)
    assert np.array_equal(ev.pid.cpu().numpy(), pid[idx])


def test_empty_event_arrays(ctx, tmp_path):
    phys = mi.physics_default(1)
    dev = torch.device("cuda", 0)
    f64 = dict(dtype=torch.float64, device=dev)
    ev = mi.Events(*[torch.empty(0, **f64) for _ in range(5)], torch.empty(0, dtype=torch.int32, device=dev),
                   torch.empty(0, dtype=torch.int64, device=dev), 0, 0.0)
    rec = ctx.build_events(phys, ev, key=1)
    assert rec.px.shape == (0, 4) and rec.n_part.numel() == 0
    rw = ctx.reweight(phys, ev.sHat, ev.tHat, ev.x1, ev.x2, ev.pid, ev.weight, [(1.0, 1.0, 0), (2.0, 1.0, 0)])
    assert tuple(rw.shape) == (0, 2)
    path = str(tmp_path / "empty.lhe")
    ctx.write_lhe(path, phys, ev, key=1)
    text = open(path).read()
    assert text.startswith("<LesHouchesEvents") and "<event>" not in text and text.rstrip().endswith(
        "</LesHouchesEvents>")
