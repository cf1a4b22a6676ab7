"""mi_integrate_noalloc (block records in the library's static device scratch, no allocation in
the call) against mi_integrate through the C ABI: the same estimate bit for bit (same kernels,
geometry and reduction), on two calls in a row, the second with a smaller input, for the flat and
VEGAS maps of the hadronic integrand (FP32 and FP64) and the partonic one; the flow sampler and a
NULL output are refused.  Runs alone in a fresh process and after the other tests."""
import ctypes as C

import pytest

import cupythia_amd as mi
from cupythia_amd import _abi as A

pytestmark = pytest.mark.gpu


def _noalloc(ctx, run):
    est = A.MiSigmaEst()
    ctx.sync_stream()
    rc = ctx.lib.mi_integrate_noalloc(ctx.h, C.byref(run), C.byref(est))
    ctx._check(rc, "mi_integrate_noalloc")
    return est


def _same(a, b):
    n = A.MI_MAX_PROC + 1
    assert list(a.sumW)[:n] == list(b.sumW)[:n]
    assert list(a.sumW2)[:n] == list(b.sumW2)[:n]
    assert list(a.wMax)[:n] == list(b.wMax)[:n]
    assert list(a.sigmaGen)[:n] == list(b.sigmaGen)[:n] and list(a.sigmaErr)[:n] == list(b.sigmaErr)[:n]
    assert a.nTry == b.nTry and a.nNonZero == b.nNonZero


@pytest.mark.parametrize("case", ["flat_fp32", "flat_fp64", "vegas_fp32", "heavy", "partonic"])
def test_noalloc_equals_integrate(case):
    with mi.Context(0) as ctx:
        sampler = mi.MI_SAMPLER_FLAT
        if case == "partonic":
            phys, prec = mi.physics_default(0), mi.MI_PREC_FP64
        elif case == "heavy":
            phys, prec = mi.physics_default(1, proc_mask=0x3FF), mi.MI_PREC_FP32
        else:
            phys = mi.physics_default(1)
            prec = mi.MI_PREC_FP64 if case == "flat_fp64" else mi.MI_PREC_FP32
        if case == "vegas_fp32":
            ctx.vegas_init(3, 64)
            ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 4, 1.5)
            sampler = mi.MI_SAMPLER_VEGAS
        for n, c0 in ((1 << 24, 0), ((1 << 20) + 12345, 1 << 30)):     # second call smaller, ragged
            run = mi.make_run(phys, n, precision=prec, sampler=sampler, key=42, counter0=c0)
            ref = ctx.integrate(run).raw
            got = _noalloc(ctx, run)
            print(f"{case} n={n}: sigma {got.sigmaGen[0]:.8e} (integrate {ref.sigmaGen[0]:.8e})")
            _same(got, ref)


def test_noalloc_refusals():
    with mi.Context(0) as ctx:
        phys = mi.physics_default(1)
        ctx.vegas_init(3, 16)
        ctx.flow_init(1)
        with pytest.raises(mi.MIError) as e:
            _noalloc(ctx, mi.make_run(phys, 1 << 16, sampler=mi.MI_SAMPLER_FLOW))
        assert e.value.status == A.MI_ERR_ARG and "flow" in str(e.value)
        rc = ctx.lib.mi_integrate_noalloc(ctx.h, C.byref(mi.make_run(phys, 1 << 16)), None)
        assert rc == A.MI_ERR_ARG
        with pytest.raises(mi.MIError) as e:
            _noalloc(ctx, mi.make_run(phys, (1 << 40) + 1))
        assert e.value.status == A.MI_ERR_RANGE
