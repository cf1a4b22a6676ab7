"""Failure detection of non-finite weights through the C ABI.  A corrupt PDF grid -- gluon
densities of 1e30 on the nodes with 1e-3 < x < 0.1, inside the FP32 range -- makes the FP32
luminosities overflow (1e60), so the FP32 weights are inf, and every FP32 entry point that sums
them returns MI_ERR_NUMERIC with a message instead of a number: integration (also without
allocation), VEGAS adaptation, unweighting, the flow's training and the time-to-precision driver.
In FP64 the same weights are finite (1e60 fits) and the integration returns them.  As in Pythia
(PhaseSpace.cc:1135 clamps only negative weights), a NaN weight is not zeroed but propagates into
sigma.  A grid value beyond the FP32 range is refused when the grid is loaded (MI_ERR_RANGE), and
a non-finite K factor up front (MI_ERR_ARG).  The context stays usable: with a sound grid loaded
again the same calls succeed."""
import ctypes as C
import math
import os
import sys

import pytest

import cupythia_amd as mi
from cupythia_amd import _abi as A

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pdfgrid_util as PG  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grids(tmp_path_factory, O):
    d = tmp_path_factory.mktemp("numeric")
    good, bad = str(d / "good.dat"), str(d / "bad.dat")
    xs, qs = PG.nnpdf_like_x_grid(30, 20, 1e-8), PG.q_grid(1.0, 1.0e4, 20)
    toy = PG.toy_fn(O)
    PG.write_lhagrid1(good, xs, [qs], toy)
    PG.write_lhagrid1(bad, xs, [qs], lambda i, x, q: 1e30 if (i == 21 and 1e-3 < x < 0.1) else toy(i, x, q))
    huge = str(d / "huge.dat")
    PG.write_lhagrid1(huge, xs, [qs], lambda i, x, q: 1e300 if (i == 21 and 1e-3 < x < 0.1) else toy(i, x, q))
    return good, bad, huge


def _status(fn):
    with pytest.raises(mi.MIError) as e:
        fn()
    return e.value.status, str(e.value)


def test_nonfinite_weights_give_numeric_errors(grids):
    good, bad, huge = grids
    with mi.Context(0) as ctx:
        phys = mi.physics_default(1, pdf_source=mi.MI_PDF_GRID)
        run = mi.make_run(phys, 1 << 18, key=3)
        st, msg = _status(lambda: ctx.pdf_load(huge))
        assert st == A.MI_ERR_RANGE and "FP32" in msg
        ctx.pdf_load(bad)
        st, msg = _status(lambda: ctx.integrate(run))
        assert st == A.MI_ERR_NUMERIC and "non-finite" in msg
        e64 = ctx.integrate(mi.make_run(phys, 1 << 18, key=3, precision=mi.MI_PREC_FP64))
        assert math.isfinite(e64.sigmaGen()) and e64.sigmaGen() > 1e40
        est = A.MiSigmaEst()
        ctx.sync_stream()
        assert ctx.lib.mi_integrate_noalloc(ctx.h, C.byref(run), C.byref(est)) == A.MI_ERR_NUMERIC
        assert _status(lambda: ctx.unweight_compact(run))[0] == A.MI_ERR_NUMERIC
        assert _status(lambda: ctx.unweight_compact(run, wmax=1.0))[0] == A.MI_ERR_NUMERIC
        assert _status(lambda: ctx.adjoint(run))[0] == A.MI_ERR_NUMERIC
        ctx.vegas_init(3, 16)
        # one iteration: the refinement from non-finite sums leaves a grid that must not be sampled
        st, msg = _status(lambda: ctx.vegas_adapt(mi.make_run(phys, 1 << 16, key=3), 1))
        assert st == A.MI_ERR_NUMERIC and "iteration 0" in msg
        # the flow on a sound grid, trained on the corrupt densities
        ctx.pdf_load(good)
        ctx.vegas_init(3, 16)
        ctx.vegas_adapt(mi.make_run(phys, 1 << 16, key=3), 2)
        ctx.flow_init(1)
        ctx.pdf_load(bad)
        # one step: its batch is drawn from the sound identity flow; the update from non-finite
        # weights leaves parameters that must not be sampled
        assert _status(lambda: ctx.flow_train(mi.make_run(phys, 1 << 16, key=4), 1, 1 << 14))[0] == A.MI_ERR_NUMERIC
        assert _status(lambda: ctx.solve(mi.make_run(phys, 1, key=5), 1e-2, mi.MI_SAMPLER_VEGAS))[0] == A.MI_ERR_NUMERIC
        # the context recovers with a sound grid
        ctx.pdf_load(good)
        e = ctx.integrate(run)
        assert math.isfinite(e.sigmaGen()) and e.sigmaGen() > 0
        ctx.vegas_init(3, 16)
        info = ctx.vegas_adapt(mi.make_run(phys, 1 << 16, key=3), 2)
        assert math.isfinite(info.sigma)
        ev = ctx.unweight_compact(mi.make_run(phys, 1 << 16, key=7))
        assert ev.n > 0


def test_nonfinite_kfactor_is_refused():
    with mi.Context(0) as ctx:
        for k in (float("nan"), float("inf"), -1.0):
            st, msg = _status(lambda: ctx.integrate(mi.make_run(mi.physics_default(1, kfactor=k), 1 << 10)))
            assert st == A.MI_ERR_ARG and "kfactor" in msg
