"""The decisions of the time-to-precision driver (cupythia_amd/csrc/mi_solve_policy.h, used by
mi_solve), CPU only: the header is HIP-free, so this test compiles it with g++ into a small
library and drives it with synthetic costs.

Scenarios (device milliseconds):
  * the point count of a final pass is 1.1 P (rel / target)^2, never below the minimum;
  * the measured 5-process case: VEGAS needs ~1e6 points at 10 ps (a final pass of ~0.01 ms)
    while one 10-step chunk of 2^16-point training batches samples >= 6.6e5 points through the
    flow at 352 ps: the flow is hopeless before training, VEGAS stays;
  * an expensive integrand (10 ns per VEGAS point, cv 1; the flow 5 ns): the flow is worth training, keeps
    training while chunks pay back against the better of the flow now and VEGAS, stops when a chunk saves
    less than it cost, and is used when its
    projected pilot + final pass is below VEGAS's final pass -- and not when it is above."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "cupythia_amd", "csrc", "mi_solve_policy.h")

SRC = r'''
#include "mi_solve_policy.h"
extern "C" {
unsigned long long final_points(double rel, unsigned long long n, double t, unsigned long long nmin)
{ return mi::solve_final_points(rel, n, t, nmin); }
double flow_rest(double cv, double ms, unsigned long long np, double t, unsigned long long nmin)
{ return mi::solve_flow_rest_ms(cv, ms, np, t, nmin); }
int hopeless(double v, double ms, int chunk, unsigned long long batch, unsigned long long np, unsigned long long nmin)
{ return mi::solve_flow_hopeless(v, ms, chunk, batch, np, nmin); }
int keep(double prev, double rest, double chunk_ms, int steps, int maxs, double v)
{ return mi::solve_keep_training(prev, rest, chunk_ms, steps, maxs, v); }
int use_flow(double rest, double v) { return mi::solve_use_flow(rest, v); }
}
'''


@pytest.fixture(scope="module")
def P(tmp_path_factory):
    d = tmp_path_factory.mktemp("policy")
    src = d / "p.cpp"
    src.write_text(SRC)
    so = str(d / "libp.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.dirname(HDR), "-o", so, str(src)],
                   check=True)
    L = C.CDLL(so)
    u64, d_, i = C.c_ulonglong, C.c_double, C.c_int
    L.final_points.argtypes, L.final_points.restype = [d_, u64, d_, u64], u64
    L.flow_rest.argtypes, L.flow_rest.restype = [d_, d_, u64, d_, u64], d_
    L.hopeless.argtypes, L.hopeless.restype = [d_, d_, i, u64, u64, u64], i
    L.keep.argtypes, L.keep.restype = [d_, d_, d_, i, i, d_], i
    L.use_flow.argtypes, L.use_flow.restype = [d_, d_], i
    return L


PILOT, NMIN = 1 << 20, 1 << 16


def test_final_points(P):
    # rel 1e-3 at 2^20 points -> target 1e-3 needs 1.1 * 2^20
    assert P.final_points(1e-3, PILOT, 1e-3, NMIN) == 1153434
    assert P.final_points(2e-3, PILOT, 1e-3, NMIN) == 4613735
    assert P.final_points(1e-5, PILOT, 1e-3, NMIN) == NMIN
    assert P.final_points(0.0, PILOT, 1e-3, NMIN) == NMIN
    # the flow's rest: pilot + final at its cv (cv 0.1 -> 1.1e4 points, below the minimum)
    assert P.flow_rest(0.1, 1e-6, PILOT, 1e-3, NMIN) == pytest.approx(1e-6 * (PILOT + NMIN))


def test_measured_library_case_falls_back_before_training(P):
    """The 5-process library on MI355X: VEGAS cv 0.96 at 10.2 ps per point, flow 352 ps."""
    n_v = P.final_points(0.96 / 1024.0, PILOT, 1e-3, NMIN)
    vegas_final = n_v * 10.2e-9
    assert vegas_final < 0.02
    assert P.hopeless(vegas_final, 352e-9, 10, 1 << 16, PILOT, NMIN) == 1


def test_expensive_integrand_trains_and_uses_the_flow(P):
    """10 ns per VEGAS point, cv 1 (final pass 1.1e6 points = 11 ms); the flow at 5 ns per
    point, chunks of 10 steps costing 3 ms, cv 1 -> 0.5 -> 0.3 -> 0.28."""
    ms_v, ms_f = 10e-6, 5e-6
    vegas_final = P.final_points(1.0 / 1024.0, PILOT, 1e-3, NMIN) * ms_v
    assert vegas_final == pytest.approx(11.0, abs=0.01)
    assert P.hopeless(vegas_final, ms_f, 10, 1 << 16, PILOT, NMIN) == 0
    rest = [P.flow_rest(cv, ms_f, PILOT, 1e-3, NMIN) for cv in (1.0, 0.5, 0.3, 0.28)]
    assert rest[0] == pytest.approx(10.743, abs=1e-3) and rest[1] == pytest.approx(6.618, abs=1e-3)
    # chunk 1 saves 4.1 ms > 3 ms, and another such chunk ends at 5.5 ms < min(6.6, 11): keep;
    # chunk 3 saves 0.07 ms: stop; no steps left: stop
    assert P.keep(rest[0], rest[1], 3.0, 10, 200, vegas_final) == 1
    assert P.keep(rest[2], rest[3], 3.0, 30, 200, vegas_final) == 0
    assert P.keep(rest[0], rest[1], 3.0, 200, 200, vegas_final) == 0
    # a chunk that saved more than it cost (4.1 > 4 ms), but another like it would end at
    # 6.618 - 4.125 + 4 = 6.49 ms, above a VEGAS final pass of 5 ms: stop
    assert P.keep(rest[0], rest[1], 4.0, 10, 200, 5.0) == 0
    assert P.keep(rest[0], rest[1], 4.0, 10, 200, 8.0) == 1
    assert P.use_flow(rest[3], vegas_final) == 1
    # the same flow on a cheaper integrand (VEGAS final 1.2 ms) is not used
    assert P.use_flow(rest[3], 1.2) == 0
