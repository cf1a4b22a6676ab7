"""The oracle's sigmaHat (Pythia's forms, massless codes 111-116 and the massive pairs) against
an independent implementation from the reference tree: the standalone MadGraph5 helicity-amplitude
matrix elements that Pythia ships as a plugin (pythia8317/plugins/mg5mes/procs_qcd_sm), compiled
from their sources by oracle/Makefile's `ref` target into oracle/_ref with the driver
oracle/ref_mg5.cpp.  d sigma / dt = |M|^2 / (16 pi sHat^2).  Skipped where the reference tree is
absent (the GPU box)."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MG5 = "/root/reference/pythia8317/plugins/mg5mes/procs_qcd_sm"
LIB = os.path.join(ROOT, "oracle", "_ref", "libmg5ref.so")


@pytest.fixture(scope="module")
def mg5():
    if not os.path.isdir(MG5):
        pytest.skip("reference tree not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    M = C.CDLL(LIB)
    M.mg5ref_init.argtypes = [C.c_char_p]
    M.mg5ref_init.restype = C.c_int
    M.mg5ref_me2.argtypes = [C.c_int] * 5 + [C.c_double] * 3 + [C.POINTER(C.c_double)]
    M.mg5ref_me2.restype = C.c_double
    assert M.mg5ref_init(os.path.join(MG5, "param_card_sm.dat").encode()) == 0

    def me(which, ids, s, c, a):
        m = C.c_double()
        v = M.mg5ref_me2(which, *ids, s, c, a, C.byref(m))
        assert v >= 0, (which, ids, v)
        return v / (16.0 * math.pi * s * s), m.value
    return me


def _points(seed, n=60):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(math.log(50.0 ** 2), math.log(5000.0 ** 2), n))
    c = rng.uniform(-0.98, 0.98, n)
    a = rng.uniform(0.08, 0.25, n)
    return zip(s, c, a)


# (MG5 class, flavours id1 id2 -> id3 id4, Pythia code, the ids Pythia's sigmaHat sees)
MASSLESS = [
    (0, (21, 21, 21, 21), 111, (21, 21)),
    (1, (21, 21, 2, -2), 112, (21, 21)),
    (1, (21, 21, 3, -3), 112, (21, 21)),
    (2, (2, 21, 2, 21), 113, (2, 21)),
    (2, (21, -1, 21, -1), 113, (21, -1)),
    (3, (2, -2, 21, 21), 115, (2, -2)),
    (3, (-3, 3, 21, 21), 115, (-3, 3)),
    (4, (2, 1, 2, 1), 114, (2, 1)),
    (4, (2, -1, 2, -1), 114, (2, -1)),
    (4, (2, 2, 2, 2), 114, (2, 2)),
    (4, (-1, -1, -1, -1), 114, (-1, -1)),
    (4, (2, -2, 1, -1), 116, (2, -2)),
    (4, (-3, 3, -2, 2), 116, (-3, 3)),
]


@pytest.mark.parametrize("which,ids,code,pids", MASSLESS)
def test_massless_sigma_hat_matches_mg5(O, mg5, which, ids, code, pids):
    L = O.lib()
    for s, c, a in _points(code + 7 * which):
        t = -0.5 * s * (1.0 - c)
        u = -s - t
        ref, m = mg5(which, ids, s, c, a)
        assert m == 0.0
        got = L.orc_sigma_hat_pythia(code, s, t, u, a, pids[0], pids[1], 1)
        assert got == pytest.approx(ref, rel=1e-12), (code, ids, s, c)


@pytest.mark.parametrize("ids", [(2, -2, 2, -2), (1, -1, 1, -1), (-4, 4, -4, 4)])
def test_same_flavour_qqbar_is_114_plus_116(O, mg5, ids):
    """q qbar -> q qbar of one flavour: Pythia splits MG5's full |M|^2 into the t-channel and
    interference part (114) and the s-channel part (116 per outgoing flavour)."""
    L = O.lib()
    for s, c, a in _points(sum(ids) + 99):
        t = -0.5 * s * (1.0 - c)
        u = -s - t
        ref, _ = mg5(4, ids, s, c, a)
        got = (L.orc_sigma_hat_pythia(114, s, t, u, a, ids[0], ids[1], 1)
               + L.orc_sigma_hat_pythia(116, s, t, u, a, ids[0], ids[1], 1))
        assert got == pytest.approx(ref, rel=1e-12)


@pytest.mark.parametrize("which,ids,code", [(5, (21, 21, 5, -5), 123), (6, (2, -2, 5, -5), 124),
                                            (6, (-1, 1, 5, -5), 124), (5, (21, 21, 5, -5), 121),
                                            (6, (3, -3, 5, -5), 122)])
def test_massive_sigma_hat_matches_mg5(O, mg5, which, ids, code):
    """The heavy pairs with MG5's b mass (4.7 GeV in its card): the c cbar codes share the same
    massive forms, evaluated at that mass; near threshold and far above it."""
    L = O.lib()
    rng = np.random.default_rng(code)
    for _ in range(60):
        m = 4.7
        s = 4 * m * m * math.exp(rng.uniform(0.002, 9.0))
        c = rng.uniform(-0.99, 0.99)
        a = rng.uniform(0.08, 0.25)
        ref, mq = mg5(which, ids, s, c, a)
        assert mq == pytest.approx(4.7)
        beta = math.sqrt(1 - 4 * m * m / s)
        t = m * m - 0.5 * s * (1 - beta * c)
        u = 2 * m * m - s - t
        got = L.orc_sigma_hat_massive(code, s, t, u, m * m, a)
        assert got == pytest.approx(ref, rel=1e-12), (code, s, c)
