"""Pins of the oracle's VEGAS refinement (oracle/oracle_vegas_refine.c, G. P. Lepage,
J. Comput. Phys. 27 (1978) 192), CPU only.  The device refinement is compared with this function
in tests/test_gpu_vegas_unweight.py; here the function itself is fixed by what the mathematics
fixes, independently of its code:

(1) alpha = 0 damps every bin's weight to 1, and a histogram whose smoothed values are equal gives
    equal weights for any alpha: both leave any grid -- here random, non-uniform ones -- unchanged;
(2) a 4-bin example worked by hand (below);
(3) linearisation: bin widths (1 + eps (-1)^i) / N under a flat integrand (per-bin sum of w^2
    proportional to width^2) grow, after one refinement, to (1 + g eps (-1)^i) / N in the interior,
    g = 1 + (2/3) kappa, kappa = d ln m / d ln r = alpha (1/(-ln r) - r/(1 - r)) at r = 1/N: the
    alternating sign turns the 3-point average into -1/3 of the perturbation (widths enter the
    w^2 sums squared, hence 2 eps), and the re-partition moves each edge by the mass excess of the
    bin it lies in;
(4) separable integrands f(x) = (1 + 3x), e^{2y}, (1 + z)^2 in the three dimensions, with the
    expected per-bin sums width * int_bin f^2: iterated refinement converges to the closed-form
    grids of equal int f (the variance-optimal density is proportional to |f|), each dimension
    from its own histogram;
(5) one iteration of the FP32-mirror adaptation equals the refinement of the histogram built
    here from the bit-exact per-point weights and the points' bins;
(6) plausible mistakes planted in copies of the source (no smoothing, damping as a power of r,
    the damping exponent inverted, shifted edge targets, the wrong per-dimension stride, the end
    bins smoothed over three) each fail at least one of (1)-(4).

Reference counterpart: the 1-D refinement cuPythia/11_vegas_importance/vegas.cu:69-85 (first
moment, 3-point smoothing, no damping), restated and pinned by its KAT in test_reference_kats.py.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
REFINE_SRC = os.path.join(ORACLE_DIR, "oracle_vegas_refine.c")


def _refine_from(lib):
    fn = lib.orc_vegas_refine
    fn.argtypes = [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    fn.restype = None

    def refine(edges, hist, alpha):
        e = np.ascontiguousarray(edges, dtype=np.float64).copy()
        h = np.ascontiguousarray(hist, dtype=np.float64)
        fn(e.shape[0], e.shape[1] - 1, alpha, h.ctypes.data_as(C.POINTER(C.c_double)),
           e.ctypes.data_as(C.POINTER(C.c_double)))
        return e
    return refine


def _random_grid(rng, ndim, nbins):
    w = rng.uniform(0.2, 3.0, (ndim, nbins))
    e = np.concatenate([np.zeros((ndim, 1)), np.cumsum(w, axis=1) / w.sum(axis=1, keepdims=True)], axis=1)
    e[:, -1] = 1.0
    return e


# ------------------------------------------------------------------------------ the pins
def pin_alpha0(refine):
    rng = np.random.default_rng(3)
    for nbins in (4, 17, 64):
        e = _random_grid(rng, 3, nbins)
        h = rng.uniform(0.01, 5.0, (3, nbins))
        out = refine(e, h, 0.0)
        assert np.max(np.abs(out - e)) < 4e-15, nbins


def pin_flat_histogram(refine):
    rng = np.random.default_rng(4)
    for alpha in (0.5, 1.5, 2.0):
        e = _random_grid(rng, 3, 64)
        out = refine(e, np.full((3, 64), 2.5), alpha)
        assert np.max(np.abs(out - e)) < 4e-15, alpha


# Worked by hand.  Edges (0, 0.1, 0.3, 0.6, 1), per-bin sums of w^2 (1, 2, 4, 1), alpha = 1.5.
#   smoothing: (1+2)/2 = 3/2, (1+2+4)/3 = 7/3, (2+4+1)/3 = 7/3, (4+1)/2 = 5/2; total 26/3;
#   r = (9/52, 7/26, 7/26, 15/52);
#   m = ((1 - r)/(-ln r))^1.5:  r = 9/52:  (43/52)/ln(52/9)  = 0.8269231/1.7540191 = 0.4714446,
#                                          ^1.5 = 0.3237026;
#                               r = 7/26:  (19/26)/ln(26/7)  = 0.7307692/1.3121864 = 0.5569131,
#                                          ^1.5 = 0.4156015 (twice);
#                               r = 15/52: (37/52)/ln(52/15) = 0.7115385/1.2432338 = 0.5723290,
#                                          ^1.5 = 0.4330016;
#   sum m = 1.5879071, one share = 0.3969768;
#   edge 1: share 1 lies past bin 0 (0.3237026): (0.3969768 - 0.3237026)/0.4156015 = 0.1763086 of
#           bin 1 -> 0.1 + 0.1763086 * 0.2 = 0.1352617;
#   edge 2: 2 shares = 0.7939535, cumulative after bins 0, 1 = 0.7393041: 0.0546494/0.4156015 =
#           0.1314949 of bin 2 -> 0.3 + 0.1314949 * 0.3 = 0.3394485;
#   edge 3: 3 shares = 1.1909303, cumulative after bins 0..2 = 1.1549056: 0.0360247/0.4330016 =
#           0.0831978 of bin 3 -> 0.6 + 0.0831978 * 0.4 = 0.6332791.
# (The same numbers to 17 digits, from 30-digit arithmetic of the steps above, are asserted.)
HAND_EDGES = [0.0, 0.13526175525785881, 0.33944847836481815, 0.63327912689495855, 1.0]


def pin_hand_example(refine):
    out = refine(np.array([[0.0, 0.1, 0.3, 0.6, 1.0]]), np.array([[1.0, 2.0, 4.0, 1.0]]), 1.5)
    assert np.max(np.abs(out[0] - HAND_EDGES)) < 1e-12, out[0]
    # the 7-digit hand values
    assert np.max(np.abs(out[0] - [0.0, 0.1352617, 0.3394485, 0.6332791, 1.0])) < 1e-7


def pin_linearisation(refine):
    eps = 1e-6
    for nbins in (16, 64, 256):
        sig = (-1.0) ** np.arange(nbins)
        w = (1.0 + eps * sig) / nbins
        e = np.concatenate([[0.0], np.cumsum(w)])[None, :]
        e[0, -1] = 1.0
        for alpha in (0.5, 1.5, 2.0):
            out = refine(e, (w ** 2)[None, :], alpha)
            amp = (nbins * np.diff(out[0]) - 1.0) * sig / eps
            r = 1.0 / nbins
            g = 1.0 + (2.0 / 3.0) * alpha * (1.0 / -math.log(r) - r / (1.0 - r))
            mid = amp[nbins // 4: 3 * nbins // 4]
            # first-order prediction; the O(eps) remainder is ~1e-6 of g
            assert np.max(np.abs(mid - g)) < 1e-5 * g, (nbins, alpha, mid.min(), mid.max(), g)


# f_d and the closed forms: F2[d](x) = int_0^x f_d^2, INV[d](u) = the edge where int_0^x f_d = u int_0^1 f_d
F2 = (lambda x: (1 + 3 * x) ** 3 / 9.0, lambda y: np.exp(4 * y) / 4.0, lambda z: (1 + z) ** 5 / 5.0)
INV = (lambda u: (-1.0 + np.sqrt(1.0 + 15.0 * u)) / 3.0,
       lambda u: np.log1p(u * (math.e ** 2 - 1.0)) / 2.0,
       lambda u: np.cbrt(1.0 + 7.0 * u) - 1.0)


def pin_separable(refine):
    nbins = 64
    e = np.tile(np.linspace(0.0, 1.0, nbins + 1), (3, 1))
    target = np.stack([INV[d](np.arange(nbins + 1) / nbins) for d in range(3)])
    errs = []
    for _ in range(8):
        a, b = e[:, :-1], e[:, 1:]
        h = np.stack([(b[d] - a[d]) * (F2[d](b[d]) - F2[d](a[d])) for d in range(3)])
        e = refine(e, h, 1.5)
        errs.append(np.max(np.abs(e - target), axis=1))
    errs = np.array(errs)
    # geometric approach (about 0.35 per iteration), then the O(1/N^2) floor of 64 bins
    assert (errs[1:5] < 0.45 * errs[:4]).all(), errs
    assert (errs[-1] < 1e-3).all(), errs[-1]
    assert np.all(np.diff(e, axis=1) > 0)


PINS = (pin_alpha0, pin_flat_histogram, pin_hand_example, pin_linearisation, pin_separable)


# ------------------------------------------------------------------------------ the oracle itself
@pytest.mark.parametrize("pin", PINS, ids=[p.__name__ for p in PINS])
def test_refinement_pins(O, pin):
    pin(_refine_from(O.lib()))


def test_adaptation_is_refinement_of_the_sampled_histogram(O):
    """One FP32-mirror adaptation iteration (orc32_vegas_adapt) from the uniform grid == the
    refinement of the per-bin sums of w^2 built here from the bit-exact per-point weights
    (had_points32) and each point's bin (the top bits of its 24-bit uniforms)."""
    cfg = O.hadcfg_default()
    nbins, n, key, c0 = 16, 1 << 14, 42, 5
    e_adapt, sums = O.vegas_adapt32(cfg, nbins, 1.5, 1, key, c0, n)
    w, _ = O.had_points32(cfg, key, c0, n)
    words = O.philox_fill(key, c0, n).reshape(n, 4)
    bins = ((words[:, :3] >> 8).astype(np.uint64) * nbins) >> 24
    t = w[:, 0].astype(np.float64)
    hist = np.zeros((3, nbins))
    for d in range(3):
        np.add.at(hist[d], bins[:, d].astype(np.int64), t * t)
    e0 = np.tile(np.linspace(0.0, 1.0, nbins + 1), (3, 1))
    assert np.allclose(O.vegas_refine(e0, hist, 1.5), e_adapt, rtol=0, atol=1e-13)
    assert sums[0, 0] == pytest.approx(t.sum(), rel=1e-12)
    assert sums[0, 1] == pytest.approx((t * t).sum(), rel=1e-12)


# ------------------------------------------------------------------------------ planted mistakes
MUTANTS = {
    "no smoothing": ("else s = (h[i - 1] + h[i] + h[i + 1]) / 3.0;", "else s = h[i];"),
    "damping as r^alpha": ("mi = (r < 1.0) ? pow((1.0 - r) / -log(r), alpha) : 1.0;", "mi = pow(r, alpha);"),
    "exponent inverted": ("pow((1.0 - r) / -log(r), alpha)", "pow((1.0 - r) / -log(r), 1.0 / alpha)"),
    "edge targets shifted": ("double target = step * j;", "double target = step * (j - 1) + 0.5 * step;"),
    "dimension stride": ("const double *h = hist + (size_t)d * nbins;",
                         "const double *h = hist + (size_t)d * (nbins - 1);"),
    "end bins over three": ("else if (i == 0) s = 0.5 * (h[0] + h[1]);", "else if (i == 0) s = (h[0] + h[1]) / 3.0;"),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    src = open(REFINE_SRC).read()
    d = tmp_path_factory.mktemp("vegas_mutants")
    libs = {}
    for name, (old, new) in MUTANTS.items():
        assert src.count(old) == 1, f"mutation site of '{name}' not found once in oracle_vegas_refine.c"
        path = d / (name.replace(" ", "_").replace("^", "") + ".c")
        path.write_text(src.replace(old, new))
        so = str(path)[:-2] + ".so"
        subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-I", ORACLE_DIR,
                        "-o", so, str(path), "-lm"], check=True)
        libs[name] = C.CDLL(so)
    return libs


@pytest.mark.parametrize("name", list(MUTANTS))
def test_planted_mistakes_fail_a_pin(mutant_libs, name):
    refine = _refine_from(mutant_libs[name])
    failed = []
    for pin in PINS:
        try:
            pin(refine)
        except AssertionError:
            failed.append(pin.__name__)
    assert failed, f"planted mistake '{name}' passes every pin"
