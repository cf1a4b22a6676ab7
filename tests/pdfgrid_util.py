"""Writers of LHAPDF6 "lhagrid1" files for the PDF-grid tests (test infrastructure).

The format is the one Pythia's LHAGrid1 reads (pythia8317/src/PartonDistributions.cc:478-598):
header lines up to "---", then per Q subgrid an x line, a Q line, a flavour-id line,
nx * nq lines of xf values (x outer, Q inner) and a "---" line.
"""
import math

import numpy as np

IDS = [-5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 21]
# slot of each id in the 12-entry value vector: g, d u s c b, dbar ubar sbar cbar bbar, photon
SLOT = {21: 0, 1: 1, 2: 2, 3: 3, 4: 4, 5: 5, -1: 6, -2: 7, -3: 8, -4: 9, -5: 10, 22: 11}


def nnpdf_like_x_grid(n_log=50, n_lin=50, xmin=1e-9):
    """x nodes like the NNPDF2.3 grid: log-spaced up to 0.1, then linear to 1."""
    lo = np.exp(np.linspace(math.log(xmin), math.log(0.1), n_log, endpoint=False))
    hi = np.linspace(0.1, 1.0, n_lin)
    return np.concatenate([lo, hi])


def q_grid(qmin=1.0, qmax=1.0e4, n=50):
    return np.exp(np.linspace(math.log(qmin), math.log(qmax), n))


def write_lhagrid1(path, xs, q_subgrids, fn, ids=IDS):
    """fn(id, x, q) -> xf.  q_subgrids: list of Q arrays (consecutive subgrids share a border)."""
    with open(path, "w") as f:
        f.write("PdfType: central\nFormat: lhagrid1\n---\n")
        for qs in q_subgrids:
            f.write(" ".join(f"{x:.17e}" for x in xs) + "\n")
            f.write(" ".join(f"{q:.17e}" for q in qs) + "\n")
            f.write(" ".join(str(i) for i in ids) + "\n")
            for x in xs:
                for q in qs:
                    f.write(" ".join(f"{fn(i, x, q):.17e}" for i in ids) + "\n")
            f.write("---\n")


def cubic_fn(coef_seed=3):
    """xf(id, x, q) = a cubic polynomial in (ln x, ln q), different per flavour: the
    interpolation reproduces it exactly."""
    rng = np.random.default_rng(coef_seed)
    table = {i: rng.normal(size=(4, 4)) for i in IDS}

    def fn(i, x, q):
        lx, lq = math.log(x), math.log(q)
        c = table[i]
        return float(sum(c[a, b] * lx ** a * lq ** b for a in range(4) for b in range(4)))
    return fn


def toy_fn(O, Qdep=0.0, q0=10.0):
    """The oracle's toy PDFs (default normalisation), optionally with a mild ln Q slope
    (xf * (1 + Qdep ln(q/q0)) for sea and gluon), as a grid."""
    import ctypes as C
    cfg = O.hadcfg_default()
    p = cfg.pdf
    L = O.lib()
    L.orc_toypdf_xf.argtypes = [C.POINTER(type(p)), C.c_int, C.c_double]
    L.orc_toypdf_xf.restype = C.c_double

    def fn(i, x, q):
        if x >= 1.0:
            return 0.0
        v = L.orc_toypdf_xf(C.byref(p), i, x)
        if Qdep:
            val = {1: L.orc_toypdf_xf(C.byref(p), 1, x) - L.orc_toypdf_xf(C.byref(p), -1, x),
                   2: L.orc_toypdf_xf(C.byref(p), 2, x) - L.orc_toypdf_xf(C.byref(p), -2, x)}.get(i, 0.0)
            v = val + (v - val) * (1.0 + Qdep * math.log(q / q0))
        return v
    return fn
