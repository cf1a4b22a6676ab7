"""ctypes binding of the CPU oracle (oracle/_build/libmcoracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this
module.  The product package ``cupythia_amd`` never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmcoracle.so")
_lib = None

CONVERT2MB = 0.389380
CONV_CUPYTHIA = 0.3893793721
NCONST32 = 42
NPROC = 10
# process code of each weight slot (bit i of procMask)
CODES = (111, 112, 113, 114, 115, 116, 121, 122, 123, 124)


class OrcAlphas(C.Structure):
    _fields_ = [("valueRef", C.c_double), ("order", C.c_int), ("nfmax", C.c_int),
                ("mc", C.c_double), ("mb", C.c_double), ("mt", C.c_double),
                ("mc2", C.c_double), ("mb2", C.c_double), ("mt2", C.c_double),
                ("L3s", C.c_double), ("L4s", C.c_double), ("L5s", C.c_double), ("L6s", C.c_double),
                ("scale2Min", C.c_double)]


class OrcToyPdf(C.Structure):
    _fields_ = [("ag", C.c_double), ("bg", C.c_double), ("av", C.c_double), ("bv", C.c_double),
                ("as_", C.c_double), ("bs", C.c_double),
                ("Ag", C.c_double), ("Auv", C.c_double), ("Adv", C.c_double), ("As", C.c_double),
                ("nQuarkIn", C.c_int)]


class OrcLhaGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nq", C.c_int), ("nqSub", C.c_int),
                ("x", C.POINTER(C.c_double)), ("lnx", C.POINTER(C.c_double)),
                ("q", C.POINTER(C.c_double)), ("lnq", C.POINTER(C.c_double)),
                ("val", C.POINTER(C.c_double)), ("nqSum", C.c_int * 8), ("qDiv", C.c_double * 8),
                ("xMin", C.c_double), ("xMax", C.c_double), ("qMin", C.c_double), ("qMax", C.c_double)]


class OrcHadCfg(C.Structure):
    _fields_ = [("eCM", C.c_double), ("pTHatMin", C.c_double), ("as_", OrcAlphas),
                ("renormFac", C.c_double), ("pdf", OrcToyPdf), ("nQuarkIn", C.c_int),
                ("nQuarkNew", C.c_int), ("procMask", C.c_int), ("kfactor", C.c_double),
                ("mQ", C.c_double * 2), ("pdfSource", C.c_int), ("factFac", C.c_double), ("grid", C.POINTER(OrcLhaGrid))]


class LhaGrid:
    """An LHAPDF6 lhagrid1 file read by the oracle (Pythia's LHAGrid1 semantics)."""

    def __init__(self, path: str):
        ptr = C.POINTER(OrcLhaGrid)()
        rc = lib().orc_lhagrid_load(path.encode(), C.byref(ptr))
        if rc != 0:
            raise ValueError(f"cannot read lhagrid1 file {path} (code {rc})")
        self.ptr = ptr
        self.g = ptr.contents

    def xf(self, x: float, Q2: float):
        """Raw interpolated xf for the 12 slots (g, d u s c b, dbar ubar sbar cbar bbar, photon)."""
        out = (C.c_double * 12)()
        lib().orc_lhagrid_xf(self.ptr, x, Q2, out)
        return list(out)

    def close(self):
        if self.ptr:
            lib().orc_lhagrid_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OrcKin(C.Structure):
    _fields_ = [("sH", C.c_double), ("tH", C.c_double), ("uH", C.c_double), ("x1", C.c_double),
                ("x2", C.c_double), ("pT2", C.c_double), ("alpS", C.c_double), ("jac", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, seconds)."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        d, i, u32, u64 = C.c_double, C.c_int, C.c_uint32, C.c_uint64
        P = C.POINTER
        L.orc_philox4x32_10.argtypes = [P(u32), P(u32), P(u32)]
        L.orc_philox_fill.argtypes = [u64, u64, u32, u64, P(u32)]
        L.orc_u53.argtypes = [u32, u32]; L.orc_u53.restype = d
        L.orc_u24.argtypes = [u32]; L.orc_u24.restype = C.c_float
        L.orc_splitmix64.argtypes = [u64]; L.orc_splitmix64.restype = u64
        L.orc_u01.argtypes = [u64]; L.orc_u01.restype = d
        L.orc_sigma_hat_pythia.argtypes = [i, d, d, d, d, i, i, i]; L.orc_sigma_hat_pythia.restype = d
        L.orc_bracket_textbook.argtypes = [i, d, d, d, i, i, i]; L.orc_bracket_textbook.restype = d
        L.orc_alphas_init.argtypes = [P(OrcAlphas), d, i, i]
        L.orc_alphas_eval.argtypes = [P(OrcAlphas), d]; L.orc_alphas_eval.restype = d
        L.orc_toypdf_init.argtypes = [P(OrcToyPdf), P(d), P(d), P(d), i]
        L.orc_toypdf_xf.argtypes = [P(OrcToyPdf), i, d]; L.orc_toypdf_xf.restype = d
        L.orc_hadcfg_default.argtypes = [P(OrcHadCfg)]
        L.orc_had_point.argtypes = [P(OrcHadCfg), P(d), P(d), P(OrcKin), i]; L.orc_had_point.restype = d
        L.orc_had_quadrature.argtypes = [P(OrcHadCfg), i, i, i, P(d)]; L.orc_had_quadrature.restype = d
        L.orc_c1_weight.argtypes = [d, d, d, d, d]; L.orc_c1_weight.restype = d
        L.orc_c1_simpson.argtypes = [i, d, d, d, d]; L.orc_c1_simpson.restype = d
        L.orc_c1_closed.argtypes = [d, d, d, d]; L.orc_c1_closed.restype = d
        L.orc_c1_mc_philox.argtypes = [u64, u64, u64, d, d, d, d, i, P(d)]
        L.orc_c1_cpumc_compat.argtypes = [u64, u64, d, d, d, d]; L.orc_c1_cpumc_compat.restype = d
        L.orc_c1_wmax_scan.argtypes = [i, d, d, d]; L.orc_c1_wmax_scan.restype = d
        L.orc_c1_unweight_compat.argtypes = [u64, u64, d, d, d, d, P(C.c_uint8), P(d)]
        L.orc_c1_unweight_compat.restype = u64
        L.orc_expf.argtypes = [C.c_float]; L.orc_expf.restype = C.c_float
        L.orc_logf.argtypes = [C.c_float]; L.orc_logf.restype = C.c_float
        L.orc32_logtab.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]; L.orc32_logtab.restype = None
        L.orc32_consts_fill.argtypes = [P(OrcHadCfg), P(C.c_float)]
        L.orc32_had_point.argtypes = [P(C.c_float), i, C.c_float, C.c_float, C.c_float,
                                      P(C.c_float), P(C.c_float)]
        L.orc32_had_point.restype = C.c_float
        _bind_extras(L)
        _lib = L
    return _lib


def _bind_extras(L):
    """Bind the batched / VEGAS / unweighting / adjoint entry points."""
    d, i, u64 = C.c_double, C.c_int, C.c_uint64
    P = C.POINTER
    f = C.c_float
    L.orc_lhagrid_load.argtypes = [C.c_char_p, C.POINTER(C.POINTER(OrcLhaGrid))]
    L.orc_lhagrid_load.restype = C.c_int
    L.orc_lhagrid_free.argtypes = [C.POINTER(OrcLhaGrid)]
    L.orc_lhagrid_free.restype = None
    L.orc_lhagrid_xf.argtypes = [C.POINTER(OrcLhaGrid), C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.orc_lhagrid_xf.restype = None
    specs = {
        "orc_had_point_ld": ([P(OrcHadCfg), P(d), P(d), P(OrcKin), i], d),
        "orc32_had_point_om": ([P(f), i, f, f, f, f, f, f, P(f), P(f)], f),
        "orc_had_points_ld": ([P(OrcHadCfg), i, P(d), u64, u64, u64, P(d), P(d)], None),
        "orc32_had_points": ([P(OrcHadCfg), i, P(d), u64, u64, u64, P(f), P(f)], None),
        "orc32_unweight": ([P(OrcHadCfg), i, P(d), u64, u64, u64, d, P(C.c_uint8), P(C.c_int32), P(f)], u64),
        "orc_vegas_refine": ([i, i, d, P(d), P(d)], None),
        "orc32_vegas_adapt": ([P(OrcHadCfg), i, d, i, u64, u64, u64, P(d), P(d)], None),
        "orc_c1_vegas_adapt": ([d, d, d, d, i, d, i, u64, u64, u64, P(d), P(d)], None),
        "orc_vegas1d_compat": ([i, i, i, u64, i, d, d, d, d, P(d)], None),
        "orc_adjoint": ([P(OrcHadCfg), u64, u64, u64, P(d)], None),
        "orc_had_mc": ([P(OrcHadCfg), u64, u64, u64, i, P(d)], None),
        "orc_sigma_hat_massive": ([i, d, d, d, d, d], d),
        "orc_sigma_tot_massive_textbook": ([i, d, d, d], d),
        "orc_build_events": ([P(OrcHadCfg), i, d, u64, u64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], None),
    }
    for name, (argt, rest) in specs.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = rest


# ------------------------------------------------------------------------- helpers
def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


EVENT_INT_FIELDS = ("pdg", "status", "col", "acol", "mo1", "mo2", "d1", "d2")


def build_events(cfg, key, index, code, sH, tH, x1=None, x2=None, partonic=False, sHat_partonic=0.0):
    """The partons of n compacted events (oracle_events.c): a dict of (n, 4) arrays for
    EVENT_INT_FIELDS, p (n, 4, 4) = (px, py, pz, e), m (n, 4), scale (n,), n_part (n,)."""
    index = np.ascontiguousarray(index, dtype=np.uint64)
    n = index.size
    code = np.ascontiguousarray(code, dtype=np.int32)
    sH = np.ascontiguousarray(sH, dtype=np.float64)
    tH = np.ascontiguousarray(tH, dtype=np.float64)
    x1 = None if x1 is None else np.ascontiguousarray(x1, dtype=np.float64)
    x2 = None if x2 is None else np.ascontiguousarray(x2, dtype=np.float64)
    ints = np.zeros((n, 4, 8), dtype=np.int32)
    mom = np.zeros((n, 4, 5))
    scale = np.zeros(n)
    npart = np.zeros(n, dtype=np.int32)
    ptr = (lambda a: None if a is None else a.ctypes.data)
    lib().orc_build_events(C.byref(cfg), int(partonic), float(sHat_partonic), key, n, ptr(index), ptr(code), ptr(sH),
                           ptr(tH), ptr(x1), ptr(x2), ptr(ints), ptr(mom), ptr(scale), ptr(npart))
    out = {f: ints[:, :, j] for j, f in enumerate(EVENT_INT_FIELDS)}
    out.update(p=mom[:, :, :4], m=mom[:, :, 4], scale=scale, n_part=npart)
    return out


def philox_fill(key: int, counter0: int, n: int, stream: int = 0) -> np.ndarray:
    out = np.zeros(4 * n, dtype=np.uint32)
    lib().orc_philox_fill(key, counter0, stream, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def philox_block(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [o[j] for j in range(4)]


def hadcfg_default(**kw) -> OrcHadCfg:
    cfg = OrcHadCfg()
    lib().orc_hadcfg_default(C.byref(cfg))
    alpha_value = kw.pop("alphaS_value", None)
    alpha_order = kw.pop("alphaS_order", None)
    alpha_nfmax = kw.pop("alphaS_nfmax", None)
    pdf_g = kw.pop("pdf_g", None)
    grid = kw.pop("grid", None)
    for k, v in kw.items():
        setattr(cfg, k, v)
    if grid is not None:
        cfg.grid = grid.ptr
        cfg.pdfSource = 1
        cfg._grid_ref = grid            # keeps the grid alive with the config
    if alpha_value is not None or alpha_order is not None or alpha_nfmax is not None:
        lib().orc_alphas_init(C.byref(cfg.as_),
                              cfg.as_.valueRef if alpha_value is None else alpha_value,
                              cfg.as_.order if alpha_order is None else alpha_order,
                              cfg.as_.nfmax if alpha_nfmax is None else alpha_nfmax)
    if pdf_g is not None or "nQuarkIn" in kw:
        g = (C.c_double * 2)(*(pdf_g if pdf_g is not None else (cfg.pdf.ag, cfg.pdf.bg)))
        v = (C.c_double * 2)(cfg.pdf.av, cfg.pdf.bv)
        s = (C.c_double * 2)(cfg.pdf.as_, cfg.pdf.bs)
        lib().orc_toypdf_init(C.byref(cfg.pdf), g, v, s, cfg.nQuarkIn)
    return cfg


def had_point(cfg: OrcHadCfg, u, literal_u: bool = False):
    uu = (C.c_double * 3)(*u)
    w = (C.c_double * NPROC)()
    k = OrcKin()
    tot = lib().orc_had_point(C.byref(cfg), uu, w, C.byref(k), int(literal_u))
    return tot, np.array(list(w)), k


def consts32(cfg: OrcHadCfg) -> np.ndarray:
    k = np.zeros(NCONST32, dtype=np.float32)
    lib().orc32_consts_fill(C.byref(cfg), k.ctypes.data_as(C.POINTER(C.c_float)))
    return k


def had_point32(k: np.ndarray, mask: int, u0: float, u1: float, u2: float):
    w = (C.c_float * NPROC)()
    kin = (C.c_float * 4)()
    tot = lib().orc32_had_point(k.ctypes.data_as(C.POINTER(C.c_float)), mask, u0, u1, u2, w, kin)
    return np.float32(tot), np.array(list(w), dtype=np.float32), np.array(list(kin), dtype=np.float32)


def had_quadrature(cfg: OrcHadCfg, n_tau=16, n_y=16, n_rho=32):
    out = (C.c_double * NPROC)()
    tot = lib().orc_had_quadrature(C.byref(cfg), n_tau, n_y, n_rho, out)
    return tot, np.array(list(out))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _edges_ptr(edges):
    if edges is None:
        return 0, None, None
    e = np.ascontiguousarray(edges, dtype=np.float64)
    return e.shape[1] - 1, e, _dp(e)


def had_points_ld(cfg, key, ctr0, n, edges=None):
    """Per-point weights (n x 11: total, then the codes of CODES) and kinematics (n x 4) from the
    long-double Pythia-form oracle."""
    nb, keep, ep = _edges_ptr(edges)
    w = np.zeros((n, NPROC + 1)); kin = np.zeros((n, 4))
    lib().orc_had_points_ld(C.byref(cfg), nb, ep, key, ctr0, n, _dp(w), _dp(kin))
    return w, kin


def had_points32(cfg, key, ctr0, n, edges=None):
    """Per-point FP32 weights (n x 11) and kinematics (n x 4), the device arithmetic bit for bit."""
    nb, keep, ep = _edges_ptr(edges)
    w = np.zeros((n, NPROC + 1), dtype=np.float32); kin = np.zeros((n, 4), dtype=np.float32)
    lib().orc32_had_points(C.byref(cfg), nb, ep, key, ctr0, n, _fp(w), _fp(kin))
    return w, kin


def unweight32(cfg, key, ctr0, n, wmax, edges=None):
    nb, keep, ep = _edges_ptr(edges)
    flags = np.zeros(n, dtype=np.uint8); pid = np.zeros(n, dtype=np.int32); w = np.zeros(n, dtype=np.float32)
    cnt = lib().orc32_unweight(C.byref(cfg), nb, ep, key, ctr0, n, wmax,
                               flags.ctypes.data_as(C.POINTER(C.c_uint8)),
                               pid.ctypes.data_as(C.POINTER(C.c_int32)), _fp(w))
    return cnt, flags, pid, w


def vegas_refine(edges, hist, alpha):
    e = np.ascontiguousarray(edges, dtype=np.float64).copy()
    h = np.ascontiguousarray(hist, dtype=np.float64)
    lib().orc_vegas_refine(e.shape[0], e.shape[1] - 1, alpha, _dp(h), _dp(e))
    return e


def vegas_adapt32(cfg, nbins, alpha, n_iter, key, ctr0, n):
    e = np.tile(np.linspace(0.0, 1.0, nbins + 1), (3, 1))
    sums = np.zeros(2 * n_iter)
    lib().orc32_vegas_adapt(C.byref(cfg), nbins, alpha, n_iter, key, ctr0, n, _dp(e), _dp(sums))
    return e, sums.reshape(n_iter, 2)


def c1_vegas_adapt(s, alpS, cMax, conv, nbins, alpha, n_iter, key, ctr0, n):
    e = np.linspace(0.0, 1.0, nbins + 1)[None, :].copy()
    sums = np.zeros(2 * n_iter)
    lib().orc_c1_vegas_adapt(s, alpS, cMax, conv, nbins, alpha, n_iter, key, ctr0, n, _dp(e), _dp(sums))
    return e, sums.reshape(n_iter, 2)
