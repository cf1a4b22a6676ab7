"""ctypes mirror of include/mcint.h (structs, constants, prototypes) and the library loader.

The HIP library is the product: this module loads ``cupythia_amd/_lib/libmcint.so`` and
raises if it is missing.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmcint.so")

MI_OK = 0
MI_ERR_ARG = -1
MI_ERR_HIP = -2
MI_ERR_STATE = -3
MI_ERR_RANGE = -4
MI_ERR_NUMERIC = -5
MI_ERR_CAPACITY = -6

MI_GG2GG, MI_GG2QQBAR, MI_QG2QG, MI_QQ2QQ, MI_QQBAR2GG, MI_QQBAR2QQBARNEW = 111, 112, 113, 114, 115, 116
MI_GG2CCBAR, MI_QQBAR2CCBAR, MI_GG2BBBAR, MI_QQBAR2BBBAR = 121, 122, 123, 124
MI_MAX_PROC = 10
# Pythia code of process slot i (bit i of proc_mask, column 1 + i of per-process arrays)
PROC_CODES = (111, 112, 113, 114, 115, 116, 121, 122, 123, 124)

MI_MODE_PARTONIC_1D, MI_MODE_HADRONIC_3D = 0, 1
MI_PREC_FP32, MI_PREC_FP64, MI_PREC_FP32_HW = 0, 1, 2
MI_SAMPLER_FLAT, MI_SAMPLER_VEGAS, MI_SAMPLER_FLOW = 0, 1, 2
MI_RNG_PHILOX, MI_RNG_SPLITMIX_COMPAT = 0, 1
MI_PDF_TOY, MI_PDF_GRID = 0, 1
MI_ABI_VERSION = 4
MI_PDF_SLOTS = 8

MI_FLOW_LAYERS, MI_FLOW_BINS, MI_FLOW_HIDDEN, MI_FLOW_NOUT = 6, 16, 64, 49
MI_FLOW_NPARAMS = 45222

D, I, U32, U64, I64 = C.c_double, C.c_int, C.c_uint32, C.c_uint64, C.c_int64


class MiPhysics(C.Structure):
    _fields_ = [("mode", I), ("proc_mask", I), ("sHat", D), ("cos_max", D), ("eCM", D),
                ("pTHatMin", D), ("alphaS_value", D), ("alphaS_order", I), ("alphaS_nfmax", I),
                ("renorm_fac", D), ("nQuarkIn", I), ("nQuarkNew", I), ("pdf_g", D * 2),
                ("pdf_val", D * 2), ("pdf_sea", D * 2), ("kfactor", D), ("pdf_source", I), ("fact_fac", D),
                ("mass_Q", D * 2)]


class MiRun(C.Structure):
    _fields_ = [("phys", MiPhysics), ("precision", I), ("sampler", I), ("rng", I),
                ("key", U64), ("counter0", U64), ("n_points", U64)]


NP1 = MI_MAX_PROC + 1


class MiSigmaEst(C.Structure):
    _fields_ = [("code", I * NP1), ("sigmaGen", D * NP1), ("sigmaErr", D * NP1),
                ("sumW", D * NP1), ("sumW2", D * NP1), ("wMax", D * NP1),
                ("nTry", I64), ("nSel", I64), ("nAcc", I64), ("nNonZero", I64), ("kernel_ms", D)]


class MiVegasInfo(C.Structure):
    _fields_ = [("n_iter", I), ("sigma", D), ("sigma_err", D), ("chi2_dof", D), ("last_sigma", D),
                ("last_err", D), ("last_wmax", D), ("last_eta", D), ("kernel_ms", D)]


class MiEvents(C.Structure):
    _fields_ = [("sHat", C.c_void_p), ("tHat", C.c_void_p), ("x1", C.c_void_p), ("x2", C.c_void_p),
                ("weight", C.c_void_p), ("pid", C.c_void_p), ("index", C.c_void_p), ("capacity", U64)]


class MiParticles(C.Structure):
    _fields_ = [("px", C.c_void_p), ("py", C.c_void_p), ("pz", C.c_void_p), ("e", C.c_void_p), ("m", C.c_void_p),
                ("pdg", C.c_void_p), ("status", C.c_void_p), ("col", C.c_void_p), ("acol", C.c_void_p),
                ("mo1", C.c_void_p), ("mo2", C.c_void_p), ("d1", C.c_void_p), ("d2", C.c_void_p),
                ("n_part", C.c_void_p), ("scale", C.c_void_p), ("capacity", U64), ("max_part", I),
                ("reserved", I)]


class MiWeightVar(C.Structure):
    _fields_ = [("renorm_fac", D), ("fact_fac", D), ("pdf_slot", I)]


class MiSigmaGrad(C.Structure):
    _fields_ = [("sigma", D), ("sigma_err", D), ("d_alphaS_MZ", D), ("d_alphaS_MZ_err", D),
                ("d_renorm_fac", D), ("d_renorm_fac_err", D), ("d_pdf_g_a", D), ("d_pdf_g_a_err", D),
                ("d_pdf_g_b", D), ("d_pdf_g_b_err", D), ("kernel_ms", D)]


class MiFlowInfo(C.Structure):
    _fields_ = [("steps", I), ("loss_first", D), ("loss_last", D), ("cv_vegas", D), ("cv_flow", D),
                ("kernel_ms", D)]


class MiSolveOpts(C.Structure):
    _fields_ = [("rel_target", D), ("sampler", I), ("chunk_steps", I), ("max_steps", I), ("force_flow", I),
                ("pilot_points", U64), ("probe_points", U64), ("batch", U64), ("min_points", U64), ("lr", D),
                ("flow_seed", U64)]


MI_SOLVE_VEGAS, MI_SOLVE_FALLBACK_PROBE, MI_SOLVE_FALLBACK_TRAINED, MI_SOLVE_FLOW, MI_SOLVE_FLOW_FORCED = range(5)


class MiSolveInfo(C.Structure):
    _fields_ = [("sampler_used", I), ("decision", I), ("train_steps", I), ("reserved", I), ("gpu_ms", D),
                ("pilot_ms", D), ("probe_ms", D), ("train_ms", D), ("flow_pilot_ms", D), ("final_ms", D),
                ("vegas_ms_per_point", D), ("flow_ms_per_point", D), ("cv_vegas", D), ("cv_flow", D),
                ("vegas_final_ms_proj", D), ("flow_rest_ms_proj", D), ("final_points", U64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


VP = C.c_void_p
P = C.POINTER

# name -> (argtypes, restype); every symbol include/mcint.h declares
PROTOTYPES = {
    "mi_version": ([], I),
    "mi_ctx_create": ([I, P(VP)], I),
    "mi_ctx_destroy": ([VP], I),
    "mi_last_error": ([VP], C.c_char_p),
    "mi_ctx_set_stream": ([VP, VP], I),
    "mi_physics_default": ([I, P(MiPhysics)], I),
    "mi_rng_philox_fill": ([VP, U64, U64, U32, U64, VP], I),
    "mi_sigma_hat": ([VP, I, U64, VP, VP, VP, VP, VP, VP, I, D, VP], I),
    "mi_points": ([VP, P(MiRun), VP, VP], I),
    "mi_integrate": ([VP, P(MiRun), P(MiSigmaEst)], I),
    "mi_integrate_noalloc": ([VP, P(MiRun), P(MiSigmaEst)], I),
    "mi_vegas_init": ([VP, I, I], I),
    "mi_vegas_adapt": ([VP, P(MiRun), I, D, P(MiVegasInfo)], I),
    "mi_vegas_get_grid": ([VP, P(D), P(I), P(I)], I),
    "mi_vegas_set_grid": ([VP, P(D), I, I], I),
    "mi_unweight_compact": ([VP, P(MiRun), D, P(MiEvents), P(U64), P(D), P(MiSigmaEst)], I),
    "mi_write_lhe": ([VP, C.c_char_p, P(MiPhysics), P(MiEvents), U64, U64, P(MiSigmaEst)], I),
    "mi_build_events": ([VP, P(MiPhysics), P(MiEvents), U64, U64, P(MiParticles)], I),
    "mi_reweight_scales": ([VP, P(MiPhysics), U64, VP, VP, VP, VP, I, P(D), VP], I),
    "mi_adjoint": ([VP, P(MiRun), P(MiSigmaGrad)], I),
    "mi_pdf_load": ([VP, C.c_char_p], I),
    "mi_pdf_load_slot": ([VP, I, C.c_char_p], I),
    "mi_reweight": ([VP, P(MiPhysics), U64, VP, VP, VP, VP, VP, VP, I, P(MiWeightVar), VP], I),
    "mi_reweight_events": ([VP, P(MiRun), U64, VP, VP, VP, I, P(MiWeightVar), VP], I),
    "mi_solve": ([VP, P(MiRun), P(MiSolveOpts), P(MiSigmaEst), P(MiSolveInfo)], I),
    "mi_pdf_xf": ([VP, U64, VP, VP, VP], I),
    "mi_flow_init": ([VP, U64], I),
    "mi_flow_train": ([VP, P(MiRun), I, U64, D, P(MiFlowInfo)], I),
    "mi_flow_forward": ([VP, U64, VP, VP, VP], I),
    "mi_flow_inverse": ([VP, U64, VP, VP, VP], I),
    "mi_flow_backward": ([VP, U64, VP, VP, VP], I),
    "mi_flow_get_params": ([VP, P(C.c_float)], I),
    "mi_flow_set_params": ([VP, P(C.c_float)], I),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the HIP library (once).  Raises LibraryMissing when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryMissing(
            f"{path} not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (argt, rest) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    _lib = lib
    return lib
