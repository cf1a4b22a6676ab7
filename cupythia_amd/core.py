"""Context, run configuration and results over the C ABI (include/mcint.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import torch

from . import _abi as A


class MIError(RuntimeError):
    """A library call returned a negative status; ``status`` holds it."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def physics_default(which: int = 1, **overrides) -> A.MiPhysics:
    """Workload defaults: which=0 partonic gg->gg at sHat=(100 GeV)^2, |cos|<0.9, alphaS=0.118;
    which=1 the 5-process library at sqrt(S)=13.6 TeV, pT>20 GeV, toy PDFs, 1-loop alphaS."""
    p = A.MiPhysics()
    rc = A.load().mi_physics_default(which, C.byref(p))
    if rc:
        raise MIError(rc, f"mi_physics_default({which})")
    for k, v in overrides.items():
        if k in ("pdf_g", "pdf_val", "pdf_sea", "mass_Q"):
            getattr(p, k)[0], getattr(p, k)[1] = v
        else:
            setattr(p, k, v)
    return p


def make_run(phys: A.MiPhysics, n_points: int, *, precision: int = A.MI_PREC_FP32,
             sampler: int = A.MI_SAMPLER_FLAT, rng: int = A.MI_RNG_PHILOX, key: int = 42,
             counter0: int = 0) -> A.MiRun:
    r = A.MiRun()
    r.phys = phys
    r.precision = precision
    r.sampler = sampler
    r.rng = rng
    r.key = key
    r.counter0 = counter0
    r.n_points = n_points
    return r


@dataclass
class SigmaEstimate:
    """Pythia's sigma-estimate view of one call (Info::sigmaGen / sigmaErr; mb)."""
    raw: A.MiSigmaEst

    def _idx(self, code: int) -> int:
        return 0 if code == 0 else 1 + A.PROC_CODES.index(code)

    def sigmaGen(self, code: int = 0) -> float:
        return self.raw.sigmaGen[self._idx(code)]

    def sigmaErr(self, code: int = 0) -> float:
        return self.raw.sigmaErr[self._idx(code)]

    def sumW(self, code: int = 0) -> float:
        return self.raw.sumW[self._idx(code)]

    def sumW2(self, code: int = 0) -> float:
        return self.raw.sumW2[self._idx(code)]

    def wMax(self, code: int = 0) -> float:
        return self.raw.wMax[self._idx(code)]

    def nTried(self) -> int:
        return self.raw.nTry

    def nSelected(self) -> int:
        return self.raw.nSel

    def nAccepted(self) -> int:
        return self.raw.nAcc

    @property
    def kernel_ms(self) -> float:
        return self.raw.kernel_ms


@dataclass
class Events:
    """Compacted unweighted events (device tensors, length n)."""
    sHat: torch.Tensor
    tHat: torch.Tensor
    x1: torch.Tensor
    x2: torch.Tensor
    weight: torch.Tensor
    pid: torch.Tensor
    index: torch.Tensor
    n: int
    wmax: float
    estimate: Optional[SigmaEstimate] = None


@dataclass
class EventRecord:
    """Device event record of n events (mi_build_events): per-particle tensors of shape
    (n, max_part), particles 0, 1 incoming (status -21), 2, 3 outgoing (status 23)."""
    px: torch.Tensor
    py: torch.Tensor
    pz: torch.Tensor
    e: torch.Tensor
    m: torch.Tensor
    pdg: torch.Tensor
    status: torch.Tensor
    col: torch.Tensor
    acol: torch.Tensor
    mo1: torch.Tensor
    mo2: torch.Tensor
    d1: torch.Tensor
    d2: torch.Tensor
    n_part: torch.Tensor
    scale: torch.Tensor

    def p4(self) -> torch.Tensor:
        """(n, max_part, 4) four-momenta (px, py, pz, E)."""
        return torch.stack([self.px, self.py, self.pz, self.e], dim=-1)


def _ptr(t: Optional[torch.Tensor]):
    """Device pointer of a tensor already checked by _arr (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def _arr(t: Optional[torch.Tensor], name: str, dtype: torch.dtype, shape, device: int, optional: bool = False):
    """Checks a bulk array against what the C ABI reads or writes through its pointer -- dtype,
    exact shape (None entries are free), contiguity, a CUDA tensor on the context's device -- and
    raises ValueError (TypeError for a non-tensor) before anything is launched."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if len(shape) != t.dim() or any(s is not None and s != d for s, d in zip(shape, t.shape)):
        want = "(" + ", ".join("*" if s is None else str(s) for s in shape) + ")"
        raise ValueError(f"{name} must have shape {want}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (bulk arrays cross the ABI as device pointers)")
    if t.device.index != device:
        raise ValueError(f"{name} is on cuda:{t.device.index}, the context on cuda:{device}")
    return t


class Context:
    """One mi_ctx on a HIP device; launches on torch's current stream of that device."""

    def __init__(self, device: int = 0):
        self.lib = A.load()
        self.device = device
        h = C.c_void_p()
        rc = self.lib.mi_ctx_create(device, C.byref(h))
        if rc:
            raise MIError(rc, f"mi_ctx_create({device})")
        self.h = h
        self.sync_stream()

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -------------------------------------------------------------- plumbing
    def sync_stream(self):
        s = torch.cuda.current_stream(self.device)
        self.lib.mi_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream))

    def _check(self, rc: int, what: str):
        if rc:
            msg = self.lib.mi_last_error(self.h)
            raise MIError(rc, f"{what}: {msg.decode() if msg else ''}")

    @property
    def dev(self):
        return torch.device("cuda", self.device)

    # -------------------------------------------------------------- RNG
    def philox_fill(self, key: int, counter0: int, n: int, stream: int = 0) -> torch.Tensor:
        out = torch.empty(4 * n, dtype=torch.int32, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_rng_philox_fill(self.h, key, counter0, stream, n, _ptr(out)),
                    "mi_rng_philox_fill")
        return out

    # -------------------------------------------------------------- sigmaHat
    def sigma_hat(self, code: int, sH, tH, alpS, uH=None, id1=None, id2=None, nQuarkNew: int = 3,
                  kfactor: float = 1.0) -> torch.Tensor:
        f64, i32, dv = torch.float64, torch.int32, self.device
        n = _arr(sH, "sH", f64, (None,), dv).shape[0]
        _arr(tH, "tH", f64, (n,), dv)
        _arr(alpS, "alpS", f64, (n,), dv)
        _arr(uH, "uH", f64, (n,), dv, optional=True)
        _arr(id1, "id1", i32, (n,), dv, optional=True)
        _arr(id2, "id2", i32, (n,), dv, optional=True)
        out = torch.empty(n, dtype=torch.float64, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_sigma_hat(self.h, code, n, _ptr(sH), _ptr(tH), _ptr(uH), _ptr(alpS),
                                          _ptr(id1), _ptr(id2), nQuarkNew, kfactor, _ptr(out)),
                    "mi_sigma_hat")
        return out

    # -------------------------------------------------------------- integration
    def points(self, run: A.MiRun, with_kin: bool = True):
        n = int(run.n_points)
        w = torch.empty((n, A.MI_MAX_PROC + 1), dtype=torch.float64, device=self.dev)
        kin = torch.empty((n, 4), dtype=torch.float64, device=self.dev) if with_kin else None
        self.sync_stream()
        self._check(self.lib.mi_points(self.h, C.byref(run), _ptr(w), _ptr(kin)), "mi_points")
        return w, kin

    def integrate(self, run: A.MiRun) -> SigmaEstimate:
        est = A.MiSigmaEst()
        self.sync_stream()
        self._check(self.lib.mi_integrate(self.h, C.byref(run), C.byref(est)), "mi_integrate")
        return SigmaEstimate(est)

    # -------------------------------------------------------------- VEGAS
    def vegas_init(self, ndim: int, nbins: int):
        self._check(self.lib.mi_vegas_init(self.h, ndim, nbins), "mi_vegas_init")

    def vegas_adapt(self, run: A.MiRun, n_iter: int, alpha: float = 1.5) -> A.MiVegasInfo:
        info = A.MiVegasInfo()
        self.sync_stream()
        self._check(self.lib.mi_vegas_adapt(self.h, C.byref(run), n_iter, alpha, C.byref(info)),
                    "mi_vegas_adapt")
        return info

    def vegas_get_grid(self):
        import numpy as np
        nd, nb = C.c_int(), C.c_int()
        self._check(self.lib.mi_vegas_get_grid(self.h, None, C.byref(nd), C.byref(nb)), "mi_vegas_get_grid")
        e = np.zeros(nd.value * (nb.value + 1), dtype=np.float64)
        self._check(self.lib.mi_vegas_get_grid(self.h, e.ctypes.data_as(C.POINTER(C.c_double)), None, None),
                    "mi_vegas_get_grid")
        return e.reshape(nd.value, nb.value + 1)

    def vegas_set_grid(self, edges):
        import numpy as np
        e = np.ascontiguousarray(edges, dtype=np.float64)
        nd, nb1 = e.shape
        self._check(self.lib.mi_vegas_set_grid(self.h, e.ctypes.data_as(C.POINTER(C.c_double)), nd, nb1 - 1),
                    "mi_vegas_set_grid")

    # -------------------------------------------------------------- unweighting
    def unweight_compact(self, run: A.MiRun, wmax: float = 0.0, capacity: Optional[int] = None) -> Events:
        n = int(run.n_points)
        cap = n if capacity is None else int(capacity)
        d = self.dev
        f64 = dict(dtype=torch.float64, device=d)
        ev = Events(torch.empty(cap, **f64), torch.empty(cap, **f64), torch.empty(cap, **f64),
                    torch.empty(cap, **f64), torch.empty(cap, **f64),
                    torch.empty(cap, dtype=torch.int32, device=d), torch.empty(cap, dtype=torch.int64, device=d),
                    0, 0.0)
        raw = A.MiEvents(_ptr(ev.sHat), _ptr(ev.tHat), _ptr(ev.x1), _ptr(ev.x2), _ptr(ev.weight), _ptr(ev.pid),
                         _ptr(ev.index), cap)
        nacc, wused = C.c_uint64(), C.c_double()
        est = A.MiSigmaEst()
        self.sync_stream()
        rc = self.lib.mi_unweight_compact(self.h, C.byref(run), wmax, C.byref(raw), C.byref(nacc),
                                          C.byref(wused), C.byref(est))
        self._check(rc, "mi_unweight_compact")
        k = nacc.value
        ev = Events(ev.sHat[:k], ev.tHat[:k], ev.x1[:k], ev.x2[:k], ev.weight[:k], ev.pid[:k], ev.index[:k], k,
                    wused.value, SigmaEstimate(est))
        return ev

    def _events_raw(self, ev: "Events") -> A.MiEvents:
        """The mi_events view of ev after checking each array (length ev.n, the SoA dtypes)."""
        f64, dv, n = torch.float64, self.device, int(ev.n)
        for name in ("sHat", "tHat", "x1", "x2", "weight"):
            _arr(getattr(ev, name), name, f64, (n,), dv)
        _arr(ev.pid, "pid", torch.int32, (n,), dv)
        _arr(ev.index, "index", torch.int64, (n,), dv)
        return A.MiEvents(_ptr(ev.sHat), _ptr(ev.tHat), _ptr(ev.x1), _ptr(ev.x2), _ptr(ev.weight), _ptr(ev.pid),
                          _ptr(ev.index), n)

    def write_lhe(self, path: str, phys: A.MiPhysics, ev: "Events", key: int) -> None:
        """LHEF 3.0 file of compacted events (flavours, colour flow, lab momenta; pb)."""
        raw = self._events_raw(ev)
        est = ev.estimate.raw if ev.estimate is not None else A.MiSigmaEst()
        self.sync_stream()
        self._check(self.lib.mi_write_lhe(self.h, path.encode(), C.byref(phys), C.byref(raw), ev.n, key,
                                          C.byref(est)), "mi_write_lhe")

    def build_events(self, phys: A.MiPhysics, ev: "Events", key: int, max_part: int = 4) -> EventRecord:
        """The partons of compacted events on the device: flavours, colour flow, lab four-momenta
        (Philox streams 3 and 4 of each event's index under key)."""
        raw = self._events_raw(ev)
        n, d = int(ev.n), self.dev
        f64, i32 = dict(dtype=torch.float64, device=d), dict(dtype=torch.int32, device=d)
        shp = (n, max_part)
        r = EventRecord(*[torch.zeros(shp, **f64) for _ in range(5)], *[torch.zeros(shp, **i32) for _ in range(8)],
                        torch.zeros(n, **i32), torch.zeros(n, **f64))
        out = A.MiParticles(_ptr(r.px), _ptr(r.py), _ptr(r.pz), _ptr(r.e), _ptr(r.m), _ptr(r.pdg), _ptr(r.status),
                            _ptr(r.col), _ptr(r.acol), _ptr(r.mo1), _ptr(r.mo2), _ptr(r.d1), _ptr(r.d2),
                            _ptr(r.n_part), _ptr(r.scale), n, max_part, 0)
        self.sync_stream()
        self._check(self.lib.mi_build_events(self.h, C.byref(phys), C.byref(raw), n, key, C.byref(out)),
                    "mi_build_events")
        return r

    # -------------------------------------------------------------- reweighting
    @staticmethod
    def _variations(variations):
        vs = [tuple(v) for v in variations]
        if not 1 <= len(vs) <= 64:
            raise ValueError("1..64 variations")
        return (A.MiWeightVar * len(vs))(*[A.MiWeightVar(float(r), float(f), int(sl)) for (r, f, sl) in vs]), len(vs)

    def reweight_scales(self, phys: A.MiPhysics, sHat: torch.Tensor, tHat: torch.Tensor, w: torch.Tensor,
                        factors, pid: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Weights at renormalisation-scale factors (n x nvar), one pass over stored events; pid (the
        events' process codes, int32) gives the heavy pairs their scale pT^2 + m^2."""
        f64, dv = torch.float64, self.device
        n = _arr(sHat, "sHat", f64, (None,), dv).shape[0]
        _arr(tHat, "tHat", f64, (n,), dv)
        _arr(w, "w", f64, (n,), dv)
        _arr(pid, "pid", torch.int32, (n,), dv, optional=True)
        factors = [float(f) for f in factors]
        if not 1 <= len(factors) <= 64:
            raise ValueError("1..64 scale factors")
        fac = (C.c_double * len(factors))(*factors)
        out = torch.empty((n, len(factors)), dtype=torch.float64, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_reweight_scales(self.h, C.byref(phys), n, _ptr(sHat), _ptr(tHat), _ptr(pid),
                                                _ptr(w), len(factors), fac, _ptr(out)), "mi_reweight_scales")
        return out

    def reweight(self, phys: A.MiPhysics, sHat: torch.Tensor, tHat: torch.Tensor, x1: torch.Tensor,
                 x2: torch.Tensor, pid: torch.Tensor, w: torch.Tensor, variations) -> torch.Tensor:
        """Weights under (renorm_fac, fact_fac, pdf_slot) variations (n x nvar) of stored events: each
        event's own process weight recomputed from its (sHat, tHat, x1, x2) at the varied scales /
        PDFs over the run's (mi_reweight)."""
        f64, dv = torch.float64, self.device
        n = _arr(sHat, "sHat", f64, (None,), dv).shape[0]
        for name, t in (("tHat", tHat), ("x1", x1), ("x2", x2), ("w", w)):
            _arr(t, name, f64, (n,), dv)
        _arr(pid, "pid", torch.int32, (n,), dv)
        vs, nvar = self._variations(variations)
        out = torch.empty((n, nvar), dtype=torch.float64, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_reweight(self.h, C.byref(phys), n, _ptr(sHat), _ptr(tHat), _ptr(x1), _ptr(x2),
                                         _ptr(pid), _ptr(w), nvar, vs, _ptr(out)), "mi_reweight")
        return out

    def reweight_events(self, run: A.MiRun, index: torch.Tensor, pid: torch.Tensor, variations,
                        w: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Weights under (renorm_fac, fact_fac, pdf_slot) variations (n x nvar) of events regenerated
        from their counters under run (mi_reweight_events): each event's own process weight in FP64
        at each variation -- w * W_k / W_0 with w given (e.g. unweighted events), else W_k itself,
        bit for bit the per-point weight of a run at those scales."""
        dv = self.device
        n = _arr(index, "index", torch.int64, (None,), dv).shape[0]
        _arr(pid, "pid", torch.int32, (n,), dv)
        _arr(w, "w", torch.float64, (n,), dv, optional=True)
        vs, nvar = self._variations(variations)
        out = torch.empty((n, nvar), dtype=torch.float64, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_reweight_events(self.h, C.byref(run), n, _ptr(index), _ptr(pid), _ptr(w), nvar, vs,
                                                _ptr(out)), "mi_reweight_events")
        return out

    # -------------------------------------------------------------- PDF grids
    def pdf_load(self, path: str, slot: int = 0) -> None:
        """Proton PDF from an LHAPDF6 lhagrid1 file (Pythia's LHAGrid1) into a PDF slot: slot 0 is
        the grid of runs with physics_default(..., pdf_source=MI_PDF_GRID); slots 1.. hold the
        grids of reweight()'s PDF variations."""
        self._check(self.lib.mi_pdf_load_slot(self.h, slot, path.encode()), "mi_pdf_load_slot")

    def pdf_xf(self, x: torch.Tensor, Q2: torch.Tensor) -> torch.Tensor:
        """Raw interpolated xf (n x 12 FP64: g, d, u, s, c, b, dbar, ubar, sbar, cbar, bbar, 0)."""
        x = torch.as_tensor(x).to(device=self.dev, dtype=torch.float64).contiguous().reshape(-1)
        Q2 = torch.as_tensor(Q2).to(device=self.dev, dtype=torch.float64).contiguous().reshape(-1)
        if Q2.numel() != x.numel():
            raise ValueError(f"x has {x.numel()} values, Q2 {Q2.numel()}")
        out = torch.empty((x.numel(), 12), dtype=torch.float64, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_pdf_xf(self.h, x.numel(), _ptr(x), _ptr(Q2), _ptr(out)), "mi_pdf_xf")
        return out

    # -------------------------------------------------------------- time to precision
    def solve(self, run: A.MiRun, rel_target: float, sampler: int = A.MI_SAMPLER_VEGAS, **opts):
        """sigma to a relative error (mi_solve): plain VEGAS, or the flow on top of VEGAS with plain
        VEGAS as the fallback.  opts: chunk_steps, max_steps, force_flow, pilot_points,
        probe_points, batch, min_points, lr, flow_seed (0 = the library's defaults).  Returns the
        final pass's SigmaEstimate and the MiSolveInfo of the decisions and costs."""
        o = A.MiSolveOpts()
        o.rel_target = float(rel_target)
        o.sampler = int(sampler)
        for k, v in opts.items():
            if k not in dict(A.MiSolveOpts._fields_):
                raise ValueError(f"unknown solve option {k}")
            setattr(o, k, v)
        est, info = A.MiSigmaEst(), A.MiSolveInfo()
        self.sync_stream()
        self._check(self.lib.mi_solve(self.h, C.byref(run), C.byref(o), C.byref(est), C.byref(info)), "mi_solve")
        return SigmaEstimate(est), info

    # -------------------------------------------------------------- adjoint
    def adjoint(self, run: A.MiRun) -> A.MiSigmaGrad:
        g = A.MiSigmaGrad()
        self.sync_stream()
        self._check(self.lib.mi_adjoint(self.h, C.byref(run), C.byref(g)), "mi_adjoint")
        return g

    # -------------------------------------------------------------- flow
    def flow_init(self, seed: int = 1):
        self._check(self.lib.mi_flow_init(self.h, seed), "mi_flow_init")

    def flow_train(self, run: A.MiRun, steps: int, batch: int, lr: float = 1e-3) -> A.MiFlowInfo:
        info = A.MiFlowInfo()
        self.sync_stream()
        self._check(self.lib.mi_flow_train(self.h, C.byref(run), steps, batch, lr, C.byref(info)), "mi_flow_train")
        return info

    def _coords(self, y: torch.Tensor, name: str) -> int:
        """n of an (n, 3) float32 coordinate array on the context's device."""
        return _arr(y, name, torch.float32, (None, 3), self.device).shape[0]

    def flow_forward(self, y: torch.Tensor):
        n = self._coords(y, "y")
        z = torch.empty_like(y)
        ld = torch.empty(n, dtype=torch.float32, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_flow_forward(self.h, n, _ptr(y), _ptr(z), _ptr(ld)), "mi_flow_forward")
        return z, ld

    def flow_inverse(self, z: torch.Tensor):
        n = self._coords(z, "z")
        y = torch.empty_like(z)
        ld = torch.empty(n, dtype=torch.float32, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_flow_inverse(self.h, n, _ptr(z), _ptr(y), _ptr(ld)), "mi_flow_inverse")
        return y, ld

    def flow_backward(self, y: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
        n = self._coords(y, "y")
        _arr(g, "g", torch.float32, (n,), self.device)
        grad = torch.empty(A.MI_FLOW_NPARAMS, dtype=torch.float32, device=self.dev)
        self.sync_stream()
        self._check(self.lib.mi_flow_backward(self.h, n, _ptr(y), _ptr(g), _ptr(grad)), "mi_flow_backward")
        return grad

    def flow_get_params(self):
        import numpy as np
        p = np.zeros(A.MI_FLOW_NPARAMS, dtype=np.float32)
        self._check(self.lib.mi_flow_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))),
                    "mi_flow_get_params")
        return p

    def flow_set_params(self, params):
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != A.MI_FLOW_NPARAMS:
            raise ValueError(f"{A.MI_FLOW_NPARAMS} parameters expected, got {p.size}")
        self._check(self.lib.mi_flow_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))),
                    "mi_flow_set_params")
