"""Process identity of the QCD 2->2 library, mirroring Pythia's SigmaProcess plugin API.

Each class carries Pythia's ``code()``, ``name()`` and ``inFlux()``
(pythia8317/include/Pythia8/SigmaQCD.h:188-392) and evaluates ``sigmaHatWrap`` -- Kfactor x
d(sigmaHat)/d(tHat) x CONVERT2MB, in mb/GeV^2 (SigmaProcess.h:516-520) -- for a batch of
phase-space points on the GPU through ``mi_sigma_hat``.  The heavy-quark pair classes
(Sigma2gg2QQbar, Sigma2qqbar2QQbar with idNew = 4, 5: codes 121..124) take the pair mass from
the invariants, m^2 = (sH + tH + uH) / 2, so they need uH.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import _abi as A
from .core import Context


class _Sigma2Process:
    _code = 0
    _name = ""
    _flux = ""

    def __init__(self, ctx: Context, nQuarkNew: int = 3, Kfactor: float = 1.0):
        self.ctx = ctx
        self.nQuarkNew = nQuarkNew
        self.Kfactor = Kfactor

    def code(self) -> int:
        return self._code

    def name(self) -> str:
        return self._name

    def inFlux(self) -> str:
        return self._flux

    def sigmaHatWrap(self, sH: torch.Tensor, tH: torch.Tensor, alpS: torch.Tensor,
                     uH: Optional[torch.Tensor] = None, id1: Optional[torch.Tensor] = None,
                     id2: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Batched sigmaHatWrap(id1, id2) for FP64 device tensors; uH defaults to -(sH + tH)."""
        return self.ctx.sigma_hat(self._code, sH, tH, alpS, uH=uH, id1=id1, id2=id2,
                                  nQuarkNew=self.nQuarkNew, kfactor=self.Kfactor)


class Sigma2gg2gg(_Sigma2Process):
    _code, _name, _flux = A.MI_GG2GG, "g g -> g g", "gg"


class Sigma2gg2qqbar(_Sigma2Process):
    _code, _name, _flux = A.MI_GG2QQBAR, "g g -> q qbar (uds)", "gg"


class Sigma2qg2qg(_Sigma2Process):
    _code, _name, _flux = A.MI_QG2QG, "q g -> q g", "qg"


class Sigma2qq2qq(_Sigma2Process):
    _code, _name, _flux = A.MI_QQ2QQ, "q q(bar)' -> q q(bar)'", "qq"


class Sigma2qqbar2gg(_Sigma2Process):
    _code, _name, _flux = A.MI_QQBAR2GG, "q qbar -> g g", "qqbarSame"


class Sigma2qqbar2qqbarNew(_Sigma2Process):
    _code, _name, _flux = A.MI_QQBAR2QQBARNEW, "q qbar -> q' qbar' (uds)", "qqbarSame"


class _Sigma2Heavy(_Sigma2Process):
    _idNew = 4

    def idNew(self) -> int:
        return self._idNew

    def sigmaHatWrap(self, sH: torch.Tensor, tH: torch.Tensor, alpS: torch.Tensor,
                     uH: Optional[torch.Tensor] = None, id1: Optional[torch.Tensor] = None,
                     id2: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Batched sigmaHatWrap for FP64 device tensors; uH is required (it carries the mass)."""
        if uH is None:
            raise ValueError(f"{self._name}: uH is required (m^2 = (sH + tH + uH) / 2)")
        return super().sigmaHatWrap(sH, tH, alpS, uH=uH, id1=id1, id2=id2)


class Sigma2gg2ccbar(_Sigma2Heavy):
    _code, _name, _flux, _idNew = A.MI_GG2CCBAR, "g g -> c cbar", "gg", 4


class Sigma2qqbar2ccbar(_Sigma2Heavy):
    _code, _name, _flux, _idNew = A.MI_QQBAR2CCBAR, "q qbar -> c cbar", "qqbarSame", 4


class Sigma2gg2bbbar(_Sigma2Heavy):
    _code, _name, _flux, _idNew = A.MI_GG2BBBAR, "g g -> b bbar", "gg", 5


class Sigma2qqbar2bbbar(_Sigma2Heavy):
    _code, _name, _flux, _idNew = A.MI_QQBAR2BBBAR, "q qbar -> b bbar", "qqbarSame", 5


PROCESSES = [Sigma2gg2gg, Sigma2gg2qqbar, Sigma2qg2qg, Sigma2qq2qq, Sigma2qqbar2gg, Sigma2qqbar2qqbarNew,
             Sigma2gg2ccbar, Sigma2qqbar2ccbar, Sigma2gg2bbbar, Sigma2qqbar2bbbar]


def process_by_code(code: int):
    for cls in PROCESSES:
        if cls._code == code:
            return cls
    raise KeyError(code)
