"""Pins of the adjoint oracle (CPU): finite differences of the long-double point.

(1) fixed alphaS: sigma ~ alphaS^2, so sum dw/dalphaS = 2 sum w / alphaS and dw/dxi = 0;
(2) 1-loop alphaS: the oracle's MC derivative agrees with the derivative of the deterministic
    quadrature (central difference of orc_had_quadrature) within the MC error, for every
    parameter -- an independent route through different variables and no random numbers."""
import ctypes as C

import numpy as np


def adj(O, cfg, n, key=42, ctr0=0):
    out = (C.c_double * 9)()
    O.lib().orc_adjoint(C.byref(cfg), key, ctr0, n, out)
    return np.array(list(out))


def test_fixed_alphas_scaling(O):
    cfg = O.hadcfg_default(alphaS_order=0, alphaS_value=0.13)
    r = adj(O, cfg, 4096)
    assert abs(r[1] - 2 * r[0] / 0.13) < 1e-8 * abs(r[1])
    assert abs(r[2]) < 1e-9 * r[0]


def test_adjoint_matches_quadrature_derivative(O):
    cfg = O.hadcfg_default(procMask=0x1F)
    n = 1 << 15
    r = adj(O, cfg, n)
    mean = r[1:5] / n
    err = np.sqrt(np.maximum(r[5:9] / n - mean ** 2, 0) / n)

    def quad(c):
        return O.had_quadrature(c, 8, 8, 16)[0]

    h = 1e-4
    fd = [
        (quad(O.hadcfg_default(alphaS_value=0.118 + h)) - quad(O.hadcfg_default(alphaS_value=0.118 - h))) / (2 * h),
        (quad(O.hadcfg_default(renormFac=1 + h)) - quad(O.hadcfg_default(renormFac=1 - h))) / (2 * h),
        (quad(O.hadcfg_default(pdf_g=(-0.1 + h, 5))) - quad(O.hadcfg_default(pdf_g=(-0.1 - h, 5)))) / (2 * h),
        (quad(O.hadcfg_default(pdf_g=(-0.1, 5 + h))) - quad(O.hadcfg_default(pdf_g=(-0.1, 5 - h)))) / (2 * h),
    ]
    for q in range(4):
        assert abs(mean[q] - fd[q]) < 4 * err[q] + 1e-6 * abs(fd[q]), (q, mean[q], err[q], fd[q])
    # signs fixed by the physics: more coupling, more sigma; a larger renormalisation scale, less
    assert fd[0] > 0 and fd[1] < 0


def test_heavy_adjoint_matches_quadrature_derivative(O):
    """The heavy-quark families alone (codes 121..124, pT > 5 GeV): the same check, so the massive
    map's scale pT^2 + m^2 and the g g -> Q Qbar gluon luminosity enter both routes."""
    kw = dict(procMask=0x3C0, pTHatMin=5.0)
    n = 1 << 15
    r = adj(O, O.hadcfg_default(**kw), n)
    mean = r[1:5] / n
    err = np.sqrt(np.maximum(r[5:9] / n - mean ** 2, 0) / n)

    def quad(**over):
        return O.had_quadrature(O.hadcfg_default(**kw, **over), 8, 8, 16)[0]

    h = 1e-4
    fd = [
        (quad(alphaS_value=0.118 + h) - quad(alphaS_value=0.118 - h)) / (2 * h),
        (quad(renormFac=1 + h) - quad(renormFac=1 - h)) / (2 * h),
        (quad(pdf_g=(-0.1 + h, 5)) - quad(pdf_g=(-0.1 - h, 5))) / (2 * h),
        (quad(pdf_g=(-0.1, 5 + h)) - quad(pdf_g=(-0.1, 5 - h))) / (2 * h),
    ]
    for q in range(4):
        assert abs(mean[q] - fd[q]) < 4 * err[q] + 1e-6 * abs(fd[q]), (q, mean[q], err[q], fd[q])
    assert fd[0] > 0 and fd[1] < 0
