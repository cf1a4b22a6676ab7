/*
 * oracle_had_body.h -- Pythia-form hadronic point, instantiated by oracle_physics.c twice:
 * REAL = double (suffix _d) and REAL = long double (suffix _ld).  TEST INFRASTRUCTURE ONLY.
 *
 * The FP64 instantiation is Pythia's arithmetic as written; near the pT cut at large sHat it
 * loses up to ~1e-10 relative in 1 - z, 1 - x and 1 - pT2Min/p2Abs (all computed by
 * subtraction, as Pythia does).  The 80-bit instantiation runs the same steps with ~1e-19
 * rounding and is the per-point reference for the FP64 kernels.
 *
 * Requires: REAL, FN(name), M_EXP, M_LOG, M_POW, M_SQRT, M_FMIN, M_FMAX.
 */

/* sigmaHat, Pythia's literal expressions (GeV^-2), massless:
 *   111 Sigma2gg2gg::sigmaKin           SigmaQCD.cc:115-129
 *   112 Sigma2gg2qqbar::sigmaKin        SigmaQCD.cc:170-189 (threshold open)
 *   113 Sigma2qg2qg::sigmaKin           SigmaQCD.cc:216-226
 *   114 Sigma2qq2qq::sigmaKin/sigmaHat  SigmaQCD.cc:257-286
 *   115 Sigma2qqbar2gg::sigmaKin        SigmaQCD.cc:315-325
 *   116 Sigma2qqbar2qqbarNew::sigmaKin  SigmaQCD.cc:364-378 (threshold open) */
static REAL FN(sigma_hat)(int code, REAL sH, REAL tH, REAL uH, REAL alpS, int id1, int id2, int nQuarkNew)
{
    const REAL pi = (REAL)3.14159265358979323846264338327950288L;
    REAL sH2 = sH * sH, tH2 = tH * tH, uH2 = uH * uH;
    REAL pre = (pi / sH2) * (alpS * alpS);
    switch (code) {
    case ORC_GG2GG: {
        REAL sigTS = ((REAL)9 / 4) * (tH2 / sH2 + 2 * tH / sH + 3 + 2 * sH / tH + sH2 / tH2);
        REAL sigUS = ((REAL)9 / 4) * (uH2 / sH2 + 2 * uH / sH + 3 + 2 * sH / uH + sH2 / uH2);
        REAL sigTU = ((REAL)9 / 4) * (tH2 / uH2 + 2 * tH / uH + 3 + 2 * uH / tH + uH2 / tH2);
        return pre * (REAL)0.5 * (sigTS + sigUS + sigTU);
    }
    case ORC_GG2QQBAR: {
        REAL sigTS = ((REAL)1 / 6) * uH / tH - ((REAL)3 / 8) * uH2 / sH2;
        REAL sigUS = ((REAL)1 / 6) * tH / uH - ((REAL)3 / 8) * tH2 / sH2;
        return pre * nQuarkNew * (sigTS + sigUS);
    }
    case ORC_QG2QG: {
        REAL sigTS = uH2 / tH2 - ((REAL)4 / 9) * uH / sH;
        REAL sigTU = sH2 / tH2 - ((REAL)4 / 9) * sH / uH;
        return pre * (sigTS + sigTU);
    }
    case ORC_QQ2QQ: {
        REAL sigT  = ((REAL)4 / 9) * (sH2 + uH2) / tH2;
        REAL sigU  = ((REAL)4 / 9) * (sH2 + tH2) / uH2;
        REAL sigTU = -((REAL)8 / 27) * sH2 / (tH * uH);
        REAL sigST = -((REAL)8 / 27) * uH2 / (sH * tH);
        REAL sigSum;
        if (id2 == id1)       sigSum = (REAL)0.5 * (sigT + sigU + sigTU);
        else if (id2 == -id1) sigSum = sigT + sigST;
        else                  sigSum = sigT;
        return pre * sigSum;
    }
    case ORC_QQBAR2GG: {
        REAL sigTS = ((REAL)32 / 27) * uH / tH - ((REAL)8 / 3) * uH2 / sH2;
        REAL sigUS = ((REAL)32 / 27) * tH / uH - ((REAL)8 / 3) * tH2 / sH2;
        return pre * (REAL)0.5 * (sigTS + sigUS);
    }
    case ORC_QQBAR2QQBARNEW:
        return pre * nQuarkNew * ((REAL)4 / 9) * (tH2 + uH2) / sH2;
    default:
        return 0;
    }
}

/* AlphaStrong::alphaS, orders 0/1 (StandardModel.cc:188-215). */
static REAL FN(alphas)(const orc_alphas *a, REAL scale2)
{
    const REAL pi = (REAL)3.14159265358979323846264338327950288L;
    if (a->order == 0) return (REAL)a->valueRef;
    if (scale2 < (REAL)a->scale2Min) scale2 = (REAL)a->scale2Min;
    if (scale2 > (REAL)a->mt2 && a->nfmax >= 6) return 12 * pi / (21 * M_LOG(scale2 / (REAL)a->L6s));
    if (scale2 > (REAL)a->mb2) return 12 * pi / (23 * M_LOG(scale2 / (REAL)a->L5s));
    if (scale2 > (REAL)a->mc2) return 12 * pi / (25 * M_LOG(scale2 / (REAL)a->L4s));
    return 12 * pi / (27 * M_LOG(scale2 / (REAL)a->L3s));
}

/* Toy xf(x): u = uv + sea, d = dv + sea, other quarks / antiquarks = sea, 21 = gluon. */
static REAL FN(xf)(const orc_toypdf *p, int id, REAL x)
{
    if (x <= 0 || x >= 1) return 0;
    int aid = id < 0 ? -id : id;
    if (id == 21) return (REAL)p->Ag * M_POW(x, (REAL)p->ag) * M_POW(1 - x, (REAL)p->bg);
    if (aid == 0 || aid > p->nQuarkIn) return 0;
    REAL sea = (REAL)p->As * M_POW(x, (REAL)p->as) * M_POW(1 - x, (REAL)p->bs);
    if (id == 2) return sea + (REAL)p->Auv * M_POW(x, (REAL)p->av) * M_POW(1 - x, (REAL)p->bv);
    if (id == 1) return sea + (REAL)p->Adv * M_POW(x, (REAL)p->av) * M_POW(1 - x, (REAL)p->bv);
    return sea;
}

/* LHAGrid1::xfxevolve (PartonDistributions.cc:648-745), PDF:extrapolate = off: the grid
 * neighbours by bisection in x and in q (within the q subgrid that holds q), 4-point Lagrange
 * weights in ln x and ln q (linear in ln q on subgrids of fewer than 4 points), frozen outside
 * the q range and below xMin.  x >= xMax (x = 1) gives 0 (Pythia keeps the previous values). */
static void FN(grid_xf)(const orc_lhagrid *g, REAL x, REAL Q2, REAL out[ORC_GRID_NF])
{
    REAL q = M_SQRT(Q2);
    int nx = g->nx, nq = g->nq;
    int inx = (x <= (REAL)g->xMin) ? -1 : ((x >= (REAL)g->xMax) ? 1 : 0);
    int inq = (q <= (REAL)g->qMin) ? -1 : ((q >= (REAL)g->qMax) ? 1 : 0);
    int minx = 0, maxx = nx - 1, m3x = 0;
    REAL wx[4] = { 1, 1, 1, 1 };
    if (inx == 0) {
        while (maxx - minx > 1) {
            int midx = (minx + maxx) / 2;
            if (x < (REAL)g->x[midx]) maxx = midx;
            else                      minx = midx;
        }
        REAL lnx = M_LOG(x);
        if      (minx == 0)      m3x = 0;
        else if (maxx == nx - 1) m3x = nx - 4;
        else                     m3x = minx - 1;
        for (int i3 = 0; i3 < 4; ++i3)
            for (int j = 0; j < 4; ++j)
                if (j != i3) wx[i3] *= (lnx - (REAL)g->lnx[m3x + j]) / ((REAL)g->lnx[m3x + i3] - (REAL)g->lnx[m3x + j]);
    }
    int iqDiv = 0;
    for (int s = 1; s < g->nqSub; ++s)
        if (q > (REAL)g->qDiv[s - 1]) iqDiv = s;
    int minS = (iqDiv == 0) ? 0 : g->nqSum[iqDiv - 1];
    int maxS = g->nqSum[iqDiv] - 1;
    int minq = minS, maxq = maxS, n3q = 4, m3q = 0;
    REAL wq[4] = { 1, 1, 1, 1 };
    if (inq == 0) {
        while (maxq - minq > 1) {
            int midq = (minq + maxq) / 2;
            if (q < (REAL)g->q[midq]) maxq = midq;
            else                      minq = midq;
        }
        REAL lnq = M_LOG(q);
        if (maxS - minS < 3) {
            n3q = 2;
            m3q = minq;
            wq[1] = (lnq - (REAL)g->lnq[minq]) / ((REAL)g->lnq[maxq] - (REAL)g->lnq[minq]);
            wq[0] = 1 - wq[1];
        } else {
            if      (minq == minS) m3q = minS;
            else if (maxq == maxS) m3q = maxS - 3;
            else                   m3q = minq - 1;
            for (int i3 = 0; i3 < 4; ++i3)
                for (int j = 0; j < 4; ++j)
                    if (j != i3) wq[i3] *= (lnq - (REAL)g->lnq[m3q + j]) / ((REAL)g->lnq[m3q + i3] - (REAL)g->lnq[m3q + j]);
        }
    } else {
        n3q = 1;
        m3q = (inq == 1) ? nq - 1 : 0;
    }
    for (int f = 0; f < ORC_GRID_NF; ++f) {
        REAL sum0 = 0;
        if (inx == 0) {
            for (int i3q = 0; i3q < n3q; ++i3q) {
                const double *p = g->val + ((size_t)(m3q + i3q) * nx + m3x) * ORC_GRID_NF + f;
                sum0 += wq[i3q] * (wx[0] * (REAL)p[0] + wx[1] * (REAL)p[ORC_GRID_NF] + wx[2] * (REAL)p[2 * ORC_GRID_NF]
                                   + wx[3] * (REAL)p[3 * ORC_GRID_NF]);
            }
        } else if (inx == -1) {
            for (int i3q = 0; i3q < n3q; ++i3q)
                sum0 += wq[i3q] * (REAL)g->val[((size_t)(m3q + i3q) * nx) * ORC_GRID_NF + f];
        }
        out[f] = sum0;
    }
}

/* Parton densities of one beam at x (and Q2 = the factorization scale for grids), indexed
 * id + nQuarkIn for quarks / antiquarks, 2 nQuarkIn + 1 for the gluon.  Grid values are clipped
 * at 0 as PDF::xf does (PartonDistributions.cc:122-133, 178-180). */
static void FN(fill_xf)(const orc_hadcfg *c, REAL x, REAL Q2, REAL *xf)
{
    int nQ = c->nQuarkIn;
    if (c->pdfSource == 1) {
        REAL v[ORC_GRID_NF];
        FN(grid_xf)(c->grid, x, Q2, v);
        for (int id = -nQ; id <= nQ; ++id) {
            if (id == 0) { xf[nQ] = 0; continue; }
            REAL t = v[id > 0 ? id : 5 - id];
            xf[id + nQ] = t > 0 ? t : 0;
        }
        xf[2 * nQ + 1] = v[0] > 0 ? v[0] : 0;
        return;
    }
    for (int id = -nQ; id <= nQ; ++id) xf[id + nQ] = (id == 0) ? 0 : FN(xf)(&c->pdf, id, x);
    xf[2 * nQ + 1] = FN(xf)(&c->pdf, 21, x);
}

/* sigmaPDF for one code: Pythia's incoming-pair list for the code's flux (initFlux,
 * SigmaProcess.cc:124-193), Kfactor * sigmaHatWrap * xfA * xfB summed (sigmaPDF :423-480). */
static REAL FN(sigma_pdf)(const orc_hadcfg *c, int code, REAL sH, REAL tH, REAL uH, REAL alpS,
                          const REAL *xfA, const REAL *xfB)
{
    int nQ = c->nQuarkIn, ig = 2 * nQ + 1;
    REAL K = (REAL)c->kfactor * (REAL)ORC_CONVERT2MB;
#define XF(arr, id) ((id) == 21 ? (arr)[ig] : (arr)[(id) + nQ])
    REAL sum = 0;
    if (code == ORC_GG2GG || code == ORC_GG2QQBAR) {                    /* "gg" */
        sum = K * FN(sigma_hat)(code, sH, tH, uH, alpS, 21, 21, c->nQuarkNew) * XF(xfA, 21) * XF(xfB, 21);
    } else if (code == ORC_QG2QG) {                                     /* "qg" */
        for (int id = -nQ; id <= nQ; ++id) {
            if (id == 0) continue;
            sum += K * FN(sigma_hat)(code, sH, tH, uH, alpS, id, 21, c->nQuarkNew) * XF(xfA, id) * XF(xfB, 21);
            sum += K * FN(sigma_hat)(code, sH, tH, uH, alpS, 21, id, c->nQuarkNew) * XF(xfA, 21) * XF(xfB, id);
        }
    } else if (code == ORC_QQ2QQ) {                                     /* "qq" */
        for (int i1 = -nQ; i1 <= nQ; ++i1) {
            if (i1 == 0) continue;
            for (int i2 = -nQ; i2 <= nQ; ++i2) {
                if (i2 == 0) continue;
                sum += K * FN(sigma_hat)(code, sH, tH, uH, alpS, i1, i2, c->nQuarkNew) * XF(xfA, i1) * XF(xfB, i2);
            }
        }
    } else if (code == ORC_QQBAR2GG || code == ORC_QQBAR2QQBARNEW) {   /* "qqbarSame" */
        for (int id = -nQ; id <= nQ; ++id) {
            if (id == 0) continue;
            sum += K * FN(sigma_hat)(code, sH, tH, uH, alpS, id, -id, c->nQuarkNew) * XF(xfA, id) * XF(xfB, -id);
        }
    }
#undef XF
    return sum;
}

/* Massive heavy-quark pair production, m3 = m4 = m (s3 = s4 = m2), Pythia's sigmaKin:
 *   121/123 Sigma2gg2QQbar::sigmaKin       SigmaQCD.cc:736-766 (openFracPair = 1, no top model)
 *   122/124 Sigma2qqbar2QQbar::sigmaKin    SigmaQCD.cc:869-891
 * with s34Avg = m2, tHQ = -(sH - tH + uH)/2, uHQ = -(sH + tH - uH)/2. */
static REAL FN(sigma_hat_massive)(int code, REAL sH, REAL tH, REAL uH, REAL m2, REAL alpS)
{
    const REAL pi = (REAL)3.14159265358979323846264338327950288L;
    REAL s3 = m2, s4 = m2;
    REAL sH2 = sH * sH;
    REAL s34Avg = (REAL)0.5 * (s3 + s4) - (REAL)0.25 * (s3 - s4) * (s3 - s4) / sH;
    REAL tHQ = -(REAL)0.5 * (sH - tH + uH);
    REAL uHQ = -(REAL)0.5 * (sH + tH - uH);
    REAL tHQ2 = tHQ * tHQ, uHQ2 = uHQ * uHQ;
    if (code == ORC_GG2CCBAR || code == ORC_GG2BBBAR) {
        REAL tumHQ = tHQ * uHQ - s34Avg * sH;
        REAL sigTS = (uHQ / tHQ - (REAL)2.25 * uHQ2 / sH2 + (REAL)4.5 * s34Avg * tumHQ / (sH * tHQ2)
                      + (REAL)0.5 * s34Avg * (tHQ + s34Avg) / tHQ2 - s34Avg * s34Avg / (sH * tHQ)) / 6;
        REAL sigUS = (tHQ / uHQ - (REAL)2.25 * tHQ2 / sH2 + (REAL)4.5 * s34Avg * tumHQ / (sH * uHQ2)
                      + (REAL)0.5 * s34Avg * (uHQ + s34Avg) / uHQ2 - s34Avg * s34Avg / (sH * uHQ)) / 6;
        return (pi / sH2) * (alpS * alpS) * (sigTS + sigUS);
    }
    REAL sigS = ((REAL)4 / 9) * ((tHQ2 + uHQ2) / sH2 + 2 * s34Avg / sH);
    return (pi / sH2) * (alpS * alpS) * sigS;
}

/* One process family (a set of processes sharing the final-state masses and hence the phase
 * space): family 0 = the massless codes 111..116, 1 = c cbar (121, 122), 2 = b bbar (123, 124).
 * Each process of Pythia has its own PhaseSpace; the family's processes share it, evaluated at
 * the same uniforms u.  Steps: selectTau (PhaseSpace.cc:1256-1337, iTau = 0, tauMin =
 * (mT3Min + mT4Min)^2 / s with mT = sqrt(m^2 + pT2Min), :1166-1172), y flat in [-yMax, yMax]
 * (limitY :1183-1200, wtY = 2 yMax), p2Abs = ((sH - s3 - s4)^2 - 4 s3 s4) / (4 sH) (:1263-1266),
 * limitZ/selectZ (:1205-1250, :1453-1568, iZ = 0 with its neg/pos split, tHat uHat from the
 * stable product incl. s3 s4 z^2), store2Kin (SigmaProcess.cc:846-925: pT2 = (tH uH - s3 s4)/sH,
 * massless Q2 = pT2, massive Q2Ren = sqrt((pT2 + s3)(pT2 + s4)) [renormScale2 = 2] and
 * Q2Fac = pT2 + min(s3, s4) [factorScale2 = 1]), weight = sigmaPDF * wtTau * wtY * wtZ
 * (PhaseSpace.cc:1074).  kin[] of the family. */
static REAL FN(had_family)(const orc_hadcfg *c, int fam, const double uin[3], REAL w[ORC_NPROC], REAL kin[8],
                           int literalU)
{
    REAL m = fam == 0 ? 0 : (REAL)c->mQ[fam - 1];
    REAL s3 = m * m, s4 = m * m;
    REAL u0 = (REAL)uin[0], u1 = (REAL)uin[1], u2 = (REAL)uin[2];
    REAL s = (REAL)c->eCM * (REAL)c->eCM;
    REAL pT2Min = (REAL)c->pTHatMin * (REAL)c->pTHatMin;
    REAL tauMin;
    if (fam == 0) tauMin = 4 * pT2Min / s;                  /* (mT3Min + mT4Min)^2 / s, massless */
    else {
        REAL mT3Min = M_SQRT(s3 + pT2Min), mT4Min = M_SQRT(s4 + pT2Min);
        tauMin = (mT3Min + mT4Min) * (mT3Min + mT4Min) / s;
    }
    REAL tauMax = 1;
    REAL intTau0 = M_LOG(tauMax / tauMin);
    REAL tau = tauMin * M_POW(tauMax / tauMin, u0);
    REAL wtTau = intTau0;

    REAL yMax = -(REAL)0.5 * M_LOG(tau);
    REAL y = yMax * (2 * u1 - 1);
    REAL wtY = 2 * yMax;
    REAL x1 = M_SQRT(tau) * M_EXP(y);
    REAL x2 = M_SQRT(tau) * M_EXP(-y);

    REAL sH = tau * s;
    REAL mHat = M_SQRT(sH);
    REAL p2Abs = fam == 0 ? (REAL)0.25 * sH * sH / sH : (REAL)0.25 * ((sH - s3 - s4) * (sH - s3 - s4) - 4 * s3 * s4) / sH;
    REAL pAbs = M_SQRT(p2Abs);
    REAL zArg = 1 - pT2Min / p2Abs;
    REAL zMax = zArg > 0 ? M_SQRT(zArg) : 0;
    REAL zMin = 0;
    REAL zNegMin = -zMax, zNegMax = -zMin, zPosMin = zMin, zPosMax = zMax;
    REAL area0Neg = zNegMax - zNegMin, area0Pos = zPosMax - zPosMin;
    REAL area0 = area0Neg + area0Pos;
    REAL z;
    if (u2 * area0 < area0Neg) z = zNegMin + (u2 * area0 / area0Neg) * area0Neg;
    else                       z = zPosMin + ((u2 * area0 - area0Neg) / area0Pos) * area0Pos;
    if (z < 0) z = M_FMIN(zNegMax, M_FMAX(zNegMin, z));
    else       z = M_FMIN(zPosMax, M_FMAX(zPosMin, z));
    REAL wtZ = mHat * pAbs * area0;                    /* zCoef[0] = 1 */

    REAL sH34 = -(REAL)0.5 * (sH - s3 - s4);
    REAL tHuH = sH34 * sH34 * (1 - z) * (1 + z) + s3 * s4 * z * z;
    REAL tH, uH;
    if (z < 0) { tH = sH34 + mHat * pAbs * z; uH = tHuH / tH; }
    else       { uH = sH34 - mHat * pAbs * z; tH = tHuH / uH; }
    if (literalU) uH = s3 + s4 - (sH + tH);

    REAL pT2 = (tH * uH - s3 * s4) / sH;
    REAL Q2Ren = fam == 0 ? pT2 : M_SQRT((pT2 + s3) * (pT2 + s4));
    REAL Q2Fac = fam == 0 ? pT2 : pT2 + (s3 < s4 ? s3 : s4);
    Q2Ren *= (REAL)c->renormFac * (REAL)c->renormFac;
    Q2Fac *= (REAL)c->factFac * (REAL)c->factFac;
    REAL alpS = FN(alphas)(&c->as, Q2Ren);

    REAL xfA[16], xfB[16];
    FN(fill_xf)(c, x1, Q2Fac, xfA);
    FN(fill_xf)(c, x2, Q2Fac, xfB);

    REAL jac = wtTau * wtY * wtZ;
    REAL total = 0;
    int nQ = c->nQuarkIn, ig = 2 * nQ + 1;
    REAL K = (REAL)c->kfactor * (REAL)ORC_CONVERT2MB;
    for (int i = 0; i < ORC_NPROC; ++i) {
        if (!(c->procMask & (1 << i))) continue;
        int code = i < 6 ? 111 + i : 121 + (i - 6);
        int f = i < 6 ? 0 : (i < 8 ? 1 : 2);
        if (f != fam) continue;
        REAL wi;
        if (fam == 0) {
            wi = FN(sigma_pdf)(c, code, sH, tH, uH, alpS, xfA, xfB) * jac;
        } else if (code == ORC_GG2CCBAR || code == ORC_GG2BBBAR) {           /* "gg" */
            wi = K * FN(sigma_hat_massive)(code, sH, tH, uH, s3, alpS) * xfA[ig] * xfB[ig] * jac;
        } else {                                                            /* "qqbarSame" */
            REAL sum = 0;
            REAL sig = K * FN(sigma_hat_massive)(code, sH, tH, uH, s3, alpS);
            for (int id = -nQ; id <= nQ; ++id) {
                if (id == 0) continue;
                sum += sig * xfA[id + nQ] * xfB[-id + nQ];
            }
            wi = sum * jac;
        }
        if (wi < 0) wi = 0;                              /* PhaseSpace.cc:1135 */
        w[i] = wi;
        total += wi;
    }
    kin[0] = sH; kin[1] = tH; kin[2] = uH; kin[3] = x1; kin[4] = x2; kin[5] = pT2; kin[6] = alpS; kin[7] = jac;
    return total;
}

/* The point: every family selected by procMask, at the same uniforms; kin[] of the first
 * selected family (massless, then c cbar, then b bbar). */
static REAL FN(had_point)(const orc_hadcfg *c, const double uin[3], REAL w[ORC_NPROC], REAL kin[8], int literalU)
{
    for (int i = 0; i < ORC_NPROC; ++i) w[i] = 0;
    REAL total = 0;
    int have_kin = 0;
    static const int famMask[3] = { 0x3F, 0xC0, 0x300 };
    for (int fam = 0; fam < 3; ++fam) {
        if (!(c->procMask & famMask[fam])) continue;
        REAL kf[8];
        total += FN(had_family)(c, fam, uin, w, kf, literalU);
        if (!have_kin) { for (int q = 0; q < 8; ++q) kin[q] = kf[q]; have_kin = 1; }
    }
    if (!have_kin) for (int q = 0; q < 8; ++q) kin[q] = 0;
    return total;
}
