/* oracle_events.c -- test infrastructure (the oracle): CPU restatement of the partons that complete a
 * compacted event (incoming and outgoing flavours, colour flow, lab four-momenta), the checker of the
 * device event record.  Not part of the library.
 *
 * Follows Pythia 8.317:
 *   flux lists           SigmaProcess::initFlux, pythia8317/src/SigmaProcess.cc:124-193
 *                        ("gg"; "qg": (q, g), (g, q) for q = -nQ..nQ; "qq": every (q1, q2);
 *                         "qqbarSame": (q, -q))
 *   incoming pair        SigmaProcess::pickInState, SigmaProcess.cc:486-508: pair i with weight
 *                        Kfactor sigmaHatWrap(id1, id2) xf1 xf2 (pdfSigma), the first i where
 *                        sigmaSum u - sum_{j <= i} pdfSigma_j <= 0 (here: the last pair when rounding
 *                        never gets there)
 *   flavours, colours    setIdColAcol: Sigma2gg2gg SigmaQCD.cc:135-148, Sigma2gg2qqbar :195-206,
 *                        Sigma2qg2qg :232-243, Sigma2qq2qq :292-303, Sigma2qqbar2gg :331-342,
 *                        Sigma2qqbar2qqbarNew :384-394, Sigma2gg2QQbar :768-781 (sigTS, sigUS of the
 *                        massive sigmaKin :736-755), Sigma2qqbar2QQbar :894-905 (no top model);
 *                        the kinematic factors as each sigmaKin computes them (:115-129, 170-189,
 *                        216-226, 257-265, 315-325)
 *   momenta              PhaseSpace::finalKin, PhaseSpace.cc:2236-2251: incoming (0, 0, +-x eCM/2,
 *                        x eCM/2), outgoing (0, 0, +-pAbs, mHat/2) rotated by (theta, phi) with
 *                        Vec4::rot (Basics.cc:367-378) and boosted by betaZ = (x1 - x2)/(x1 + x2) with
 *                        Vec4::bst (Basics.cc:433-444); sH = x1 x2 eCM^2, pAbs = sqrt(sH/4 - m^2),
 *                        cos theta from the record's (sHat, tHat): (1 - 2 (m^2 - tHat)/sHat) / beta;
 *                        evaluated in long double (see vbst_z)
 *   hard scale           renormFac^2 (pT^2 + m^2) = Q2Ren of SigmaProcess::store2Kin with
 *                        renormScale2 = 2 (SigmaProcess.cc:846-925); the record holds its root.
 * Random numbers (this library's definition): Philox4x32-10 with counter (index lo, index hi,
 * stream, 0); stream 3 words (0, 1) the pair and (2, 3) the colour topology, stream 4 words (0, 1)
 * the azimuth, as 53-bit uniforms orc_u53; stream 4 word 2 the g g -> g g orientation and word 3
 * the new flavour of 112 / 116 as 24-bit uniforms (w >> 8) 2^-24.
 * Pinned by tests/test_oracle_events.py: colour-topology and flavour frequencies at fixed
 * kinematics against Pythia's probability ratios, colour-line validity, four-momentum conservation,
 * on-shell masses and the invariants (p1 + p2)^2 = x1 x2 S, (p1 - p3)^2 = tHat. */
#include "mcint_oracle.h"

#include <math.h>
#include <string.h>

static double uni24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

static void stream_words(uint64_t key, uint64_t index, uint32_t stream, uint32_t out[4])
{
    const uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), stream, 0u};
    const uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    orc_philox4x32_10(ctr, k, out);
}

/* xf of one beam, clipped at 0 (PDF::xf); the grid at Q2F = factFac^2 (pT^2 + m^2). */
static double beam_xf(const orc_hadcfg *c, int id, double x, double Q2F)
{
    if (c->pdfSource == 1) {
        double v[ORC_GRID_NF];
        orc_lhagrid_xf(c->grid, x, Q2F, v);
        const double t = v[id == 21 ? 0 : (id > 0 ? id : 5 - id)];
        return t > 0 ? t : 0;
    }
    return orc_toypdf_xf(&c->pdf, id, x);
}

static void set_col(orc_event4 *E, int c1, int a1, int c2, int a2, int c3, int a3, int c4, int a4)
{
    const int c[4] = {c1, c2, c3, c4}, a[4] = {a1, a2, a3, a4};
    for (int k = 0; k < 4; ++k) { E->col[k] = c[k]; E->acol[k] = a[k]; }
}

static void swap_col_acol(orc_event4 *E)
{
    for (int k = 0; k < 4; ++k) { int t = E->col[k]; E->col[k] = E->acol[k]; E->acol[k] = t; }
}

/* Vec4::rot(theta, phi) and Vec4::bst(0, 0, betaZ) on (x, y, z, t), in long double: Pythia's
 * gamma = 1 / sqrt(1 - beta^2) loses ~1e-16 / (1 - beta^2) (5e-11 at x1 / x2 = 2e6), which long
 * double brings below the double rounding of the result. */
static void vrot(long double *p, long double theta, long double phi)
{
    const long double cthe = cosl(theta), sthe = sinl(theta), cphi = cosl(phi), sphi = sinl(phi);
    const long double tx = cthe * cphi * p[0] - sphi * p[1] + sthe * cphi * p[2];
    const long double ty = cthe * sphi * p[0] + cphi * p[1] + sthe * sphi * p[2];
    const long double tz = -sthe * p[0] + cthe * p[2];
    p[0] = tx; p[1] = ty; p[2] = tz;
}
static void vbst_z(long double *p, long double betaZ)
{
    const long double beta2 = betaZ * betaZ;
    if (beta2 >= 1.0L) return;
    const long double gamma = 1.0L / sqrtl(1.0L - beta2);
    const long double prod1 = betaZ * p[2];
    const long double prod2 = gamma * (gamma * prod1 / (1.0L + gamma) + p[3]);
    p[2] += prod2 * betaZ;
    p[3] = gamma * (p[3] + prod1);
}

void orc_build_event(const orc_hadcfg *c, int partonic, double sHatPartonic, uint64_t key, uint64_t index, int code,
                     double sH, double tH, double x1, double x2, orc_event4 *E)
{
    memset(E, 0, sizeof *E);
    const int heavy = code >= 121 && code <= 124;
    if (!((code >= 111 && code <= 116) || heavy) || (partonic && code != 111)) return;   /* nPart = 0 */
    uint32_t r3[4], r4[4];
    stream_words(key, index, 3u, r3);
    stream_words(key, index, 4u, r4);
    const int idQ = (code == 121 || code == 122) ? 4 : 5;
    const double mQ = heavy ? c->mQ[idQ - 4] : 0.0, m2 = mQ * mQ;
    const double s = sH, t = tH, u = 2.0 * m2 - (s + t);
    const double Q2 = (m2 - t) * (m2 - u) / s;
    const double X1 = partonic ? 1.0 : x1, X2 = partonic ? 1.0 : x2;
    const int nQ = c->nQuarkIn;

    /* ---- incoming pair (pickInState) */
    int id1 = 21, id2 = 21;
    if (!partonic && (code == 113 || code == 114 || code == 115 || code == 116 || code == 122 || code == 124)) {
        const double Q2F = c->factFac * c->factFac * Q2;
        int ia[128], ib[128], np = 0;
        double ps[128], sum = 0.0;
        for (int q = -nQ; q <= nQ; ++q) {
            if (q == 0) continue;
            if (code == 113) {
                ia[np] = q; ib[np++] = 21;
                ia[np] = 21; ib[np++] = q;
            } else if (code == 114) {
                for (int q2 = -nQ; q2 <= nQ; ++q2)
                    if (q2 != 0) { ia[np] = q; ib[np++] = q2; }
            } else {
                ia[np] = q; ib[np++] = -q;
            }
        }
        for (int i = 0; i < np; ++i) {
            const double sig = heavy ? orc_sigma_hat_massive(code, s, t, u, m2, 0.118)
                                     : orc_sigma_hat_pythia(code, s, t, u, 0.118, ia[i], ib[i], c->nQuarkNew);
            ps[i] = c->kfactor * ORC_CONVERT2MB * sig * beam_xf(c, ia[i], X1, Q2F) * beam_xf(c, ib[i], X2, Q2F);
            sum += ps[i];
        }
        double r = sum * orc_u53(r3[0], r3[1]);
        int pick = np - 1;
        for (int i = 0; i < np; ++i) {
            r -= ps[i];
            if (r <= 0.0) { pick = i; break; }
        }
        id1 = ia[pick];
        id2 = ib[pick];
    }

    /* ---- outgoing flavours and colour flow (setIdColAcol) */
    const double uc = orc_u53(r3[2], r3[3]), us = uni24(r4[2]), un = uni24(r4[3]);
    const double sH2 = s * s, tH2 = t * t, uH2 = u * u;
    int id3 = 21, id4 = 21;
    if (code == 111) {
        const double sigTS = (9. / 4.) * (tH2 / sH2 + 2. * t / s + 3. + 2. * s / t + sH2 / tH2);
        const double sigUS = (9. / 4.) * (uH2 / sH2 + 2. * u / s + 3. + 2. * s / u + sH2 / uH2);
        const double sigTU = (9. / 4.) * (tH2 / uH2 + 2. * t / u + 3. + 2. * u / t + uH2 / tH2);
        const double sigRand = (sigTS + sigUS + sigTU) * uc;
        if (sigRand < sigTS) set_col(E, 1, 2, 2, 3, 1, 4, 4, 3);
        else if (sigRand < sigTS + sigUS) set_col(E, 1, 2, 3, 1, 3, 4, 4, 2);
        else set_col(E, 1, 2, 3, 4, 1, 4, 3, 2);
        if (us > 0.5) swap_col_acol(E);
    } else if (code == 112 || code == 121 || code == 123) {
        double sigTS, sigUS;
        int idNew;
        if (code == 112) {
            idNew = 1 + (int)(c->nQuarkNew * un);
            sigTS = (1. / 6.) * u / t - (3. / 8.) * uH2 / sH2;
            sigUS = (1. / 6.) * t / u - (3. / 8.) * tH2 / sH2;
        } else {
            idNew = idQ;
            const double s34Avg = m2, tHQ = -0.5 * (s - t + u), uHQ = -0.5 * (s + t - u);
            const double tHQ2 = tHQ * tHQ, uHQ2 = uHQ * uHQ, tumHQ = tHQ * uHQ - s34Avg * s;
            sigTS = (uHQ / tHQ - 2.25 * uHQ2 / sH2 + 4.5 * s34Avg * tumHQ / (s * tHQ2) +
                     0.5 * s34Avg * (tHQ + s34Avg) / tHQ2 - s34Avg * s34Avg / (s * tHQ)) / 6.;
            sigUS = (tHQ / uHQ - 2.25 * tHQ2 / sH2 + 4.5 * s34Avg * tumHQ / (s * uHQ2) +
                     0.5 * s34Avg * (uHQ + s34Avg) / uHQ2 - s34Avg * s34Avg / (s * uHQ)) / 6.;
        }
        id3 = idNew; id4 = -idNew;
        if ((sigTS + sigUS) * uc < sigTS) set_col(E, 1, 2, 2, 3, 1, 0, 0, 3);
        else set_col(E, 1, 2, 3, 1, 3, 0, 0, 2);
    } else if (code == 113) {
        const double sigTS = uH2 / tH2 - (4. / 9.) * u / s, sigTU = sH2 / tH2 - (4. / 9.) * s / u;
        id3 = id1; id4 = id2;
        if ((sigTS + sigTU) * uc < sigTS) set_col(E, 1, 0, 2, 1, 3, 0, 2, 3);
        else set_col(E, 1, 0, 2, 3, 2, 0, 1, 3);
        if (id1 == 21) {                                   /* swapCol1234: 1 <-> 2, 3 <-> 4 */
            int t0;
            t0 = E->col[0]; E->col[0] = E->col[1]; E->col[1] = t0;
            t0 = E->acol[0]; E->acol[0] = E->acol[1]; E->acol[1] = t0;
            t0 = E->col[2]; E->col[2] = E->col[3]; E->col[3] = t0;
            t0 = E->acol[2]; E->acol[2] = E->acol[3]; E->acol[3] = t0;
        }
        if (id1 < 0 || id2 < 0) swap_col_acol(E);
    } else if (code == 114) {
        const double sigT = (4. / 9.) * (sH2 + uH2) / tH2, sigU = (4. / 9.) * (sH2 + tH2) / uH2;
        id3 = id1; id4 = id2;
        if (id1 * id2 > 0) set_col(E, 1, 0, 2, 0, 2, 0, 1, 0);
        else set_col(E, 1, 0, 0, 1, 2, 0, 0, 2);
        if (id2 == id1 && (sigT + sigU) * uc > sigT) set_col(E, 1, 0, 2, 0, 1, 0, 2, 0);
        if (id1 < 0) swap_col_acol(E);
    } else if (code == 115) {
        const double sigTS = (32. / 27.) * u / t - (8. / 3.) * uH2 / sH2;
        const double sigUS = (32. / 27.) * t / u - (8. / 3.) * tH2 / sH2;
        if ((sigTS + sigUS) * uc < sigTS) set_col(E, 1, 0, 0, 2, 1, 3, 3, 2);
        else set_col(E, 1, 0, 0, 2, 3, 2, 1, 3);
        if (id1 < 0) swap_col_acol(E);
    } else {                                               /* 116, 122, 124 */
        const int idNew = heavy ? idQ : 1 + (int)(c->nQuarkNew * un);
        id3 = id1 > 0 ? idNew : -idNew;
        id4 = -id3;
        set_col(E, 1, 0, 0, 2, 1, 0, 0, 2);
        if (id1 < 0) swap_col_acol(E);
    }

    /* ---- momenta (finalKin), long double */
    const long double eCM = partonic ? sqrtl((long double)sHatPartonic) : (long double)c->eCM;
    const long double sHk = (long double)X1 * (long double)X2 * eCM * eCM, mHat = sqrtl(sHk);
    const long double pAbs = sqrtl(fmaxl(0.0L, 0.25L * sHk - m2));
    const long double betaQ = sqrtl(fmaxl(0.0L, 1.0L - 4.0L * m2 / (long double)s));
    long double z = (1.0L - 2.0L * ((long double)m2 - t) / (long double)s) / (heavy ? betaQ : 1.0L);
    z = fminl(1.0L, fmaxl(-1.0L, z));
    const long double theta = acosl(z), phi = 6.283185307179586476925286766559L * orc_u53(r4[0], r4[1]);   /* 2 pi u */
    const long double betaZ = ((long double)X1 - X2) / ((long double)X1 + X2);
    long double p[4][4] = {{0.0L, 0.0L, 0.5L * eCM * X1, 0.5L * eCM * X1},
                           {0.0L, 0.0L, -0.5L * eCM * X2, 0.5L * eCM * X2},
                           {0.0L, 0.0L, pAbs, 0.5L * mHat},
                           {0.0L, 0.0L, -pAbs, 0.5L * mHat}};
    for (int k = 2; k < 4; ++k) {
        vrot(p[k], theta, phi);
        vbst_z(p[k], betaZ);
    }
    const int ids[4] = {id1, id2, id3, id4};
    for (int k = 0; k < 4; ++k) {
        E->pdg[k] = ids[k];
        E->status[k] = k < 2 ? -21 : 23;
        E->col[k] = E->col[k] ? 500 + E->col[k] : 0;
        E->acol[k] = E->acol[k] ? 500 + E->acol[k] : 0;
        E->mo1[k] = k < 2 ? -1 : 0;
        E->mo2[k] = k < 2 ? -1 : 1;
        E->d1[k] = k < 2 ? 2 : -1;
        E->d2[k] = k < 2 ? 3 : -1;
        for (int j = 0; j < 4; ++j) E->p[k][j] = (double)p[k][j];
        E->p[k][4] = k < 2 ? 0.0 : mQ;
    }
    E->scale = c->renormFac * sqrt(fmax(Q2, 0.0));
    E->nPart = 4;
}

void orc_build_events(const orc_hadcfg *c, int partonic, double sHatPartonic, uint64_t key, uint64_t n,
                      const uint64_t *index, const int32_t *code, const double *sH, const double *tH,
                      const double *x1, const double *x2, int32_t *ints, double *mom, double *scale, int32_t *nPart)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        orc_event4 E;
        orc_build_event(c, partonic, sHatPartonic, key, index[i], code[i], sH[i], tH[i], x1 ? x1[i] : 1.0,
                        x2 ? x2[i] : 1.0, &E);
        int32_t *o = ints + (size_t)i * 32;
        for (int k = 0; k < 4; ++k) {
            o[8 * k + 0] = E.pdg[k]; o[8 * k + 1] = E.status[k]; o[8 * k + 2] = E.col[k]; o[8 * k + 3] = E.acol[k];
            o[8 * k + 4] = E.mo1[k]; o[8 * k + 5] = E.mo2[k]; o[8 * k + 6] = E.d1[k]; o[8 * k + 7] = E.d2[k];
            for (int j = 0; j < 5; ++j) mom[(size_t)i * 20 + 5 * k + j] = E.p[k][j];
        }
        scale[i] = E.scale;
        nPart[i] = E.nPart;
    }
}
