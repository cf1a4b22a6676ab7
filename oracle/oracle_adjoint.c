/*
 * oracle_adjoint.c -- CPU oracle for the adjoint pass.  TEST INFRASTRUCTURE ONLY (see
 * mcint_oracle.h).
 *
 * Independent of the device's analytic derivatives: every point's weight derivative is a
 * central finite difference (relative step 1e-5) of the long-double Pythia-form point
 * (orc_had_point_ld) with the configuration rebuilt at theta +- h -- alphaS(MZ) re-runs the
 * AlphaStrong initialisation (Lambda matching, StandardModel.cc:58-62), the gluon exponents
 * re-run the toy-PDF normalisation (gluon momentum fraction 0.5).  Same events (flat map,
 * Philox stream 0, three 24-bit uniforms), so the sums are the derivative of the same MC sum.
 *
 * out[0] = sum w, out[1..4] = sum d w / d {alphaS(MZ), renorm_fac, gluon a, gluon b},
 * out[5..8] = sums of their squares.
 */
#include "mcint_oracle.h"
#include <math.h>
#include <string.h>

static void with_param(const orc_hadcfg *c, int which, double delta, orc_hadcfg *o)
{
    *o = *c;
    if (which == 0) {
        orc_alphas_init(&o->as, c->as.valueRef + delta, c->as.order, c->as.nfmax);
    } else if (which == 1) {
        o->renormFac = c->renormFac + delta;
    } else {
        double g[2] = { c->pdf.ag, c->pdf.bg };
        double v[2] = { c->pdf.av, c->pdf.bv };
        double s[2] = { c->pdf.as, c->pdf.bs };
        g[which - 2] += delta;
        orc_toypdf_init(&o->pdf, g, v, s, c->nQuarkIn);
    }
}

static double param_value(const orc_hadcfg *c, int which)
{
    switch (which) {
    case 0: return c->as.valueRef;
    case 1: return c->renormFac;
    case 2: return c->pdf.ag;
    default: return c->pdf.bg;
    }
}

void orc_adjoint(const orc_hadcfg *c, uint64_t key, uint64_t ctr0, uint64_t n, double out[9])
{
    orc_hadcfg cp[4], cm[4];
    double h[4];
    for (int q = 0; q < 4; ++q) {
        double v = param_value(c, q);
        h[q] = 1e-5 * (fabs(v) > 0.1 ? fabs(v) : 0.1);
        with_param(c, q, +h[q], &cp[q]);
        with_param(c, q, -h[q], &cm[q]);
    }
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    double acc[9];
    memset(acc, 0, sizeof acc);
#pragma omp parallel
    {
        double loc[9];
        memset(loc, 0, sizeof loc);
#pragma omp for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            uint64_t e = ctr0 + (uint64_t)i;
            uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
            orc_philox4x32_10(c4, kk, o);
            double u[3] = { (double)(o[0] >> 8) * (1.0 / 16777216.0), (double)(o[1] >> 8) * (1.0 / 16777216.0),
                            (double)(o[2] >> 8) * (1.0 / 16777216.0) };
            double w[ORC_NPROC];
            double w0 = orc_had_point_ld(c, u, w, NULL, 0);
            loc[0] += w0;
            for (int q = 0; q < 4; ++q) {
                double wp = orc_had_point_ld(&cp[q], u, w, NULL, 0);
                double wm = orc_had_point_ld(&cm[q], u, w, NULL, 0);
                double d = (wp - wm) / (2.0 * h[q]);
                loc[1 + q] += d;
                loc[5 + q] += d * d;
            }
        }
#pragma omp critical
        for (int q = 0; q < 9; ++q) acc[q] += loc[q];
    }
    for (int q = 0; q < 9; ++q) out[q] = acc[q];
}
