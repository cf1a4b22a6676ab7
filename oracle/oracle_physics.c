/*
 * oracle_physics.c -- CPU oracle: QCD 2->2 sigmaHat, alphaS, toy PDFs, the (tau, y, z)
 * phase-space map, the partonic gg->gg integrand and deterministic quadratures.
 * TEST INFRASTRUCTURE ONLY (see mcint_oracle.h).  FP64 throughout.
 *
 * Each function restates the reference's algorithm from the physics; the reference
 * file:line it follows is given beside it.
 */
#include "mcint_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double pow2(double x) { return x * x; }

/* ------------------------------------------------------------------------------------
 * Pythia-form point evaluation, instantiated in FP64 (_d) and 80-bit long double (_ld).
 * ---------------------------------------------------------------------------------- */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUFFIX)

#define REAL double
#define SUFFIX _d
#define M_EXP exp
#define M_LOG log
#define M_POW pow
#define M_SQRT sqrt
#define M_FMIN fmin
#define M_FMAX fmax
#include "oracle_had_body.h"
#undef REAL
#undef SUFFIX
#undef M_EXP
#undef M_LOG
#undef M_POW
#undef M_SQRT
#undef M_FMIN
#undef M_FMAX

#define REAL long double
#define SUFFIX _ld
#define M_EXP expl
#define M_LOG logl
#define M_POW powl
#define M_SQRT sqrtl
#define M_FMIN fminl
#define M_FMAX fmaxl
#include "oracle_had_body.h"
#undef REAL
#undef SUFFIX
#undef M_EXP
#undef M_LOG
#undef M_POW
#undef M_SQRT
#undef M_FMIN
#undef M_FMAX

double orc_sigma_hat_pythia(int code, double sH, double tH, double uH, double alpS,
                            int id1, int id2, int nQuarkNew)
{
    return sigma_hat_d(code, sH, tH, uH, alpS, id1, id2, nQuarkNew);
}

/* Textbook brackets (Ellis-Stirling-Webber table 7.1 / Combridge, as restated in
 * cuPythia/08_qcd_process_library/qcd_library.cu:25-60); flavour combination for 114 and
 * the nQuarkNew factor follow Pythia so the two forms describe the same quantity. */
double orc_bracket_textbook(int code, double s, double t, double u, int id1, int id2, int nQuarkNew)
{
    double s2 = s * s, t2 = t * t, u2 = u * u;
    switch (code) {
    case ORC_GG2GG:    return (9. / 4.) * (3. - t * u / s2 - s * u / t2 - s * t / u2);
    case ORC_GG2QQBAR: return nQuarkNew * ((1. / 6.) * (t2 + u2) / (t * u) - (3. / 8.) * (t2 + u2) / s2);
    case ORC_QG2QG:    return (s2 + u2) / t2 - (4. / 9.) * (s2 + u2) / (s * u);
    case ORC_QQ2QQ: {
        double tch = (4. / 9.) * (s2 + u2) / t2;
        if (id2 == id1) {          /* identical quarks: t + u channels + interference, 1/2 */
            double uch = (4. / 9.) * (s2 + t2) / u2;
            return 0.5 * (tch + uch - (8. / 27.) * s2 / (t * u));
        }
        if (id2 == -id1)           /* q qbar -> q qbar, elastic: t + s channels + interference */
            return tch - (8. / 27.) * u2 / (s * t);
        return tch;
    }
    case ORC_QQBAR2GG: return (16. / 27.) * (t2 + u2) / (t * u) - (4. / 3.) * (t2 + u2) / s2;
    case ORC_QQBAR2QQBARNEW: return nQuarkNew * (4. / 9.) * (t2 + u2) / s2;
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------------------------
 * AlphaStrong, orders 0 and 1.  StandardModel.cc:22-62 (init), 188-215 (alphaS):
 * Lambda5 from alphaS(MZ) with MZ = 91.188, matched at mc = 1.5, mb = 4.8, mt = 171.0,
 * frozen below (1.07 Lambda3)^2.
 * ---------------------------------------------------------------------------------- */
void orc_alphas_init(orc_alphas *a, double valueRef, int order, int nfmax)
{
    const double MZ = 91.188;
    memset(a, 0, sizeof *a);
    a->valueRef = valueRef;
    a->order = order < 0 ? 0 : (order > 1 ? 1 : order);
    a->nfmax = nfmax < 5 ? 5 : (nfmax > 6 ? 6 : nfmax);
    a->mc = 1.5; a->mb = 4.8; a->mt = 171.0;
    a->mc2 = a->mc * a->mc; a->mb2 = a->mb * a->mb; a->mt2 = a->mt * a->mt;
    if (a->order == 1) {
        double L5 = MZ * exp(-6. * M_PI / (23. * valueRef));
        double L6 = L5 * pow(L5 / a->mt, 2. / 21.);
        double L4 = L5 * pow(a->mb / L5, 2. / 25.);
        double L3 = L4 * pow(a->mc / L4, 2. / 27.);
        a->L3s = L3 * L3; a->L4s = L4 * L4; a->L5s = L5 * L5; a->L6s = L6 * L6;
        a->scale2Min = pow2(1.07 * L3);
    }
}

double orc_alphas_eval(const orc_alphas *a, double scale2) { return alphas_d(a, scale2); }

/* ------------------------------------------------------------------------------------
 * Toy PDFs xf = A x^a (1-x)^b (the parametrisation is fixed by the workload; normalisation
 * and flavour content are this project's definition, see mcint_oracle.h).
 * ---------------------------------------------------------------------------------- */
static double beta_fn(double p, double q) { return exp(lgamma(p) + lgamma(q) - lgamma(p + q)); }

void orc_toypdf_init(orc_toypdf *p, const double g_ab[2], const double v_ab[2],
                     const double s_ab[2], int nQuarkIn)
{
    p->ag = g_ab[0]; p->bg = g_ab[1];
    p->av = v_ab[0]; p->bv = v_ab[1];
    p->as = s_ab[0]; p->bs = s_ab[1];
    p->nQuarkIn = nQuarkIn;
    double bnum = beta_fn(p->av, p->bv + 1.);          /* int x^(a-1)(1-x)^b */
    p->Auv = 2.0 / bnum;
    p->Adv = 1.0 / bnum;
    double momV = (p->Auv + p->Adv) * beta_fn(p->av + 1., p->bv + 1.);
    p->Ag = 0.5 / beta_fn(p->ag + 1., p->bg + 1.);
    double momS = 1.0 - 0.5 - momV;
    p->As = momS / (2.0 * nQuarkIn * beta_fn(p->as + 1., p->bs + 1.));
}

double orc_toypdf_xf(const orc_toypdf *p, int id, double x) { return xf_d(p, id, x); }

/* ------------------------------------------------------------------------------------
 * Hadronic integrand, Pythia form.
 * ---------------------------------------------------------------------------------- */
void orc_hadcfg_default(orc_hadcfg *c)
{
    memset(c, 0, sizeof *c);
    c->eCM = 13600.0;
    c->pTHatMin = 20.0;
    orc_alphas_init(&c->as, 0.118, 1, 6);
    c->renormFac = 1.0;
    c->nQuarkIn = 5;
    c->nQuarkNew = 3;
    double g[2] = { -0.1, 5.0 }, v[2] = { 0.5, 3.0 }, s[2] = { -0.2, 7.0 };
    orc_toypdf_init(&c->pdf, g, v, s, c->nQuarkIn);
    c->procMask = 0x1F;                       /* 111..115 */
    c->kfactor = 1.0;
    c->mQ[0] = 1.5;                           /* ParticleData m0 of c and b */
    c->mQ[1] = 4.8;
    c->pdfSource = 0;
    c->factFac = 1.0;
    c->grid = NULL;
}

void orc_lhagrid_xf(const orc_lhagrid *g, double x, double Q2, double out[12]) { grid_xf_d(g, x, Q2, out); }

static double had_point_out(double total, const double *w, const double *kin, double wout[ORC_NPROC], orc_kin *k)
{
    for (int i = 0; i < ORC_NPROC; ++i) wout[i] = w[i];
    if (k) {
        k->sH = kin[0]; k->tH = kin[1]; k->uH = kin[2]; k->x1 = kin[3]; k->x2 = kin[4];
        k->pT2 = kin[5]; k->alpS = kin[6]; k->jac = kin[7];
    }
    return total;
}

double orc_had_point(const orc_hadcfg *c, const double u[3], double w[ORC_NPROC], orc_kin *k, int literalU)
{
    double ww[ORC_NPROC], kin[8];
    double tot = had_point_d(c, u, ww, kin, literalU);
    return had_point_out(tot, ww, kin, w, k);
}

double orc_had_point_ld(const orc_hadcfg *c, const double u[3], double w[ORC_NPROC], orc_kin *k, int literalU)
{
    long double ww[ORC_NPROC], kin[8];
    long double tot = had_point_ld(c, u, ww, kin, literalU);
    double wd[ORC_NPROC], kd[8];
    for (int i = 0; i < ORC_NPROC; ++i) wd[i] = (double)ww[i];
    for (int i = 0; i < 8; ++i) kd[i] = (double)kin[i];
    return had_point_out((double)tot, wd, kd, w, k);
}

/* ------------------------------------------------------------------------------------
 * Gauss-Legendre nodes (Newton on P_n, textbook).  Composite rules below.
 * ---------------------------------------------------------------------------------- */
static void gauss_legendre(int n, double *x, double *wt)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) { double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-15) break;
        }
        x[i] = -z; x[n - 1 - i] = z;
        wt[i] = wt[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

#define GL_ORDER 8

/* Deterministic quadrature of the hadronic sigma over the same cut, in the variables
 * v = ln tau (d tau / tau = dv), y, and rho = ln(a/b) with a = -tQ/sH, b = -uQ/sH, a + b = 1
 * (tQ = t - m^2, uQ = u - m^2; |dt| = sH a b d rho): the integrand is smooth in all three, so
 * composite Gauss-Legendre converges fast.  The pT cut is sH a b >= mT^2 = m^2 + pT2Min, i.e.
 * |rho| <= ln((1 + zM)/(1 - zM)) with zM = sqrt(1 - 4 mT^2 / sH); per family (massless, c cbar,
 * b bbar).  Independent of the MC map and of the RNG. */
static void quad_family(const orc_hadcfg *c, int fam, int n_tau, int n_y, int n_rho, double acc[ORC_NPROC])
{
    double gx[GL_ORDER], gw[GL_ORDER];
    gauss_legendre(GL_ORDER, gx, gw);
    double s = c->eCM * c->eCM;
    double m2 = fam == 0 ? 0.0 : c->mQ[fam - 1] * c->mQ[fam - 1];
    double mT2 = c->pTHatMin * c->pTHatMin + m2;
    double vMin = log(4.0 * mT2 / s), vMax = 0.0;
    int nQ = c->nQuarkIn, ig = 2 * nQ + 1;
    double K = c->kfactor * ORC_CONVERT2MB;
#pragma omp parallel
    {
        double loc[ORC_NPROC] = { 0 };
        double xfA[16], xfB[16];
#pragma omp for schedule(dynamic, 1)
        for (int it = 0; it < n_tau; ++it) {
            /* v = vMin + lam^2 removes the sqrt(v - vMin) threshold behaviour of the z range */
            double lMax = sqrt(vMax - vMin);
            double la = lMax * it / n_tau, lb = lMax * (it + 1) / n_tau;
            for (int gi = 0; gi < GL_ORDER; ++gi) {
                double lam = 0.5 * (la + lb) + 0.5 * (lb - la) * gx[gi];
                double v = vMin + lam * lam;
                double wv = 0.5 * (lb - la) * gw[gi] * 2.0 * lam;
                double tau = exp(v), sH = tau * s;
                double yMax = -0.5 * v;
                double r = 4.0 * mT2 / sH;
                if (r >= 1.0) continue;
                double zMax = sqrt(1.0 - r);
                double rhoM = log((1.0 + zMax) / (1.0 - zMax));
                for (int iy = 0; iy < n_y; ++iy) {
                    double ya = -yMax + 2.0 * yMax * iy / n_y, yb = -yMax + 2.0 * yMax * (iy + 1) / n_y;
                    for (int gj = 0; gj < GL_ORDER; ++gj) {
                        double y = 0.5 * (ya + yb) + 0.5 * (yb - ya) * gx[gj];
                        double wy = 0.5 * (yb - ya) * gw[gj];
                        double x1 = sqrt(tau) * exp(y), x2 = sqrt(tau) * exp(-y);
                        if (c->pdfSource != 1) {                  /* toy PDFs: no scale dependence */
                            fill_xf_d(c, x1, 0.0, xfA);
                            fill_xf_d(c, x2, 0.0, xfB);
                        }
                        for (int ir = 0; ir < n_rho; ++ir) {
                            double ra = -rhoM + 2.0 * rhoM * ir / n_rho, rb = -rhoM + 2.0 * rhoM * (ir + 1) / n_rho;
                            for (int gk = 0; gk < GL_ORDER; ++gk) {
                                double rho = 0.5 * (ra + rb) + 0.5 * (rb - ra) * gx[gk];
                                double wr = 0.5 * (rb - ra) * gw[gk];
                                /* a = e^rho / (1 + e^rho), b = 1 / (1 + e^rho), stably */
                                double a, b;
                                if (rho > 0) { double e = exp(-rho); a = 1.0 / (1.0 + e); b = e / (1.0 + e); }
                                else         { double e = exp(rho);  a = e / (1.0 + e);   b = 1.0 / (1.0 + e); }
                                double tH = -sH * a + m2, uH = -sH * b + m2;
                                double Q2 = sH * a * b;                   /* pT2 + m^2 (massless: pT2) */
                                double alpS = alphas_d(&c->as, c->renormFac * c->renormFac * Q2);
                                if (c->pdfSource == 1) {
                                    double Q2Fac = c->factFac * c->factFac * Q2;
                                    fill_xf_d(c, x1, Q2Fac, xfA);
                                    fill_xf_d(c, x2, Q2Fac, xfB);
                                }
                                double jac = wv * wy * wr * sH * a * b;
                                for (int p = 0; p < ORC_NPROC; ++p) {
                                    if (!(c->procMask & (1 << p))) continue;
                                    int f = p < 6 ? 0 : (p < 8 ? 1 : 2);
                                    if (f != fam) continue;
                                    int code = p < 6 ? 111 + p : 121 + (p - 6);
                                    double val;
                                    if (fam == 0) {
                                        val = sigma_pdf_d(c, code, sH, tH, uH, alpS, xfA, xfB);
                                    } else {
                                        double sig = K * orc_sigma_hat_massive(code, sH, tH, uH, m2, alpS);
                                        if (code == ORC_GG2CCBAR || code == ORC_GG2BBBAR) {
                                            val = sig * xfA[ig] * xfB[ig];
                                        } else {
                                            val = 0.0;
                                            for (int id = -nQ; id <= nQ; ++id)
                                                if (id) val += sig * xfA[id + nQ] * xfB[-id + nQ];
                                        }
                                    }
                                    loc[p] += jac * val;
                                }
                            }
                        }
                    }
                }
            }
        }
#pragma omp critical
        for (int p = 0; p < ORC_NPROC; ++p) acc[p] += loc[p];
    }
}

double orc_had_quadrature(const orc_hadcfg *c, int n_tau, int n_y, int n_rho, double out[ORC_NPROC])
{
    double acc[ORC_NPROC] = { 0 };
    if (c->procMask & 0x3F) quad_family(c, 0, n_tau, n_y, n_rho, acc);
    if (c->procMask & 0xC0) quad_family(c, 1, n_tau, n_y, n_rho, acc);
    if (c->procMask & 0x300) quad_family(c, 2, n_tau, n_y, n_rho, acc);
    double tot = 0.0;
    for (int p = 0; p < ORC_NPROC; ++p) { out[p] = acc[p]; tot += acc[p]; }
    return tot;
}

double orc_sigma_hat_massive(int code, double sH, double tH, double uH, double m2, double alpS)
{
    return sigma_hat_massive_d(code, sH, tH, uH, m2, alpS);
}

double orc_sigma_tot_massive_textbook(int code, double sH, double m2, double alpS)
{
    double rho = 4.0 * m2 / sH, beta = sqrt(1.0 - rho);
    if (code == ORC_QQBAR2CCBAR || code == ORC_QQBAR2BBBAR)
        return 8.0 * M_PI * alpS * alpS / (27.0 * sH) * beta * (1.0 + 0.5 * rho);
    return M_PI * alpS * alpS / (3.0 * sH) *
           ((1.0 + rho + rho * rho / 16.0) * log((1.0 + beta) / (1.0 - beta)) - beta * (7.0 / 4.0 + 31.0 * rho / 16.0));
}

/* ------------------------------------------------------------------------------------
 * Partonic gg->gg at fixed sHat (fused_gg2gg.cu:40-74).
 * ---------------------------------------------------------------------------------- */
double orc_c1_weight(double s, double alpS, double cMax, double c, double conv)
{
    double t = -0.5 * s * (1.0 - c);
    double u = -0.5 * s * (1.0 + c);
    return conv * orc_sigma_hat_pythia(ORC_GG2GG, s, t, u, alpS, 21, 21, 0) * (2.0 * cMax) * (s / 2.0);
}

/* Reference-form weight (uHat = -s - t, reciprocal-free Pythia literal), used by the
 * reference-compat checks of kernels 03/10. */
static double c1_weight_ref(double s, double alpS, double c)
{
    double t = -0.5 * s * (1.0 - c);
    return orc_sigma_hat_pythia(ORC_GG2GG, s, t, -s - t, alpS, 21, 21, 0);
}

double orc_c1_simpson(int N, double s, double alpS, double cMax, double conv)
{
    double a = -cMax, b = cMax, h = (b - a) / N, sum = 0.0;
    for (int i = 0; i <= N; ++i) {
        double c = a + i * h;
        double f = c1_weight_ref(s, alpS, c);
        double w = (i == 0 || i == N) ? 1.0 : (i % 2 ? 4.0 : 2.0);
        sum += w * f;
    }
    return (sum * h / 3.0) * (s / 2.0) * conv;
}

/* Closed form: with t = -s(1-c)/2, u = -s(1+c)/2 the bracket is
 * 9/4 (3 - (1-c^2)/4 + 2(1+c)/(1-c)^2 + 2(1-c)/(1+c)^2), integrated over [-c0, c0]. */
double orc_c1_closed(double s, double alpS, double cMax, double conv)
{
    double c0 = cMax;
    double I = 5.5 * c0 + c0 * c0 * c0 / 6.0
             + 4.0 * (4.0 * c0 / (1.0 - c0 * c0) - log((1.0 + c0) / (1.0 - c0)));
    I *= 9.0 / 4.0;
    return conv * M_PI * alpS * alpS / (2.0 * s) * I;
}

/* Neumaier-compensated accumulation. */
typedef struct { double s, c; } nsum;
static void nadd(nsum *a, double x)
{
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x; else a->c += (x - t) + a->s;
    a->s = t;
}

void orc_c1_mc_philox(uint64_t key, uint64_t counter0, uint64_t n, double s, double alpS,
                      double cMax, double conv, int nthreads, double sums[2])
{
    if (nthreads < 1) nthreads = 1;
    uint32_t k[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    double S1 = 0.0, S2 = 0.0;
    /* fixed partition into nthreads contiguous chunks, combined in chunk order */
    double *p1 = (double *)calloc((size_t)nthreads, sizeof(double));
    double *p2 = (double *)calloc((size_t)nthreads, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int t = 0; t < nthreads; ++t) {
        uint64_t lo = n * (uint64_t)t / (uint64_t)nthreads, hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        nsum a = { 0, 0 }, b = { 0, 0 };
        for (uint64_t i = lo; i < hi; ++i) {
            uint64_t e = counter0 + i;
            uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
            orc_philox4x32_10(c4, k, o);
            double uu = orc_u53(o[0], o[1]);
            double c = cMax * (2.0 * uu - 1.0);
            double w = orc_c1_weight(s, alpS, cMax, c, conv);
            nadd(&a, w);
            nadd(&b, w * w);
        }
        p1[t] = a.s + a.c;
        p2[t] = b.s + b.c;
    }
    for (int t = 0; t < nthreads; ++t) { S1 += p1[t]; S2 += p2[t]; }
    free(p1); free(p2);
    sums[0] = S1; sums[1] = S2;
}

double orc_c1_cpumc_compat(uint64_t seed, uint64_t total, double s, double alpS, double cMax, double conv)
{
    uint64_t ctr = seed;
    double sum = 0.0;
    for (uint64_t i = 0; i < total; ++i) {
        double c = (2.0 * orc_u01(orc_splitmix64(ctr++)) - 1.0) * cMax;
        sum += c1_weight_ref(s, alpS, c);
    }
    return (2.0 * cMax) * (s / 2.0) * (sum / (double)total) * conv;
}

double orc_c1_wmax_scan(int N, double s, double alpS, double cMax)
{
    double m = 0.0;
    for (int i = 0; i <= N; ++i) {
        double c = -cMax + (2.0 * cMax) * i / N;
        m = fmax(m, c1_weight_ref(s, alpS, c));
    }
    return m;
}

uint64_t orc_c1_unweight_compat(uint64_t seed, uint64_t n, double s, double alpS, double cMax,
                                double wMax, uint8_t *flags, double *cos_out)
{
    uint64_t acc = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t ctr = seed + i * 0x100000001B3ULL;
        double c = (2.0 * orc_u01(orc_splitmix64(ctr++)) - 1.0) * cMax;
        double w = c1_weight_ref(s, alpS, c);
        double u = orc_u01(orc_splitmix64(ctr++));
        uint8_t f = (u * wMax < w) ? 1 : 0;
        if (flags) flags[i] = f;
        if (cos_out) cos_out[i] = c;
        acc += f;
    }
    return acc;
}
