/*
 * oracle_fp32.c -- bit-for-bit CPU restatement of the FP32 device arithmetic of the
 * hadronic integrand.  TEST INFRASTRUCTURE ONLY (see mcint_oracle.h).
 *
 * Purpose: the unweighting decision u * wMax < w must come out identically on the CPU and
 * on the GPU for the compacted event indices to be bit-exact, so the FP32 weight has to be
 * reproducible bit for bit.  Both sides therefore use only correctly rounded IEEE
 * operations (+ - * / sqrt fma rint and exact bit manipulation), the same operation order,
 * no contraction (this file is compiled with -ffp-contract=off), and this project's own
 * exp/log:
 *   exp: Cody-Waite reduction x = n ln2 + r (ln2 split hi/lo; W. J. Cody, W. Waite, "Software
 *        Manual for the Elementary Functions", 1980), e^r on |r| <= ln2/2 by a degree-6
 *        polynomial fitted for minimum relative error with Lawson's iteratively reweighted
 *        least squares (coefficients rounded to FP32 one at a time, the rest refitted;
 *        relative error 3.1e-9), scaling by 2^n with ldexpf.
 *   1 - e^x near 0: -x q(x), q ~ (e^x - 1)/x of degree 5 by the same fit (2.5e-10).
 *   log: table-driven (P. T. P. Tang, ACM TOMS 16 (1990) 378-400): x = 2^e m, m in [1, 2),
 *        j = top 7 bits of m, G_j = RN(1/(1 + (j + 1/2)/128)), L_j = RN(-log G_j - k_j ln2)
 *        with k_j = [j >= 53] folding the upper entries into [0.707, 1); r = fma(m, G_j, -1),
 *        |r| <= 2^-8, log1p r ~ r - r^2/2 + r^3/3; log x = (e + k_j) ln2 + L_j + log1p r.
 * The accuracy of the whole FP32 weight is pinned against the FP64 Pythia-form oracle
 * (orc_had_point) in tests/test_oracle_physics.py.
 */
#include "mcint_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static float i2f(int32_t i) { float f; memcpy(&f, &i, 4); return f; }
static int32_t f2i(float f) { int32_t i; memcpy(&i, &f, 4); return i; }

float orc_expf(float x)
{
    if (x < -86.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float n = rintf(x * 1.4426950408889634f);
    float r = fmaf(n, -0.693145751953125f, x);
    r = fmaf(n, -1.4286068203094172e-6f, r);
    float p = 0.001382148f;
    p = fmaf(p, r, 0.008368663f);
    p = fmaf(p, r, 0.041668255f);
    p = fmaf(p, r, 0.16666521f);
    p = fmaf(p, r, 0.49999994f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)n);
}

/* 1 - e^x for x <= 0: -x q(x) when x > -0.35 (no cancellation near x = 0), plain 1 - exp(x)
 * otherwise (then 1 - e^x >= 0.295 and the subtraction is benign). */
static float one_minus_exp(float x, float ex)
{
    if (x > -0.35f) {
        float p = fmaf(x, 0.0011900244f, 0.008254852f);
        p = fmaf(x, p, 0.041652385f);
        p = fmaf(x, p, 0.16666552f);
        p = fmaf(x, p, 0.49999997f);
        p = fmaf(x, p, 1.0f);
        return -(x * p);
    }
    return 1.0f - ex;
}

void orc32_logtab(float G[128], float L[128])
{
    for (int j = 0; j < 128; ++j) {
        G[j] = (float)(1.0 / (1.0 + (j + 0.5) / 128.0));
        L[j] = (float)(-log((double)G[j]) - (j >= 53 ? log(2.0) : 0.0));
    }
}

static float logG[128], logL[128];
__attribute__((constructor)) static void init_logtab(void) { orc32_logtab(logG, logL); }

float orc_logf(float x)
{
    if (!(x > 0.0f)) return -1.0e30f;           /* only reached for 1 - x == 0 (x rounds to 1) */
    uint32_t ix = (uint32_t)f2i(x);
    float fe = (float)((ix + 0x004B0000u) >> 23) - 127.0f;
    uint32_t j = (ix >> 16) & 127u;
    float m = i2f((int32_t)((ix & 0x007fffffu) | 0x3f800000u));
    float r = fmaf(m, logG[j], -1.0f);
    float r2 = r * r;
    float p = fmaf(r2, fmaf(r, 0.333333343f, -0.5f), r);
    float u = logL[j] + fmaf(fe, 1.4286068203094172e-6f, p);
    return fmaf(fe, 0.693145751953125f, u);
}

/* Constant block (all FP32):
 *  0 LT = ln(1/tauMin)       1 fourPT2min = 4 pTmin^2     2 pref = CONV pi LT K / (4 pTmin^2)
 *  3 xiR2 = renormFac^2      4 as_order (0/1)             5 as_fixed (order 0 value)
 *  6 scale2Min               7 mt2  8 mb2  9 mc2          10 nf6 (1 if nfmax >= 6)
 * 11 lnL6s 12 lnL5s 13 lnL4s 14 lnL3s                     15 c6 16 c5 17 c4 18 c3 (12 pi / b0)
 * 19 lnAg 20 ag 21 bg        22 lnAs 23 as 24 bs          25 lnAuv 26 av 27 bv 28 rdu = Adv/Auv
 * 29 nSea = 2 nQuarkIn       30 nQuarkNew                 31 unused
 * 32..36 the c cbar family: LT, 4 mT^2 (mT^2 = m_c^2 + pTmin^2), pref with 4 mT^2, m_c^2,
 *        m_c^2 / (4 mT^2)
 * 37..41 the b bbar family: the same with m_b
 */
void orc32_consts_fill(const orc_hadcfg *c, float k[ORC32_NCONST])
{
    double s = c->eCM * c->eCM;
    double fourPT2 = 4.0 * c->pTHatMin * c->pTHatMin;
    double LT = log(s / fourPT2);
    for (int i = 0; i < ORC32_NCONST; ++i) k[i] = 0.0f;
    k[0] = (float)LT;
    k[1] = (float)fourPT2;
    k[2] = (float)(ORC_CONVERT2MB * M_PI * LT * c->kfactor / fourPT2);
    k[3] = (float)(c->renormFac * c->renormFac);
    k[4] = (float)c->as.order;
    k[5] = (float)c->as.valueRef;
    k[6] = (float)c->as.scale2Min;
    k[7] = (float)c->as.mt2; k[8] = (float)c->as.mb2; k[9] = (float)c->as.mc2;
    k[10] = (c->as.nfmax >= 6) ? 1.0f : 0.0f;
    if (c->as.order == 1) {
        k[11] = (float)log(c->as.L6s); k[12] = (float)log(c->as.L5s);
        k[13] = (float)log(c->as.L4s); k[14] = (float)log(c->as.L3s);
    }
    k[15] = (float)(12.0 * M_PI / 21.0); k[16] = (float)(12.0 * M_PI / 23.0);
    k[17] = (float)(12.0 * M_PI / 25.0); k[18] = (float)(12.0 * M_PI / 27.0);
    k[19] = (float)log(c->pdf.Ag);  k[20] = (float)c->pdf.ag; k[21] = (float)c->pdf.bg;
    k[22] = (float)log(c->pdf.As);  k[23] = (float)c->pdf.as; k[24] = (float)c->pdf.bs;
    k[25] = (float)log(c->pdf.Auv); k[26] = (float)c->pdf.av; k[27] = (float)c->pdf.bv;
    k[28] = (float)(c->pdf.Adv / c->pdf.Auv);
    k[29] = (float)(2 * c->nQuarkIn);
    k[30] = (float)c->nQuarkNew;
    for (int f = 0; f < 2; ++f) {
        double m2 = c->mQ[f] * c->mQ[f];
        double fourMT2 = 4.0 * (m2 + c->pTHatMin * c->pTHatMin);
        double LTf = log(s / fourMT2);
        k[32 + 5 * f] = (float)LTf;
        k[33 + 5 * f] = (float)fourMT2;
        k[34 + 5 * f] = (float)(ORC_CONVERT2MB * M_PI * LTf * c->kfactor / fourMT2);
        k[35 + 5 * f] = (float)m2;
        k[36 + 5 * f] = (float)(m2 / fourMT2);
    }
}

/* The device kernel's per-point arithmetic, restated operation by operation. */
float orc32_had_point(const float k[ORC32_NCONST], int procMask, float u0, float u1, float u2,
                      float w[ORC_NPROC], float kin[4])
{
    return orc32_had_point_om(k, procMask, u0, 1.0f - u0, u1, 1.0f - u1, u2, 1.0f - u2, w, kin);
}

/* ------------------------------------------------------------------ FP32 PDF grid */
static float *f32copy(const double *a, size_t n)
{
    float *o = (float *)malloc(sizeof(float) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) o[i] = (float)a[i];
    return o;
}

static float *lagrange_inv32(const double *t, int n)
{
    float *inv = (float *)calloc((size_t)ORC_GRID_NF * n, sizeof(float));
    for (int m = 0; m + 3 < n; ++m) {
        int k = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                if (j != i) inv[(size_t)ORC_GRID_NF * m + k++] = (float)(1.0 / (t[m + i] - t[m + j]));
    }
    return inv;
}

void orc32_grid_build(const orc_lhagrid *g, double factFac, orc32_grid *o)
{
    memset(o, 0, sizeof *o);
    o->nx = g->nx; o->nq = g->nq; o->nqSub = g->nqSub;
    o->xk = f32copy(g->x, (size_t)g->nx);
    o->lnx = f32copy(g->lnx, (size_t)g->nx);
    o->qk = f32copy(g->q, (size_t)g->nq);
    o->lnq = f32copy(g->lnq, (size_t)g->nq);
    o->val = f32copy(g->val, (size_t)g->nq * g->nx * ORC_GRID_NF);
    o->ixinv = lagrange_inv32(g->lnx, g->nx);
    o->iqinv = lagrange_inv32(g->lnq, g->nq);
    for (int s = 0; s < g->nqSub; ++s) {
        int minS = s == 0 ? 0 : g->nqSum[s - 1], maxS = g->nqSum[s] - 1;
        if (maxS - minS < 3)
            for (int m = minS; m < maxS; ++m) o->iqinv[(size_t)ORC_GRID_NF * m] = (float)(1.0 / (g->lnq[m + 1] - g->lnq[m]));
    }
    for (int s = 0; s < ORC_GRID_MAXSUB; ++s) { o->nqSum[s] = g->nqSum[s]; o->qDiv[s] = (float)g->qDiv[s]; }
    o->xMin = (float)g->x[0]; o->xMax = (float)g->x[g->nx - 1];
    o->qMin = (float)g->q[0]; o->qMax = (float)g->q[g->nq - 1];
    o->factFac2 = (float)(factFac * factFac);
}

void orc32_grid_free(orc32_grid *g)
{
    free(g->xk); free(g->lnx); free(g->qk); free(g->lnq); free(g->val); free(g->ixinv); free(g->iqinv);
    memset(g, 0, sizeof *g);
}

/* The device's FP32 grid evaluation, operation for operation (GridPdf, grid_qw, grid_xf,
 * grid_lumi in the device header): Lagrange weights by reciprocals, slots 0..10. */
static void lagrange4_32(float t, const float *n, const float *inv, float *w)
{
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        w[i] = 1.0f;
        for (int j = 0; j < 4; ++j)
            if (j != i) { w[i] *= (t - n[j]) * inv[k]; ++k; }
    }
}

static void grid_qw32(const orc32_grid *g, float q, float lnq, int *m3q, int *n3q, float *wq)
{
    wq[0] = wq[1] = wq[2] = wq[3] = 1.0f;
    if (q <= g->qMin) { *n3q = 1; *m3q = 0; return; }
    if (q >= g->qMax) { *n3q = 1; *m3q = g->nq - 1; return; }
    int iqDiv = 0;
    for (int s = 1; s < g->nqSub; ++s)
        if (q > g->qDiv[s - 1]) iqDiv = s;
    int minS = iqDiv == 0 ? 0 : g->nqSum[iqDiv - 1], maxS = g->nqSum[iqDiv] - 1;
    int minq = minS, maxq = maxS;
    while (maxq - minq > 1) {
        int mid = (minq + maxq) >> 1;
        if (q < g->qk[mid]) maxq = mid;
        else                minq = mid;
    }
    if (maxS - minS < 3) {
        *n3q = 2;
        *m3q = minq;
        wq[1] = (lnq - g->lnq[minq]) * g->iqinv[ORC_GRID_NF * minq];
        wq[0] = 1.0f - wq[1];
    } else {
        *n3q = 4;
        *m3q = (minq == minS) ? minS : ((maxq == maxS) ? maxS - 3 : minq - 1);
        lagrange4_32(lnq, g->lnq + *m3q, g->iqinv + ORC_GRID_NF * *m3q, wq);
    }
}

static void grid_xf32(const orc32_grid *g, float x, float lnx, int m3q, int n3q, const float *wq, float *out)
{
    for (int f = 0; f < 11; ++f) out[f] = 0.0f;
    if (x <= g->xMin) {
        for (int i3q = 0; i3q < n3q; ++i3q) {
            const float *p = g->val + (size_t)(m3q + i3q) * g->nx * ORC_GRID_NF;
            for (int f = 0; f < 11; ++f) out[f] = fmaf(wq[i3q], p[f], out[f]);
        }
        return;
    }
    if (x >= g->xMax) return;
    int minx = 0, maxx = g->nx - 1;
    while (maxx - minx > 1) {
        int mid = (minx + maxx) >> 1;
        if (x < g->xk[mid]) maxx = mid;
        else                minx = mid;
    }
    int m3x = (minx == 0) ? 0 : ((maxx == g->nx - 1) ? g->nx - 4 : minx - 1);
    float wx[4];
    lagrange4_32(lnx, g->lnx + m3x, g->ixinv + ORC_GRID_NF * m3x, wx);
    /* node by node with the combined weight wq * wx, one fma per flavour */
    for (int i3q = 0; i3q < n3q; ++i3q) {
        const float *p = g->val + ((size_t)(m3q + i3q) * g->nx + m3x) * ORC_GRID_NF;
        for (int ix = 0; ix < 4; ++ix) {
            float w = wq[i3q] * wx[ix];
            for (int f = 0; f < 11; ++f) out[f] = fmaf(w, p[ix * ORC_GRID_NF + f], out[f]);
        }
    }
}

/* luminosities gg, qg, QQ, same, opp from the grid at (x1, x2), q (grid_lumi) */
static void grid_lumi32(const orc32_grid *g, int nQ, float x1, float lx1, float x2, float lx2, float q, float lnq,
                        float L[5])
{
    int m3q, n3q;
    float wq[4], f1[11], f2[11];
    grid_qw32(g, q, lnq, &m3q, &n3q, wq);
    grid_xf32(g, x1, lx1, m3q, n3q, wq, f1);
    grid_xf32(g, x2, lx2, m3q, n3q, wq, f2);
    for (int f = 0; f < 11; ++f) {
        f1[f] = f1[f] > 0.0f ? f1[f] : 0.0f;
        f2[f] = f2[f] > 0.0f ? f2[f] : 0.0f;
    }
    float Q1 = 0.0f, Q2 = 0.0f, same = 0.0f, opp = 0.0f;
    for (int k = 1; k <= nQ; ++k) {
        float q1 = f1[k], a1 = f1[5 + k], q2 = f2[k], a2 = f2[5 + k];
        Q1 = Q1 + (q1 + a1);
        Q2 = Q2 + (q2 + a2);
        same = same + (q1 * q2 + a1 * a2);
        opp = opp + (q1 * a2 + a1 * q2);
    }
    L[0] = f1[0] * f2[0];
    L[1] = Q1 * f2[0] + f1[0] * Q2;
    L[2] = Q1 * Q2;
    L[3] = same;
    L[4] = opp;
}

float orc32_had_point_om(const float k[ORC32_NCONST], int procMask, float u0, float om0, float u1,
                         float om1, float u2, float om2, float w[ORC_NPROC], float kin[4])
{
    return orc32_had_point_g(k, NULL, procMask, u0, om0, u1, om1, u2, om2, w, kin);
}

/* One family: 0 = the massless codes 111..116, 1 = c cbar (121, 122), 2 = b bbar (123, 124).
 * The massive map is the massless one with pTmin^2 replaced by mT^2 = m^2 + pTmin^2 (a, b and
 * the Jacobian are unchanged), Q^2 = sH a b for both scales, tH = m^2 - sH a, mu = m^2 / sH. */
static float family32(const float *k, const orc32_grid *g, int fam, int procMask, float u0, float om0,
                      float u1, float om1, float u2, float om2, float w[ORC_NPROC], float kin[4])
{
    const float *kf = fam == 0 ? k : k + 27 + 5 * fam;
    const float LT = kf[0];
    const float m2 = fam == 0 ? 0.0f : kf[3];
    float r = orc_expf(-(u0 * LT));                    /* tauMin / tau */
    float yMax = 0.5f * (LT * om0);
    /* ln x1 = y - yMax = -2 yMax (1 - u1), ln x2 = -y - yMax = -2 yMax u1: no cancellation */
    float twoYMax = yMax + yMax;
    float lx1 = -(twoYMax * om1), lx2 = -(twoYMax * u1);
    float x1 = orc_expf(lx1), x2 = orc_expf(lx2);
    float zMax = sqrtf(one_minus_exp(-(u0 * LT), r));
    /* one division for 1/r and 1/(1 + zMax) */
    float opz = 1.0f + zMax;
    float R1 = 1.0f / (r * opz);
    float omz = (r * r) * R1;                           /* r/(1 + zMax) = 1 - zMax, cancellation-free */
    float sH = (kf[1] * opz) * R1;                      /* 4 mT^2 / r */
    float twoz = zMax + zMax;
    float A = fmaf(twoz, om2, omz);                     /* 1 - z */
    float B = fmaf(twoz, u2, omz);                      /* 1 + z */
    float a = 0.5f * A, b = 0.5f * B;                   /* -(tH - m^2)/s, -(uH - m^2)/s */
    float ab = a * b;
    float pT2 = sH * ab;                                /* pT^2 + m^2 */

    float alpS, inv_ab;
    if (k[4] == 0.0f) {
        alpS = k[5];
        inv_ab = 1.0f / ab;
    } else {
        float Q2 = k[3] * pT2;
        if (Q2 < k[6]) Q2 = k[6];
        float lq = orc_logf(Q2);
        float lam, cf;
        if (Q2 > k[7] && k[10] != 0.0f) { lam = k[11]; cf = k[15]; }
        else if (Q2 > k[8])             { lam = k[12]; cf = k[16]; }
        else if (Q2 > k[9])             { lam = k[13]; cf = k[17]; }
        else                            { lam = k[14]; cf = k[18]; }
        /* one division for 1/(a b) and cf/(lq - lam) */
        float den = lq - lam;
        float R2 = 1.0f / (ab * den);
        inv_ab = den * R2;
        alpS = (cf * ab) * R2;
    }
    float ia = b * inv_ab, ib = a * inv_ab;

    /* parton densities at x1 and x2 */
    float Lgg, Lqg, LQQ, Lsame, Lopp;
    if (g) {
        float Q2F = g->factFac2 * pT2;
        float L[5];
        grid_lumi32(g, (int)(0.5f * k[29]), x1, lx1, x2, lx2, sqrtf(Q2F), 0.5f * orc_logf(Q2F), L);
        Lgg = L[0]; Lqg = L[1]; LQQ = L[2]; Lsame = L[3]; Lopp = L[4];
    } else {
        float l1a = orc_logf(one_minus_exp(lx1, x1)), l1b = orc_logf(one_minus_exp(lx2, x2));
        float g1 = orc_expf(fmaf(k[21], l1a, fmaf(k[20], lx1, k[19])));
        float g2 = orc_expf(fmaf(k[21], l1b, fmaf(k[20], lx2, k[19])));
        float S1 = orc_expf(fmaf(k[24], l1a, fmaf(k[23], lx1, k[22])));
        float S2 = orc_expf(fmaf(k[24], l1b, fmaf(k[23], lx2, k[22])));
        float V1 = orc_expf(fmaf(k[27], l1a, fmaf(k[26], lx1, k[25])));
        float V2 = orc_expf(fmaf(k[27], l1b, fmaf(k[26], lx2, k[25])));
        float Vd1 = k[28] * V1, Vd2 = k[28] * V2;
        float U1 = V1 + S1, U2 = V2 + S2, D1 = Vd1 + S1, D2 = Vd2 + S2;
        float Q1 = fmaf(k[29], S1, V1 + Vd1), Q2s = fmaf(k[29], S2, V2 + Vd2);
        float SS = S1 * S2;
        Lgg = g1 * g2;
        Lqg = fmaf(Q1, g2, g1 * Q2s);
        LQQ = Q1 * Q2s;
        Lsame = fmaf(U1, U2, fmaf(D1, D2, (k[29] - 2.0f) * SS));
        Lopp = fmaf(U1 + D1, S2, fmaf(S1, U2 + D2, (k[29] - 4.0f) * SS));
    }

    /* brackets: d sigma / dt = pi alpS^2 / sH^2 * bracket */
    float q2 = fmaf(a, a, b * b);
    float common = kf[2] * (alpS * alpS) * (LT * om0) * zMax * r;
    float ws[ORC_NPROC];
    int lo, hi;
    if (fam == 0) {
        lo = 0; hi = 6;
        ws[0] = Lgg * (2.25f * fmaf(b * ia, ia, fmaf(a * ib, ib, 3.0f - ab)));
        ws[1] = Lgg * (k[30] * q2 * fmaf(inv_ab, 0.16666666666666667f, -0.375f));
        ws[2] = Lqg * (fmaf(b, b, 1.0f) * fmaf(ia, ia, 0.44444444444444444f * ib));
        {
            float sigT = 0.44444444444444444f * fmaf(b, b, 1.0f) * (ia * ia);
            float sigU = 0.44444444444444444f * fmaf(a, a, 1.0f) * (ib * ib);
            float sigTU = -0.29629629629629630f * inv_ab;
            float sigST = 0.29629629629629630f * (b * b) * ia;
            float csame = 0.5f * (sigU + sigTU - sigT);
            ws[3] = fmaf(sigT, LQQ, fmaf(csame, Lsame, sigST * Lopp));
        }
        ws[4] = Lopp * (q2 * fmaf(inv_ab, 0.59259259259259259f, -1.3333333333333333f));
        ws[5] = Lopp * (k[30] * 0.44444444444444444f * q2);
    } else {
        /* Pythia's massive forms (SigmaQCD.cc Sigma2gg2QQbar::sigmaKin, Sigma2qqbar2QQbar::
         * sigmaKin) with tHQ = -sH a, uHQ = -sH b and mu = m^2 / sH = r m^2 / (4 mT^2):
         * gg:   [b/a - 9/4 b^2 + mu/a^2 (a (9/2 b - 1/2 + mu) - 4 mu)  +  (a <-> b)] / 6
         * qqb:  4/9 (a^2 + b^2 + 2 mu) */
        lo = 4 + 2 * fam; hi = lo + 2;
        float mu = kf[4] * r;
        float cT = fmaf(a, fmaf(4.5f, b, -0.5f) + mu, -4.0f * mu);
        float cU = fmaf(b, fmaf(4.5f, a, -0.5f) + mu, -4.0f * mu);
        float sTS = fmaf(b, ia, fmaf(-2.25f * b, b, (mu * (ia * ia)) * cT));
        float sUS = fmaf(a, ib, fmaf(-2.25f * a, a, (mu * (ib * ib)) * cU));
        ws[lo] = Lgg * (0.16666666666666667f * (sTS + sUS));
        ws[lo + 1] = Lopp * (0.44444444444444444f * fmaf(2.0f, mu, q2));
    }
    float tot = 0.0f;
    for (int i = lo; i < hi; ++i) {
        float wi = (procMask & (1 << i)) ? common * ws[i] : 0.0f;
        wi = (wi > 0.0f || wi != wi) ? wi : 0.0f;              /* negative -> 0, NaN kept */
        w[i] = wi;
        tot += wi;
    }
    if (kin) { kin[0] = sH; kin[1] = fmaf(-sH, a, m2); kin[2] = x1; kin[3] = x2; }
    return tot;
}

float orc32_had_point_g(const float k[ORC32_NCONST], const orc32_grid *g, int procMask, float u0, float om0,
                        float u1, float om1, float u2, float om2, float w[ORC_NPROC], float kin[4])
{
    static const int famMask[3] = { 0x3F, 0xC0, 0x300 };
    for (int i = 0; i < ORC_NPROC; ++i) w[i] = 0.0f;
    float tot = 0.0f;
    int have_kin = 0;
    for (int fam = 0; fam < 3; ++fam) {
        if (!(procMask & famMask[fam])) continue;
        float kf[4];
        tot += family32(k, g, fam, procMask, u0, om0, u1, om1, u2, om2, w, kf);
        if (kin && !have_kin) { for (int q = 0; q < 4; ++q) kin[q] = kf[q]; have_kin = 1; }
    }
    if (kin && !have_kin) for (int q = 0; q < 4; ++q) kin[q] = 0.0f;
    return tot;
}
