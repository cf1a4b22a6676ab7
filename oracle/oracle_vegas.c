/*
 * oracle_vegas.c -- CPU oracle: batched per-point evaluation through the flat or VEGAS map,
 * VEGAS adaptation loops, and the FP32 unweighting decisions.  TEST INFRASTRUCTURE ONLY (see
 * mcint_oracle.h).
 *
 * VEGAS: the adaptation loops sample through the grid, accumulate per-bin sums of (f J)^2 and
 * call the Lepage refinement of oracle_vegas_refine.c.  The reference's 1-D variant
 * (cuPythia/11_vegas_importance/vegas.cu:40-124: first moment, 3-point smoothing, no damping) is
 * restated here too, for its KAT.
 * Unweighting: accept iff u * wMax < w with u from a disjoint stream
 * (cuPythia/pipeline/generate.cu:56-62; Pythia ProcessContainer.cc:437).
 */
#include "mcint_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Map of one 24-bit uniform k through a 1-D grid (edges e[0..nbins]); FP64 point x, 1 - x of
 * the same point, Jacobian nbins * width. */
static void map64(uint32_t k, int nbins, const double *e, double *x, double *omx, double *jac)
{
    if (!e) {
        *x = (double)k * (1.0 / 16777216.0);
        *omx = (double)(16777216u - k) * (1.0 / 16777216.0);
        *jac = 1.0;
        return;
    }
    uint32_t kb = k * (uint32_t)nbins;
    int b = (int)(kb >> 24);
    double frac = (double)(kb & 0xFFFFFFu) * (1.0 / 16777216.0);
    double lo = e[b], wd = e[b + 1] - e[b];
    *x = fma(frac, wd, lo);
    *omx = 1.0 - *x;
    *jac = (double)nbins * wd;
}

void orc_had_points_ld(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                       uint64_t n, double *w_out, double *kin_out)
{
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t e = ctr0 + (uint64_t)i;
        uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
        orc_philox4x32_10(c4, kk, o);
        double u[3], om[3], jac = 1.0;
        for (int d = 0; d < 3; ++d) {
            double jd;
            map64(o[d] >> 8, nbins, edges ? edges + (size_t)d * (nbins + 1) : NULL, &u[d], &om[d], &jd);
            jac = jac * jd;
        }
        double w[ORC_NPROC];
        orc_kin k;
        double tot = orc_had_point_ld(c, u, w, &k, 0);
        w_out[(ORC_NPROC + 1) * i] = tot * jac;
        for (int p = 0; p < ORC_NPROC; ++p) w_out[(ORC_NPROC + 1) * i + 1 + p] = w[p] * jac;
        if (kin_out) {
            kin_out[4 * i + 0] = k.sH; kin_out[4 * i + 1] = k.tH;
            kin_out[4 * i + 2] = k.x1; kin_out[4 * i + 3] = k.x2;
        }
    }
}

/* ---------------------------------------------------------------- FP32 mirror, batched */
/* FP32 tables {lo, width, 1 - lo} from the FP64 edges, as the device builds them. */
static void tables32(int nbins, const double *edges, float *t)
{
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < nbins; ++b) {
            const double *e = edges + (size_t)d * (nbins + 1);
            float *q = t + 3 * ((size_t)d * nbins + b);
            q[0] = (float)e[b];
            q[1] = (float)(e[b + 1] - e[b]);
            q[2] = (float)(1.0 - e[b]);
        }
}

static float point32(const float *k, const orc32_grid *g, int mask, const uint32_t o[4], int nbins, const float *tab,
                     float w[ORC_NPROC], float kin[4], int bins[3])
{
    float u[3], om[3], jd[3];
    const float nbf = (float)nbins;
    for (int d = 0; d < 3; ++d) {
        uint32_t kk = o[d] >> 8;
        if (!tab) {
            u[d] = (float)kk * (1.0f / 16777216.0f);
            om[d] = (float)(16777216u - kk) * (1.0f / 16777216.0f);
            jd[d] = 1.0f;
            bins[d] = 0;
        } else {
            uint32_t kb = kk * (uint32_t)nbins;
            int b = (int)(kb >> 24);
            float frac = (float)(kb & 0xFFFFFFu) * (1.0f / 16777216.0f);
            const float *q = tab + 3 * ((size_t)d * nbins + b);
            u[d] = fmaf(frac, q[1], q[0]);
            om[d] = fmaf(-frac, q[1], q[2]);
            jd[d] = nbf * q[1];
            bins[d] = b;
        }
    }
    float tot = orc32_had_point_g(k, g, mask, u[0], om[0], u[1], om[1], u[2], om[2], w, kin);
    if (tab) {
        float jac = 1.0f;
        jac = jac * jd[0];
        jac = jac * jd[1];
        jac = jac * jd[2];
        for (int i = 0; i < ORC_NPROC; ++i) w[i] = w[i] * jac;
        tot = tot * jac;
    }
    return tot;
}

void orc32_had_points(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                      uint64_t n, float *w_out, float *kin_out)
{
    float k[ORC32_NCONST];
    orc32_consts_fill(c, k);
    orc32_grid g32;
    const orc32_grid *gp = NULL;
    if (c->pdfSource == 1 && c->grid) { orc32_grid_build(c->grid, c->factFac, &g32); gp = &g32; }
    float *tab = NULL;
    if (edges) { tab = (float *)malloc(sizeof(float) * 9 * nbins); tables32(nbins, edges, tab); }
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t e = ctr0 + (uint64_t)i;
        uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
        orc_philox4x32_10(c4, kk, o);
        float w[ORC_NPROC], kin[4];
        int bins[3];
        float tot = point32(k, gp, c->procMask, o, nbins, tab, w, kin, bins);
        w_out[(ORC_NPROC + 1) * i] = tot;
        for (int p = 0; p < ORC_NPROC; ++p) w_out[(ORC_NPROC + 1) * i + 1 + p] = w[p];
        if (kin_out) for (int q = 0; q < 4; ++q) kin_out[4 * i + q] = kin[q];
    }
    free(tab);
    if (gp) orc32_grid_free(&g32);
}

uint64_t orc32_unweight(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                        uint64_t n, double wmax, uint8_t *flags, int32_t *pid, float *w_out)
{
    float k[ORC32_NCONST];
    orc32_consts_fill(c, k);
    orc32_grid g32;
    const orc32_grid *gp = NULL;
    if (c->pdfSource == 1 && c->grid) { orc32_grid_build(c->grid, c->factFac, &g32); gp = &g32; }
    float *tab = NULL;
    if (edges) { tab = (float *)malloc(sizeof(float) * 9 * nbins); tables32(nbins, edges, tab); }
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    uint64_t acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (long long i = 0; i < (long long)n; ++i) {
        uint64_t e = ctr0 + (uint64_t)i;
        uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4], a[4];
        orc_philox4x32_10(c4, kk, o);
        float w[ORC_NPROC], kin[4];
        int bins[3];
        float tot = point32(k, gp, c->procMask, o, nbins, tab, w, kin, bins);
        uint32_t c5[4] = { (uint32_t)e, (uint32_t)(e >> 32), 1u, 0u };
        orc_philox4x32_10(c5, kk, a);
        double u = orc_u53(a[0], a[1]);
        int f = u * wmax < (double)tot;
        /* process choice: first code whose cumulative FP32 weight exceeds u24 * total */
        float target = (float)(a[2] >> 8) * (1.0f / 16777216.0f) * tot;
        float cum = 0.0f;
        int p = 0;
        for (int q = 0; q < ORC_NPROC; ++q) {
            if (!((c->procMask >> q) & 1) || !(w[q] > 0.0f)) continue;
            p = q < 6 ? 111 + q : 121 + (q - 6);
            cum = cum + w[q];
            if (cum > target) break;
        }
        if (flags) flags[i] = (uint8_t)f;
        if (pid) pid[i] = p;
        if (w_out) w_out[i] = tot;
        acc += (uint64_t)f;
    }
    free(tab);
    if (gp) orc32_grid_free(&g32);
    return acc;
}

/* FP32-mirror VEGAS adaptation of the hadronic integrand: iteration it samples events
 * ctr0 + it*n ... through the current grid, accumulates per-bin sums of w^2 (FP64) and refines.
 * sums_out[2*it] = sum w, [2*it+1] = sum w^2 of each iteration. */
void orc32_vegas_adapt(const orc_hadcfg *c, int nbins, double alpha, int n_iter, uint64_t key, uint64_t ctr0,
                       uint64_t n, double *edges, double *sums_out)
{
    float k[ORC32_NCONST];
    orc32_consts_fill(c, k);
    orc32_grid g32;
    const orc32_grid *gp = NULL;
    if (c->pdfSource == 1 && c->grid) { orc32_grid_build(c->grid, c->factFac, &g32); gp = &g32; }
    float *tab = (float *)malloc(sizeof(float) * 9 * nbins);
    double *hist = (double *)malloc(sizeof(double) * 3 * nbins);
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    for (int it = 0; it < n_iter; ++it) {
        tables32(nbins, edges, tab);
        memset(hist, 0, sizeof(double) * 3 * nbins);
        double sw = 0.0, sw2 = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t e = ctr0 + (uint64_t)it * n + i;
            uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
            orc_philox4x32_10(c4, kk, o);
            float w[ORC_NPROC], kin[4];
            int bins[3];
            double t = (double)point32(k, gp, c->procMask, o, nbins, tab, w, kin, bins);
            sw += t;
            sw2 += t * t;
            for (int d = 0; d < 3; ++d) hist[d * nbins + bins[d]] += t * t;
        }
        sums_out[2 * it] = sw;
        sums_out[2 * it + 1] = sw2;
        orc_vegas_refine(3, nbins, alpha, hist, edges);
    }
    free(tab);
    free(hist);
    if (gp) orc32_grid_free(&g32);
}

/* 1-D VEGAS on the partonic gg -> gg integrand (FP64, Philox stream 0, 53-bit uniform). */
void orc_c1_vegas_adapt(double s, double alpS, double cMax, double conv, int nbins, double alpha, int n_iter,
                        uint64_t key, uint64_t ctr0, uint64_t n, double *edges, double *sums_out)
{
    double *hist = (double *)malloc(sizeof(double) * nbins);
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    for (int it = 0; it < n_iter; ++it) {
        memset(hist, 0, sizeof(double) * nbins);
        double sw = 0.0, sw2 = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t e = ctr0 + (uint64_t)it * n + i;
            uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
            orc_philox4x32_10(c4, kk, o);
            uint64_t x53 = ((((uint64_t)o[1]) << 32) | o[0]) >> 11;
            unsigned __int128 kb = (unsigned __int128)x53 * (unsigned)nbins;
            int b = (int)(kb >> 53);
            double frac = (double)(uint64_t)(kb & ((((unsigned __int128)1) << 53) - 1)) * (1.0 / 9007199254740992.0);
            double lo = edges[b], wd = edges[b + 1] - edges[b];
            double u = fma(frac, wd, lo);
            double cth = cMax * (2.0 * u - 1.0);
            double w = orc_c1_weight(s, alpS, cMax, cth, conv) * ((double)nbins * wd);
            sw += w;
            sw2 += w * w;
            hist[b] += w * w;
        }
        sums_out[2 * it] = sw;
        sums_out[2 * it + 1] = sw2;
        orc_vegas_refine(1, nbins, alpha, hist, edges);
    }
    free(hist);
}

/* The reference's own 1-D VEGAS (cuPythia/11_vegas_importance/vegas.cu:40-124): gg -> gg over
 * cos(theta) in [-cMax, cMax], K bins, SplitMix64 streams ctr = seed_it + tid*0x100000001B3 with
 * two draws per sample (bin, position), importance sum f*d per bin, 3-point smoothing +
 * 1e-300, equal-share rebin, no damping; blocks x threads x Tper samples per iteration.
 * Returns eta of the first and last iteration and the last integral (GeV^-2 * conv). */
void orc_vegas1d_compat(int K, int blocks, int threads, uint64_t Tper, int n_iter, double s, double alpS,
                        double cMax, double conv, double out[3])
{
    double *grid = (double *)malloc(sizeof(double) * (K + 1));
    double *dacc = (double *)malloc(sizeof(double) * K);
    double *d = (double *)malloc(sizeof(double) * K);
    double *ng = (double *)malloc(sizeof(double) * (K + 1));
    for (int i = 0; i <= K; ++i) grid[i] = -cMax + (2.0 * cMax) * i / K;
    uint64_t nthr = (uint64_t)blocks * threads;
    double M = (double)nthr * (double)Tper;
    for (int it = 0; it < n_iter; ++it) {
        memset(dacc, 0, sizeof(double) * K);
        double sumW = 0.0, maxW = 0.0;
#pragma omp parallel
        {
            double *ld = (double *)calloc((size_t)K, sizeof(double));
            double ls = 0.0, lm = 0.0;
#pragma omp for schedule(static)
            for (long long tid = 0; tid < (long long)nthr; ++tid) {
                uint64_t ctr = 0xBEEF0000ULL + (uint64_t)it + (uint64_t)tid * 0x100000001B3ULL;
                for (uint64_t i = 0; i < Tper; ++i) {
                    double xi = orc_u01(orc_splitmix64(ctr++));
                    int b = (int)(xi * K);
                    if (b >= K) b = K - 1;
                    double lo = grid[b], hi = grid[b + 1], dd = hi - lo;
                    double c = lo + orc_u01(orc_splitmix64(ctr++)) * dd;
                    double t = -0.5 * s * (1.0 - c);
                    double f = orc_sigma_hat_pythia(ORC_GG2GG, s, t, -s - t, alpS, 21, 21, 0);
                    double w = f * (double)K * dd;
                    ls += w;
                    if (w > lm) lm = w;
                    ld[b] += f * dd;
                }
            }
#pragma omp critical
            {
                sumW += ls;
                if (lm > maxW) maxW = lm;
                for (int b = 0; b < K; ++b) dacc[b] += ld[b];
            }
            free(ld);
        }
        double meanW = sumW / M, eta = meanW / maxW;
        if (it == 0) out[0] = eta;
        out[1] = eta;
        out[2] = meanW * (s / 2.0) * conv;
        /* rebin, vegas.cu:69-85 */
        for (int i = 0; i < K; ++i) {
            double l = (i > 0) ? dacc[i - 1] : dacc[i], r = (i < K - 1) ? dacc[i + 1] : dacc[i];
            d[i] = (l + dacc[i] + r) / 3.0 + 1e-300;
        }
        double tot = 0.0;
        for (int i = 0; i < K; ++i) tot += d[i];
        double step = tot / K, cur = 0.0;
        int oj = 0;
        ng[0] = grid[0];
        ng[K] = grid[K];
        for (int k = 1; k < K; ++k) {
            double tgt = k * step;
            while (oj < K - 1 && cur + d[oj] < tgt) { cur += d[oj]; oj++; }
            double frac = (tgt - cur) / d[oj];
            if (frac < 0) frac = 0;
            if (frac > 1) frac = 1;
            ng[k] = grid[oj] + frac * (grid[oj + 1] - grid[oj]);
        }
        for (int i = 0; i <= K; ++i) grid[i] = ng[i];
    }
    free(grid); free(dacc); free(d); free(ng);
}

/* CPU baseline of the hadronic workload: the Pythia-form FP64 point (flux lists, literal
 * sigmaHat) over events ctr0 .. ctr0+n-1 with the flat map, Philox stream 0; nthreads
 * contiguous chunks (OpenMP), combined in chunk order.  sums = {sum w, sum w^2}. */
void orc_had_mc(const orc_hadcfg *c, uint64_t key, uint64_t ctr0, uint64_t n, int nthreads, double sums[2])
{
    if (nthreads < 1) nthreads = 1;
    uint32_t kk[2] = { (uint32_t)key, (uint32_t)(key >> 32) };
    double *p1 = (double *)calloc((size_t)nthreads, sizeof(double));
    double *p2 = (double *)calloc((size_t)nthreads, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int t = 0; t < nthreads; ++t) {
        uint64_t lo = n * (uint64_t)t / (uint64_t)nthreads, hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        double a = 0.0, b = 0.0;
        for (uint64_t i = lo; i < hi; ++i) {
            uint64_t e = ctr0 + i;
            uint32_t c4[4] = { (uint32_t)e, (uint32_t)(e >> 32), 0u, 0u }, o[4];
            orc_philox4x32_10(c4, kk, o);
            double u[3] = { (double)(o[0] >> 8) * (1.0 / 16777216.0), (double)(o[1] >> 8) * (1.0 / 16777216.0),
                            (double)(o[2] >> 8) * (1.0 / 16777216.0) };
            double w[ORC_NPROC];
            double tot = orc_had_point(c, u, w, NULL, 1);
            a += tot;
            b += tot * tot;
        }
        p1[t] = a;
        p2[t] = b;
    }
    sums[0] = sums[1] = 0.0;
    for (int t = 0; t < nthreads; ++t) { sums[0] += p1[t]; sums[1] += p2[t]; }
    free(p1);
    free(p2);
}
