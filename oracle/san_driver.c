/*
 * san_driver.c -- exercises the CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
 * (oracle/Makefile target `sanitize`, run by tests/test_oracle_sanitize.py).  TEST INFRASTRUCTURE
 * ONLY: host code, built with gcc, never part of the product.
 *
 * Small sizes of every batched entry point: Philox and SplitMix streams, the Pythia-form and
 * long-double hadronic points (all ten codes, flat and VEGAS maps, toy and grid PDFs), the FP32
 * mirror, unweighting, VEGAS adaptation and refinement, the adjoint, quadrature, the partonic
 * reference paths and the event record.  argv[1]: an lhagrid1 file (optional).  Exit status 0 and
 * a sanitizer-clean run are the result; the printed checksum only keeps the calls alive.
 */
#include "mcint_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static double sum(const double *a, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i];
    return s;
}

int main(int argc, char **argv)
{
    double chk = 0.0;
    enum { N = 512, NB = 16 };
    uint32_t words[4 * N];
    orc_philox_fill(42, 7, 0, N, words);
    for (int i = 0; i < N; ++i) chk += orc_u53(words[4 * i], words[4 * i + 1]) + orc_u24(words[4 * i + 2]);
    chk += orc_u01(orc_splitmix64(123));

    orc_lhagrid *grid = NULL;
    if (argc > 1 && orc_lhagrid_load(argv[1], &grid) != 0) {
        fprintf(stderr, "cannot load %s\n", argv[1]);
        return 2;
    }
    double *edges = malloc(sizeof(double) * 3 * (NB + 1));
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b <= NB; ++b) edges[d * (NB + 1) + b] = pow((double)b / NB, 1.0 + 0.3 * d);
    double *w = malloc(sizeof(double) * N * (ORC_NPROC + 1)), *kin = malloc(sizeof(double) * N * 4);
    float *w32 = malloc(sizeof(float) * N * (ORC_NPROC + 1)), *k32 = malloc(sizeof(float) * N * 4);
    uint8_t flags[N];
    int32_t pid[N];
    float wu[N];
    for (int src = 0; src < (grid ? 2 : 1); ++src) {
        orc_hadcfg c;
        orc_hadcfg_default(&c);
        c.procMask = 0x3FF;
        c.pTHatMin = 10.0;
        c.renormFac = 0.7;
        if (src == 1) { c.pdfSource = 1; c.grid = grid; c.factFac = 1.3; }
        for (int vg = 0; vg < 2; ++vg) {
            const double *e = vg ? edges : NULL;
            const int nb = vg ? NB : 0;
            orc_had_points_ld(&c, nb, e, 42, 0, N, w, kin);
            chk += sum(w, (size_t)N * (ORC_NPROC + 1));
            orc32_had_points(&c, nb, e, 42, 0, N, w32, k32);
            for (int i = 0; i < N * (ORC_NPROC + 1); ++i) chk += w32[i];
            chk += (double)orc32_unweight(&c, nb, e, 42, 0, N, 1e-3, flags, pid, wu);
        }
        double u[3] = {0.3, 0.6, 0.9}, wp[ORC_NPROC];
        orc_kin k;
        chk += orc_had_point(&c, u, wp, &k, 1) + orc_had_point_ld(&c, u, wp, &k, 0);
        double e2[3 * (NB + 1)], sums[4];
        for (int i = 0; i < 3 * (NB + 1); ++i) e2[i] = (double)(i % (NB + 1)) / NB;
        orc32_vegas_adapt(&c, NB, 1.5, 2, 42, 0, N, e2, sums);
        chk += sum(e2, 3 * (NB + 1)) + sum(sums, 4);
        double q[ORC_NPROC];
        chk += orc_had_quadrature(&c, 3, 3, 4, q);
        double adj[9];
        orc_adjoint(&c, 42, 0, 64, adj);
        chk += sum(adj, 9);
        uint64_t idx[8];
        int32_t code[8];
        double sH[8], tH[8], x1[8], x2[8];
        for (int i = 0; i < 8; ++i) {
            idx[i] = (uint64_t)(3 * i + 1);
            code[i] = i < 6 ? 111 + i : 121 + (i - 6);
            x1[i] = 0.01 + 0.02 * i; x2[i] = 0.02 + 0.01 * i;
            sH[i] = x1[i] * x2[i] * c.eCM * c.eCM;
            tH[i] = -0.4 * sH[i];
        }
        int32_t ints[8 * 4 * 8], np[8];
        double mom[8 * 4 * 5], sc[8];
        orc_build_events(&c, 0, 0.0, 5, 8, idx, code, sH, tH, x1, x2, ints, mom, sc, np);
        chk += sum(mom, 8 * 4 * 5) + sum(sc, 8);
        double hs[3 * NB];
        for (int i = 0; i < 3 * NB; ++i) hs[i] = 1.0 + (i % 5);
        orc_vegas_refine(3, NB, 1.5, hs, e2);
        chk += sum(e2, 3 * (NB + 1));
    }
    if (grid) {
        double xf[12];
        orc_lhagrid_xf(grid, 1e-3, 100.0, xf);
        chk += sum(xf, 12);
    }
    double s2[2];
    orc_c1_mc_philox(42, 0, N, 1e4, 0.118, 0.9, ORC_CONVERT2MB, 2, s2);
    chk += s2[0] + orc_c1_simpson(1000, 1e4, 0.118, 0.9, ORC_CONVERT2MB) + orc_c1_closed(1e4, 0.118, 0.9, ORC_CONVERT2MB);
    chk += orc_c1_cpumc_compat(0xABC, 1000, 1e4, 0.118, 0.9, ORC_CONV_CUPYTHIA);
    chk += (double)orc_c1_unweight_compat(0xC0B0, 1000, 1e4, 0.118, 0.9, orc_c1_wmax_scan(1000, 1e4, 0.118, 0.9),
                                          NULL, NULL);
    double e1[NB + 1], s1[4];
    for (int b = 0; b <= NB; ++b) e1[b] = (double)b / NB;
    orc_c1_vegas_adapt(1e4, 0.118, 0.9, ORC_CONVERT2MB, NB, 1.5, 2, 42, 0, N, e1, s1);
    double o3[3];
    orc_vegas1d_compat(16, 2, 32, 4, 2, 1e4, 0.118, 0.9, ORC_CONV_CUPYTHIA, o3);
    chk += sum(e1, NB + 1) + sum(o3, 3);
    printf("oracle sanitizer run: checksum %.6e\n", chk);
    free(edges); free(w); free(kin); free(w32); free(k32);
    if (grid) orc_lhagrid_free(grid);
    return isfinite(chk) ? 0 : 1;
}
