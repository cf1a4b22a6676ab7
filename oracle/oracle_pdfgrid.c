/*
 * oracle_pdfgrid.c -- reader of LHAPDF6 "lhagrid1" PDF data files, restating Pythia's
 * LHAGrid1::init (pythia8317/src/PartonDistributions.cc:478-598).  TEST INFRASTRUCTURE ONLY
 * (see mcint_oracle.h).  The interpolation (LHAGrid1::xfxevolve, :648-745) is in
 * oracle_had_body.h so that it exists in FP64 and in long double; orc_lhagrid_xf below is its
 * FP64 instance.
 *
 * File layout: header lines up to a line containing "---", then for every Q subgrid: one line
 * of x values (the same for every subgrid), one line of Q values, one line of flavour ids, nx *
 * nq_sub lines of xf values (x outer, Q inner, one column per flavour id), and a "---" line.
 */
#define _POSIX_C_SOURCE 200809L
#include "mcint_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int read_doubles(const char *line, double **arr, int *n, int *cap)
{
    const char *p = line;
    char *end;
    int cnt = 0;
    for (;;) {
        double v = strtod(p, &end);
        if (end == p) break;
        if (*n >= *cap) {
            *cap = *cap ? 2 * *cap : 64;
            double *na = (double *)realloc(*arr, sizeof(double) * (size_t)*cap);
            if (!na) return -1;
            *arr = na;
        }
        (*arr)[(*n)++] = v;
        ++cnt;
        p = end;
    }
    return cnt;
}

void orc_lhagrid_free(orc_lhagrid *g)
{
    if (!g) return;
    free(g->x); free(g->lnx); free(g->q); free(g->lnq); free(g->val);
    free(g);
}

/* Flavour id -> slot: 21 or 0 -> 0 (gluon), 1..5 -> 1..5, -1..-5 -> 6..10, 22 -> 11 (photon),
 * anything else (e.g. +-6) is dropped, as in Pythia. */
static int flavour_slot(int id)
{
    if (id == 21 || id == 0) return 0;
    if (id > 0 && id < 6) return id;
    if (id < 0 && id > -6) return 5 - id;
    if (id == 22) return 11;
    return -1;
}

int orc_lhagrid_load(const char *path, orc_lhagrid **out)
{
    *out = NULL;
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t lcap = 0;
    int rc = -2;
    orc_lhagrid *g = (orc_lhagrid *)calloc(1, sizeof *g);
    double *xs = NULL, *qs = NULL, *vals = NULL, *tmp = NULL;
    int nx = 0, xcap = 0, nq = 0, qcap = 0, nval = 0, vcap = 0, ntmp = 0, tcap = 0;
    int *idmap = NULL, nid = 0;
    int subStart[ORC_GRID_MAXSUB + 1], nsub = 0;
    /* value blocks per subgrid: slot-mapped, kept as [sub][ix][iq_local][12] in vals */
    while (getline(&line, &lcap, f) >= 0)
        if (strstr(line, "---")) break;
    for (;;) {
        if (getline(&line, &lcap, f) < 0) break;                 /* x line (or EOF) */
        if (nsub >= ORC_GRID_MAXSUB) goto fail;
        ntmp = 0;
        if (read_doubles(line, &tmp, &ntmp, &tcap) <= 0) break;
        if (nsub == 0) {
            for (int i = 0; i < ntmp; ++i) {
                if (nx >= xcap) { xcap = xcap ? 2 * xcap : 128; xs = (double *)realloc(xs, sizeof(double) * (size_t)xcap); }
                xs[nx++] = tmp[i];
            }
        } else {
            if (ntmp != nx) goto fail;
            for (int i = 0; i < nx; ++i)
                if (fabs(log(tmp[i]) - log(xs[i])) > 1e-5) goto fail;   /* x nodes differ between Q subgrids */
        }
        if (getline(&line, &lcap, f) < 0) goto fail;              /* Q line */
        int q0 = nq;
        int nqNow = read_doubles(line, &qs, &nq, &qcap);
        if (nqNow <= 0) goto fail;
        if (nsub > 0) {
            if (fabs(qs[q0] / qs[q0 - 1] - 1.0) > 1e-5) goto fail;   /* a Q subgrid does not start where the previous ended */
            /* qGrid border averaged; the log grid keeps the values as read (Pythia :539-544) */
        }
        subStart[nsub] = q0;
        if (getline(&line, &lcap, f) < 0) goto fail;              /* flavour ids */
        {
            char *p = line, *end;
            nid = 0;
            free(idmap);
            idmap = (int *)malloc(sizeof(int) * 64);
            for (;;) {
                long v = strtol(p, &end, 10);
                if (end == p) break;
                if (nid >= 64) goto fail;
                idmap[nid++] = flavour_slot((int)v);
                p = end;
            }
            if (nid == 0) goto fail;
        }
        /* nx * nqNow value lines, x outer, Q inner */
        for (int ix = 0; ix < nx; ++ix)
            for (int iq = 0; iq < nqNow; ++iq) {
                if (getline(&line, &lcap, f) < 0) goto fail;
                ntmp = 0;
                if (read_doubles(line, &tmp, &ntmp, &tcap) < nid) goto fail;
                for (int k = 0; k < ORC_GRID_NF; ++k) {
                    if (nval >= vcap) { vcap = vcap ? 2 * vcap : 4096; vals = (double *)realloc(vals, sizeof(double) * (size_t)vcap); }
                    vals[nval++] = 0.0;
                }
                double *dst = vals + nval - ORC_GRID_NF;
                for (int k = 0; k < nid; ++k) if (idmap[k] >= 0) dst[idmap[k]] = tmp[k];
            }
        if (getline(&line, &lcap, f) < 0) { ++nsub; break; }      /* "---" separator */
        ++nsub;
    }
    if (nsub == 0 || nx < 4) goto fail;
    subStart[nsub] = nq;
    g->nx = nx; g->nq = nq; g->nqSub = nsub;
    g->x = (double *)malloc(sizeof(double) * (size_t)nx);
    g->lnx = (double *)malloc(sizeof(double) * (size_t)nx);
    g->q = (double *)malloc(sizeof(double) * (size_t)nq);
    g->lnq = (double *)malloc(sizeof(double) * (size_t)nq);
    g->val = (double *)malloc(sizeof(double) * (size_t)nq * nx * ORC_GRID_NF);
    for (int i = 0; i < nx; ++i) { g->x[i] = xs[i]; g->lnx[i] = log(xs[i]); }
    for (int i = 0; i < nq; ++i) { g->q[i] = qs[i]; g->lnq[i] = log(qs[i]); }
    for (int s = 0; s < nsub; ++s) {
        g->nqSum[s] = subStart[s + 1];
        g->qDiv[s] = qs[subStart[s + 1] - 1];                     /* last Q of the subgrid as read */
        if (s > 0) {
            int b = subStart[s];
            double avg = 0.5 * (qs[b - 1] + qs[b]);
            g->q[b - 1] = avg;
            g->q[b] = avg;
        }
    }
    /* reorder values into [iq][ix][12] */
    {
        size_t off = 0;
        for (int s = 0; s < nsub; ++s) {
            int nqs = subStart[s + 1] - subStart[s];
            for (int ix = 0; ix < nx; ++ix)
                for (int iq = 0; iq < nqs; ++iq) {
                    memcpy(g->val + ((size_t)(subStart[s] + iq) * nx + ix) * ORC_GRID_NF, vals + off,
                           sizeof(double) * ORC_GRID_NF);
                    off += ORC_GRID_NF;
                }
        }
    }
    g->xMin = g->x[0]; g->xMax = g->x[nx - 1];
    g->qMin = g->q[0]; g->qMax = g->q[nq - 1];
    rc = 0;
    *out = g;
    g = NULL;
fail:
    if (g) orc_lhagrid_free(g);
    free(line); free(xs); free(qs); free(vals); free(tmp); free(idmap);
    fclose(f);
    return rc;
}
