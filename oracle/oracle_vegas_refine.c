/*
 * oracle_vegas_refine.c -- CPU oracle: the VEGAS grid refinement alone.  TEST INFRASTRUCTURE
 * ONLY (see mcint_oracle.h).  Kept in a file of its own so that a test can compile copies of it
 * with a planted mistake and show that the pins of tests/test_oracle_vegas.py reject them.
 *
 * G. P. Lepage, J. Comput. Phys. 27 (1978) 192 (the algorithm of the VEGAS program):
 *   1. per dimension, per-bin sums d_i of (f J)^2 over the points that fell in bin i;
 *   2. neighbour smoothing: (d_0 + d_1)/2 at the first bin, (d_{i-1} + d_i + d_{i+1})/3 inside,
 *      (d_{N-2} + d_{N-1})/2 at the last bin;
 *   3. normalisation r_i = d_i / sum_j d_j and damping m_i = ((1 - r_i) / (-ln r_i))^alpha
 *      (m_i = 0 for r_i = 0, 1 for r_i = 1);
 *   4. re-partition: the new edges split sum_i m_i into N equal shares, each old bin's m_i spread
 *      uniformly over its width (linear interpolation inside the old bin); the outer edges stay.
 * The reference's refinement (cuPythia/11_vegas_importance/vegas.cu:69-85) is the 1-D special case
 * with the first moment f J, step 2 alone and no damping; its restatement is orc_vegas1d_compat.
 */
#include "mcint_oracle.h"
#include <math.h>
#include <stdlib.h>

void orc_vegas_refine(int ndim, int nbins, double alpha, const double *hist, double *edges)
{
    double *m = (double *)malloc(sizeof(double) * nbins);
    double *ne = (double *)malloc(sizeof(double) * (nbins + 1));
    for (int d = 0; d < ndim; ++d) {
        const double *h = hist + (size_t)d * nbins;
        double *x = edges + (size_t)d * (nbins + 1);
        double tot = 0.0;
        for (int i = 0; i < nbins; ++i) {
            double s;
            if (nbins == 1) s = h[0];
            else if (i == 0) s = 0.5 * (h[0] + h[1]);
            else if (i == nbins - 1) s = 0.5 * (h[nbins - 2] + h[nbins - 1]);
            else s = (h[i - 1] + h[i] + h[i + 1]) / 3.0;
            m[i] = s;
            tot += s;
        }
        if (!(tot > 0.0)) continue;
        double msum = 0.0;
        for (int i = 0; i < nbins; ++i) {
            double r = m[i] / tot, mi = 0.0;
            if (r > 0.0) mi = (r < 1.0) ? pow((1.0 - r) / -log(r), alpha) : 1.0;
            m[i] = mi;
            msum += mi;
        }
        if (!(msum > 0.0)) continue;
        double step = msum / nbins, acc = 0.0;
        int old = 0;
        ne[0] = x[0];
        ne[nbins] = x[nbins];
        for (int j = 1; j < nbins; ++j) {
            double target = step * j;
            while (old < nbins - 1 && acc + m[old] < target) { acc += m[old]; ++old; }
            double frac = (m[old] > 0.0) ? (target - acc) / m[old] : 0.0;
            frac = fmin(fmax(frac, 0.0), 1.0);
            ne[j] = x[old] + frac * (x[old + 1] - x[old]);
        }
        for (int j = 0; j <= nbins; ++j) x[j] = ne[j];
    }
    free(m);
    free(ne);
}
