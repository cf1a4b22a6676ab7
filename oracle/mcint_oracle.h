/*
 * mcint_oracle.h -- CPU oracle for the Monte-Carlo matrix-element integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, not the product: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.  The
 * product library (cupythia_amd/) never links, calls or falls back to anything here.
 *
 * It restates, in plain C (FP64 unless stated), the algorithm of the reference path:
 *   - SplitMix64 counter RNG + 53-bit u01           cuPythia/common/rng.cuh:47-57
 *   - Philox4x32-10 (Salmon et al., SC'11, "Random123") -- no reference implementation;
 *     pinned by the published known-answer vectors (tests/test_oracle_rng.py).
 *   - QCD 2->2 sigmaHat, Pythia form                pythia8317/src/SigmaQCD.cc:115-378
 *     textbook (ESW / Combridge) forms               cuPythia/08_qcd_process_library/qcd_library.cu:25-60
 *   - sigmaHatWrap x CONVERT2MB                     pythia8317/include/Pythia8/SigmaProcess.h:516-520,
 *                                                   pythia8317/src/SigmaProcess.cc:24
 *   - AlphaStrong order 0/1 with flavour thresholds pythia8317/src/StandardModel.cc:22-62,188-215
 *   - parton flux lists and sigmaPDF                pythia8317/src/SigmaProcess.cc:124-193,423-480
 *   - 2->2 (tau, y, z) map, weights, stable tHat/uHat
 *                                                   pythia8317/src/PhaseSpace.cc:1149-1250,1256-1568
 *   - store2Kin scales (pT2, renormMultFac)         pythia8317/src/SigmaProcess.cc:846-925
 *   - sigma estimate (sigmaDelta)                   pythia8317/src/ProcessContainer.cc:1416-1472
 *   - flat gg->gg MC, Simpson quadrature            cuPythia/03_fused_resident_me/fused_gg2gg.cu:53-74
 *   - hit-or-miss unweighting + stable compaction   cuPythia/10_cub_compaction/cub_compaction.cu:36-63
 *
 * Parity pins (which check fixes which function) are listed next to each declaration.
 */
#ifndef MCINT_ORACLE_H
#define MCINT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pythia's GeV^-2 -> mb conversion (SigmaProcess.cc:24). */
#define ORC_CONVERT2MB 0.389380
/* cuPythia's conversion (fused_gg2gg.cu:87), used only for the reference-compat KATs. */
#define ORC_CONV_CUPYTHIA 0.3893793721

/* ---------------------------------------------------------------- RNG ---- */
/* Philox4x32-10 block.  Pinned: Random123 KATs + key=42 vectors. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n blocks for counters (e = counter0 + i) with ctr = (e_lo, e_hi, stream, 0), key = (key_lo, key_hi).
   out holds 4*n words. */
void orc_philox_fill(uint64_t key, uint64_t counter0, uint32_t stream, uint64_t n, uint32_t *out);
/* Uniforms derived from Philox words (this project's frozen definition). */
double orc_u53(uint32_t w0, uint32_t w1);   /* (((w1:w0) >> 11) * 2^-53), in [0,1) */
float  orc_u24(uint32_t w);                 /* ((w >> 8) * 2^-24), in [0,1)          */
/* SplitMix64 and u01, rng.cuh:47-57.  Pinned: 1,680,778 KAT (10_cub_compaction/README.md:12). */
uint64_t orc_splitmix64(uint64_t x);
double   orc_u01(uint64_t x);

/* --------------------------------------------------------- sigma-hat ---- */
/* Pythia process codes 111..116 (SigmaQCD.h). */
enum { ORC_GG2GG = 111, ORC_GG2QQBAR = 112, ORC_QG2QG = 113, ORC_QQ2QQ = 114,
       ORC_QQBAR2GG = 115, ORC_QQBAR2QQBARNEW = 116,
       ORC_GG2CCBAR = 121, ORC_QQBAR2CCBAR = 122, ORC_GG2BBBAR = 123, ORC_QQBAR2BBBAR = 124 };
/* Process slots: i = 0..5 <-> codes 111..116 (massless), i = 6..9 <-> 121..124 (c cbar, b bbar). */
#define ORC_NPROC 10
/* d(sigmaHat)/d(tHat) in GeV^-2, Pythia's literal expression (sigmaKin + sigmaHat), massless.
   id1/id2 matter only for 114.  nQuarkNew multiplies 112/116.
   Pinned: textbook identities (< 1e-12 per point), t<->u symmetry, cuPythia 08 sigma_cut KATs. */
double orc_sigma_hat_pythia(int code, double sH, double tH, double uH, double alpS,
                            int id1, int id2, int nQuarkNew);
/* Independent textbook (ESW / Combridge) brackets; d sigma/dt = pi alpS^2/sH^2 * bracket. */
double orc_bracket_textbook(int code, double s, double t, double u, int id1, int id2, int nQuarkNew);
/* Pythia's massive sigmaKin (GeV^-2) of codes 121..124 for m3 = m4 = m (m2 = m^2), and the
   textbook total partonic cross sections over the full t range (Combridge, Nucl. Phys. B151
   (1979) 429; Gluck, Owens, Reya, Phys. Rev. D17 (1978) 2324), beta = sqrt(1 - 4 m2 / s):
   qqbar: 8 pi alpS^2 / (27 s) beta (1 + 2 m2 / s);
   gg:    pi alpS^2 / (3 s) [(1 + rho + rho^2/16) ln((1+beta)/(1-beta)) - beta (7/4 + 31 rho / 16)], rho = 4 m2 / s. */
double orc_sigma_hat_massive(int code, double sH, double tH, double uH, double m2, double alpS);
double orc_sigma_tot_massive_textbook(int code, double sH, double m2, double alpS);

/* ------------------------------------------------------------ alphaS ---- */
typedef struct {
  double valueRef; int order; int nfmax;
  double mc, mb, mt, mc2, mb2, mt2;
  double L3s, L4s, L5s, L6s;   /* Lambda_nf^2 */
  double scale2Min;
} orc_alphas;
/* Pinned: alphaS(MZ^2) = valueRef exactly (order 1, nf=5 region); continuity at mb, mt. */
void   orc_alphas_init(orc_alphas *a, double valueRef, int order, int nfmax);
double orc_alphas_eval(const orc_alphas *a, double scale2);

/* ----------------------------------------------------------- toy PDFs ---- */
/* xf(x) = A x^a (1-x)^b, Q^2-independent.  Flavours: u = uv + sea, d = dv + sea, every other
   quark / antiquark = sea, g = gluon.  Default normalisation: int uv = 2, int dv = 1,
   gluon momentum fraction 0.5, sea carries the rest of the momentum sum rule.
   Pinned: number + momentum sum rules by quadrature (tests/test_oracle_physics.py). */
typedef struct {
  double ag, bg, av, bv, as, bs;     /* exponents */
  double Ag, Auv, Adv, As;            /* normalisations */
  int nQuarkIn;
} orc_toypdf;
void   orc_toypdf_init(orc_toypdf *p, const double g_ab[2], const double v_ab[2],
                       const double s_ab[2], int nQuarkIn);
double orc_toypdf_xf(const orc_toypdf *p, int id, double x);

/* -------------------------------------------- PDF grids (LHAGrid1) ---- */
/* LHAPDF6 "lhagrid1" data file read as Pythia's LHAGrid1::init does
   (pythia8317/src/PartonDistributions.cc:478-598): one x grid shared by all Q subgrids, the
   subgrids' Q values concatenated with each shared border averaged in q (the log grid keeps
   the values as read), flavour columns mapped to slots 0 = g, 1..5 = d u s c b,
   6..10 = dbar ubar sbar cbar bbar, 11 = photon (others dropped).  Values flavour-interleaved:
   val[(iq * nx + ix) * 12 + slot].
   Pinned: node exactness and cubic exactness of the interpolation on synthetic files, the
   subgrid borders, and the momentum / valence sum rules of the NNPDF2.3 LO file
   (tests/test_oracle_pdfgrid.py). */
#define ORC_GRID_NF 12
#define ORC_GRID_MAXSUB 8
typedef struct {
  int nx, nq, nqSub;
  double *x, *lnx, *q, *lnq;
  double *val;
  int nqSum[ORC_GRID_MAXSUB];        /* cumulative Q count after subgrid s */
  double qDiv[ORC_GRID_MAXSUB];      /* last Q of subgrid s, as read */
  double xMin, xMax, qMin, qMax;
} orc_lhagrid;
int  orc_lhagrid_load(const char *path, orc_lhagrid **out);   /* 0 on success */
void orc_lhagrid_free(orc_lhagrid *g);
/* LHAGrid1::xfxevolve (PartonDistributions.cc:648-745) with PDF:extrapolate = off: 4-point
   Lagrange interpolation in ln x and ln q (linear in ln q on 2-3 point subgrids), frozen at the
   q borders and at x <= xMin, 0 at x >= xMax; raw values (Pythia's xf clips them at 0). */
void orc_lhagrid_xf(const orc_lhagrid *g, double x, double Q2, double out[12]);

/* ------------------------------------------------ hadronic integrand ---- */
typedef struct {
  double eCM, pTHatMin;
  orc_alphas as;
  double renormFac;          /* mu_R = renormFac * pT  (Q2Ren = renormFac^2 * pT2) */
  orc_toypdf pdf;
  int nQuarkIn, nQuarkNew;
  int procMask;              /* bit i <-> code 111+i (i < 6), 121+(i-6) (i = 6..9) */
  double kfactor;
  double mQ[2];              /* c and b masses of the heavy-quark pair codes 121..124 */
  int pdfSource;             /* 0: toy PDFs (pdf); 1: the grid below */
  double factFac;            /* mu_F = factFac * pT  (Q2Fac = factFac^2 * pT2) */
  const orc_lhagrid *grid;
} orc_hadcfg;
void orc_hadcfg_default(orc_hadcfg *c);   /* the 5-process library at sqrt(S) = 13.6 TeV */

typedef struct { double sH, tH, uH, x1, x2, pT2, alpS, jac; } orc_kin;
/* Pythia-form evaluation of one point u in [0,1)^3:
   tau = tauMin (1/tauMin)^u0, y = yMax (2 u1 - 1), z = zMax (2 u2 - 1);
   weight_p = sigmaPDF_p * wtTau * wtY * wtZ (PhaseSpace.cc:1074), in mb.
   w[0..5] = codes 111..116 (0 where masked out); returns the sum.
   If literalU is non-zero, uHat is recomputed as -(sH + tH) like store2Kin does
   (SigmaProcess.cc:877) instead of keeping PhaseSpace's stable value. */
double orc_had_point(const orc_hadcfg *c, const double u[3], double w[ORC_NPROC], orc_kin *k, int literalU);
/* The same steps in 80-bit long double (rounding ~1e-19): the per-point reference for the FP64
   kernels, free of the ~1e-10 cancellation that FP64 Pythia arithmetic shows near the pT cut. */
double orc_had_point_ld(const orc_hadcfg *c, const double u[3], double w[ORC_NPROC], orc_kin *k, int literalU);

/* Deterministic quadrature of sigma_p over the same cut (Gauss-Legendre, variables
   (ln tau, y, rho = ln(-t/-u))).  out[0..5] per code, returns the sum. */
double orc_had_quadrature(const orc_hadcfg *c, int n_tau, int n_y, int n_rho, double out[ORC_NPROC]);

/* -------------------------------------------------- partonic gg->gg ---- */
/* w(c) = CONV * dsigma/dt(s, t(c), u(c), alpS) * (2 cMax) * (s/2); stable t = -s(1-c)/2, u = -s(1+c)/2 */
double orc_c1_weight(double s, double alpS, double cMax, double c, double conv);
/* Simpson (fused_gg2gg.cu:65-74) and the closed-form integral, in units of conv. */
double orc_c1_simpson(int N, double s, double alpS, double cMax, double conv);
double orc_c1_closed(double s, double alpS, double cMax, double conv);
/* Philox flat MC: event e uses block (key, (e_lo,e_hi,0,0)), c = cMax (2 u53 - 1).
   Fills sums[0]=sum w, sums[1]=sum w^2 (fixed order, compensated). nthreads >= 1. */
void orc_c1_mc_philox(uint64_t key, uint64_t counter0, uint64_t n, double s, double alpS,
                      double cMax, double conv, int nthreads, double sums[2]);
/* Reference-compat flat MC: cpuMC of fused_gg2gg.cu:53-62 (ctr = seed, ctr++). */
double orc_c1_cpumc_compat(uint64_t seed, uint64_t total, double s, double alpS, double cMax,
                           double conv);
/* Reference kernel 10 accept test: wMaxScan (cub_compaction.cu:48-50) */
double orc_c1_wmax_scan(int N, double s, double alpS, double cMax);
/* Reference kernel 10 flags: candidate i uses ctr = seed + i*0x100000001B3, draws c then u.
   Writes flags (may be NULL) and returns the accepted count. */
uint64_t orc_c1_unweight_compat(uint64_t seed, uint64_t n, double s, double alpS, double cMax,
                                double wMax, uint8_t *flags, double *cos_out);

/* -------------------------------------------- FP32 kernel-arithmetic mirror ---- */
/* Bit-for-bit restatement of the FP32 device arithmetic (own exp/log, same operation order,
   no contraction).  Used to check compaction indices bit-exactly; its accuracy is pinned
   against the FP64 Pythia-form oracle above. */
float orc_expf(float x);
float orc_logf(float x);
/* the 128-entry table of orc_logf (and of the device log): G_j, L_j */
void orc32_logtab(float G[128], float L[128]);
/* constant block shared by host and device (see orc32_consts_fill for the layout). */
#define ORC32_NCONST 42
void  orc32_consts_fill(const orc_hadcfg *c, float k[ORC32_NCONST]);
/* One point from three 24-bit uniforms; w[0..9] per code (111..116, 121..124); returns total
   (FP32); kin = {sHat, tHat, x1, x2} of the first selected family. */
float orc32_had_point(const float k[ORC32_NCONST], int procMask, float u0, float u1, float u2,
                      float w[ORC_NPROC], float kin[4]);
/* Same with the complements 1 - u_i given explicitly (as the VEGAS map provides them). */
float orc32_had_point_om(const float k[ORC32_NCONST], int procMask, float u0, float om0, float u1,
                         float om1, float u2, float om2, float w[ORC_NPROC], float kin[4]);
/* FP32 tables of a PDF grid exactly as the library builds them for its FP32 kernels: nodes,
   logs and values rounded from FP64, reciprocals of the Lagrange denominators computed in FP64
   and rounded (12 per stencil start; [12 m] = 1 / (lnq[m+1] - lnq[m]) on 2-3 point Q subgrids). */
typedef struct {
  int nx, nq, nqSub;
  float *xk, *lnx, *qk, *lnq, *val, *ixinv, *iqinv;
  int nqSum[ORC_GRID_MAXSUB];
  float qDiv[ORC_GRID_MAXSUB];
  float xMin, xMax, qMin, qMax, factFac2;
} orc32_grid;
void orc32_grid_build(const orc_lhagrid *g, double factFac, orc32_grid *out);
void orc32_grid_free(orc32_grid *g);
/* orc32_had_point_om with the parton densities from the FP32 grid g (NULL: the toy PDFs). */
float orc32_had_point_g(const float k[ORC32_NCONST], const orc32_grid *g, int procMask, float u0, float om0,
                        float u1, float om1, float u2, float om2, float w[ORC_NPROC], float kin[4]);

/* ------------------------------------------------ batched points, VEGAS, unweighting ---- */
/* Events ctr0 .. ctr0+n-1 (Philox stream 0, three 24-bit uniforms) through the flat map
   (edges == NULL) or a 3 x (nbins + 1) grid; w_out is n x 11 (total, codes 111..116, 121..124),
   Jacobian applied; kin_out (may be NULL) n x 4 (sHat, tHat, x1, x2).
   _ld: the Pythia-form point in long double with the FP64 map of the device.
   32 : the FP32 device arithmetic, bit for bit. */
void orc_had_points_ld(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                       uint64_t n, double *w_out, double *kin_out);
void orc32_had_points(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                      uint64_t n, float *w_out, float *kin_out);
/* FP32 unweighting decisions of the same candidates (accept stream 1): flags, chosen process
   code and total weight per candidate; returns the accepted count. */
uint64_t orc32_unweight(const orc_hadcfg *c, int nbins, const double *edges, uint64_t key, uint64_t ctr0,
                        uint64_t n, double wmax, uint8_t *flags, int32_t *pid, float *w_out);
/* Lepage refinement of ndim x (nbins + 1) edges from per-bin sums of w^2 (oracle_vegas_refine.c).
   Pinned (tests/test_oracle_vegas.py): alpha = 0 and a flat histogram leave a random grid
   unchanged; a 4-bin example worked by hand; the growth factor of a small alternating
   perturbation of the bin widths against its linearisation; convergence toward the closed-form
   grid of separable integrands in each of 3 dimensions; planted mistakes fail these pins. */
void orc_vegas_refine(int ndim, int nbins, double alpha, const double *hist, double *edges);
void orc32_vegas_adapt(const orc_hadcfg *c, int nbins, double alpha, int n_iter, uint64_t key, uint64_t ctr0,
                       uint64_t n, double *edges, double *sums_out);
void orc_c1_vegas_adapt(double s, double alpS, double cMax, double conv, int nbins, double alpha, int n_iter,
                        uint64_t key, uint64_t ctr0, uint64_t n, double *edges, double *sums_out);
/* The reference's 1-D VEGAS (vegas.cu:40-124); out = {eta first, eta last, last integral}.
   Pinned: eta 10.0% -> 76.0%, integral relerr 1.2e-5 (11_vegas_importance/README.md:13-16). */
void orc_vegas1d_compat(int K, int blocks, int threads, uint64_t Tper, int n_iter, double s, double alpS,
                        double cMax, double conv, double out[3]);
/* CPU baseline: Pythia-form FP64 MC of the hadronic integrand (flat map, Philox stream 0) on
   nthreads cores; sums = {sum w, sum w^2}. */
void orc_had_mc(const orc_hadcfg *c, uint64_t key, uint64_t ctr0, uint64_t n, int nthreads, double sums[2]);

/* ------------------------------------------------------------------ adjoint ---- */
/* Central finite differences (relative step 1e-5) of the long-double point, per event, with the
   configuration rebuilt at theta +- h.  out[0] = sum w, out[1..4] = sum dw/d{alphaS(MZ),
   renorm_fac, gluon a, gluon b}, out[5..8] = sums of squares.
   Pinned: FD of a closed form (fixed alphaS: d ln sigma / d ln alphaS = 2 exactly). */
void orc_adjoint(const orc_hadcfg *c, uint64_t key, uint64_t ctr0, uint64_t n, double out[9]);

/* ------------------------------------------------------------- event record ---- */
/* The partons completing a compacted event (oracle_events.c): incoming pair by pickInState over
   the code's flux list, outgoing flavours and colour flow by setIdColAcol, lab momenta by
   finalKin (rotation + boost), hard scale renormFac sqrt(pT^2 + m^2).  Colour tags 501.. (0 none),
   status -21 / 23, mothers / daughters as slots (-1 none); p[k] = (px, py, pz, e, m).  partonic:
   g g -> g g at sqrt(sHatPartonic) per collision, x1 = x2 = 1.  An unknown code gives nPart = 0.
   Pinned: tests/test_oracle_events.py. */
typedef struct {
  int pdg[4], status[4], col[4], acol[4], mo1[4], mo2[4], d1[4], d2[4];
  double p[4][5];
  double scale;
  int nPart;
} orc_event4;
void orc_build_event(const orc_hadcfg *c, int partonic, double sHatPartonic, uint64_t key, uint64_t index, int code,
                     double sH, double tH, double x1, double x2, orc_event4 *out);
/* n events; ints[i][k][8] = (pdg, status, col, acol, mo1, mo2, d1, d2), mom[i][k][5]. */
void orc_build_events(const orc_hadcfg *c, int partonic, double sHatPartonic, uint64_t key, uint64_t n,
                      const uint64_t *index, const int32_t *code, const double *sH, const double *tH,
                      const double *x1, const double *x2, int32_t *ints, double *mom, double *scale, int32_t *nPart);

#ifdef __cplusplus
}
#endif
#endif
