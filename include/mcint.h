/*
 * mcint.h -- C ABI of the MI355X (gfx950) Monte-Carlo matrix-element integration library
 * (cupythia_amd/_lib/libmcint.so).
 *
 * Plain C: no C++ types, no exceptions, no torch types cross this boundary.  Every entry point
 * returns an int status (MI_OK = 0, negative = error); mi_last_error(ctx) returns the message
 * of the last failure on that context.  Bulk arrays are DEVICE pointers (memory resident in
 * HBM); configuration and result structs are host memory owned by the caller.  One mi_ctx per
 * host thread; contexts are independent.  Results are deterministic for a given
 * (key, counter range, configuration).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   process identity + per-point d(sigmaHat)/d(tHat)   pythia8317/include/Pythia8/SigmaProcess.h:88-402,
 *       (Sigma2Process::set2Kin / sigmaKin / sigmaHatWrap) SigmaProcess.h:487-546, :516-520;
 *       codes 111..116                                   pythia8317/include/Pythia8/SigmaQCD.h:188-392
 *   sigma estimate (sigmaGen / sigmaErr / nTry / nAcc)  pythia8317/src/ProcessContainer.cc:1416-1472,
 *                                                        pythia8317/include/Pythia8/Info.h:316-330,
 *                                                        pythia8317/src/ProcessLevel.cc:428-472
 *   counter RNG                                          cuPythia/common/rng.cuh:47-57 (SplitMix64);
 *                                                        Pythia hook RndmEngine::flat, Basics.h:360-371
 *   fused flat MC of gg->gg                              cuPythia/03_fused_resident_me/fused_gg2gg.cu:40-51
 *   VEGAS sample + rebin                                 cuPythia/11_vegas_importance/vegas.cu:40-85
 *   unweighting + stable compaction                      cuPythia/10_cub_compaction/cub_compaction.cu:36-80,
 *                                                        cuPythia/pipeline/generate.cu:56-62,98-103
 *   same-sample scale variations (adjoint precursor)    cuPythia/pipeline/reweight.cu:36-51
 */
#ifndef MCINT_H
#define MCINT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_ABI_VERSION 4

typedef struct mi_ctx mi_ctx;   /* opaque */

/* status codes */
#define MI_OK            0
#define MI_ERR_ARG      -1   /* invalid argument / configuration */
#define MI_ERR_HIP      -2   /* HIP runtime failure */
#define MI_ERR_STATE    -3   /* call out of order (e.g. no VEGAS grid) */
#define MI_ERR_RANGE    -4   /* size beyond a documented maximum */
#define MI_ERR_NUMERIC  -5   /* non-finite result */
#define MI_ERR_CAPACITY -6   /* output buffer too small (count still reported) */

/* Pythia process codes (SigmaQCD.h).  Slot i of per-process arrays and bit i of proc_mask:
 * i = 0..5 <-> codes 111..116 (massless), i = 6..9 <-> codes 121..124 (heavy-quark pairs with
 * m3 = m4 = mass_Q: HardQCD:gg2ccbar, qqbar2ccbar, gg2bbbar, qqbar2bbbar). */
#define MI_GG2GG          111
#define MI_GG2QQBAR       112
#define MI_QG2QG          113
#define MI_QQ2QQ          114
#define MI_QQBAR2GG       115
#define MI_QQBAR2QQBARNEW 116
#define MI_GG2CCBAR       121
#define MI_QQBAR2CCBAR    122
#define MI_GG2BBBAR       123
#define MI_QQBAR2BBBAR    124
#define MI_MAX_PROC       10
#define MI_PROC_CODE(i)   ((i) < 6 ? 111 + (i) : 121 + (i) - 6)

enum mi_mode      { MI_MODE_PARTONIC_1D = 0, MI_MODE_HADRONIC_3D = 1 };
/* MI_PREC_FP32: ME and phase space in FP32 with this library's own exp / log and IEEE division
 * and square root -- every FP32 weight reproducible bit for bit on the host, which makes the
 * unweighting decisions and compacted indices bit-exact.  MI_PREC_FP32_HW: the same FP32 algebra
 * with the hardware exp2 / log2 / reciprocal / square root (v_exp_f32, v_log_f32, v_rcp_f32,
 * v_sqrt_f32; ~1 ulp each) in mi_integrate, mi_integrate_noalloc, mi_points, mi_vegas_adapt and
 * mi_solve's VEGAS passes; per-point weights within 1e-5 of the FP64 form, not bit-reproducible on
 * the host.  Every other entry point runs an FP32_HW run as FP32.  Sums are FP64 in all modes. */
enum mi_precision { MI_PREC_FP32 = 0, MI_PREC_FP64 = 1, MI_PREC_FP32_HW = 2 };
enum mi_sampler   { MI_SAMPLER_FLAT = 0, MI_SAMPLER_VEGAS = 1, MI_SAMPLER_FLOW = 2 };
enum mi_pdf_source { MI_PDF_TOY = 0, MI_PDF_GRID = 1 };
#define MI_PDF_SLOTS      8    /* PDF grids a context holds: slot 0 (runs) + variation slots 1..7 */
enum mi_rng       { MI_RNG_PHILOX = 0, MI_RNG_SPLITMIX_COMPAT = 1 };

/* Physics configuration.  Defaults (mi_physics_default) mirror Pythia's settings where the
 * workload does not fix them: SigmaProcess:alphaSorder = 1, StandardModel:alphaSnfmax = 6,
 * SigmaProcess:renormScale2 = 2 (mu_R^2 = pT^2 for massless 2->2), HardQCD:nQuarkNew = 3,
 * PDFinProcess:nQuarkIn = 5, SigmaProcess:Kfactor = 1. */
typedef struct mi_physics {
    int    mode;            /* enum mi_mode */
    int    proc_mask;       /* bit i <-> code MI_PROC_CODE(i) (hadronic); partonic mode is gg->gg only */
    /* partonic 1-D: fixed sHat, fixed alphaS, cos(theta) flat in [-cos_max, cos_max] */
    double sHat;            /* GeV^2 */
    double cos_max;
    /* hadronic 3-D: (tau, y, z) with the pT cut */
    double eCM;             /* sqrt(S), GeV */
    double pTHatMin;        /* GeV */
    double alphaS_value;    /* alphaS(MZ) for order 1; the fixed value for order 0 */
    int    alphaS_order;    /* 0 fixed, 1 one-loop with flavour thresholds */
    int    alphaS_nfmax;    /* 5 or 6 */
    double renorm_fac;      /* mu_R = renorm_fac * pT */
    int    nQuarkIn;        /* incoming quark flavours (2..5) */
    int    nQuarkNew;       /* outgoing flavours for 112 / 116 */
    double pdf_g[2];        /* toy xf = A x^a (1-x)^b exponents (a, b): gluon */
    double pdf_val[2];      /* valence (u_v, d_v) */
    double pdf_sea[2];      /* sea (every quark and antiquark) */
    double kfactor;
    int    pdf_source;      /* enum mi_pdf_source: toy PDFs above, or the grid of mi_pdf_load */
    double fact_fac;        /* mu_F = fact_fac * pT (PDF scale with a grid); heavy pairs: mu_F^2 = fact_fac^2 (pT^2 + m^2) */
    double mass_Q[2];       /* c and b masses (GeV) of codes 121..124; defaults 1.5, 4.8 (Pythia's m0) */
} mi_physics;

typedef struct mi_run {
    mi_physics phys;
    int      precision;     /* enum mi_precision (FP32: ME in FP32, sums in FP64) */
    int      sampler;       /* enum mi_sampler */
    int      rng;           /* enum mi_rng */
    uint64_t key;           /* Philox key (64 bit: key_lo, key_hi) or SplitMix seed */
    uint64_t counter0;      /* first event index */
    uint64_t n_points;      /* events (points) in this call */
} mi_run;

/* Pythia's sigma-estimate semantics.  Index 0 = sum over processes (Info::sigmaGen(0)),
 * index 1 + i = code MI_PROC_CODE(i) (code[] holds them).  For weighted integration nSel = nAcc = nTry; sigmaErr follows
 * ProcessContainer::sigmaDelta.  The sum's error uses the per-point total weight (the
 * processes share points), not the quadrature sum of independent containers. */
typedef struct mi_sigma_est {
    int      code[MI_MAX_PROC + 1];
    double   sigmaGen[MI_MAX_PROC + 1];   /* mb */
    double   sigmaErr[MI_MAX_PROC + 1];   /* mb */
    double   sumW[MI_MAX_PROC + 1];
    double   sumW2[MI_MAX_PROC + 1];
    double   wMax[MI_MAX_PROC + 1];
    int64_t  nTry, nSel, nAcc;
    int64_t  nNonZero;                    /* points with total weight > 0 */
    double   kernel_ms;                   /* device time of the integration kernel (hipEvent pair on the stream) */
} mi_sigma_est;

/* VEGAS adaptation summary (cumulative inverse-variance combination over the iterations). */
typedef struct mi_vegas_info {
    int      n_iter;
    double   sigma, sigma_err, chi2_dof;  /* combined over the iterations */
    double   last_sigma, last_err;        /* last iteration alone */
    double   last_wmax, last_eta;         /* last iteration's max weight and <w>/wMax */
    double   kernel_ms;
} mi_vegas_info;

/* Compacted unweighted events, structure of arrays, DEVICE pointers of length `capacity`.
 * Accepted events keep the candidate order (stable filter). */
typedef struct mi_events {
    double   *sHat, *tHat, *x1, *x2, *weight;
    int32_t  *pid;          /* Pythia process code */
    uint64_t *index;        /* candidate event index (counter) */
    uint64_t capacity;
} mi_events;

/* Gradients of sigma (sum over the selected processes), same points (common random numbers). */
typedef struct mi_sigma_grad {
    double sigma, sigma_err;
    double d_alphaS_MZ, d_alphaS_MZ_err;
    double d_renorm_fac, d_renorm_fac_err;
    double d_pdf_g_a, d_pdf_g_a_err;
    double d_pdf_g_b, d_pdf_g_b_err;
    double kernel_ms;
} mi_sigma_grad;

/* Flow training / evaluation summary. */
typedef struct mi_flow_info {
    int      steps;
    double   loss_first, loss_last;      /* variance loss <w^2>/<w>^2 - 1 on the training batches */
    double   cv_vegas, cv_flow;          /* coefficient of variation of w before / after */
    double   kernel_ms;
} mi_flow_info;

/* Flow architecture (fixed): 3-D, 6 rational-quadratic spline couplings with 16 bins on
 * [0,1], each transforming one coordinate conditioned on the other two through an MLP
 * 2 -> 64 -> 64 -> 49 (16 widths, 16 heights, 17 knot derivatives), ReLU, BF16 weights for
 * the 64-wide layers with FP32 accumulation.  The flow maps the unit cube onto the VEGAS
 * grid's unit cube: x_phys = VEGAS(flow(z)), z ~ U[0,1)^3. */
#define MI_FLOW_LAYERS   6
#define MI_FLOW_BINS     16
#define MI_FLOW_HIDDEN   64
#define MI_FLOW_NOUT     49
#define MI_FLOW_NPARAMS  45222           /* 6 x (2*64+64 + 64*64+64 + 64*49+49) */

/* ------------------------------------------------------------------ context ---- */
int         mi_version(void);
/* Create a context on HIP device `device`; owns its scratch buffers, VEGAS grid and flow. */
int         mi_ctx_create(int device, mi_ctx **out);
int         mi_ctx_destroy(mi_ctx *ctx);
/* Message of the last failed call on ctx (never NULL). */
const char *mi_last_error(const mi_ctx *ctx);
/* Launch on this HIP stream (hipStream_t passed as void*; NULL = the null stream). */
int         mi_ctx_set_stream(mi_ctx *ctx, void *stream);
/* Fill *p with the workload defaults: which = 0 partonic gg->gg at sHat = (100 GeV)^2,
 * |cos| < 0.9, alphaS = 0.118; which = 1 the 5-process library at sqrt(S) = 13.6 TeV,
 * pT > 20 GeV, toy PDFs, 1-loop alphaS(MZ) = 0.118. */
int         mi_physics_default(int which, mi_physics *p);

/* --------------------------------------------------------------------- RNG ---- */
/* Philox4x32-10 blocks for events counter0 .. counter0+n-1: ctr = (e_lo, e_hi, stream, 0),
 * key = (key_lo, key_hi).  out_dev receives 4*n uint32 words.  (Test hook; the integration
 * kernels generate the same words in registers.) */
int mi_rng_philox_fill(mi_ctx *ctx, uint64_t key, uint64_t counter0, uint32_t stream,
                       uint64_t n, uint32_t *out_dev);

/* ---------------------------------------------------------------- sigmaHat ---- */
/* Batched Kfactor * sigmaHatWrap(id1, id2) (mb/GeV^2) for code 111..116 (massless) and the
 * heavy-quark pairs 121..124, FP64.  uH_dev may be NULL for 111..116 (then uH = -(sH + tH) as in
 * Sigma2Process::store2Kin); 121..124 need it and take m3^2 = m4^2 = (sH + tH + uH) / 2.
 * id1/id2 may be NULL (only 114 depends on them; NULL means q q' with different flavours).
 * With alpS given per point this is also Pythia's multiparton-interaction entry,
 * Sigma2Process::set2KinMPI(x1, x2, sH, tH, uH, alpS, alpEM, needMasses, m3, m4) then sigmaHatWrap
 * (pythia8317/include/Pythia8/SigmaProcess.h:504-508, SigmaProcess.cc:931-983): x1, x2 and alpEM
 * do not enter these codes' sigmaHat, and needMasses with m3 = m4 is the heavy pairs' uH. */
int mi_sigma_hat(mi_ctx *ctx, int code, uint64_t n, const double *sH_dev, const double *tH_dev,
                 const double *uH_dev, const double *alpS_dev, const int32_t *id1_dev,
                 const int32_t *id2_dev, int nQuarkNew, double kfactor, double *out_dev);

/* ----------------------------------------------------------- integration ---- */
/* Per-point weights (mb) of events counter0 .. counter0+n_points-1: w_dev is
 * n x (MI_MAX_PROC + 1) doubles (column 0 = total, 1 + i = code MI_PROC_CODE(i)); kin_dev (may
 * be NULL) is n x 4 doubles (sHat, tHat, x1, x2) of the first family proc_mask selects
 * (massless, c cbar, b bbar: each has its own phase-space map at the same uniforms).  Uses run->sampler's map (flat or the
 * context's VEGAS grid).  Parity hook for the integration kernels. */
int mi_points(mi_ctx *ctx, const mi_run *run, double *w_dev, double *kin_dev);
/* Sigma estimate over run->n_points events; fused RNG -> map -> PDFs x sigmaHat -> sum w,
 * sum w^2, reduced deterministically on the device. */
int mi_integrate(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out);

/* The same estimate as mi_integrate for the flat and VEGAS samplers (bit for bit: the same kernels,
 * launch geometry and reduction), with its block records and reduced record in a static device
 * array of the library instead of the context's buffers: the call makes no device allocation (no
 * hipMalloc, hipFree or allocating library routine; the context, the VEGAS grid and any PDF grid
 * are set up beforehand).  No counterpart in the reference: it is for callers that cannot
 * allocate -- tight serving loops, harnesses that forbid hipMalloc.  Limit: one launch of at most
 * 4096 blocks (the scratch is 1,179,936 B per GPU, held until the process ends; MI355X launches
 * at most 8 blocks per CU, 2048); MI_ERR_RANGE beyond it.  Every call on a GPU shares that GPU's
 * scratch: calls that overlap on one GPU (several streams or threads) race.  The flow sampler is
 * refused (MI_ERR_ARG): it allocates its buffers. */
int mi_integrate_noalloc(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out);

/* ----------------------------------------------------------------- VEGAS ---- */
/* Uniform grid with nbins (<= 256) bins per dimension (1 for partonic, 3 for hadronic). */
int mi_vegas_init(mi_ctx *ctx, int ndim, int nbins);
/* n_iter iterations of n_points each (counters counter0 + it * n_points ...): sample,
 * accumulate per-bin sum w^2, refine on the device with damping alpha (Lepage 1978).
 * The grid is used by every later call with sampler = MI_SAMPLER_VEGAS. */
int mi_vegas_adapt(mi_ctx *ctx, const mi_run *run, int n_iter, double alpha, mi_vegas_info *out);
/* Host copies of the grid edges: ndim x (nbins + 1) doubles. */
int mi_vegas_get_grid(mi_ctx *ctx, double *edges_host, int *ndim, int *nbins);
int mi_vegas_set_grid(mi_ctx *ctx, const double *edges_host, int ndim, int nbins);

/* ------------------------------------------------------- unweighting ---- */
/* Hit-or-miss unweighting of run->n_points candidates: accept iff u * wMax < w with u from
 * the disjoint stream 1 of the same counter; accepted events carry weight max(w, wMax) and a
 * process id drawn with probability w_p / w.  If wmax <= 0 the exact maximum over the
 * candidates is used (one extra pass).  Writes the accepted events in candidate order;
 * *n_acc is the accepted count even when it exceeds out->capacity (then MI_ERR_CAPACITY).
 * est (may be NULL) receives the weighted sigma estimate of the same candidates with
 * nAcc = accepted count. */
int mi_unweight_compact(mi_ctx *ctx, const mi_run *run, double wmax, const mi_events *out,
                        uint64_t *n_acc, double *wmax_used, mi_sigma_est *est);

/* ------------------------------------------------------ device event record ---- */
/* Partons of the compacted events on the device, structure of arrays in caller-owned device
 * memory: particle p of event e at e * max_part + p (max_part >= 4; slots beyond n_part[e] are
 * left untouched, headroom for later stages).  Replaces cuPythia/pipeline/event.cuh:20-31
 * (DeviceEvents); x1, x2 and the weight stay in mi_events. */
typedef struct mi_particles {
    double   *px, *py, *pz, *e, *m;        /* [capacity * max_part], GeV, lab frame */
    int32_t  *pdg, *status;                /* PDG id; -21 incoming, 23 outgoing (Pythia's codes) */
    int32_t  *col, *acol;                  /* colour tags 501, 502, ... (0 = none) */
    int32_t  *mo1, *mo2, *d1, *d2;         /* mothers / daughters as record slots, -1 = none */
    int32_t  *n_part;                      /* [capacity] particles written (4; 0 for an unknown code) */
    double   *scale;                       /* [capacity] renorm_fac sqrt(pT^2 + m^2), GeV */
    uint64_t capacity;                     /* events */
    int32_t  max_part;
    int32_t  reserved;
} mi_particles;

/* Builds the record of n compacted events (ev, e.g. from mi_unweight_compact) on the device, one
 * thread per event, as Pythia completes a hard process:
 *   incoming flavours: a pair from the code's flux list (SigmaProcess::initFlux) drawn with
 *     weight xf1 xf2 dsigmaHat/dtHat (SigmaProcess::pickInState, SigmaProcess.cc:486-508);
 *   outgoing flavours and colour flow: the code's setIdColAcol (SigmaQCD.cc:135-394, 768-781,
 *     894-905);
 *   momenta: incoming partons along +-z with E = x sqrt(S)/2, the outgoing pair back to back in
 *     the partonic frame at the record's tHat and a uniform azimuth, boosted to the lab
 *     (PhaseSpace.cc:2236-2251); partonic mode: sqrt(sHat)/2 per beam, x = 1.
 * Random numbers: Philox streams 3 and 4 of ev->index[i] under `key` (stream 3: the pair and the
 * colour topology as 53-bit uniforms; stream 4: the azimuth, the g g -> g g orientation, the new
 * flavour of 112 / 116), so the record is a function of (key, event index).  phys: the run's physics (PDFs for the
 * flavour draw, masses, renorm_fac).  Replaces cuPythia/pipeline/build_events.cu:22-41 and
 * cuPythia/12_phase_space/phase_space.cu:37-53. */
int mi_build_events(mi_ctx *ctx, const mi_physics *phys, const mi_events *ev, uint64_t n, uint64_t key,
                    const mi_particles *out);

/* -------------------------------------------------------------- LHE output ---- */
/* LHEF 3.0 file of n compacted events (device SoA, e.g. from mi_unweight_compact): the partons
 * of mi_build_events (same key: flavours, colour flow, lab momenta), cross sections
 * (est->sigmaGen per process code) in pb; unweighted events (IDWTUP = 3) each with XWGTUP =
 * est->sigmaGen[0] / n in pb, SCALUP the hard scale, AQCDUP alphaS at it, AQEDUP alphaEM(MZ).  Replaces cuPythia/pipeline/lhe_writer.cu:46-70 and
 * cuPythia/07_unweighting_lhe/unweight_lhe.cu:74-102. */
int mi_write_lhe(mi_ctx *ctx, const char *path, const mi_physics *phys, const mi_events *ev, uint64_t n,
                 uint64_t key, const mi_sigma_est *est);

/* ------------------------------------------------------------ reweighting ---- */
/* Renormalisation-scale variations of events in one pass: out[i * nvar + k] =
 * w[i] * (alphaS(f_k^2 Q^2) / alphaS(renorm_fac^2 Q^2))^2 with the event's Pythia scale
 * Q^2 = tHat uHat / sHat for codes 111..116 and (m^2 - tHat)(m^2 - uHat) / sHat = pT^2 + m^2 for
 * the heavy pairs 121..124 (m = phys->mass_Q, uHat = 2 m^2 - sHat - tHat); alphaS per phys
 * (order 0/1, FP64).  pid_dev (may be NULL: all massless) is the events' process codes, as
 * mi_unweight_compact writes them.  Replaces the one-pass multi-weight stage of
 * cuPythia/pipeline/reweight.cu:36-51 and generate.cu:49-54.  nvar <= 64. */
int mi_reweight_scales(mi_ctx *ctx, const mi_physics *phys, uint64_t n, const double *sHat_dev,
                       const double *tHat_dev, const int32_t *pid_dev, const double *w_dev, int nvar,
                       const double *factors_host, double *out_dev);

/* One weight variation: mu_R = renorm_fac * Q, mu_F = fact_fac * Q (Q^2 = pT^2, or pT^2 + m^2 for
 * the heavy pairs) and the PDFs of a slot (0 = the grid of mi_pdf_load; slots 1..7 from
 * mi_pdf_load_slot, e.g. the members of an LHAPDF set). */
typedef struct mi_weight_var {
    double renorm_fac;
    double fact_fac;
    int    pdf_slot;
} mi_weight_var;

/* Scale and PDF variations of events in one pass: out[i * nvar + k] = w[i] * W_k / W_0, where W is
 * the event's own process weight -- alphaS(mu_R)^2 times its sigmaHat bracket times its flux sum
 * over the flavours (Pythia's sigmaPDF for that code) -- at the event's (sHat, tHat, x1, x2),
 * W_0 at phys's renorm_fac, fact_fac and slot 0.  FP64.  With the toy PDFs (phys->pdf_source =
 * MI_PDF_TOY) fact_fac has no effect and pdf_slot must be 0.  pid_dev: the events' codes (as
 * mi_unweight_compact writes them); nvar <= 64.  Replaces the multi-weight stage of
 * cuPythia/pipeline/reweight.cu:36-67 (there: mu_R only) for scales and PDF members. */
int mi_reweight(mi_ctx *ctx, const mi_physics *phys, uint64_t n, const double *sHat_dev, const double *tHat_dev,
                const double *x1_dev, const double *x2_dev, const int32_t *pid_dev, const double *w_dev,
                int nvar, const mi_weight_var *vars_host, double *out_dev);

/* Scale and PDF variations of events regenerated from their counters, as the reference reweights
 * (cuPythia/pipeline/reweight.cu:36-67: every weight recomputed from the event's own counter seed,
 * one pass equal to pinned re-runs).  Event i is the point of counter index_dev[i] under run (its
 * key, physics and map: flat, or the context's VEGAS grid, which must still be the one the events
 * were drawn with; run->n_points and counter0 are not used), evaluated in FP64 at each variation k:
 * run->phys with vars[k].renorm_fac, vars[k].fact_fac and the PDFs of slot vars[k].pdf_slot (grid
 * PDFs only; with the toy PDFs fact_fac has no effect and the slot must be 0).  W_k is the FP64
 * weight of the event's own process pid_dev[i]: bit for bit mi_points (FP64, same map) of the varied
 * run, column of that code.  out[i * nvar + k] = W_k when w_dev is NULL, else w_dev[i] * W_k / W_0,
 * W_0 at run->phys itself (e.g. the weights of unweighted events).  Codes the run's proc_mask does
 * not select (or 0) give 0.  FP64 for either run precision: with FP32 runs the ratio is that of the
 * FP64 point at the same uniforms.  nvar <= 64. */
int mi_reweight_events(mi_ctx *ctx, const mi_run *run, uint64_t n, const uint64_t *index_dev,
                       const int32_t *pid_dev, const double *w_dev, int nvar, const mi_weight_var *vars_host,
                       double *out_dev);

/* ------------------------------------------------------- time to precision ---- */
/* sigma of run's processes to the relative error rel_target, with plain VEGAS or with the flow on
 * top of VEGAS and plain VEGAS as the fallback.  The context's VEGAS grid is used as it is (adapt it
 * first with mi_vegas_adapt).  Passes, on consecutive counter ranges from run->counter0 (training on
 * counters from run->counter0 + 2^41):
 *   VEGAS pilot (pilot_points) -> the points N_V a VEGAS final pass needs;
 *   sampler = MI_SAMPLER_FLOW, each test falling back to VEGAS with no further flow work:
 *     - one training chunk (chunk_steps x batch points through the flow) plus the flow's pilot and
 *       smallest final pass would cost more than the VEGAS final pass even at the pilot's cost per
 *       point (a flow point evaluates the same integrand through the same map, plus the flow);
 *     - a VEGAS pass of batch points: with the pilot, VEGAS's affine pass cost t0 + c n (a pass
 *       has a fixed cost of launches and reductions); the same test with chunk_steps such passes;
 *     - mi_flow_init(flow_seed) and an identity-flow probe (probe_points) for the flow's pass cost
 *       (t0 + its own cost per point); the same test at that cost;
 *     - one training step: the rest of the first chunk at its cost, plus the flow's pilot and
 *       smallest final pass, against VEGAS's final pass;
 *     then training in chunks while the flow's projected pilot + final pass reaches below both that
 *     projection and VEGAS's final pass within the steps left.  The projection starts from the
 *     rate the last (up to three) chunks show and lets each further chunk gain 0.8 of the
 *     previous chunk's gain.  Then the flow (pilot, final pass) if its projection is below VEGAS's
 *     final pass, otherwise VEGAS;
 *   the final pass of max(min_points, 1.1 P (rel_pilot / rel_target)^2) points: its estimate is out.
 * The reference has no such driver: it integrates fixed sizes (cuPythia/11_vegas_importance/
 * vegas.cu:106-124); the flow is its roadmap item (cuPythia/README.md:75).  Zero fields of opts take
 * the defaults in brackets. */
typedef struct mi_solve_opts {
    double   rel_target;     /* sigmaErr / sigmaGen of the final pass, in (0, 1) */
    int      sampler;        /* MI_SAMPLER_VEGAS or MI_SAMPLER_FLOW (flow + VEGAS, VEGAS fallback) */
    int      chunk_steps;    /* training steps between decisions [10] */
    int      max_steps;      /* [400] */
    int      force_flow;     /* 1: train max_steps and finish with the flow regardless (measurement) */
    uint64_t pilot_points;   /* [2^20] */
    uint64_t probe_points;   /* [2^20] */
    uint64_t batch;          /* training batch [2^16] */
    uint64_t min_points;     /* smallest final pass [2^16] */
    double   lr;             /* Adam learning rate [1e-2] */
    uint64_t flow_seed;
} mi_solve_opts;

#define MI_SOLVE_VEGAS            0   /* plain VEGAS was asked for */
#define MI_SOLVE_FALLBACK_PROBE   1   /* flow asked for; VEGAS's final pass is cheaper than any continuation of the flow
                                         (from VEGAS's cost per point, or the probe's: probe_ms tells which) */
#define MI_SOLVE_FALLBACK_TRAINED 2   /* flow trained (one step or more), then no continuation was projected to
                                         finish below VEGAS's final pass */
#define MI_SOLVE_FLOW             3   /* the flow finished the solve */
#define MI_SOLVE_FLOW_FORCED      4   /* force_flow */

typedef struct mi_solve_info {
    int      sampler_used;         /* sampler of the final pass */
    int      decision;             /* MI_SOLVE_* */
    int      train_steps;
    int      reserved;
    double   gpu_ms;               /* device time of every pass of the solve (adaptation not included) */
    double   pilot_ms, probe_ms, train_ms, flow_pilot_ms, final_ms;
    double   vegas_ms_per_point, flow_ms_per_point;   /* per-point part of the pass costs (flow arm: affine) */
    double   cv_vegas, cv_flow;    /* per-point coefficients of variation (flow: its pilot, else the last training batch) */
    double   vegas_final_ms_proj;  /* projected VEGAS final pass */
    double   flow_rest_ms_proj;    /* projected flow pilot + final pass after training (0 if not trained) */
    uint64_t final_points;
} mi_solve_info;

int mi_solve(mi_ctx *ctx, const mi_run *run, const mi_solve_opts *opts, mi_sigma_est *out, mi_solve_info *info);

/* ---------------------------------------------------------------- adjoint ---- */
/* d sigma / d {alphaS(MZ), renorm_fac, gluon a, gluon b} in the same pass as sigma, on the
 * same points, FP64 accumulation.  Hadronic mode, flat sampler, any proc_mask within bits 0..9
 * (each heavy-quark family at its own scale pT^2 + m^2), alphaS_order 1 (or 0 for the fixed-
 * coupling alphaS^2 check).  The gluon shape parameters a, b are the toy PDFs' own: with
 * pdf_source = MI_PDF_GRID their derivatives (and errors) are 0. */
int mi_adjoint(mi_ctx *ctx, const mi_run *run, mi_sigma_grad *out);

/* ------------------------------------------------------------- PDF grids ---- */
/* Loads a proton PDF from an LHAPDF6 "lhagrid1" data file into the context, read as Pythia's
 * LHAGrid1 does (pythia8317/src/PartonDistributions.cc:478-598, e.g. the default
 * NNPDF23_lo_as_0130_qed_0000.dat); runs with phys.pdf_source = MI_PDF_GRID then take their
 * parton densities from it at mu_F^2 = fact_fac^2 pT^2, interpolated as LHAGrid1::xfxevolve
 * (:648-745) with PDF:extrapolate = off.  Replaces cuPythia's grid fill and interpolator
 * (cuPythia/pipeline/genpdf.cc:18-70, pdf.cuh:21-49).  A later call replaces the grid. */
int mi_pdf_load(mi_ctx *ctx, const char *path);
/* The same into PDF slot `slot` (0..MI_PDF_SLOTS-1; 0 is mi_pdf_load's): grids for mi_reweight's
 * PDF variations.  A later call replaces the slot's grid. */
int mi_pdf_load_slot(mi_ctx *ctx, int slot, const char *path);
/* Batched PDF evaluation (FP64, device arrays): xf_dev[i * 12 + slot] for slots g, d, u, s, c,
 * b, dbar, ubar, sbar, cbar, bbar (11 = photon, written 0), raw interpolated values (Pythia's
 * PDF::xf clips them at 0, PartonDistributions.cc:122-133).  The counterpart of Pythia's
 * PDF::xf(id, x, Q2) for all flavours at once. */
int mi_pdf_xf(mi_ctx *ctx, uint64_t n, const double *x_dev, const double *Q2_dev, double *xf_dev);

/* ------------------------------------------------------------------- flow ---- */
/* Identity initialisation (last layer of every conditioner zero) with hidden layers drawn
 * from Philox(seed); composes with the context's VEGAS grid (uniform grid if none). */
int mi_flow_init(mi_ctx *ctx, uint64_t seed);
/* Train with the variance loss on batches of `batch` self-generated points (counters from
 * run->counter0), Adam with learning rate lr, `steps` steps, entirely on the device. */
int mi_flow_train(mi_ctx *ctx, const mi_run *run, int steps, uint64_t batch, double lr,
                  mi_flow_info *out);
/* Normalizing direction y -> z (y, z in [0,1)^3, n x 3 floats) with log|det dz/dy|. */
int mi_flow_forward(mi_ctx *ctx, uint64_t n, const float *y_dev, float *z_dev, float *logdet_dev);
/* Sampling direction z -> y with log|det dy/dz|. */
int mi_flow_inverse(mi_ctx *ctx, uint64_t n, const float *z_dev, float *y_dev, float *logdet_dev);
/* grad_dev (MI_FLOW_NPARAMS floats, overwritten) = d/dtheta sum_i g_i * logdet_fwd(y_i). */
int mi_flow_backward(mi_ctx *ctx, uint64_t n, const float *y_dev, const float *g_dev,
                     float *grad_dev);
/* FP32 master copy of the parameters, layer-major: per layer W1[64][2], b1[64], W2[64][64],
 * b2[64], W3[49][64], b3[49]. */
int mi_flow_get_params(mi_ctx *ctx, float *params_host);
int mi_flow_set_params(mi_ctx *ctx, const float *params_host);

#ifdef __cplusplus
}
#endif
#endif
