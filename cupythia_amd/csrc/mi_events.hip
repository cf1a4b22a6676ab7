// mi_events.hip -- the device event record: the partons of the compacted events with flavours,
// colour flow and four-momenta, one thread per event.
//
// What the compacted record (sHat, tHat, x1, x2, process code, event index) leaves open is assigned
// as Pythia completes a hard process:
//   incoming flavours  -- a pair from the code's flux list in initFlux order
//                         (pythia8317/src/SigmaProcess.cc:124-193: "gg", "qg" = (q, g), (g, q) per
//                         q = -nQ..nQ, "qq" = all (q1, q2), "qqbarSame" = (q, -q)) drawn with weight
//                         xf1 xf2 dsigmaHat/dtHat (pickInState, SigmaProcess.cc:486-508; only the
//                         flavour-dependent factor is kept: the q q' bracket of Sigma2qq2qq::sigmaHat,
//                         SigmaQCD.cc:257-286, is the only one);
//   outgoing flavours and colour flow -- setIdColAcol of the code (SigmaQCD.cc:135-148, 195-206,
//                         232-243, 292-303, 331-342, 384-394; the heavy pairs as Sigma2gg2QQbar
//                         :768-781 with the massive sigTS / sigUS of :736-755, and Sigma2qqbar2QQbar
//                         :894-905, no top model);
//   azimuth            -- uniform.
// Random numbers: Philox stream 3 of the event index (word 0 the pair, 1 the colour topology,
// 2 the g g -> g g orientation, 3 the new flavour of 112 / 116) and stream 4 (word 0 the azimuth),
// 24-bit uniforms, so the record is a function of (key, event index).
// Momenta (PhaseSpace.cc:2236-2251): incoming partons along +-z with E = x sqrt(S)/2, the outgoing
// pair back to back in the partonic frame, |p| = sqrt(x1 x2 S / 4 - m^2) and cos theta =
// (1 - 2 a) / beta with a = (m^2 - tHat) / sHat and beta = sqrt(1 - 4 m^2 / sHat) from the record,
// then boosted along z by (x1 - x2) / (x1 + x2) (gamma = (x1 + x2) / (2 sqrt(x1 x2)), free of the
// cancellation in 1 - beta^2 that Vec4::bst has).  The record's sHat agrees with x1 x2 S to the
// precision it was computed in; the momenta use x1 x2 S, which conserves four-momentum to FP64
// rounding.
// Reference counterpart: cuPythia/pipeline/event.cuh:20-71 and build_events.cu:22-41 (g g -> g g
// at fixed sHat with one colour flow), cuPythia/12_phase_space/phase_space.cu:37-53.
#include "mi_internal.h"

#include <cmath>

namespace mi {

namespace {

// Slot of a PDG id in the 11 parton densities of a beam: 0 gluon, 1..5 d u s c b, 6..10 their
// antiquarks (the PDF grid's order).
__device__ __forceinline__ int xf_slot(int id) { return id == 21 ? 0 : (id > 0 ? id : 5 - id); }

__device__ __forceinline__ double uni24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

// Densities of one beam at x: the toy PDFs (Q^2-independent), or the grid at the q stencil of
// the event's factorization scale (clipped at 0, as PDF::xf).
__device__ void beam_xf(const EvConst &c, const GridPdf<double> *g, int m3q, int n3q, const double *wq, double x,
                        double f[11])
{
    if (g) {
        grid_xf<double>(*g, x, log(x), m3q, n3q, wq, f);
        for (int i = 0; i < 11; ++i) f[i] = f[i] > 0.0 ? f[i] : 0.0;
        return;
    }
    for (int i = 0; i < 11; ++i) f[i] = 0.0;
    if (!(x > 0.0 && x < 1.0)) return;
    const double omx = 1.0 - x;
    f[0] = c.Ag * pow(x, c.ag) * pow(omx, c.bg);
    const double sea = c.As * pow(x, c.as) * pow(omx, c.bs);
    const double pv = pow(x, c.av), pw = pow(omx, c.bv);
    for (int q = 1; q <= c.nQ; ++q) { f[q] = sea; f[5 + q] = sea; }
    f[1] = sea + c.Adv * pv * pw;
    f[2] = sea + c.Auv * pv * pw;
}

// The flux-list draw: weights in initFlux order, summed in that order, then the first pair where
// u * sum minus the running weights reaches 0 (the last pair if rounding never gets there).
// kind 0: "qg", 1: "qq" (bracket per flavour relation), 2: "qqbarSame".
__device__ void pick_pair(int kind, int nQ, const double *f1, const double *f2, const double br[3], double u,
                          int &id1, int &id2)
{
    double tot = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        double target = u * tot;
        for (int q = -nQ; q <= nQ; ++q) {
            if (q == 0) continue;
            const int n2 = kind == 1 ? 2 * nQ : (kind == 0 ? 2 : 1);
            for (int j = 0; j < n2; ++j) {
                int a, b;
                double w;
                if (kind == 0) {
                    a = j == 0 ? q : 21;
                    b = j == 0 ? 21 : q;
                    w = f1[xf_slot(a)] * f2[xf_slot(b)];
                } else if (kind == 1) {
                    a = q;
                    b = j < nQ ? j - nQ : j - nQ + 1;
                    w = f1[xf_slot(a)] * f2[xf_slot(b)] * br[b == a ? 0 : (b == -a ? 1 : 2)];
                } else {
                    a = q;
                    b = -q;
                    w = f1[xf_slot(a)] * f2[xf_slot(b)];
                }
                if (pass == 0) {
                    tot += w;
                } else {
                    id1 = a;
                    id2 = b;
                    target -= w;
                    if (target <= 0.0) return;
                }
            }
        }
    }
}

struct Flow { int col[4], acol[4]; };

__device__ __forceinline__ void set_flow(Flow &F, int c1, int a1, int c2, int a2, int c3, int a3, int c4, int a4)
{
    F.col[0] = c1; F.acol[0] = a1; F.col[1] = c2; F.acol[1] = a2;
    F.col[2] = c3; F.acol[2] = a3; F.col[3] = c4; F.acol[3] = a4;
}

__device__ __forceinline__ void swap_col_acol(Flow &F)
{
    for (int k = 0; k < 4; ++k) { const int t = F.col[k]; F.col[k] = F.acol[k]; F.acol[k] = t; }
}

__global__ __launch_bounds__(kBlock) void k_build_events(EvConst c, GridPdf<double> gpdf, uint64_t n, uint64_t key,
                                                         const double *__restrict__ sHat,
                                                         const double *__restrict__ tHat,
                                                         const double *__restrict__ x1v,
                                                         const double *__restrict__ x2v,
                                                         const int32_t *__restrict__ pid,
                                                         const uint64_t *__restrict__ index, mi_particles out)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int code = pid[i];
    const bool heavy = code >= 121 && code <= 124;
    if (!((code >= 111 && code <= 116) || heavy) || (c.partonic && code != 111)) {
        out.n_part[i] = 0;
        return;
    }
    const U4 r3 = philox_event(key, index[i], 3u);
    const U4 r4 = philox_event(key, index[i], 4u);
    const int idQ = (code == 121 || code == 122) ? 4 : 5;
    const double mQ = heavy ? c.mQ[idQ - 4] : 0.0, m2 = mQ * mQ;
    const double s = sHat[i], t = tHat[i], u = 2.0 * m2 - (s + t);
    const double Q2 = (m2 - t) * (m2 - u) / s;                       // pT^2 + m^2
    const double X1 = c.partonic ? 1.0 : x1v[i], X2 = c.partonic ? 1.0 : x2v[i];

    // ---- incoming flavours
    int id1 = 21, id2 = 21;
    if (!c.partonic && code != 111 && code != 112 && code != 121 && code != 123) {
        const GridPdf<double> *g = c.grid ? &gpdf : nullptr;
        int m3q = 0, n3q = 0;
        double wq[4] = {0.0, 0.0, 0.0, 0.0};
        if (g) {
            const double q = sqrt(gpdf.factFac2 * Q2);
            grid_qw<double>(gpdf, q, log(q), m3q, n3q, wq);
        }
        double f1[11], f2[11];
        beam_xf(c, g, m3q, n3q, wq, X1, f1);
        beam_xf(c, g, m3q, n3q, wq, X2, f2);
        double br[3] = {1.0, 1.0, 1.0};
        if (code == 114) {                                           // Sigma2qq2qq::sigmaHat brackets
            const double sigT = (4. / 9.) * (s * s + u * u) / (t * t);
            br[0] = 0.5 * (sigT + (4. / 9.) * (s * s + t * t) / (u * u) - (8. / 27.) * s * s / (t * u));
            br[1] = sigT - (8. / 27.) * u * u / (s * t);
            br[2] = sigT;
        }
        pick_pair(code == 113 ? 0 : (code == 114 ? 1 : 2), c.nQ, f1, f2, br, uni24(r3.x), id1, id2);
    }

    // ---- outgoing flavours and colour flow
    Flow F;
    int id3 = 21, id4 = 21;
    const double uc = uni24(r3.y), us = uni24(r3.z), un = uni24(r3.w);
    const double s2 = s * s, t2 = t * t, u2 = u * u;
    switch (code) {
    case 111: {
        const double a = (9. / 4.) * (t2 / s2 + 2. * t / s + 3. + 2. * s / t + s2 / t2);
        const double b = (9. / 4.) * (u2 / s2 + 2. * u / s + 3. + 2. * s / u + s2 / u2);
        const double d = (9. / 4.) * (t2 / u2 + 2. * t / u + 3. + 2. * u / t + u2 / t2);
        const double r = (a + b + d) * uc;
        if (r < a) set_flow(F, 1, 2, 2, 3, 1, 4, 4, 3);
        else if (r < a + b) set_flow(F, 1, 2, 3, 1, 3, 4, 4, 2);
        else set_flow(F, 1, 2, 3, 4, 1, 4, 3, 2);
        if (us > 0.5) swap_col_acol(F);
        break;
    }
    case 112: {
        const int idNew = 1 + (int)(c.nQuarkNew * un);
        const double a = (1. / 6.) * u / t - (3. / 8.) * u2 / s2, b = (1. / 6.) * t / u - (3. / 8.) * t2 / s2;
        id3 = idNew; id4 = -idNew;
        if ((a + b) * uc < a) set_flow(F, 1, 2, 2, 3, 1, 0, 0, 3);
        else set_flow(F, 1, 2, 3, 1, 3, 0, 0, 2);
        break;
    }
    case 113: {
        const double a = u2 / t2 - (4. / 9.) * u / s, b = s2 / t2 - (4. / 9.) * s / u;
        id3 = id1; id4 = id2;
        if ((a + b) * uc < a) set_flow(F, 1, 0, 2, 1, 3, 0, 2, 3);
        else set_flow(F, 1, 0, 2, 3, 2, 0, 1, 3);
        if (id1 == 21) {                                             // swapCol1234
            int tmp = F.col[0]; F.col[0] = F.col[1]; F.col[1] = tmp;
            tmp = F.col[2]; F.col[2] = F.col[3]; F.col[3] = tmp;
            tmp = F.acol[0]; F.acol[0] = F.acol[1]; F.acol[1] = tmp;
            tmp = F.acol[2]; F.acol[2] = F.acol[3]; F.acol[3] = tmp;
        }
        if (id1 < 0 || id2 < 0) swap_col_acol(F);
        break;
    }
    case 114: {
        id3 = id1; id4 = id2;
        if (id1 * id2 > 0) set_flow(F, 1, 0, 2, 0, 2, 0, 1, 0);
        else set_flow(F, 1, 0, 0, 1, 2, 0, 0, 2);
        const double sigT = (4. / 9.) * (s2 + u2) / t2, sigU = (4. / 9.) * (s2 + t2) / u2;
        if (id2 == id1 && (sigT + sigU) * uc > sigT) set_flow(F, 1, 0, 2, 0, 1, 0, 2, 0);
        if (id1 < 0) swap_col_acol(F);
        break;
    }
    case 115: {
        const double a = (32. / 27.) * u / t - (8. / 3.) * u2 / s2, b = (32. / 27.) * t / u - (8. / 3.) * t2 / s2;
        if ((a + b) * uc < a) set_flow(F, 1, 0, 0, 2, 1, 3, 3, 2);
        else set_flow(F, 1, 0, 0, 2, 3, 2, 1, 3);
        if (id1 < 0) swap_col_acol(F);
        break;
    }
    case 121:
    case 123: {                                                      // Sigma2gg2QQbar
        const double tq = t - m2, uq = u - m2, tq2 = tq * tq, uq2 = uq * uq;
        const double tum = tq * uq - m2 * s;
        const double a = (uq / tq - 2.25 * uq2 / s2 + 4.5 * m2 * tum / (s * tq2) + 0.5 * m2 * (tq + m2) / tq2 -
                          m2 * m2 / (s * tq)) / 6.;
        const double b = (tq / uq - 2.25 * tq2 / s2 + 4.5 * m2 * tum / (s * uq2) + 0.5 * m2 * (uq + m2) / uq2 -
                          m2 * m2 / (s * uq)) / 6.;
        id3 = idQ; id4 = -idQ;
        if ((a + b) * uc < a) set_flow(F, 1, 2, 2, 3, 1, 0, 0, 3);
        else set_flow(F, 1, 2, 3, 1, 3, 0, 0, 2);
        break;
    }
    default: {                                                       // 116; 122 / 124 (Sigma2qqbar2QQbar)
        const int idNew = heavy ? idQ : 1 + (int)(c.nQuarkNew * un);
        id3 = id1 > 0 ? idNew : -idNew;
        id4 = -id3;
        set_flow(F, 1, 0, 0, 2, 1, 0, 0, 2);
        if (id1 < 0) swap_col_acol(F);
        break;
    }
    }

    // ---- momenta
    const double pz1 = X1 * c.eBeam, pz2 = X2 * c.eBeam;
    const double sLab = 4.0 * pz1 * pz2;
    const double eCMp = 0.5 * sqrt(sLab);
    const double pabs = sqrt(fmax(0.0, 0.25 * sLab - m2));
    const double bQ = sqrt(fmax(0.0, 1.0 - 4.0 * m2 / s));
    const double cth = fmin(1.0, fmax(-1.0, (1.0 - 2.0 * (m2 - t) / s) / (heavy ? bQ : 1.0)));
    const double sth = sqrt(fmax(0.0, (1.0 - cth) * (1.0 + cth)));   // no cancellation at |cth| -> 1
    double sphi, cphi;
    sincos(6.283185307179586 * uni24(r4.x), &sphi, &cphi);
    // boost along z by beta = (x1 - x2) / (x1 + x2): gamma and gamma beta as ratios of the x's,
    // exact where 1 - beta^2 would cancel (x1 >> x2)
    const double r12 = 0.5 / sqrt(X1 * X2);
    const double gam = (X1 + X2) * r12, gb = (X1 - X2) * r12;
    const double px3 = pabs * sth * cphi, py3 = pabs * sth * sphi, pzc = pabs * cth;
    const double pz3 = gam * pzc + gb * eCMp, e3 = gam * eCMp + gb * pzc;
    const double pz4 = gb * eCMp - gam * pzc, e4 = gam * eCMp - gb * pzc;

    const uint64_t b0 = i * (uint64_t)c.max_part;
    const int ids[4] = {id1, id2, id3, id4};
    const double P[4][5] = {{0.0, 0.0, pz1, pz1, 0.0},
                            {0.0, 0.0, -pz2, pz2, 0.0},
                            {px3, py3, pz3, e3, mQ},
                            {-px3, -py3, pz4, e4, mQ}};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t j = b0 + k;
        out.px[j] = P[k][0]; out.py[j] = P[k][1]; out.pz[j] = P[k][2]; out.e[j] = P[k][3]; out.m[j] = P[k][4];
        out.pdg[j] = ids[k];
        out.status[j] = k < 2 ? -21 : 23;
        out.col[j] = F.col[k] ? 500 + F.col[k] : 0;
        out.acol[j] = F.acol[k] ? 500 + F.acol[k] : 0;
        out.mo1[j] = k < 2 ? -1 : 0;
        out.mo2[j] = k < 2 ? -1 : 1;
        out.d1[j] = k < 2 ? 2 : -1;
        out.d2[j] = k < 2 ? 3 : -1;
    }
    out.n_part[i] = 4;
    out.scale[i] = c.renorm_fac * sqrt(fmax(Q2, 0.0));
}

double beta_fn(double p, double q) { return std::exp(std::lgamma(p) + std::lgamma(q) - std::lgamma(p + q)); }

}  // namespace

int event_consts(mi_ctx *ctx, const mi_physics &p, int max_part, EvConst *c, GridPdf<double> *g)
{
    *c = EvConst{};
    *g = GridPdf<double>{};
    c->partonic = p.mode == MI_MODE_PARTONIC_1D;
    if (!c->partonic && p.mode != MI_MODE_HADRONIC_3D) { ctx->err = "unknown mode"; return MI_ERR_ARG; }
    if (c->partonic ? !(p.sHat > 0.0) : !(p.eCM > 0.0)) { ctx->err = "sHat / eCM must be > 0"; return MI_ERR_ARG; }
    if (!(p.renorm_fac > 0.0)) { ctx->err = "renorm_fac must be > 0"; return MI_ERR_ARG; }
    if (!(p.mass_Q[0] >= 0.0) || !(p.mass_Q[1] >= 0.0)) { ctx->err = "mass_Q must be >= 0"; return MI_ERR_ARG; }
    if (max_part < 4) { ctx->err = "max_part must be >= 4"; return MI_ERR_ARG; }
    c->eBeam = 0.5 * (c->partonic ? std::sqrt(p.sHat) : p.eCM);
    c->renorm_fac = p.renorm_fac;
    c->mQ[0] = p.mass_Q[0];
    c->mQ[1] = p.mass_Q[1];
    c->max_part = max_part;
    if (c->partonic) return MI_OK;
    if (p.nQuarkIn < 2 || p.nQuarkIn > 5) { ctx->err = "nQuarkIn must be 2..5"; return MI_ERR_ARG; }
    if (p.nQuarkNew < 1 || p.nQuarkNew > 5) { ctx->err = "nQuarkNew must be 1..5"; return MI_ERR_ARG; }
    c->nQ = p.nQuarkIn;
    c->nQuarkNew = p.nQuarkNew;
    if (p.pdf_source == MI_PDF_GRID) {
        if (!(p.fact_fac > 0.0)) { ctx->err = "fact_fac must be > 0"; return MI_ERR_ARG; }
        const int rc = pdf_views(ctx, p, nullptr, g);
        if (rc) return rc;
        c->grid = 1;
        return MI_OK;
    }
    if (p.pdf_source != MI_PDF_TOY) { ctx->err = "unknown pdf_source"; return MI_ERR_ARG; }
    // toy normalisation: int u_v = 2, int d_v = 1, gluon momentum fraction 1/2, sea the rest of
    // the momentum sum rule, shared by 2 nQuarkIn quarks and antiquarks
    if (!(p.pdf_g[0] > -1.0) || !(p.pdf_val[0] > 0.0) || !(p.pdf_sea[0] > -1.0)) {
        ctx->err = "toy PDF exponents out of range";
        return MI_ERR_ARG;
    }
    c->ag = p.pdf_g[0]; c->bg = p.pdf_g[1];
    c->av = p.pdf_val[0]; c->bv = p.pdf_val[1];
    c->as = p.pdf_sea[0]; c->bs = p.pdf_sea[1];
    const double bnum = beta_fn(c->av, c->bv + 1.0);
    c->Auv = 2.0 / bnum;
    c->Adv = 1.0 / bnum;
    const double momV = (c->Auv + c->Adv) * beta_fn(c->av + 1.0, c->bv + 1.0);
    c->Ag = 0.5 / beta_fn(c->ag + 1.0, c->bg + 1.0);
    c->As = (0.5 - momV) / (2.0 * c->nQ * beta_fn(c->as + 1.0, c->bs + 1.0));
    return MI_OK;
}

hipError_t launch_build_events(hipStream_t s, const EvConst &c, const GridPdf<double> &g, uint64_t n, uint64_t key,
                               const mi_events &ev, uint64_t first, const mi_particles &out)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_build_events, grid, dim3(kBlock), 0, s, c, g, n, key, ev.sHat + first, ev.tHat + first,
                       ev.x1 + first, ev.x2 + first, ev.pid + first, ev.index + first, out);
    return hipGetLastError();
}

}  // namespace mi

using namespace mi;

extern "C" int mi_build_events(mi_ctx *ctx, const mi_physics *phys, const mi_events *ev, uint64_t n, uint64_t key,
                               const mi_particles *out)
{
    if (!ctx) return MI_ERR_ARG;
    if (!phys || !ev || !out) { ctx->err = "NULL argument"; return MI_ERR_ARG; }
    if (n > ev->capacity || n > out->capacity) { ctx->err = "n exceeds the event or record capacity"; return MI_ERR_ARG; }
    if (n > (uint64_t)0x7FFFFFFF * kBlock) { ctx->err = "n too large"; return MI_ERR_RANGE; }
    if (n && (!ev->sHat || !ev->tHat || !ev->x1 || !ev->x2 || !ev->pid || !ev->index)) {
        ctx->err = "NULL event array";
        return MI_ERR_ARG;
    }
    if (n && (!out->px || !out->py || !out->pz || !out->e || !out->m || !out->pdg || !out->status || !out->col ||
              !out->acol || !out->mo1 || !out->mo2 || !out->d1 || !out->d2 || !out->n_part || !out->scale)) {
        ctx->err = "NULL record array";
        return MI_ERR_ARG;
    }
    EvConst c;
    GridPdf<double> g;
    int rc = event_consts(ctx, *phys, out->max_part, &c, &g);
    if (rc) return rc;
    if (n == 0) return MI_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return MI_ERR_HIP; }
    hipError_t e = launch_build_events(ctx->stream, c, g, n, key, *ev, 0, *out);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->err = std::string("event build failed: ") + hipGetErrorString(e);
        return MI_ERR_HIP;
    }
    return MI_OK;
}
