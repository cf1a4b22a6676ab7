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
// Random numbers: Philox streams 3 and 4 of the event index -- stream 3 words 0, 1 the pair and
// words 2, 3 the colour topology (53-bit uniforms: a decision compared with a computed threshold
// flips with the threshold's last-bit rounding only with probability ~1e-16), stream 4 words 0, 1
// the azimuth (53 bits), word 2 the g g -> g g orientation and word 3 the new flavour of
// 112 / 116 (24 bits, compared with 1/2 or scaled by nQuarkNew exactly) -- so the record is a
// function of (key, event index).
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
// antiquarks (the PDF grid's order).  Used with compile-time ids only, so the densities stay in
// registers.
__host__ __device__ constexpr int xf_slot(int id) { return id == 21 ? 0 : (id > 0 ? id : 5 - id); }

__device__ __forceinline__ double uni24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

// Densities of one beam at x: the toy PDFs (Q^2-independent, A exp(a ln x + b ln(1 - x)) with
// the two logarithms shared), or the grid at the q stencil of the event's factorization scale
// (clipped at 0, as PDF::xf).
template <bool GRID>
__device__ __forceinline__ void beam_xf(const EvConst &c, const GridPdf<double> &g, int m3q, int n3q,
                                        const double *wq, double x, double (&f)[11])
{
    if (GRID) {
        grid_xf<double>(g, x, log(x), m3q, n3q, wq, f);
#pragma unroll
        for (int i = 0; i < 11; ++i) f[i] = f[i] > 0.0 ? f[i] : 0.0;
        return;
    }
#pragma unroll
    for (int i = 0; i < 11; ++i) f[i] = 0.0;
    if (!(x > 0.0 && x < 1.0)) return;
    const double lx = log(x), l1 = log(1.0 - x);
    const double sea = c.As * exp(fma(c.as, lx, c.bs * l1));
    const double val = exp(fma(c.av, lx, c.bv * l1));
    f[0] = c.Ag * exp(fma(c.ag, lx, c.bg * l1));
#pragma unroll
    for (int q = 1; q <= 5; ++q) {
        const double v = q <= c.nQ ? sea : 0.0;
        f[q] = v;
        f[5 + q] = v;
    }
    f[1] = fma(c.Adv, val, sea);
    f[2] = fma(c.Auv, val, sea);
}

// Pair j of the flux list of KIND (0 "qg", 1 "qq", 2 "qqbarSame") in initFlux order, for quark
// index q = -5..5 (0 and |q| > nQ skipped by the caller) and, for "qq", the second quark q2.
// The draw: weights summed in list order, then the first pair where u * sum minus the running
// weights reaches 0 (the last pair if rounding never gets there).
template <int KIND>
__device__ __forceinline__ void pick_pair(int nQ, const double (&f1)[11], const double (&f2)[11],
                                          const double (&br)[3], double u, int &id1, int &id2)
{
    double tot = 0.0;
#pragma unroll
    for (int q = -5; q <= 5; ++q) {
        if (q == 0 || q > nQ || -q > nQ) continue;
        if (KIND == 0) {
            tot += f1[xf_slot(q)] * f2[0];
            tot += f1[0] * f2[xf_slot(q)];
        } else if (KIND == 1) {
#pragma unroll
            for (int q2 = -5; q2 <= 5; ++q2) {
                if (q2 == 0 || q2 > nQ || -q2 > nQ) continue;
                tot += f1[xf_slot(q)] * f2[xf_slot(q2)] * br[q2 == q ? 0 : (q2 == -q ? 1 : 2)];
            }
        } else {
            tot += f1[xf_slot(q)] * f2[xf_slot(-q)];
        }
    }
    double target = u * tot;
    bool done = false;
#pragma unroll
    for (int q = -5; q <= 5; ++q) {
        if (q == 0 || q > nQ || -q > nQ) continue;
        if (KIND == 0) {
            if (!done) { id1 = q; id2 = 21; target -= f1[xf_slot(q)] * f2[0]; done = target <= 0.0; }
            if (!done) { id1 = 21; id2 = q; target -= f1[0] * f2[xf_slot(q)]; done = target <= 0.0; }
        } else if (KIND == 1) {
#pragma unroll
            for (int q2 = -5; q2 <= 5; ++q2) {
                if (q2 == 0 || q2 > nQ || -q2 > nQ) continue;
                if (!done) {
                    id1 = q; id2 = q2;
                    target -= f1[xf_slot(q)] * f2[xf_slot(q2)] * br[q2 == q ? 0 : (q2 == -q ? 1 : 2)];
                    done = target <= 0.0;
                }
            }
        } else {
            if (!done) { id1 = q; id2 = -q; target -= f1[xf_slot(q)] * f2[xf_slot(-q)]; done = target <= 0.0; }
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
#pragma unroll
    for (int k = 0; k < 4; ++k) { const int t = F.col[k]; F.col[k] = F.acol[k]; F.acol[k] = t; }
}

__device__ __forceinline__ int tag(int c) { return c ? 500 + c : 0; }

// VEC: max_part is a multiple of 4 and every record array 16-byte aligned, so each thread writes
// its 4 partons as 16-byte vectors (one int4 per int32 array, two double2 per FP64 array) and a
// wave's stores cover whole lines.  GRID: parton densities from the PDF grid (else the toy PDFs).
template <bool VEC, bool GRID>
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
    // one reciprocal each; the brackets below are products of the ratios
    const double rs = 1.0 / s, rt = 1.0 / t, ru = 1.0 / u;
    const double ts = t * rs, us = u * rs, st = s * rt, su = s * ru, tu = t * ru, ut = u * rt;

    // ---- incoming flavours
    int id1 = 21, id2 = 21;
    if (!c.partonic && code != 111 && code != 112 && code != 121 && code != 123) {
        int m3q = 0, n3q = 0;
        double wq[4] = {0.0, 0.0, 0.0, 0.0};
        if (GRID) {
            const double q = sqrt(gpdf.factFac2 * Q2);
            grid_qw<double>(gpdf, q, log(q), m3q, n3q, wq);
        }
        double f1[11], f2[11];
        beam_xf<GRID>(c, gpdf, m3q, n3q, wq, X1, f1);
        beam_xf<GRID>(c, gpdf, m3q, n3q, wq, X2, f2);
        const double up = u53(r3.x, r3.y);
        if (code == 113) {
            const double br[3] = {1.0, 1.0, 1.0};
            pick_pair<0>(c.nQ, f1, f2, br, up, id1, id2);
        } else if (code == 114) {                                    // Sigma2qq2qq::sigmaHat brackets
            const double sigT = (4. / 9.) * fma(st, st, ut * ut);
            const double br[3] = {0.5 * (sigT + (4. / 9.) * fma(su, su, tu * tu) - (8. / 27.) * (st * su)),
                                  sigT - (8. / 27.) * (us * ut), sigT};
            pick_pair<1>(c.nQ, f1, f2, br, up, id1, id2);
        } else {
            const double br[3] = {1.0, 1.0, 1.0};
            pick_pair<2>(c.nQ, f1, f2, br, up, id1, id2);
        }
    }

    // ---- outgoing flavours and colour flow (the topology weights up to common factors)
    Flow F;
    int id3 = 21, id4 = 21;
    const double uc = u53(r3.z, r3.w);
    switch (code) {
    case 111: {
        const double a = fma(ts, ts + 2.0, fma(st, st + 2.0, 3.0));
        const double b = fma(us, us + 2.0, fma(su, su + 2.0, 3.0));
        const double d = fma(tu, tu + 2.0, fma(ut, ut + 2.0, 3.0));
        const double r = (a + b + d) * uc;
        if (r < a) set_flow(F, 1, 2, 2, 3, 1, 4, 4, 3);
        else if (r < a + b) set_flow(F, 1, 2, 3, 1, 3, 4, 4, 2);
        else set_flow(F, 1, 2, 3, 4, 1, 4, 3, 2);
        if (uni24(r4.z) > 0.5) swap_col_acol(F);
        break;
    }
    case 113: {
        const double a = fma(ut, ut, -(4. / 9.) * us), b = fma(st, st, -(4. / 9.) * su);
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
        const double sigT = fma(st, st, ut * ut), sigU = fma(su, su, tu * tu);
        if (id2 == id1 && (sigT + sigU) * uc > sigT) set_flow(F, 1, 0, 2, 0, 1, 0, 2, 0);
        if (id1 < 0) swap_col_acol(F);
        break;
    }
    case 115: {
        const double a = fma(32. / 27., ut, -(8. / 3.) * (us * us)), b = fma(32. / 27., tu, -(8. / 3.) * (ts * ts));
        if ((a + b) * uc < a) set_flow(F, 1, 0, 0, 2, 1, 3, 3, 2);
        else set_flow(F, 1, 0, 0, 2, 3, 2, 1, 3);
        if (id1 < 0) swap_col_acol(F);
        break;
    }
    case 112:
    case 121:
    case 123: {
        // gg -> q qbar: Sigma2gg2qqbar's sigTS / sigUS (times 6); gg -> Q Qbar: Sigma2gg2QQbar's
        // massive forms with tHQ = t - m^2, uHQ = u - m^2 (times 6), equal to the former at m = 0
        const double tq = t - m2, uq = u - m2;
        const double rtq = heavy ? 1.0 / tq : rt, ruq = heavy ? 1.0 / uq : ru;
        const double mus = m2 * rs, tum = fma(tq, uq, -m2 * s);
        const double a = uq * rtq - 2.25 * (uq * rs) * (uq * rs) +
                         m2 * rtq * rtq * fma(4.5 * rs, tum, 0.5 * (tq + m2)) - m2 * mus * rtq;
        const double b = tq * ruq - 2.25 * (tq * rs) * (tq * rs) +
                         m2 * ruq * ruq * fma(4.5 * rs, tum, 0.5 * (uq + m2)) - m2 * mus * ruq;
        const int idNew = heavy ? idQ : 1 + (int)(c.nQuarkNew * uni24(r4.w));
        id3 = idNew; id4 = -idNew;
        if ((a + b) * uc < a) set_flow(F, 1, 2, 2, 3, 1, 0, 0, 3);
        else set_flow(F, 1, 2, 3, 1, 3, 0, 0, 2);
        break;
    }
    default: {                                                       // 116; 122 / 124 (Sigma2qqbar2QQbar)
        const int idNew = heavy ? idQ : 1 + (int)(c.nQuarkNew * uni24(r4.w));
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
    const double bQ = heavy ? sqrt(fmax(0.0, 1.0 - 4.0 * m2 * rs)) : 1.0;
    const double cth = fmin(1.0, fmax(-1.0, (1.0 - 2.0 * (m2 - t) * rs) / bQ));
    const double sth = sqrt(fmax(0.0, (1.0 - cth) * (1.0 + cth)));   // no cancellation at |cth| -> 1
    double sphi, cphi;
    sincos(6.283185307179586 * u53(r4.x, r4.y), &sphi, &cphi);
    // boost along z by beta = (x1 - x2) / (x1 + x2): gamma and gamma beta as ratios of the x's,
    // exact where 1 - beta^2 would cancel (x1 >> x2)
    const double r12 = 0.5 / sqrt(X1 * X2);
    const double gam = (X1 + X2) * r12, gb = (X1 - X2) * r12;
    const double px3 = pabs * sth * cphi, py3 = pabs * sth * sphi, pzc = pabs * cth;
    const double pz3 = gam * pzc + gb * eCMp, e3 = gam * eCMp + gb * pzc;
    const double pz4 = gb * eCMp - gam * pzc, e4 = gam * eCMp - gb * pzc;

    const uint64_t b0 = i * (uint64_t)c.max_part;
    if (VEC) {
        auto st2 = [](double *p, uint64_t j, double a, double b, double d, double e) {
            reinterpret_cast<double2 *>(p + j)[0] = make_double2(a, b);
            reinterpret_cast<double2 *>(p + j)[1] = make_double2(d, e);
        };
        auto st4 = [](int32_t *p, uint64_t j, int a, int b, int d, int e) {
            *reinterpret_cast<int4 *>(p + j) = make_int4(a, b, d, e);
        };
        st2(out.px, b0, 0.0, 0.0, px3, -px3);
        st2(out.py, b0, 0.0, 0.0, py3, -py3);
        st2(out.pz, b0, pz1, -pz2, pz3, pz4);
        st2(out.e, b0, pz1, pz2, e3, e4);
        st2(out.m, b0, 0.0, 0.0, mQ, mQ);
        st4(out.pdg, b0, id1, id2, id3, id4);
        st4(out.status, b0, -21, -21, 23, 23);
        st4(out.col, b0, tag(F.col[0]), tag(F.col[1]), tag(F.col[2]), tag(F.col[3]));
        st4(out.acol, b0, tag(F.acol[0]), tag(F.acol[1]), tag(F.acol[2]), tag(F.acol[3]));
        st4(out.mo1, b0, -1, -1, 0, 0);
        st4(out.mo2, b0, -1, -1, 1, 1);
        st4(out.d1, b0, 2, 2, -1, -1);
        st4(out.d2, b0, 3, 3, -1, -1);
    } else {
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
            out.col[j] = tag(F.col[k]);
            out.acol[j] = tag(F.acol[k]);
            out.mo1[j] = k < 2 ? -1 : 0;
            out.mo2[j] = k < 2 ? -1 : 1;
            out.d1[j] = k < 2 ? 2 : -1;
            out.d2[j] = k < 2 ? 3 : -1;
        }
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
    bool vec = out.max_part % 4 == 0;
    const void *arrs[] = {out.px, out.py, out.pz, out.e, out.m, out.pdg, out.status, out.col, out.acol, out.mo1,
                          out.mo2, out.d1, out.d2};
    for (const void *p : arrs) vec = vec && (reinterpret_cast<uintptr_t>(p) % 16 == 0);
    auto kern = vec ? (c.grid ? k_build_events<true, true> : k_build_events<true, false>)
                    : (c.grid ? k_build_events<false, true> : k_build_events<false, false>);
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, s, c, g, n, key, ev.sHat + first, ev.tHat + first, ev.x1 + first,
                       ev.x2 + first, ev.pid + first, ev.index + first, out);
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
