// mi_lhe.hip -- Les Houches Event File (LHEF 3.0) output of compacted events.
//
// For every accepted event (sHat, tHat, x1, x2, process code, weight) the host assigns what the
// compacted record leaves open, as Pythia does for a hard process:
//   incoming flavours  -- a pair from the process's flux list with probability
//                         xf1 xf2 dsigmaHat/dtHat (SigmaProcess::pickInState, SigmaProcess.cc:486-508);
//   outgoing flavours and colour flow -- SigmaXxx::setIdColAcol of the process
//                         (pythia8317/src/SigmaQCD.cc:132-389; the heavy pairs 121..124 as
//                         Sigma2gg2QQbar / Sigma2qqbar2QQbar, c cbar or b bbar);
//   azimuth            -- uniform;
// from the event's own Philox stream 3 (words 0..3) and stream 4 (word 0), so the file is a pure
// function of (key, event index).  Momenta: the partonic centre-of-mass frame
// (|p| = sqrt(sHat/4 - m^2), cos theta = (1 - 2a) / beta with a = (m^2 - tHat)/sHat and
// beta = sqrt(1 - 4 m^2/sHat); massless: 1 + 2 tHat / sHat) boosted along z to the lab
// (y = ln(x1/x2)/2).
// Colour tags are offset by 500 (the LHEF habit).  Weights in pb (IDWTUP = 3).
// Reference counterpart: cuPythia/pipeline/lhe_writer.cu:46-70, cuPythia/07_unweighting_lhe/
// unweight_lhe.cu:74-102 (gg -> gg only, fixed colour flow).
#include "mi_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace mi;

namespace {

void philox_h(uint64_t key, uint64_t e, uint32_t s, uint32_t o[4])
{
    uint32_t c0 = (uint32_t)e, c1 = (uint32_t)(e >> 32), c2 = s, c3 = 0u;
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}
double uni(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

double beta_fn(double p, double q) { return std::exp(std::lgamma(p) + std::lgamma(q) - std::lgamma(p + q)); }

struct Toy {
    double ag, bg, av, bv, as, bs, Ag, Auv, Adv, As;
    int nQ;
    double xf(int id, double x) const
    {
        if (x <= 0.0 || x >= 1.0) return 0.0;
        if (id == 21) return Ag * std::pow(x, ag) * std::pow(1.0 - x, bg);
        const int a = id < 0 ? -id : id;
        if (a == 0 || a > nQ) return 0.0;
        const double sea = As * std::pow(x, as) * std::pow(1.0 - x, bs);
        if (id == 2) return sea + Auv * std::pow(x, av) * std::pow(1.0 - x, bv);
        if (id == 1) return sea + Adv * std::pow(x, av) * std::pow(1.0 - x, bv);
        return sea;
    }
};

Toy make_toy(const mi_physics &p)
{
    Toy t;
    t.ag = p.pdf_g[0]; t.bg = p.pdf_g[1]; t.av = p.pdf_val[0]; t.bv = p.pdf_val[1];
    t.as = p.pdf_sea[0]; t.bs = p.pdf_sea[1]; t.nQ = p.nQuarkIn;
    const double bnum = beta_fn(t.av, t.bv + 1.);
    t.Auv = 2.0 / bnum; t.Adv = 1.0 / bnum;
    const double momV = (t.Auv + t.Adv) * beta_fn(t.av + 1., t.bv + 1.);
    t.Ag = 0.5 / beta_fn(t.ag + 1., t.bg + 1.);
    t.As = (0.5 - momV) / (2.0 * t.nQ * beta_fn(t.as + 1., t.bs + 1.));
    return t;
}

// Flavour-dependent part of the qq'->qq' bracket (Sigma2qq2qq::sigmaHat); other codes have none.
double qq_bracket(double s, double t, double u, int i1, int i2)
{
    const double sigT = (4. / 9.) * (s * s + u * u) / (t * t);
    if (i2 == i1) return 0.5 * (sigT + (4. / 9.) * (s * s + t * t) / (u * u) - (8. / 27.) * s * s / (t * u));
    if (i2 == -i1) return sigT - (8. / 27.) * u * u / (s * t);
    return sigT;
}

struct Ev4 { int id[4]; int col[4]; int acol[4]; };

}  // namespace

extern "C" int mi_write_lhe(mi_ctx *ctx, const char *path, const mi_physics *phys, const mi_events *ev, uint64_t n,
                            uint64_t key, const mi_sigma_est *est)
{
    if (!ctx) return MI_ERR_ARG;
    if (!path || !phys || !ev || !est) { ctx->err = "NULL argument"; return MI_ERR_ARG; }
    if (n > ev->capacity) { ctx->err = "n exceeds the event capacity"; return MI_ERR_ARG; }
    if (n && (!ev->sHat || !ev->tHat || !ev->x1 || !ev->x2 || !ev->weight || !ev->pid || !ev->index)) {
        ctx->err = "NULL event array";
        return MI_ERR_ARG;
    }
    const bool partonic = phys->mode == MI_MODE_PARTONIC_1D;
    if (!partonic && (phys->nQuarkIn < 2 || phys->nQuarkIn > 5)) { ctx->err = "nQuarkIn must be 2..5"; return MI_ERR_ARG; }
    if (!partonic && phys->pdf_source == MI_PDF_GRID && !ctx->pdf) {
        ctx->err = "pdf_source = MI_PDF_GRID but no grid loaded (mi_pdf_load)";
        return MI_ERR_STATE;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return MI_ERR_HIP; }
    std::vector<double> sH(n), tH(n), x1(n), x2(n), w(n);
    std::vector<int32_t> pid(n);
    std::vector<uint64_t> idx(n);
    hipError_t e = hipSuccess;
    if (n) {
        const size_t d = sizeof(double) * n;
        if (e == hipSuccess) e = hipMemcpyAsync(sH.data(), ev->sHat, d, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tH.data(), ev->tHat, d, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(x1.data(), ev->x1, d, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(x2.data(), ev->x2, d, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w.data(), ev->weight, d, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pid.data(), ev->pid, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(idx.data(), ev->index, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    if (e != hipSuccess) { ctx->err = std::string("copy of events failed: ") + hipGetErrorString(e); return MI_ERR_HIP; }
    FILE *f = std::fopen(path, "w");
    if (!f) { ctx->err = std::string("cannot open ") + path; return MI_ERR_ARG; }

    const double eBeam = partonic ? 0.5 * std::sqrt(phys->sHat) : 0.5 * phys->eCM;
    const int beamId = partonic ? 21 : 2212;
    const Toy toy = make_toy(*phys);
    GridPdf<double> hg;
    const bool grid = !partonic && pdf_host_view(ctx, *phys, &hg);
    const int nQ = partonic ? 0 : phys->nQuarkIn;
    const double mb2pb = 1e9;
    std::fprintf(f, "<LesHouchesEvents version=\"3.0\">\n<header>\n"
                    "  MI355X QCD 2->2 events: unweighted, compacted on the device (Philox key %llu)\n"
                    "</header>\n<init>\n", (unsigned long long)key);
    int nproc = 0;
    for (int c = 1; c <= MI_MAX_PROC; ++c) nproc += est->sigmaGen[c] > 0.0 ? 1 : 0;
    std::fprintf(f, "%d %d %.10E %.10E 0 0 0 0 3 %d\n", beamId, beamId, eBeam, eBeam, nproc);
    for (int c = 1; c <= MI_MAX_PROC; ++c)           // XSECUP XERRUP XMAXUP LPRUP (Pythia code)
        if (est->sigmaGen[c] > 0.0)
            std::fprintf(f, "%.10E %.10E %.10E %d\n", est->sigmaGen[c] * mb2pb, est->sigmaErr[c] * mb2pb,
                         est->wMax[c] * mb2pb, MI_PROC_CODE(c - 1));
    std::fprintf(f, "</init>\n");

    const double pi = 3.14159265358979323846;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t r3[4], r4[4];
        philox_h(key, idx[i], 3u, r3);
        philox_h(key, idx[i], 4u, r4);
        const int code = pid[i];
        const bool heavy = code >= 121 && code <= 124;
        const int idQ = (code == 121 || code == 122) ? 4 : 5;
        const double mQ = heavy ? phys->mass_Q[idQ - 4] : 0.0, m2 = mQ * mQ;
        const double s = sH[i], t = tH[i], u = 2.0 * m2 - (s + t);
        const double Q2 = (m2 - t) * (m2 - u) / s;                   // pT^2 + m^2
        Ev4 E;
        std::memset(&E, 0, sizeof E);
        // ---- incoming flavours: pick from the flux list with weight xf1 xf2 dsigma/dt(ids)
        int id1 = 21, id2 = 21;
        if (!partonic && code != 111 && code != 112 && code != 121 && code != 123) {
            std::vector<std::pair<int, int>> pairs;
            std::vector<double> wt;
            const double X1 = x1[i], X2 = x2[i];
            // parton densities: the toy PDFs, or the grid at mu_F^2 = fact_fac^2 (pT^2 + m^2) (clipped at 0)
            double g1[11] = {}, g2[11] = {};
            if (grid) {
                const double Q2F = hg.factFac2 * Q2, q = std::sqrt(Q2F);
                int m3q, n3q;
                double wq[4];
                grid_qw<double>(hg, q, std::log(q), m3q, n3q, wq);
                grid_xf<double>(hg, X1, std::log(X1), m3q, n3q, wq, g1);
                grid_xf<double>(hg, X2, std::log(X2), m3q, n3q, wq, g2);
            }
            auto xfa = [&](const double *gv, int id, double X) {
                if (!grid) return toy.xf(id, X);
                const double v = gv[id == 21 ? 0 : (id > 0 ? id : 5 - id)];
                return v > 0.0 ? v : 0.0;
            };
            auto add = [&](int a, int b, double br) {
                pairs.push_back({a, b});
                wt.push_back(xfa(g1, a, X1) * xfa(g2, b, X2) * br);
            };
            for (int q = -nQ; q <= nQ; ++q) {
                if (q == 0) continue;
                if (code == 113) { add(q, 21, 1.0); add(21, q, 1.0); }
                else if (code == 114) { for (int q2 = -nQ; q2 <= nQ; ++q2) if (q2) add(q, q2, qq_bracket(s, t, u, q, q2)); }
                else add(q, -q, 1.0);
            }
            double tot = 0.0;
            for (double x : wt) tot += x;
            double target = uni(r3[0]) * tot;
            size_t j = 0;
            for (; j + 1 < wt.size(); ++j) { target -= wt[j]; if (target <= 0.0) break; }
            id1 = pairs[j].first;
            id2 = pairs[j].second;
        }
        // ---- outgoing flavours and colour flow (setIdColAcol)
        auto setc = [&](int c1, int a1, int c2, int a2, int c3, int a3, int c4, int a4) {
            E.col[0] = c1; E.acol[0] = a1; E.col[1] = c2; E.acol[1] = a2;
            E.col[2] = c3; E.acol[2] = a3; E.col[3] = c4; E.acol[3] = a4;
        };
        auto swapCA = [&]() { for (int k = 0; k < 4; ++k) std::swap(E.col[k], E.acol[k]); };
        const double uc = uni(r3[1]), us = uni(r3[2]), un = uni(r3[3]);
        const double s2 = s * s, t2 = t * t, u2 = u * u;
        E.id[0] = id1; E.id[1] = id2;
        switch (code) {
        case 111: {
            const double a = (9. / 4.) * (t2 / s2 + 2. * t / s + 3. + 2. * s / t + s2 / t2);
            const double b = (9. / 4.) * (u2 / s2 + 2. * u / s + 3. + 2. * s / u + s2 / u2);
            const double c = (9. / 4.) * (t2 / u2 + 2. * t / u + 3. + 2. * u / t + u2 / t2);
            const double r = (a + b + c) * uc;
            E.id[2] = 21; E.id[3] = 21;
            if (r < a) setc(1, 2, 2, 3, 1, 4, 4, 3);
            else if (r < a + b) setc(1, 2, 3, 1, 3, 4, 4, 2);
            else setc(1, 2, 3, 4, 1, 4, 3, 2);
            if (us > 0.5) swapCA();
            break;
        }
        case 112: {
            const int idNew = 1 + (int)(phys->nQuarkNew * un);
            const double a = (1. / 6.) * u / t - (3. / 8.) * u2 / s2, b = (1. / 6.) * t / u - (3. / 8.) * t2 / s2;
            E.id[2] = idNew; E.id[3] = -idNew;
            if ((a + b) * uc < a) setc(1, 2, 2, 3, 1, 0, 0, 3);
            else setc(1, 2, 3, 1, 3, 0, 0, 2);
            break;
        }
        case 113: {
            const double a = u2 / t2 - (4. / 9.) * u / s, b = s2 / t2 - (4. / 9.) * s / u;
            E.id[2] = id1; E.id[3] = id2;
            if ((a + b) * uc < a) setc(1, 0, 2, 1, 3, 0, 2, 3);
            else setc(1, 0, 2, 3, 2, 0, 1, 3);
            if (id1 == 21) {                                       // swapCol1234
                std::swap(E.col[0], E.col[1]); std::swap(E.col[2], E.col[3]);
                std::swap(E.acol[0], E.acol[1]); std::swap(E.acol[2], E.acol[3]);
            }
            if (id1 < 0 || id2 < 0) swapCA();
            break;
        }
        case 114: {
            E.id[2] = id1; E.id[3] = id2;
            if (id1 * id2 > 0) setc(1, 0, 2, 0, 2, 0, 1, 0);
            else setc(1, 0, 0, 1, 2, 0, 0, 2);
            const double sigT = (4. / 9.) * (s2 + u2) / t2, sigU = (4. / 9.) * (s2 + t2) / u2;
            if (id2 == id1 && (sigT + sigU) * uc > sigT) setc(1, 0, 2, 0, 1, 0, 2, 0);
            if (id1 < 0) swapCA();
            break;
        }
        case 115: {
            const double a = (32. / 27.) * u / t - (8. / 3.) * u2 / s2, b = (32. / 27.) * t / u - (8. / 3.) * t2 / s2;
            E.id[2] = 21; E.id[3] = 21;
            if ((a + b) * uc < a) setc(1, 0, 0, 2, 1, 3, 3, 2);
            else setc(1, 0, 0, 2, 3, 2, 1, 3);
            if (id1 < 0) swapCA();
            break;
        }
        case 121:
        case 123: {                                                  // Sigma2gg2QQbar
            // sigTS, sigUS of the massive sigmaKin with tHQ = t - m^2, uHQ = u - m^2
            const double tq = t - m2, uq = u - m2, tq2 = tq * tq, uq2 = uq * uq;
            const double tum = tq * uq - m2 * s;
            const double a = (uq / tq - 2.25 * uq2 / s2 + 4.5 * m2 * tum / (s * tq2) + 0.5 * m2 * (tq + m2) / tq2 -
                              m2 * m2 / (s * tq)) / 6.;
            const double b = (tq / uq - 2.25 * tq2 / s2 + 4.5 * m2 * tum / (s * uq2) + 0.5 * m2 * (uq + m2) / uq2 -
                              m2 * m2 / (s * uq)) / 6.;
            E.id[2] = idQ; E.id[3] = -idQ;
            if ((a + b) * uc < a) setc(1, 2, 2, 3, 1, 0, 0, 3);
            else setc(1, 2, 3, 1, 3, 0, 0, 2);
            break;
        }
        default: {                                                   // 116, and 122 / 124 (Sigma2qqbar2QQbar)
            const int idNew = heavy ? idQ : 1 + (int)(phys->nQuarkNew * un);
            const int id3 = id1 > 0 ? idNew : -idNew;
            E.id[2] = id3; E.id[3] = -id3;
            setc(1, 0, 0, 2, 1, 0, 0, 2);
            if (id1 < 0) swapCA();
            break;
        }
        }
        // ---- momenta: x1, x2 fix the incoming partons; the record's sHat agrees with x1 x2 S only
        // to its FP32 precision, so the final state uses sHat' = x1 x2 S and the record's cos theta,
        // which conserves four-momentum to double precision.
        const double X1 = partonic ? 1.0 : x1[i], X2 = partonic ? 1.0 : x2[i];
        const double pz1 = X1 * eBeam, pz2 = X2 * eBeam;
        const double sLab = 4.0 * pz1 * pz2;
        const double eCMp = 0.5 * std::sqrt(sLab);
        const double pabs = std::sqrt(std::fmax(0.0, 0.25 * sLab - m2));
        const double bQ = std::sqrt(std::fmax(0.0, 1.0 - 4.0 * m2 / s));
        const double cth = std::fmin(1.0, std::fmax(-1.0, (1.0 - 2.0 * (m2 - t) / s) / (heavy ? bQ : 1.0)));
        const double sth = std::sqrt(std::fmax(0.0, 1.0 - cth * cth));
        const double phi = 2.0 * pi * uni(r4[0]);
        const double beta = (X1 - X2) / (X1 + X2), gam = 1.0 / std::sqrt(1.0 - beta * beta);
        double p3[4] = {pabs * sth * std::cos(phi), pabs * sth * std::sin(phi), pabs * cth, eCMp};
        double p4[4] = {-p3[0], -p3[1], -p3[2], eCMp};
        for (double *p : {p3, p4}) {
            const double ez = gam * (p[3] + beta * p[2]), pz = gam * (p[2] + beta * p[3]);
            p[2] = pz; p[3] = ez;
        }
        const double scale = phys->renorm_fac * std::sqrt(std::fmax(Q2, 0.0));
        std::fprintf(f, "<event>\n4 %d %.10E %.10E -1.0E+00 -1.0E+00\n", code, w[i] * mb2pb, scale);
        auto tag = [](int c) { return c ? c + 500 : 0; };
        std::fprintf(f, "%d -1 0 0 %d %d 0 0 %.10E %.10E 0 0 9\n", E.id[0], tag(E.col[0]), tag(E.acol[0]), pz1, pz1);
        std::fprintf(f, "%d -1 0 0 %d %d 0 0 %.10E %.10E 0 0 9\n", E.id[1], tag(E.col[1]), tag(E.acol[1]), -pz2, pz2);
        std::fprintf(f, "%d 1 1 2 %d %d %.10E %.10E %.10E %.10E %.10E 0 9\n", E.id[2], tag(E.col[2]), tag(E.acol[2]),
                     p3[0], p3[1], p3[2], p3[3], mQ);
        std::fprintf(f, "%d 1 1 2 %d %d %.10E %.10E %.10E %.10E %.10E 0 9\n</event>\n", E.id[3], tag(E.col[3]),
                     tag(E.acol[3]), p4[0], p4[1], p4[2], p4[3], mQ);
    }
    std::fprintf(f, "</LesHouchesEvents>\n");
    const bool ok = std::fclose(f) == 0;
    if (!ok) { ctx->err = "write failed"; return MI_ERR_ARG; }
    return MI_OK;
}
