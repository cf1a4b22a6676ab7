// mi_lhe.hip -- Les Houches Event File (LHEF 3.0) output of compacted events.
//
// The partons of each event (flavours, colour flow, lab momenta, the hard scale) come from the
// device event record (mi_events.hip), built in chunks and copied back; this file only formats
// them (host threads format contiguous slices of a chunk, written in event order, so the file is
// the same whatever the thread count).  Colour tags as in the record (501, ...); LHEF status -1 / 1,
// mothers 1 2 for the outgoing pair; cross sections in pb, LPRUP = the Pythia process code.
// Unweighted events (IDWTUP = 3): every event carries XWGTUP = sigma_tot / N in pb (so the weights
// sum to sigma; Pythia's strategy 3 reads the sign only and sigma from XSECUP), SCALUP the hard
// scale, AQCDUP alphaS at it (the run's running coupling), AQEDUP alphaEM(MZ).
// Reference counterpart: cuPythia/pipeline/lhe_writer.cu:46-70, cuPythia/07_unweighting_lhe/
// unweight_lhe.cu:74-102 (gg -> gg only, fixed colour flow).
#include "mi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace mi;

namespace {

constexpr uint64_t kLheChunk = 1ULL << 20;   // events per device build

// Device record of up to kLheChunk events (4 particles each), freed with the object.
struct ChunkRecord {
    mi_particles r{};
    void *mem = nullptr;
    hipError_t alloc(uint64_t cap)
    {
        const size_t np = (size_t)cap * 4;
        const size_t bytes = np * (5 * sizeof(double) + 8 * sizeof(int32_t)) + cap * (sizeof(int32_t) + sizeof(double));
        const hipError_t e = hipMalloc(&mem, bytes);
        if (e != hipSuccess) { mem = nullptr; return e; }
        char *p = static_cast<char *>(mem);
        auto take = [&p](size_t b) { char *q = p; p += b; return q; };
        double **dv[5] = {&r.px, &r.py, &r.pz, &r.e, &r.m};
        for (double **d : dv) *d = reinterpret_cast<double *>(take(np * sizeof(double)));
        r.scale = reinterpret_cast<double *>(take(cap * sizeof(double)));
        int32_t **iv[8] = {&r.pdg, &r.status, &r.col, &r.acol, &r.mo1, &r.mo2, &r.d1, &r.d2};
        for (int32_t **q : iv) *q = reinterpret_cast<int32_t *>(take(np * sizeof(int32_t)));
        r.n_part = reinterpret_cast<int32_t *>(take(cap * sizeof(int32_t)));
        r.capacity = cap;
        r.max_part = 4;
        return hipSuccess;
    }
    ~ChunkRecord() { if (mem) (void)hipFree(mem); }
};

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
    EvConst c;
    GridPdf<double> g;
    int rc = event_consts(ctx, *phys, 4, &c, &g);
    if (rc) return rc;
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return MI_ERR_HIP; }
    ChunkRecord rec;
    const uint64_t cap = n < kLheChunk ? n : kLheChunk;
    if (cap && rec.alloc(cap) != hipSuccess) { ctx->err = "hipMalloc failed"; return MI_ERR_HIP; }

    FILE *f = std::fopen(path, "w");
    if (!f) { ctx->err = std::string("cannot open ") + path; return MI_ERR_ARG; }
    const bool partonic = phys->mode == MI_MODE_PARTONIC_1D;
    const double eBeam = c.eBeam;
    const int beamId = partonic ? 21 : 2212;
    const double mb2pb = 1e9;
    std::fprintf(f, "<LesHouchesEvents version=\"3.0\">\n<header>\n"
                    "  MI355X QCD 2->2 events: unweighted, compacted on the device (Philox key %llu)\n"
                    "</header>\n<init>\n", (unsigned long long)key);
    int nproc = 0;
    for (int k = 1; k <= MI_MAX_PROC; ++k) nproc += est->sigmaGen[k] > 0.0 ? 1 : 0;
    std::fprintf(f, "%d %d %.10E %.10E 0 0 0 0 3 %d\n", beamId, beamId, eBeam, eBeam, nproc);
    for (int k = 1; k <= MI_MAX_PROC; ++k)           // XSECUP XERRUP XMAXUP LPRUP (Pythia code)
        if (est->sigmaGen[k] > 0.0)
            std::fprintf(f, "%.10E %.10E %.10E %d\n", est->sigmaGen[k] * mb2pb, est->sigmaErr[k] * mb2pb,
                         est->wMax[k] * mb2pb, MI_PROC_CODE(k - 1));
    std::fprintf(f, "</init>\n");

    // per-event weight of unweighted events, alphaS at the event's scale, alphaEM(MZ) (Pythia's
    // StandardModel:alphaEMmZ)
    const double xw = (est->sigmaGen[0] > 0.0 && n > 0) ? est->sigmaGen[0] * mb2pb / (double)n : 1.0;
    const double aqed = 0.00781751;
    HadConst<double> kd{};
    if (!partonic) fill_had_consts_f64(*phys, kd);
    const size_t np = (size_t)cap * 4;
    std::vector<double> px(np), py(np), pz(np), pe(np), pm(np), scale(cap);
    std::vector<int32_t> pdg(np), col(np), acol(np), npart(cap), pid(cap);
    hipError_t e = hipSuccess;
    for (uint64_t first = 0; first < n && e == hipSuccess; first += cap) {
        const uint64_t m = (n - first) < cap ? (n - first) : cap;
        e = launch_build_events(ctx->stream, c, g, m, key, *ev, first, rec.r);
        const size_t dp = sizeof(double) * m * 4, ip = sizeof(int32_t) * m * 4;
        struct { void *dst; const void *src; size_t b; } cp[] = {
            {px.data(), rec.r.px, dp}, {py.data(), rec.r.py, dp}, {pz.data(), rec.r.pz, dp}, {pe.data(), rec.r.e, dp},
            {pm.data(), rec.r.m, dp}, {pdg.data(), rec.r.pdg, ip}, {col.data(), rec.r.col, ip},
            {acol.data(), rec.r.acol, ip}, {npart.data(), rec.r.n_part, sizeof(int32_t) * m},
            {scale.data(), rec.r.scale, sizeof(double) * m},
            {pid.data(), ev->pid + first, sizeof(int32_t) * m}};
        for (auto &q : cp)
            if (e == hipSuccess) e = hipMemcpyAsync(q.dst, q.src, q.b, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) break;
        // text of events [a, b) of the chunk
        auto format = [&](uint64_t a, uint64_t b, std::string &out) {
            char buf[512];
            out.clear();
            out.reserve((size_t)(b - a) * 600);
            for (uint64_t i = a; i < b; ++i) {
                if (npart[i] != 4) continue;                 // unknown code: no partons
                const size_t j0 = i * 4;
                const double aqcd = partonic ? phys->alphaS_value : alphas_f64(kd, scale[i] * scale[i]);
                int len = std::snprintf(buf, sizeof buf, "<event>\n4 %d %.10E %.10E %.10E %.10E\n", pid[i], xw,
                                        scale[i], aqed, aqcd);
                out.append(buf, (size_t)len);
                for (int k = 0; k < 4; ++k) {
                    const size_t j = j0 + k;
                    if (k < 2)
                        len = std::snprintf(buf, sizeof buf, "%d -1 0 0 %d %d 0 0 %.10E %.10E 0 0 9\n", pdg[j], col[j],
                                            acol[j], pz[j], pe[j]);
                    else
                        len = std::snprintf(buf, sizeof buf, "%d 1 1 2 %d %d %.10E %.10E %.10E %.10E %.10E 0 9\n",
                                            pdg[j], col[j], acol[j], px[j], py[j], pz[j], pe[j], pm[j]);
                    out.append(buf, (size_t)len);
                }
                out.append("</event>\n");
            }
        };
        // formatted by host threads over contiguous slices, written in event order
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>(hw, m / 4096));
        std::vector<std::string> text(nt);
        std::vector<std::thread> pool;
        for (uint64_t t = 1; t < nt; ++t) pool.emplace_back(format, m * t / nt, m * (t + 1) / nt, std::ref(text[t]));
        format(0, m / nt, text[0]);
        for (std::thread &th : pool) th.join();
        for (const std::string &t : text) std::fwrite(t.data(), 1, t.size(), f);
    }
    std::fprintf(f, "</LesHouchesEvents>\n");
    const bool ok = std::fclose(f) == 0;
    if (e != hipSuccess) {
        ctx->err = std::string("event build failed: ") + hipGetErrorString(e);
        return MI_ERR_HIP;
    }
    if (!ok) { ctx->err = "write failed"; return MI_ERR_ARG; }
    return MI_OK;
}
