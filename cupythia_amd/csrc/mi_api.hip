// mi_api.hip -- host C++ behind the C ABI (include/mcint.h): context, configuration ->
// kernel constants, launch geometry, device-resident driver loops (integration, VEGAS
// adaptation, unweighting) and the sigma estimate with Pythia's semantics
// (pythia8317/src/ProcessContainer.cc:1416-1472, pythia8317/src/ProcessLevel.cc:428-472).
#include "mi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

using namespace mi;

// Scratch of mi_integrate_noalloc: the block records of one launch (at most kScratchBlocks
// blocks) and the reduced record after them, 1,179,936 B.  One copy per GPU, held until the
// process ends; every call on a GPU shares it (calls that overlap on one GPU race).  Host code
// uses the address hipGetSymbolAddress returns on each device, never the array's name.
namespace mi {
constexpr uint32_t kScratchBlocks = 4096;
constexpr size_t kScratchDoubles = ((size_t)kScratchBlocks + 1) * kRec;
}
static __device__ __align__(256) double scratch[mi::kScratchDoubles];

namespace {

int fail(mi_ctx *ctx, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
int fail(mi_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), MI_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                                       \
    } while (0)

template <typename T>
int ensure(mi_ctx *ctx, T **buf, size_t *cap, size_t n)
{
    if (n <= *cap) return MI_OK;
    if (*buf) HIPCHK(ctx, hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    HIPCHK(ctx, hipMalloc((void **)buf, n * sizeof(T)));
    *cap = n;
    return MI_OK;
}

double beta_fn(double p, double q) { return std::exp(std::lgamma(p) + std::lgamma(q) - std::lgamma(p + q)); }

// The toy-PDF normalisation: int u_v = 2, int d_v = 1, gluon momentum fraction 0.5, the sea
// (2 nQuarkIn distributions of equal shape) carries the rest of the momentum sum rule.
struct ToyPdfNorm { double Ag, Auv, Adv, As; };
ToyPdfNorm pdf_norm(const mi_physics &p)
{
    ToyPdfNorm n;
    const double bnum = beta_fn(p.pdf_val[0], p.pdf_val[1] + 1.);
    n.Auv = 2.0 / bnum;
    n.Adv = 1.0 / bnum;
    const double momV = (n.Auv + n.Adv) * beta_fn(p.pdf_val[0] + 1., p.pdf_val[1] + 1.);
    n.Ag = 0.5 / beta_fn(p.pdf_g[0] + 1., p.pdf_g[1] + 1.);
    const double momS = 1.0 - 0.5 - momV;
    n.As = momS / (2.0 * p.nQuarkIn * beta_fn(p.pdf_sea[0] + 1., p.pdf_sea[1] + 1.));
    return n;
}

// AlphaStrong order 1 (StandardModel.cc:58-62): Lambda_nf from alphaS(MZ), MZ = 91.188, flavour
// thresholds mc = 1.5, mb = 4.8, mt = 171.
struct AsLambda { double L3s, L4s, L5s, L6s, scale2Min; };
AsLambda as_lambda(double valueRef)
{
    const double MZ = 91.188, mc = 1.5, mb = 4.8, mt = 171.0;
    const double L5 = MZ * std::exp(-6. * kPi / (23. * valueRef));
    const double L6 = L5 * std::pow(L5 / mt, 2. / 21.);
    const double L4 = L5 * std::pow(mb / L5, 2. / 25.);
    const double L3 = L4 * std::pow(mc / L4, 2. / 27.);
    AsLambda a;
    a.L3s = L3 * L3; a.L4s = L4 * L4; a.L5s = L5 * L5; a.L6s = L6 * L6;
    a.scale2Min = (1.07 * L3) * (1.07 * L3);
    return a;
}

template <typename T>
void fill_consts(const mi_physics &p, HadConst<T> &k)
{
    std::memset(&k, 0, sizeof k);
    const double s = p.eCM * p.eCM;
    const double fourPT2 = 4.0 * p.pTHatMin * p.pTHatMin;
    const double LT = std::log(s / fourPT2);
    k.LT = (T)LT;
    k.fourPT2 = (T)fourPT2;
    k.pref = (T)(kConvert2mb * kPi * LT * p.kfactor / fourPT2);
    k.xiR2 = (T)(p.renorm_fac * p.renorm_fac);
    const int order = p.alphaS_order <= 0 ? 0 : 1;
    const int nfmax = p.alphaS_nfmax < 5 ? 5 : (p.alphaS_nfmax > 6 ? 6 : p.alphaS_nfmax);
    k.as_order = (T)order;
    k.as_fixed = (T)p.alphaS_value;
    k.mt2 = (T)(171.0 * 171.0); k.mb2 = (T)(4.8 * 4.8); k.mc2 = (T)(1.5 * 1.5);
    k.nf6 = (nfmax >= 6) ? (T)1 : (T)0;
    if (order == 1) {
        const AsLambda a = as_lambda(p.alphaS_value);
        k.scale2Min = (T)a.scale2Min;
        k.lnL6s = (T)std::log(a.L6s); k.lnL5s = (T)std::log(a.L5s);
        k.lnL4s = (T)std::log(a.L4s); k.lnL3s = (T)std::log(a.L3s);
        // Q2 = max(renorm_fac^2 pT2, scale2Min) with pT2 >= pTmin^2 by the cut (heavy codes: pT2 +
        // m^2); the 1% margin covers the rounding of pT2 in FP32
        k.above_mb = (p.renorm_fac * p.renorm_fac * p.pTHatMin * p.pTHatMin > 1.01 * 4.8 * 4.8) ? (T)1 : (T)0;
    }
    k.c6 = (T)(12.0 * kPi / 21.0); k.c5 = (T)(12.0 * kPi / 23.0);
    k.c4 = (T)(12.0 * kPi / 25.0); k.c3 = (T)(12.0 * kPi / 27.0);
    const ToyPdfNorm n = pdf_norm(p);
    k.lnAg = (T)std::log(n.Ag);  k.ag = (T)p.pdf_g[0];   k.bg = (T)p.pdf_g[1];
    k.lnAs = (T)std::log(n.As);  k.as = (T)p.pdf_sea[0]; k.bs = (T)p.pdf_sea[1];
    k.lnAuv = (T)std::log(n.Auv); k.av = (T)p.pdf_val[0]; k.bv = (T)p.pdf_val[1];
    k.rdu = (T)(n.Adv / n.Auv);
    k.nSea = (T)(2 * p.nQuarkIn);
    k.nSeaM2 = (T)(2 * p.nQuarkIn - 2);
    k.nSeaM4 = (T)(2 * p.nQuarkIn - 4);
    k.nQuarkNew = (T)p.nQuarkNew;
    for (int f = 0; f < 2; ++f) {
        const double m2 = p.mass_Q[f] * p.mass_Q[f];
        const double fourMT2 = 4.0 * (m2 + p.pTHatMin * p.pTHatMin);
        const double LTf = std::log(s / fourMT2);
        k.fam[f].LT = (T)LTf;
        k.fam[f].fourMT2 = (T)fourMT2;
        k.fam[f].pref = (T)(kConvert2mb * kPi * LTf * p.kfactor / fourMT2);
        k.fam[f].m2 = (T)m2;
        k.fam[f].mu = (T)(m2 / fourMT2);
        k.b2.famLT[f] = (T)(LTf / kLn2);
        k.b2.famPref[f] = (T)(kConvert2mb * kPi * LTf * p.kfactor / fourMT2 * kLn2);
    }
    k.b2.LT = (T)(LT / kLn2);
    k.b2.pref = (T)(kConvert2mb * kPi * LT * p.kfactor / fourPT2 * kLn2);
    k.b2.lnAg = (T)std::log2(n.Ag); k.b2.lnAs = (T)std::log2(n.As); k.b2.lnAuv = (T)std::log2(n.Auv);
    if (order == 1) {
        const AsLambda a = as_lambda(p.alphaS_value);
        k.b2.lam6 = (T)std::log2(a.L6s); k.b2.lam5 = (T)std::log2(a.L5s);
        k.b2.lam4 = (T)std::log2(a.L4s); k.b2.lam3 = (T)std::log2(a.L3s);
    }
    k.b2.c6 = (T)(12.0 * kPi / 21.0 / kLn2); k.b2.c5 = (T)(12.0 * kPi / 23.0 / kLn2);
    k.b2.c4 = (T)(12.0 * kPi / 25.0 / kLn2); k.b2.c3 = (T)(12.0 * kPi / 27.0 / kLn2);
}

int check_physics(mi_ctx *ctx, const mi_physics &p)
{
    if (p.mode == MI_MODE_PARTONIC_1D) {
        if (!(p.sHat > 0.0) || !(p.cos_max > 0.0 && p.cos_max < 1.0))
            return fail(ctx, MI_ERR_ARG, "partonic mode needs sHat > 0 and 0 < cos_max < 1");
        return MI_OK;
    }
    if (p.mode != MI_MODE_HADRONIC_3D) return fail(ctx, MI_ERR_ARG, "unknown mode %d", p.mode);
    if (!(p.eCM > 0.0) || !(p.pTHatMin > 0.0) || 2.0 * p.pTHatMin >= p.eCM)
        return fail(ctx, MI_ERR_ARG, "need eCM > 2 pTHatMin > 0");
    if (p.nQuarkIn < 2 || p.nQuarkIn > 5) return fail(ctx, MI_ERR_ARG, "nQuarkIn must be 2..5");
    if (p.nQuarkNew < 0 || p.nQuarkNew > 5) return fail(ctx, MI_ERR_ARG, "nQuarkNew must be 0..5");
    if (p.proc_mask == 0 || (p.proc_mask & ~0x3FF))
        return fail(ctx, MI_ERR_ARG, "proc_mask must select codes among 111..116, 121..124 (bits 0..9)");
    for (int f = 0; f < 2; ++f) {
        if (!(p.proc_mask & kFamMask[1 + f])) continue;
        const double m = p.mass_Q[f];
        if (!(m > 0.0) || !(p.eCM * p.eCM > 4.0 * (m * m + p.pTHatMin * p.pTHatMin)))
            return fail(ctx, MI_ERR_ARG, "heavy-quark codes need mass_Q[%d] > 0 and eCM > 2 mT", f);
    }
    if (!(p.alphaS_value > 0.0)) return fail(ctx, MI_ERR_ARG, "alphaS_value must be > 0");
    if (!(p.kfactor >= 0.0) || !std::isfinite(p.kfactor)) return fail(ctx, MI_ERR_ARG, "kfactor must be finite and >= 0");
    if (!(p.renorm_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "renorm_fac must be > 0");
    if (!(p.pdf_g[0] > -1.0) || !(p.pdf_val[0] > 0.0) || !(p.pdf_sea[0] > -1.0) || !(p.pdf_g[1] > -1.0) ||
        !(p.pdf_val[1] > -1.0) || !(p.pdf_sea[1] > -1.0))
        return fail(ctx, MI_ERR_ARG, "PDF exponents out of range (need integrable densities)");
    if (p.pdf_source != MI_PDF_TOY && p.pdf_source != MI_PDF_GRID) return fail(ctx, MI_ERR_ARG, "unknown pdf_source");
    if (p.pdf_source == MI_PDF_GRID) {
        if (!ctx->pdf) return fail(ctx, MI_ERR_STATE, "pdf_source = MI_PDF_GRID but no grid loaded (mi_pdf_load)");
        if (!(p.fact_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "fact_fac must be > 0");
    }
    return MI_OK;
}

int check_run(mi_ctx *ctx, const mi_run *run)
{
    if (!ctx) return MI_ERR_ARG;
    if (!run) return fail(ctx, MI_ERR_ARG, "run is NULL");
    int rc = check_physics(ctx, run->phys);
    if (rc) return rc;
    if (run->precision != MI_PREC_FP32 && run->precision != MI_PREC_FP64 && run->precision != MI_PREC_FP32_HW)
        return fail(ctx, MI_ERR_ARG, "precision must be MI_PREC_FP32, MI_PREC_FP64 or MI_PREC_FP32_HW");
    if (run->sampler < MI_SAMPLER_FLAT || run->sampler > MI_SAMPLER_FLOW)
        return fail(ctx, MI_ERR_ARG, "unknown sampler %d", run->sampler);
    if (run->n_points == 0) return fail(ctx, MI_ERR_ARG, "n_points must be > 0");
    if (run->n_points > (1ULL << 40)) return fail(ctx, MI_ERR_RANGE, "n_points above 2^40");
    if (run->rng != MI_RNG_PHILOX && run->rng != MI_RNG_SPLITMIX_COMPAT)
        return fail(ctx, MI_ERR_ARG, "unknown rng %d", run->rng);
    if (run->sampler == MI_SAMPLER_VEGAS) {
        const int need = run->phys.mode == MI_MODE_HADRONIC_3D ? 3 : 1;
        if (ctx->grid.ndim != need)
            return fail(ctx, MI_ERR_STATE, "sampler VEGAS needs a %d-D grid (mi_vegas_init)", need);
    }
    return MI_OK;
}

// Device address of `scratch` on the current device `device`, looked up once per device.
int scratch_address(int device, double **ptr)
{
    static std::mutex mu;
    static double *addr[64] = {};
    if (device < 0 || device >= 64) return MI_ERR_RANGE;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MI_ERR_HIP;
        addr[device] = static_cast<double *>(p);
    }
    *ptr = addr[device];
    return MI_OK;
}

double c1_pref(const mi_physics &p)
{
    return kConvert2mb * kPi * p.alphaS_value * p.alphaS_value * p.cos_max / p.sHat * p.kfactor;
}

void fill_estimate(const double *r, int64_t n, int partonic, int mask, double ms, mi_sigma_est *o)
{
    std::memset(o, 0, sizeof *o);
    o->code[0] = 0;
    for (int i = 0; i < MI_MAX_PROC; ++i) o->code[1 + i] = proc_code(i);
    for (int i = 0; i <= MI_MAX_PROC; ++i) {
        const bool on = (i == 0) || (partonic ? i == 1 : ((mask >> (i - 1)) & 1));
        if (!on) continue;
        const double sw = r[kRS + i], sw2 = r[kRQ + i];
        const double avg = sw / (double)n;
        o->sumW[i] = sw;
        o->sumW2[i] = sw2;
        o->wMax[i] = r[kRM + i];
        o->sigmaGen[i] = avg;
        // ProcessContainer::sigmaDelta with nSel = nAcc = nTry (weighted integration)
        double d2 = 0.0;
        if (avg != 0.0) d2 = (sw2 / (double)n - avg * avg) / (double)n / (avg * avg);
        o->sigmaErr[i] = (d2 > 0.0 ? std::sqrt(d2) : 0.0) * avg;
    }
    o->nTry = o->nSel = o->nAcc = n;
    o->nNonZero = (int64_t)r[kRNz];
    o->kernel_ms = ms;
}

}  // namespace

namespace mi {

void fill_had_consts_f32(const mi_physics &p, HadConst<float> &k) { fill_consts<float>(p, k); }
void fill_had_consts_f64(const mi_physics &p, HadConst<double> &k) { fill_consts<double>(p, k); }

Geom geom_for(uint64_t n, uint32_t min_ppt, uint32_t max_blocks, uint32_t nt)
{
    uint64_t blocks = (n + (uint64_t)nt * min_ppt - 1) / ((uint64_t)nt * min_ppt);
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks == 0) blocks = 1;
    uint64_t ppt = (n + blocks * nt - 1) / (blocks * nt);
    if (ppt == 0) ppt = 1;
    return Geom{(uint32_t)blocks, (uint32_t)ppt};
}

// The LDS window of an FP32 grid-PDF integration: fills g's window and the launch's dynamic LDS
// size; false when the run has none or the window does not fit one block's LDS after the VEGAS
// table and histograms (one copy per wave) and the reduction's static LDS.
bool had_window(const mi_ctx *ctx, const mi_run *run, bool vegas, bool hist, GridPdf<float> *g, size_t *smem)
{
    if ((run->precision != MI_PREC_FP32 && run->precision != MI_PREC_FP32_HW) || run->phys.mode != MI_MODE_HADRONIC_3D ||
        run->phys.pdf_source != MI_PDF_GRID || ctx->lds_max <= 0)
        return false;
    const size_t tab = vegas ? (sizeof(float) * 9 * ctx->grid.nbins + 7) & ~(size_t)7 : 0;
    const size_t hb = (vegas && hist) ? (kWinBlock / 64) * wave_hist_bytes(3, ctx->grid.nbins) : 0;
    const size_t off = (tab + hb + 15) & ~(size_t)15;
    const size_t stat = (size_t)(kWinBlock / 64) * kRec * sizeof(double) + 1024;   // block_record + slack
    if ((size_t)ctx->lds_max <= stat) return false;
    size_t bytes = 0;
    if (!pdf_window(ctx, run->phys, off, (size_t)ctx->lds_max - stat, g, &bytes)) return false;
    *smem = off + bytes;
    return true;
}

// One wave of equal-work blocks over all CUs: every block takes ceil(n / (blocks * 256))
// points per thread, so the grid drains together and the block records stay few.  Small runs
// (VEGAS iterations, pilots) still fill the resident-block slots down to 4 points per thread:
// one 256-thread block per CU would leave each SIMD one wave and no latency hiding.  With the
// LDS window: one kWinBlock-thread block per CU.
Geom geom_integrate(const mi_ctx *ctx, const mi_run *run, uint64_t n, bool vegas, bool hist)
{
    {
        GridPdf<float> g;
        size_t sm = 0;
        if (had_window(ctx, run, vegas, hist, &g, &sm))
            return geom_for(n, 16, (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256), kWinBlock);
    }
    const int nb = ctx->grid.nbins;
    const int occ = run->phys.mode == MI_MODE_HADRONIC_3D
                        ? occ_integrate_had(run->precision, vegas, hist, nb, run->phys.pdf_source == MI_PDF_GRID,
                                            has_heavy(run->phys.proc_mask))
                        : occ_integrate_c1(vegas, hist, nb);
    const uint32_t cap = (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * (uint32_t)occ;
    return geom_for(n, 4, cap);
}

// Integration driver shared by mi_integrate, VEGAS and unweighting: launches the integrand
// over [ctr0, ctr0 + n) and the reduction into out_dev (kRec doubles); reduce = false leaves the
// block records in ctx->partials for the caller's own reduction.
int run_integrand(mi_ctx *ctx, const mi_run *run, uint64_t ctr0, uint64_t n, bool vegas, double *hist_parts,
                  Geom g, double *out_dev, hipEvent_t after_kernel = nullptr, bool reduce = true,
                  double *partials = nullptr)
{
    int rc = MI_OK;
    if (!partials) {                        // the context's records buffer, grown on demand
        rc = ensure(ctx, &ctx->partials, &ctx->partials_cap, (size_t)g.blocks * kRec);
        if (rc) return rc;
        partials = ctx->partials;
    }
    const mi_physics &p = run->phys;
    if (p.mode == MI_MODE_HADRONIC_3D) {
        HadLaunch L{};
        L.hw = run->precision == MI_PREC_FP32_HW;
        L.prec = L.hw ? MI_PREC_FP32 : run->precision; L.vegas = vegas; L.mask = p.proc_mask;
        L.key = run->key; L.ctr0 = ctr0; L.n = n; L.g = g;
        fill_had_consts_f32(p, L.kf); fill_had_consts_f64(p, L.kd);
        rc = pdf_views(ctx, p, &L.gf, &L.gd);
        if (rc) return rc;
        L.win = had_window(ctx, run, vegas, hist_parts != nullptr, &L.gf, &L.win_smem);
        L.grid = &ctx->grid; L.partials = partials; L.hist = hist_parts;
        HIPCHK(ctx, launch_integrate_had(ctx->stream, L));
    } else {
        C1Launch L{};
        L.vegas = vegas; L.key = run->key; L.ctr0 = ctr0; L.n = n; L.g = g;
        L.pref = c1_pref(p); L.cMax = p.cos_max;
        L.grid = &ctx->grid; L.partials = partials; L.hist = hist_parts;
        HIPCHK(ctx, launch_integrate_c1(ctx->stream, L));
    }
    if (after_kernel) HIPCHK(ctx, hipEventRecord(after_kernel, ctx->stream));
    if (reduce) HIPCHK(ctx, launch_reduce_records(ctx->stream, partials, g.blocks, out_dev));
    return MI_OK;
}

}  // namespace mi

// ============================================================================ C ABI
extern "C" {

int mi_version(void) { return MI_ABI_VERSION; }

int mi_ctx_create(int device, mi_ctx **out)
{
    if (!out) return MI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MI_ERR_HIP;
    if (device < 0 || device >= ndev) return MI_ERR_ARG;
    mi_ctx *ctx = new (std::nothrow) mi_ctx();
    if (!ctx) return MI_ERR_ARG;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
        hipEventCreate(&ctx->ev1) != hipSuccess || hipEventCreate(&ctx->ev2) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipDeviceGetAttribute(&ctx->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) {
        delete ctx;
        return MI_ERR_HIP;
    }
    if (ensure(ctx, &ctx->red, &ctx->red_cap, 64 * kRec) != MI_OK) {
        delete ctx;
        return MI_ERR_HIP;
    }
    // the non-allocating entry point's scratch on this device, and room for any error message
    // (fail() formats at most 511 characters), so that its error path allocates nothing either
    const int src = scratch_address(device, &ctx->scratch);
    if (src != MI_OK) {
        (void)hipFree(ctx->red);
        delete ctx;
        return src;
    }
    ctx->err.reserve(512);
    *out = ctx;
    return MI_OK;
}

void mi_flow_release(mi_ctx *ctx);   // mi_flow.hip

int mi_ctx_destroy(mi_ctx *ctx)
{
    if (!ctx) return MI_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    mi_flow_release(ctx);
    pdf_release(ctx);
    (void)hipFree(ctx->partials); (void)hipFree(ctx->hist); (void)hipFree(ctx->red); (void)hipFree(ctx->counts);
    (void)hipFree(ctx->masks); (void)hipFree(ctx->wbuf);
    (void)hipFree(ctx->grid.edges); (void)hipFree(ctx->grid.tab32); (void)hipFree(ctx->grid.tab64);
    (void)hipEventDestroy(ctx->ev0); (void)hipEventDestroy(ctx->ev1); (void)hipEventDestroy(ctx->ev2);
    delete ctx;
    return MI_OK;
}

const char *mi_last_error(const mi_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

int mi_ctx_set_stream(mi_ctx *ctx, void *stream)
{
    if (!ctx) return MI_ERR_ARG;
    ctx->stream = (hipStream_t)stream;
    return MI_OK;
}

int mi_physics_default(int which, mi_physics *p)
{
    if (!p) return MI_ERR_ARG;
    std::memset(p, 0, sizeof *p);
    p->alphaS_value = 0.118;
    p->alphaS_order = 1;
    p->alphaS_nfmax = 6;
    p->renorm_fac = 1.0;
    p->nQuarkIn = 5;
    p->nQuarkNew = 3;
    p->pdf_g[0] = -0.1; p->pdf_g[1] = 5.0;
    p->pdf_val[0] = 0.5; p->pdf_val[1] = 3.0;
    p->pdf_sea[0] = -0.2; p->pdf_sea[1] = 7.0;
    p->kfactor = 1.0;
    p->pdf_source = MI_PDF_TOY;
    p->fact_fac = 1.0;
    p->mass_Q[0] = 1.5;     // ParticleData m0 of c and b (ParticleData.xml)
    p->mass_Q[1] = 4.8;
    p->eCM = 13600.0;
    p->pTHatMin = 20.0;
    p->sHat = 100.0 * 100.0;
    p->cos_max = 0.9;
    if (which == 0) {
        p->mode = MI_MODE_PARTONIC_1D;
        p->proc_mask = 0x1;
        p->alphaS_order = 0;
    } else if (which == 1) {
        p->mode = MI_MODE_HADRONIC_3D;
        p->proc_mask = 0x1F;
    } else {
        return MI_ERR_ARG;
    }
    return MI_OK;
}

int mi_rng_philox_fill(mi_ctx *ctx, uint64_t key, uint64_t counter0, uint32_t stream, uint64_t n,
                       uint32_t *out_dev)
{
    if (!ctx) return MI_ERR_ARG;
    if (n && !out_dev) return fail(ctx, MI_ERR_ARG, "out_dev is NULL");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_philox_fill(ctx->stream, key, counter0, stream, n, out_dev));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_sigma_hat(mi_ctx *ctx, int code, uint64_t n, const double *sH_dev, const double *tH_dev,
                 const double *uH_dev, const double *alpS_dev, const int32_t *id1_dev, const int32_t *id2_dev,
                 int nQuarkNew, double kfactor, double *out_dev)
{
    if (!ctx) return MI_ERR_ARG;
    const bool heavy = code >= MI_GG2CCBAR && code <= MI_QQBAR2BBBAR;
    if ((code < MI_GG2GG || code > MI_QQBAR2QQBARNEW) && !heavy)
        return fail(ctx, MI_ERR_ARG, "unknown process code %d", code);
    if (n && (!sH_dev || !tH_dev || !alpS_dev || !out_dev)) return fail(ctx, MI_ERR_ARG, "NULL input/output array");
    if (n && heavy && !uH_dev) return fail(ctx, MI_ERR_ARG, "codes 121..124 need uH (the mass is (s + t + u) / 2)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, launch_sigma_hat(ctx->stream, code, n, sH_dev, tH_dev, uH_dev, alpS_dev, id1_dev, id2_dev,
                                 nQuarkNew, kfactor, out_dev));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_points(mi_ctx *ctx, const mi_run *run, double *w_dev, double *kin_dev)
{
    int rc = check_run(ctx, run);
    if (rc) return rc;
    if (!w_dev) return fail(ctx, MI_ERR_ARG, "w_dev is NULL");
    if (run->sampler == MI_SAMPLER_FLOW) return fail(ctx, MI_ERR_ARG, "mi_points supports the flat and VEGAS maps");
    if (run->rng != MI_RNG_PHILOX) return fail(ctx, MI_ERR_ARG, "mi_points uses the Philox stream");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const bool vegas = run->sampler == MI_SAMPLER_VEGAS;
    const mi_physics &p = run->phys;
    if (p.mode == MI_MODE_HADRONIC_3D) {
        HadLaunch L{};
        L.hw = run->precision == MI_PREC_FP32_HW;
        L.prec = L.hw ? MI_PREC_FP32 : run->precision; L.vegas = vegas; L.mask = p.proc_mask;
        L.key = run->key; L.ctr0 = run->counter0; L.n = run->n_points;
        fill_had_consts_f32(p, L.kf); fill_had_consts_f64(p, L.kd);
        rc = pdf_views(ctx, p, &L.gf, &L.gd);
        if (rc) return rc;
        L.grid = &ctx->grid;
        HIPCHK(ctx, launch_points_had(ctx->stream, L, w_dev, kin_dev));
    } else {
        C1Launch L{};
        L.vegas = vegas; L.key = run->key; L.ctr0 = run->counter0; L.n = run->n_points;
        L.pref = c1_pref(p); L.cMax = p.cos_max; L.grid = &ctx->grid;
        HIPCHK(ctx, launch_points_c1(ctx->stream, L, p.sHat, w_dev, kin_dev));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_flow_integrate_internal(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out);   // mi_flow.hip

int mi_integrate(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out)
{
    int rc = check_run(ctx, run);
    if (rc) return rc;
    if (!out) return fail(ctx, MI_ERR_ARG, "out is NULL");
    if (run->rng != MI_RNG_PHILOX) return fail(ctx, MI_ERR_ARG, "integration uses the Philox stream");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (run->sampler == MI_SAMPLER_FLOW) return mi_flow_integrate_internal(ctx, run, out);
    const bool vegas = run->sampler == MI_SAMPLER_VEGAS;
    const Geom g = geom_integrate(ctx, run, run->n_points, vegas, false);
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = run_integrand(ctx, run, run->counter0, run->n_points, vegas, nullptr, g, ctx->red, ctx->ev1);
    if (rc) return rc;
    double r[kRec];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->red, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    fill_estimate(r, (int64_t)run->n_points, run->phys.mode == MI_MODE_PARTONIC_1D, run->phys.proc_mask, ms, out);
    if (!std::isfinite(out->sigmaGen[0])) return fail(ctx, MI_ERR_NUMERIC, "non-finite sigma");
    return MI_OK;
}

int mi_integrate_noalloc(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out)
{
    int rc = check_run(ctx, run);
    if (rc) return rc;
    if (!out) return fail(ctx, MI_ERR_ARG, "out is NULL");
    if (run->rng != MI_RNG_PHILOX) return fail(ctx, MI_ERR_ARG, "integration uses the Philox stream");
    if (run->sampler == MI_SAMPLER_FLOW)
        return fail(ctx, MI_ERR_ARG, "mi_integrate_noalloc samples the flat and VEGAS maps (the flow sampler "
                                     "allocates its buffers)");
    if (!ctx->scratch) return fail(ctx, MI_ERR_STATE, "no scratch address on this device");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const bool vegas = run->sampler == MI_SAMPLER_VEGAS;
    const Geom g = geom_integrate(ctx, run, run->n_points, vegas, false);
    if (g.blocks > kScratchBlocks)
        return fail(ctx, MI_ERR_RANGE, "the launch needs %u block records, the scratch holds %u", g.blocks,
                    kScratchBlocks);
    double *rec = ctx->scratch + (size_t)kScratchBlocks * kRec;
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = run_integrand(ctx, run, run->counter0, run->n_points, vegas, nullptr, g, rec, ctx->ev1, true, ctx->scratch);
    if (rc) return rc;
    double r[kRec];
    HIPCHK(ctx, hipMemcpyAsync(r, rec, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    fill_estimate(r, (int64_t)run->n_points, run->phys.mode == MI_MODE_PARTONIC_1D, run->phys.proc_mask, ms, out);
    if (!std::isfinite(out->sigmaGen[0])) return fail(ctx, MI_ERR_NUMERIC, "non-finite sigma");
    return MI_OK;
}

// ---------------------------------------------------------------------------- VEGAS
int mi_vegas_init(mi_ctx *ctx, int ndim, int nbins)
{
    if (!ctx) return MI_ERR_ARG;
    if (ndim != 1 && ndim != 3) return fail(ctx, MI_ERR_ARG, "ndim must be 1 or 3");
    if (nbins < 1 || nbins > kMaxBins) return fail(ctx, MI_ERR_RANGE, "nbins must be 1..%d", kMaxBins);
    std::vector<double> e((size_t)ndim * (nbins + 1));
    for (int d = 0; d < ndim; ++d)
        for (int i = 0; i <= nbins; ++i) e[(size_t)d * (nbins + 1) + i] = (double)i / nbins;
    return mi_vegas_set_grid(ctx, e.data(), ndim, nbins);
}

int mi_vegas_set_grid(mi_ctx *ctx, const double *edges_host, int ndim, int nbins)
{
    if (!ctx) return MI_ERR_ARG;
    if (!edges_host) return fail(ctx, MI_ERR_ARG, "edges is NULL");
    if (ndim != 1 && ndim != 3) return fail(ctx, MI_ERR_ARG, "ndim must be 1 or 3");
    if (nbins < 1 || nbins > kMaxBins) return fail(ctx, MI_ERR_RANGE, "nbins must be 1..%d", kMaxBins);
    for (int d = 0; d < ndim; ++d) {
        const double *x = edges_host + (size_t)d * (nbins + 1);
        if (x[0] != 0.0 || x[nbins] != 1.0) return fail(ctx, MI_ERR_ARG, "grid must span [0,1] in every dimension");
        for (int i = 0; i < nbins; ++i)
            if (!(x[i + 1] > x[i])) return fail(ctx, MI_ERR_ARG, "grid edges must increase strictly");
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Grid &g = ctx->grid;
    if (g.ndim != ndim || g.nbins != nbins) {
        (void)hipFree(g.edges); (void)hipFree(g.tab32); (void)hipFree(g.tab64);
        g.edges = nullptr; g.tab32 = nullptr; g.tab64 = nullptr; g.ndim = 0; g.nbins = 0;
        HIPCHK(ctx, hipMalloc(&g.edges, sizeof(double) * ndim * (nbins + 1)));
        HIPCHK(ctx, hipMalloc(&g.tab64, sizeof(double) * ndim * nbins * 3));
        HIPCHK(ctx, hipMalloc(&g.tab32, sizeof(float) * ndim * nbins * 3));
        g.ndim = ndim; g.nbins = nbins;
    }
    HIPCHK(ctx, hipMemcpyAsync(g.edges, edges_host, sizeof(double) * ndim * (nbins + 1), hipMemcpyHostToDevice,
                               ctx->stream));
    HIPCHK(ctx, launch_vegas_tables(ctx->stream, g));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_vegas_get_grid(mi_ctx *ctx, double *edges_host, int *ndim, int *nbins)
{
    if (!ctx) return MI_ERR_ARG;
    if (ctx->grid.ndim == 0) return fail(ctx, MI_ERR_STATE, "no VEGAS grid (mi_vegas_init)");
    if (ndim) *ndim = ctx->grid.ndim;
    if (nbins) *nbins = ctx->grid.nbins;
    if (edges_host) {
        HIPCHK(ctx, hipSetDevice(ctx->device));
        HIPCHK(ctx, hipMemcpyAsync(edges_host, ctx->grid.edges,
                                   sizeof(double) * ctx->grid.ndim * (ctx->grid.nbins + 1), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return MI_OK;
}

// Iteration it samples counters counter0 + it * n_points ...; the per-iteration records stay
// on the device until the end (no host round trip inside the loop).
int mi_vegas_adapt(mi_ctx *ctx, const mi_run *run, int n_iter, double alpha, mi_vegas_info *out)
{
    if (!ctx) return MI_ERR_ARG;
    mi_run r = *run;
    r.sampler = MI_SAMPLER_VEGAS;
    int rc = check_run(ctx, &r);
    if (rc) return rc;
    if (r.rng != MI_RNG_PHILOX) return fail(ctx, MI_ERR_ARG, "VEGAS uses the Philox stream");
    if (n_iter < 1 || n_iter > 1000) return fail(ctx, MI_ERR_ARG, "n_iter must be 1..1000");
    if (!(alpha >= 0.0)) return fail(ctx, MI_ERR_ARG, "alpha must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    Grid &gr = ctx->grid;
    const uint32_t ncol = (uint32_t)(gr.ndim * gr.nbins);
    const Geom g = geom_integrate(ctx, &r, r.n_points, true, true);
    rc = ensure(ctx, &ctx->hist, &ctx->hist_cap, (size_t)g.blocks * ncol + ncol);
    if (rc) return rc;
    rc = ensure(ctx, &ctx->red, &ctx->red_cap, (size_t)(n_iter + 1) * kRec);
    if (rc) return rc;
    double *hsum = ctx->hist + (size_t)g.blocks * ncol;
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int it = 0; it < n_iter; ++it) {
        rc = run_integrand(ctx, &r, r.counter0 + (uint64_t)it * r.n_points, r.n_points, true, ctx->hist, g,
                           nullptr, nullptr, false);
        if (rc) return rc;
        HIPCHK(ctx, launch_vegas_step(ctx->stream, ctx->partials, g.blocks, ctx->red + (size_t)it * kRec, ctx->hist,
                                      gr, hsum, alpha));
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> rec((size_t)n_iter * kRec);
    HIPCHK(ctx, hipMemcpyAsync(rec.data(), ctx->red, sizeof(double) * rec.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int it = 0; it < n_iter; ++it)
        if (!std::isfinite(rec[(size_t)it * kRec + kRS]) || !std::isfinite(rec[(size_t)it * kRec + kRQ]))
            return fail(ctx, MI_ERR_NUMERIC,
                        "non-finite weights in VEGAS iteration %d: the grid was refined from them (mi_vegas_init "
                        "starts a new one)", it);
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (out) {
        std::memset(out, 0, sizeof *out);
        out->n_iter = n_iter;
        const double n = (double)r.n_points;
        double sw = 0.0, swi = 0.0, chi = 0.0;
        std::vector<double> I(n_iter), V(n_iter);
        for (int it = 0; it < n_iter; ++it) {
            const double *q = &rec[(size_t)it * kRec];
            const double avg = q[kRS] / n;
            const double var = (q[kRQ] / n - avg * avg) / n;
            I[it] = avg;
            V[it] = var;
            if (var > 0.0) { sw += avg / var; swi += 1.0 / var; }
            if (it == n_iter - 1) {
                out->last_sigma = avg;
                out->last_err = var > 0.0 ? std::sqrt(var) : 0.0;
                out->last_wmax = q[kRM];
                out->last_eta = q[kRM] > 0.0 ? avg / q[kRM] : 0.0;
            }
        }
        if (swi > 0.0) {
            out->sigma = sw / swi;
            out->sigma_err = 1.0 / std::sqrt(swi);
            for (int it = 0; it < n_iter; ++it)
                if (V[it] > 0.0) chi += (I[it] - out->sigma) * (I[it] - out->sigma) / V[it];
            out->chi2_dof = n_iter > 1 ? chi / (n_iter - 1) : 0.0;
        }
        out->kernel_ms = ms;
    }
    return MI_OK;
}

// ---------------------------------------------------------------------------- unweighting
int mi_unweight_compact(mi_ctx *ctx, const mi_run *run, double wmax, const mi_events *out, uint64_t *n_acc,
                        double *wmax_used, mi_sigma_est *est)
{
    int rc = check_run(ctx, run);
    if (rc) return rc;
    if (!out) return fail(ctx, MI_ERR_ARG, "out is NULL");
    if (out->capacity && (!out->sHat || !out->tHat || !out->x1 || !out->x2 || !out->weight || !out->pid || !out->index))
        return fail(ctx, MI_ERR_ARG, "NULL event array with capacity > 0");
    const bool partonic = run->phys.mode == MI_MODE_PARTONIC_1D;
    const bool compat = run->rng == MI_RNG_SPLITMIX_COMPAT;
    if (compat && !partonic) return fail(ctx, MI_ERR_ARG, "the SplitMix64 compat stream exists for the partonic mode only");
    if (partonic && run->sampler != MI_SAMPLER_FLAT) return fail(ctx, MI_ERR_ARG, "partonic unweighting samples flat");
    if (run->sampler == MI_SAMPLER_FLOW) return fail(ctx, MI_ERR_ARG, "unweighting supports the flat and VEGAS maps");
    if (std::isnan(wmax)) return fail(ctx, MI_ERR_ARG, "wmax is NaN");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // Candidates in chunks of at most kUnwChunk (the stored weights of a chunk: 4 or 8 B each).
    const uint64_t n = run->n_points;
    // the unweighting decisions use the bit-reproducible FP32 math also for an FP32_HW run
    const int prec = run->precision == MI_PREC_FP32_HW ? MI_PREC_FP32 : run->precision;
    const size_t tsz = (partonic || prec == MI_PREC_FP64) ? sizeof(double) : sizeof(float);
    const uint64_t chunk = std::min<uint64_t>(n, kUnwChunk);
    const uint64_t nch = (n + chunk - 1) / chunk;
    const uint32_t ncu = (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256);
    rc = ensure(ctx, &ctx->wbuf, &ctx->wbuf_cap, (size_t)chunk * tsz);
    if (rc) return rc;
    UnwLaunch L{};
    L.partonic = partonic;
    L.compat = compat;
    L.out = out;
    L.mat_blocks = ncu * 16;
    const mi_physics &p = run->phys;
    if (partonic) {
        L.c.vegas = false; L.c.key = run->key;
        L.c.pref = c1_pref(p); L.c.cMax = p.cos_max; L.c.grid = &ctx->grid; L.c.partials = ctx->partials;
        L.sHat = p.sHat;
    } else {
        L.h.prec = prec; L.h.vegas = run->sampler == MI_SAMPLER_VEGAS; L.h.mask = p.proc_mask;
        L.h.key = run->key;
        fill_had_consts_f32(p, L.h.kf); fill_had_consts_f64(p, L.h.kd);
        rc = pdf_views(ctx, p, &L.h.gf, &L.h.gd);
        if (rc) return rc;
        L.h.grid = &ctx->grid; L.h.partials = ctx->partials;
        // FP32 grid candidates: the eval kernel reads the reachable part of the grid from LDS
        if (prec == MI_PREC_FP32 && L.h.gf.active && ctx->lds_max > 0) {
            const size_t tab = L.h.vegas ? (sizeof(float) * 9 * ctx->grid.nbins + 15) & ~(size_t)15 : 0;
            const size_t stat = (size_t)(kWinBlock / 64) * kRec * sizeof(double) + 1024;
            GridPdf<float> gw = L.h.gf;
            size_t bytes = 0;
            if ((size_t)ctx->lds_max > stat && pdf_window(ctx, p, tab, (size_t)ctx->lds_max - stat, &gw, &bytes)) {
                L.win = true;
                L.win_smem = tab + bytes;
                L.wgrid = gw;
            }
        }
    }
    // one wave of equal-work blocks of the eval kernel; flags and compact share its layout
    const uint32_t maxb = ncu * (uint32_t)occ_unweight_eval(L);
    const Geom g0 = geom_for(chunk, 16, maxb);
    rc = ensure(ctx, &ctx->partials, &ctx->partials_cap, (size_t)g0.blocks * kRec);
    if (rc) return rc;
    rc = ensure(ctx, &ctx->counts, &ctx->counts_cap, (size_t)g0.blocks + 1);
    if (rc) return rc;
    rc = ensure(ctx, &ctx->masks, &ctx->masks_cap, (size_t)g0.blocks * g0.ppt * (kBlock / 64));
    if (rc) return rc;
    L.masks = ctx->masks;
    L.counts = ctx->counts;
    if (partonic) L.c.partials = ctx->partials;
    else L.h.partials = ctx->partials;
    auto set_chunk = [&](uint64_t k) -> Geom {
        const uint64_t c0 = run->counter0 + k * chunk, nc = std::min<uint64_t>(chunk, n - k * chunk);
        const Geom g = geom_for(nc, 16, maxb);
        if (partonic) { L.c.ctr0 = c0; L.c.n = nc; L.c.g = g; }
        else          { L.h.ctr0 = c0; L.h.n = nc; L.h.g = g; }
        return g;
    };
    // The estimate record over all candidates: chunk records folded in chunk order.
    double R[kRec];
    bool haveR = false;
    auto fold = [&](const double *r) {
        for (int i = 0; i < kRec; ++i) R[i] = !haveR ? r[i] : (rec_is_max(i) ? std::fmax(R[i], r[i]) : R[i] + r[i]);
        haveR = true;
    };
    double r[kRec];
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    bool stored = false;                       // chunk 0's weights already in wbuf
    if (!(wmax > 0.0)) {
        // the exact maximum over all candidates first (a single chunk keeps its weights)
        for (uint64_t k = 0; k < nch; ++k) {
            const Geom g = set_chunk(k);
            L.w = (nch == 1) ? (void *)ctx->wbuf : nullptr;
            HIPCHK(ctx, launch_unweight_eval(ctx->stream, L));
            HIPCHK(ctx, launch_reduce_records(ctx->stream, ctx->partials, g.blocks, ctx->red));
            HIPCHK(ctx, hipMemcpyAsync(r, ctx->red, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
            fold(r);
        }
        stored = nch == 1;
        wmax = R[kRM];
        if (!std::isfinite(wmax) || !std::isfinite(R[kRS]))
            return fail(ctx, MI_ERR_NUMERIC, "non-finite candidate weights");
        if (!(wmax > 0.0)) return fail(ctx, MI_ERR_NUMERIC, "all candidate weights are zero");
    }
    const bool recs_done = haveR;
    L.wmax = wmax;
    L.w = ctx->wbuf;
    uint64_t total = 0;
    for (uint64_t k = 0; k < nch; ++k) {
        const Geom g = set_chunk(k);
        L.base = total;
        if (!stored) {
            HIPCHK(ctx, launch_unweight_eval(ctx->stream, L));
            if (!recs_done) {
                HIPCHK(ctx, launch_reduce_records(ctx->stream, ctx->partials, g.blocks, ctx->red));
                HIPCHK(ctx, hipMemcpyAsync(r, ctx->red, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
            }
        }
        HIPCHK(ctx, launch_unweight_flags(ctx->stream, L));
        HIPCHK(ctx, launch_scan_counts(ctx->stream, ctx->counts, g.blocks));
        HIPCHK(ctx, launch_unweight_compact(ctx->stream, L));
        HIPCHK(ctx, launch_unweight_materialize(ctx->stream, L));
        if (k + 1 == nch) HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        uint64_t ck = 0;
        HIPCHK(ctx, hipMemcpyAsync(&ck, ctx->counts + g.blocks, sizeof ck, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (!stored && !recs_done) fold(r);
        total += ck;
    }
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (n_acc) *n_acc = total;
    if (wmax_used) *wmax_used = wmax;
    if (!std::isfinite(R[kRS]) || !std::isfinite(R[kRQ]))
        return fail(ctx, MI_ERR_NUMERIC, "non-finite candidate weights (the accepted events are not meaningful)");
    if (est) {
        fill_estimate(R, (int64_t)run->n_points, partonic, p.proc_mask, ms, est);
        est->nAcc = (int64_t)total;
    }
    if (total > out->capacity)
        return fail(ctx, MI_ERR_CAPACITY, "%llu events accepted, capacity %llu", (unsigned long long)total,
                    (unsigned long long)out->capacity);
    return MI_OK;
}

// ---------------------------------------------------------------------------- reweighting
int mi_reweight_scales(mi_ctx *ctx, const mi_physics *phys, uint64_t n, const double *sHat_dev,
                       const double *tHat_dev, const int32_t *pid_dev, const double *w_dev, int nvar,
                       const double *factors_host, double *out_dev)
{
    if (!ctx) return MI_ERR_ARG;
    if (!phys) return fail(ctx, MI_ERR_ARG, "phys is NULL");
    if (nvar < 1 || nvar > 64) return fail(ctx, MI_ERR_RANGE, "nvar must be 1..64");
    if (!factors_host) return fail(ctx, MI_ERR_ARG, "factors is NULL");
    for (int v = 0; v < nvar; ++v)
        if (!(factors_host[v] > 0.0)) return fail(ctx, MI_ERR_ARG, "scale factors must be > 0");
    if (n && (!sHat_dev || !tHat_dev || !w_dev || !out_dev)) return fail(ctx, MI_ERR_ARG, "NULL array");
    if (!(phys->alphaS_value > 0.0) || !(phys->renorm_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "bad alphaS / renorm_fac");
    if (n == 0) return MI_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HadConst<double> k;
    mi_physics p = *phys;
    if (p.eCM <= 2.0 * p.pTHatMin || p.pTHatMin <= 0.0) { p.eCM = 13600.0; p.pTHatMin = 20.0; }
    if (p.nQuarkIn < 2 || p.nQuarkIn > 5) p.nQuarkIn = 5;
    fill_had_consts_f64(p, k);
    int rc = ensure(ctx, &ctx->hist, &ctx->hist_cap, (size_t)nvar);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->hist, factors_host, sizeof(double) * nvar, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, launch_reweight(ctx->stream, k, n, sHat_dev, tHat_dev, pid_dev, w_dev, nvar, ctx->hist, out_dev));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_reweight(mi_ctx *ctx, const mi_physics *phys, uint64_t n, const double *sHat_dev, const double *tHat_dev,
                const double *x1_dev, const double *x2_dev, const int32_t *pid_dev, const double *w_dev, int nvar,
                const mi_weight_var *vars_host, double *out_dev)
{
    if (!ctx) return MI_ERR_ARG;
    if (!phys) return fail(ctx, MI_ERR_ARG, "phys is NULL");
    if (nvar < 1 || nvar > 64) return fail(ctx, MI_ERR_RANGE, "nvar must be 1..64");
    if (!vars_host) return fail(ctx, MI_ERR_ARG, "vars is NULL");
    if (n && (!sHat_dev || !tHat_dev || !x1_dev || !x2_dev || !pid_dev || !w_dev || !out_dev))
        return fail(ctx, MI_ERR_ARG, "NULL array");
    if (!(phys->alphaS_value > 0.0) || !(phys->renorm_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "bad alphaS / renorm_fac");
    const bool grid = phys->pdf_source == MI_PDF_GRID;
    if (!grid && phys->pdf_source != MI_PDF_TOY) return fail(ctx, MI_ERR_ARG, "unknown pdf_source");
    if (grid && !(phys->fact_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "fact_fac must be > 0");
    std::vector<RwVar> vars((size_t)nvar);
    GridSet64 gs{};
    if (grid) {
        const int rc = pdf_slot_view(ctx, 0, &gs.g[0]);
        if (rc) return rc;
    }
    for (int v = 0; v < nvar; ++v) {
        const mi_weight_var &q = vars_host[v];
        if (!(q.renorm_fac > 0.0) || !(q.fact_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "scale factors must be > 0");
        if (q.pdf_slot < 0 || q.pdf_slot >= MI_PDF_SLOTS) return fail(ctx, MI_ERR_ARG, "pdf_slot out of range");
        if (!grid && q.pdf_slot != 0) return fail(ctx, MI_ERR_ARG, "PDF slots other than 0 need pdf_source = MI_PDF_GRID");
        if (grid && !gs.g[q.pdf_slot].active) {
            const int rc = pdf_slot_view(ctx, q.pdf_slot, &gs.g[q.pdf_slot]);
            if (rc) return rc;
        }
        vars[v] = RwVar{q.renorm_fac * q.renorm_fac, q.fact_fac * q.fact_fac, q.pdf_slot, 0};
    }
    if (n == 0) return MI_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HadConst<double> k;
    mi_physics p = *phys;
    if (p.eCM <= 2.0 * p.pTHatMin || p.pTHatMin <= 0.0) { p.eCM = 13600.0; p.pTHatMin = 20.0; }
    if (p.nQuarkIn < 2 || p.nQuarkIn > 5) return fail(ctx, MI_ERR_ARG, "nQuarkIn must be 2..5");
    fill_had_consts_f64(p, k);
    const size_t nd = (sizeof(RwVar) * vars.size() + sizeof(double) - 1) / sizeof(double);
    int rc = ensure(ctx, &ctx->hist, &ctx->hist_cap, nd);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->hist, vars.data(), sizeof(RwVar) * vars.size(), hipMemcpyHostToDevice, ctx->stream));
    const double ff2 = grid ? phys->fact_fac * phys->fact_fac : 1.0;
    HIPCHK(ctx, launch_reweight_vars(ctx->stream, k, gs, grid, n, sHat_dev, tHat_dev, x1_dev, x2_dev, pid_dev, w_dev,
                                     nvar, reinterpret_cast<const RwVar *>(ctx->hist), p.renorm_fac * p.renorm_fac, ff2,
                                     out_dev));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_reweight_events(mi_ctx *ctx, const mi_run *run, uint64_t n, const uint64_t *index_dev, const int32_t *pid_dev,
                       const double *w_dev, int nvar, const mi_weight_var *vars_host, double *out_dev)
{
    int rc = check_run(ctx, run);
    if (rc) return rc;
    const mi_physics &p = run->phys;
    if (p.mode != MI_MODE_HADRONIC_3D) return fail(ctx, MI_ERR_ARG, "mi_reweight_events needs the hadronic mode");
    if (run->sampler == MI_SAMPLER_FLOW) return fail(ctx, MI_ERR_ARG, "mi_reweight_events regenerates the flat and VEGAS maps");
    if (run->rng != MI_RNG_PHILOX) return fail(ctx, MI_ERR_ARG, "mi_reweight_events regenerates the Philox stream");
    if (nvar < 1 || nvar > 64) return fail(ctx, MI_ERR_RANGE, "nvar must be 1..64");
    if (!vars_host) return fail(ctx, MI_ERR_ARG, "vars is NULL");
    if (n && (!index_dev || !pid_dev || !out_dev)) return fail(ctx, MI_ERR_ARG, "NULL array");
    const bool grid = p.pdf_source == MI_PDF_GRID;
    std::vector<RwRegen> vars((size_t)nvar + 1);
    for (int v = 0; v <= nvar; ++v) {
        mi_physics q = p;
        int slot = 0;
        if (v > 0) {
            const mi_weight_var &wv = vars_host[v - 1];
            if (!(wv.renorm_fac > 0.0) || !(wv.fact_fac > 0.0)) return fail(ctx, MI_ERR_ARG, "scale factors must be > 0");
            if (wv.pdf_slot < 0 || wv.pdf_slot >= MI_PDF_SLOTS) return fail(ctx, MI_ERR_ARG, "pdf_slot out of range");
            if (!grid && wv.pdf_slot != 0)
                return fail(ctx, MI_ERR_ARG, "PDF slots other than 0 need pdf_source = MI_PDF_GRID");
            q.renorm_fac = wv.renorm_fac;
            q.fact_fac = wv.fact_fac;
            slot = wv.pdf_slot;
        }
        fill_had_consts_f64(q, vars[(size_t)v].k);
        vars[(size_t)v].g = GridPdf<double>{};
        if (grid) {
            rc = pdf_slot_view(ctx, slot, &vars[(size_t)v].g);
            if (rc) return rc;
            vars[(size_t)v].g.factFac2 = q.fact_fac * q.fact_fac;
        }
    }
    if (n == 0) return MI_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t bytes = sizeof(RwRegen) * vars.size();
    rc = ensure(ctx, &ctx->hist, &ctx->hist_cap, (bytes + sizeof(double) - 1) / sizeof(double));
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->hist, vars.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    const bool vegas = run->sampler == MI_SAMPLER_VEGAS;
    HIPCHK(ctx, launch_reweight_regen(ctx->stream, reinterpret_cast<const RwRegen *>(ctx->hist), nvar, p.proc_mask,
                                      vegas, run->key, n, index_dev, pid_dev, w_dev,
                                      vegas ? ctx->grid.tab64 : nullptr, vegas ? ctx->grid.nbins : 0, out_dev));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

}  // extern "C"
