// mi_adjoint.hip -- derivatives of sigma in the same pass as sigma, on the same points.
//
// The phase-space map does not depend on the parameters, so with common random numbers
// d sigma / d theta = E[d w / d theta] and the pathwise derivative of every point's weight is
// accumulated next to w itself (FP64 sums of d and d^2 for the error):
//   alphaS(MZ):  1-loop running with thresholds gives d alphaS(Q) / d alphaS(MZ) =
//                alphaS(Q)^2 / alphaS(MZ)^2, so d w / d alphaS(MZ) = 2 w alphaS / alphaS(MZ)^2
//                (fixed coupling: 2 w / alphaS, the alphaS^2 scaling);
//   renorm_fac:  Q^2 = xi^2 pT^2, d alphaS / d ln Q^2 = -alphaS^2 / c_nf (c_nf = 12 pi / b0), so
//                d w / d xi = -4 w alphaS / (c_nf xi);
//   gluon (a, b): xg = A x^a (1-x)^b with A fixed by the gluon momentum fraction
//                (A = 0.5 / B(a+1, b+1)): d ln g / d a = ln x + d ln A / d a, likewise for b
//                with ln(1-x); only the gg and qg luminosities depend on them.
// Reference precursor: same-sample scale variations, cuPythia/pipeline/reweight.cu:36-51.
#include "mi_kernel_common.h"

#include <cmath>
#include <cstring>

namespace mi {

template <typename T>
struct AdjConst {
    T dlnAg_da, dlnAg_db;  // derivatives of the gluon normalisation
    T as_fac;              // order 1: 2 / alphaS(MZ)^2 (times alphaS); order 0: 2 (divided by alphaS)
    T xi;
};

// One family's weights and the intermediate values of its point (its own map, x1, x2, alphaS).
template <typename T, int F> struct FamAux;
template <int F> struct FamAux<float, F> {
    static __device__ __forceinline__ float eval(const HadConst<float> &k, int mask, float u0, float o0, float u1,
                                                 float o1, float u2, float o2, float *ws, PointAux<float> *ax,
                                                 const GridPdf<float> *gp, bool grid)
    { return had_family_f32<F>(k, mask, u0, o0, u1, o1, u2, o2, ws, nullptr, nullptr, nullptr, nullptr, ax, gp, grid); }
};
template <int F> struct FamAux<double, F> {
    static __device__ __forceinline__ double eval(const HadConst<double> &k, int mask, double u0, double o0,
                                                  double u1, double o1, double u2, double o2, double *ws,
                                                  PointAux<double> *ax, const GridPdf<double> *gp, bool grid)
    { return had_family_f64<F>(k, mask, u0, o0, u1, o1, u2, o2, ws, nullptr, nullptr, nullptr, nullptr, ax, gp, grid); }
};

// Derivatives of one family's weights at a point: the coupling terms hold for every code
// (Q^2 = sH a b for all of them); the gluon-shape terms need the toy gluon (GRID = false):
// g g channels (111, 112, and 121 / 123) scale with g1 g2, q g (113) with Q1 g2 + g1 Q2.
template <typename T, int F, bool GRID>
__device__ __forceinline__ void fam_derivs(const HadConst<T> &k, const AdjConst<T> &ac, int mask, const MapOut<T> &m0,
                                           const MapOut<T> &m1, const MapOut<T> &m2, const GridPdf<T> *gp, T &tot,
                                           T &d_as, T &d_xi, T &d_a, T &d_b)
{
    T ws[kMaxProc];
    PointAux<T> ax;
    const T tf = FamAux<T, F>::eval(k, mask, m0.x, m0.omx, m1.x, m1.omx, m2.x, m2.omx, ws, &ax, gp, GRID);
    tot += tf;
    if (k.as_order == (T)0) {
        d_as += ac.as_fac * tf / ax.alpS;
    } else {
        d_as += ac.as_fac * tf * ax.alpS;
        d_xi += -(T)4 * tf * ax.alpS / (ax.cf * ac.xi);
    }
    if (GRID) return;
    const T la1 = ax.lx1 + ac.dlnAg_da, la2 = ax.lx2 + ac.dlnAg_da;
    const T lb1 = ax.l1a + ac.dlnAg_db, lb2 = ax.l1b + ac.dlnAg_db;
    const T wgg = F == 0 ? ws[0] + ws[1] : ws[4 + 2 * F];
    d_a += wgg * (la1 + la2);
    d_b += wgg * (lb1 + lb2);
    if (F == 0) {
        const T Lqg = ax.Q1 * ax.g2 + ax.g1 * ax.Q2;
        if (ws[2] > (T)0 && Lqg > (T)0) {
            const T fa = (ax.Q1 * ax.g2 * la2 + ax.g1 * ax.Q2 * la1) / Lqg;
            const T fb = (ax.Q1 * ax.g2 * lb2 + ax.g1 * ax.Q2 * lb1) / Lqg;
            d_a += ws[2] * fa;
            d_b += ws[2] * fb;
        }
    }
}

// Record layout: [kRS] sum w, [kRQ] sum w^2, [kRS + 1 + k] sum d_k, [kRQ + 1 + k] sum d_k^2,
// [kRNz] nonzero.  HEAVY: the families of the heavy-quark codes as well; GRID: parton densities
// from the PDF grid (the gluon-shape derivatives are then 0).
template <typename T, bool HEAVY, bool GRID>
__global__ __launch_bounds__(kBlock) void k_adjoint(HadConst<T> k, AdjConst<T> ac, int mask, uint64_t key,
                                                    uint64_t ctr0, uint64_t n, uint32_t ppt,
                                                    double *__restrict__ partials, GridPdf<T> gpdf)
{
    double v[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) v[i] = 0.0;
    const GridPdf<T> *gp = &gpdf;
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * ppt;
    for (uint32_t j = 0; j < ppt; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        if (i >= n) break;
        const U4 r = philox_event(key, ctr0 + i, 0u);
        const MapOut<T> m0 = map_flat<T>(k24(r.x)), m1 = map_flat<T>(k24(r.y)), m2 = map_flat<T>(k24(r.z));
        T tot = (T)0, d_as = (T)0, d_xi = (T)0, d_a = (T)0, d_b = (T)0;
        if (!HEAVY || (mask & kFamMask[0]))
            fam_derivs<T, 0, GRID>(k, ac, mask, m0, m1, m2, gp, tot, d_as, d_xi, d_a, d_b);
        if (HEAVY && (mask & kFamMask[1]))
            fam_derivs<T, 1, GRID>(k, ac, mask, m0, m1, m2, gp, tot, d_as, d_xi, d_a, d_b);
        if (HEAVY && (mask & kFamMask[2]))
            fam_derivs<T, 2, GRID>(k, ac, mask, m0, m1, m2, gp, tot, d_as, d_xi, d_a, d_b);
        const double dt = (double)tot;
        v[kRS] += dt;
        v[kRQ] = fma(dt, dt, v[kRQ]);
        v[kRNz] += (tot > (T)0) ? 1.0 : 0.0;
        const double d[4] = {(double)d_as, (double)d_xi, (double)d_a, (double)d_b};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[kRS + 1 + q] += d[q];
            v[kRQ + 1 + q] = fma(d[q], d[q], v[kRQ + 1 + q]);
        }
    }
    block_record(v, partials + (size_t)blockIdx.x * kRec);
}

template <typename T, bool HEAVY, bool GRID>
int occ_adjoint_t()
{
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<T, HEAVY, GRID>, kBlock, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

template <typename T>
int occ_adjoint_p(bool heavy, bool grid)
{
    if (heavy) return grid ? occ_adjoint_t<T, true, true>() : occ_adjoint_t<T, true, false>();
    return grid ? occ_adjoint_t<T, false, true>() : occ_adjoint_t<T, false, false>();
}

int occ_adjoint(int prec, bool heavy, bool grid)
{
    return prec == MI_PREC_FP32 ? occ_adjoint_p<float>(heavy, grid) : occ_adjoint_p<double>(heavy, grid);
}

template <typename T>
void launch_adjoint_t(hipStream_t s, dim3 g, const HadConst<T> &k, const AdjConst<T> &ac, int mask, uint64_t key,
                      uint64_t ctr0, uint64_t n, uint32_t ppt, double *partials, const GridPdf<T> &gp)
{
    const bool hv = has_heavy(mask), gr = gp.active != 0;
    if (hv && gr)
        hipLaunchKernelGGL((k_adjoint<T, true, true>), g, dim3(kBlock), 0, s, k, ac, mask, key, ctr0, n, ppt, partials, gp);
    else if (hv)
        hipLaunchKernelGGL((k_adjoint<T, true, false>), g, dim3(kBlock), 0, s, k, ac, mask, key, ctr0, n, ppt, partials, gp);
    else if (gr)
        hipLaunchKernelGGL((k_adjoint<T, false, true>), g, dim3(kBlock), 0, s, k, ac, mask, key, ctr0, n, ppt, partials, gp);
    else
        hipLaunchKernelGGL((k_adjoint<T, false, false>), g, dim3(kBlock), 0, s, k, ac, mask, key, ctr0, n, ppt, partials, gp);
}

}  // namespace mi

using namespace mi;

namespace {

// psi(x) for x > 0: recurrence up to x >= 6, then the asymptotic series
// ln x - 1/(2x) - sum_k B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18).
double digamma(double x)
{
    double r = 0.0;
    while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
    const double f = 1.0 / (x * x);
    const double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132)))));
    return r + std::log(x) - 0.5 / x + t;
}

}  // namespace

extern "C" int mi_adjoint(mi_ctx *ctx, const mi_run *run, mi_sigma_grad *out)
{
    if (!ctx) return MI_ERR_ARG;
    if (!run || !out) { ctx->err = "run / out is NULL"; return MI_ERR_ARG; }
    const mi_physics &p = run->phys;
    if (p.mode != MI_MODE_HADRONIC_3D) { ctx->err = "the adjoint needs the hadronic mode"; return MI_ERR_ARG; }
    if (run->sampler != MI_SAMPLER_FLAT) { ctx->err = "the adjoint samples the flat map"; return MI_ERR_ARG; }
    if (run->precision != MI_PREC_FP32 && run->precision != MI_PREC_FP64 && run->precision != MI_PREC_FP32_HW) {
        ctx->err = "bad precision";
        return MI_ERR_ARG;
    }
    const int prec = run->precision == MI_PREC_FP64 ? MI_PREC_FP64 : MI_PREC_FP32;   // FP32_HW runs as FP32
    if (run->n_points == 0 || run->n_points > (1ULL << 40)) { ctx->err = "n_points must be 1..2^40"; return MI_ERR_RANGE; }
    if (p.proc_mask == 0 || (p.proc_mask & ~0x3FF)) {
        ctx->err = "proc_mask must select codes among 111..116, 121..124 (bits 0..9)";
        return MI_ERR_ARG;
    }
    for (int f = 0; f < 2; ++f)
        if ((p.proc_mask & kFamMask[1 + f]) &&
            (!(p.mass_Q[f] > 0.0) || !(p.eCM * p.eCM > 4.0 * (p.mass_Q[f] * p.mass_Q[f] + p.pTHatMin * p.pTHatMin)))) {
            ctx->err = "heavy-quark codes need mass_Q > 0 and eCM > 2 mT";
            return MI_ERR_ARG;
        }
    if (p.pdf_source != MI_PDF_TOY && p.pdf_source != MI_PDF_GRID) { ctx->err = "unknown pdf_source"; return MI_ERR_ARG; }
    if (p.pdf_source == MI_PDF_GRID && !(p.fact_fac > 0.0)) { ctx->err = "fact_fac must be > 0"; return MI_ERR_ARG; }
    GridPdf<float> gf;
    GridPdf<double> gd;
    {
        const int rc = pdf_views(ctx, p, &gf, &gd);
        if (rc) return rc;
    }
    if (!(p.alphaS_value > 0.0) || !(p.renorm_fac > 0.0) || !(p.eCM > 2.0 * p.pTHatMin) || !(p.pTHatMin > 0.0) ||
        p.nQuarkIn < 2 || p.nQuarkIn > 5 || !(p.pdf_g[0] > -1.0) || !(p.pdf_g[1] > -1.0)) {
        ctx->err = "physics configuration out of range";
        return MI_ERR_ARG;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return MI_ERR_HIP; }
    const Geom g = geom_for(run->n_points, 16, (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) *
                                                   occ_adjoint(prec, has_heavy(p.proc_mask), gf.active != 0));
    if (g.blocks * (size_t)kRec > ctx->partials_cap) {
        (void)hipFree(ctx->partials);
        ctx->partials = nullptr;
        ctx->partials_cap = 0;
        if (hipMalloc(&ctx->partials, sizeof(double) * g.blocks * kRec) != hipSuccess) {
            ctx->err = "hipMalloc failed";
            return MI_ERR_HIP;
        }
        ctx->partials_cap = (size_t)g.blocks * kRec;
    }
    const double a = p.pdf_g[0], b = p.pdf_g[1];
    const double psab = digamma(a + b + 2.0);
    const double dA_da = -(digamma(a + 1.0) - psab), dA_db = -(digamma(b + 1.0) - psab);
    const bool fixed = p.alphaS_order <= 0;
    const double as_fac = fixed ? 2.0 : 2.0 / (p.alphaS_value * p.alphaS_value);
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    if (prec == MI_PREC_FP32) {
        HadConst<float> k;
        fill_had_consts_f32(p, k);
        AdjConst<float> ac{(float)dA_da, (float)dA_db, (float)as_fac, (float)p.renorm_fac};
        launch_adjoint_t<float>(ctx->stream, dim3(g.blocks), k, ac, p.proc_mask, run->key, run->counter0, run->n_points,
                                g.ppt, ctx->partials, gf);
    } else {
        HadConst<double> k;
        fill_had_consts_f64(p, k);
        AdjConst<double> ac{dA_da, dA_db, as_fac, p.renorm_fac};
        launch_adjoint_t<double>(ctx->stream, dim3(g.blocks), k, ac, p.proc_mask, run->key, run->counter0, run->n_points,
                                 g.ppt, ctx->partials, gd);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = launch_reduce_records(ctx->stream, ctx->partials, g.blocks, ctx->red);
    double r[kRec];
    if (e == hipSuccess) e = hipMemcpyAsync(r, ctx->red, sizeof r, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) {
        ctx->err = std::string("adjoint launch failed: ") + hipGetErrorString(e);
        return MI_ERR_HIP;
    }
    const double n = (double)run->n_points;
    auto mean_err = [n](double s, double s2, double *m, double *err) {
        *m = s / n;
        const double v = (s2 / n - (*m) * (*m)) / n;
        *err = v > 0.0 ? std::sqrt(v) : 0.0;
    };
    std::memset(out, 0, sizeof *out);
    mean_err(r[kRS], r[kRQ], &out->sigma, &out->sigma_err);
    mean_err(r[kRS + 1], r[kRQ + 1], &out->d_alphaS_MZ, &out->d_alphaS_MZ_err);
    mean_err(r[kRS + 2], r[kRQ + 2], &out->d_renorm_fac, &out->d_renorm_fac_err);
    mean_err(r[kRS + 3], r[kRQ + 3], &out->d_pdf_g_a, &out->d_pdf_g_a_err);
    mean_err(r[kRS + 4], r[kRQ + 4], &out->d_pdf_g_b, &out->d_pdf_g_b_err);
    out->kernel_ms = ms;
    if (!std::isfinite(out->sigma) || !std::isfinite(out->d_alphaS_MZ) || !std::isfinite(out->d_pdf_g_a)) {
        ctx->err = "non-finite adjoint result";
        return MI_ERR_NUMERIC;
    }
    return MI_OK;
}
