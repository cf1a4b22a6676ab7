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

// Record layout: [kRS] sum w, [kRQ] sum w^2, [kRS + 1 + k] sum d_k, [kRQ + 1 + k] sum d_k^2,
// [kRNz] nonzero.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_adjoint(HadConst<T> k, AdjConst<T> ac, int mask, uint64_t key,
                                                    uint64_t ctr0, uint64_t n, uint32_t ppt,
                                                    double *__restrict__ partials)
{
    double v[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) v[i] = 0.0;
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * ppt;
    for (uint32_t j = 0; j < ppt; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        if (i >= n) break;
        const U4 r = philox_event(key, ctr0 + i, 0u);
        const MapOut<T> m0 = map_flat<T>(k24(r.x)), m1 = map_flat<T>(k24(r.y)), m2 = map_flat<T>(k24(r.z));
        T ws[kMaxProc];
        PointAux<T> ax;
        const T tot = HadPointAux<T>::eval(k, mask, m0.x, m0.omx, m1.x, m1.omx, m2.x, m2.omx, ws, &ax);
        // coupling
        T d_as, d_xi;
        if (k.as_order == (T)0) {
            d_as = ac.as_fac * tot / ax.alpS;
            d_xi = (T)0;
        } else {
            d_as = ac.as_fac * tot * ax.alpS;
            d_xi = -(T)4 * tot * ax.alpS / (ax.cf * ac.xi);
        }
        // gluon shape: gg channels (111, 112) scale with g1 g2, qg (113) with Q1 g2 + g1 Q2
        const T la1 = ax.lx1 + ac.dlnAg_da, la2 = ax.lx2 + ac.dlnAg_da;
        const T lb1 = ax.l1a + ac.dlnAg_db, lb2 = ax.l1b + ac.dlnAg_db;
        const T wgg = ws[0] + ws[1];
        T d_a = wgg * (la1 + la2), d_b = wgg * (lb1 + lb2);
        const T Lqg = ax.Q1 * ax.g2 + ax.g1 * ax.Q2;
        if (ws[2] > (T)0 && Lqg > (T)0) {
            const T fa = (ax.Q1 * ax.g2 * la2 + ax.g1 * ax.Q2 * la1) / Lqg;
            const T fb = (ax.Q1 * ax.g2 * lb2 + ax.g1 * ax.Q2 * lb1) / Lqg;
            d_a += ws[2] * fa;
            d_b += ws[2] * fb;
        }
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

int occ_adjoint(int prec)
{
    int nb = 0;
    hipError_t e = (prec == MI_PREC_FP32)
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<float>, kBlock, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<double>, kBlock, 0);
    return (e == hipSuccess && nb > 0) ? nb : 1;
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
    if (run->precision != MI_PREC_FP32 && run->precision != MI_PREC_FP64) { ctx->err = "bad precision"; return MI_ERR_ARG; }
    if (run->n_points == 0 || run->n_points > (1ULL << 40)) { ctx->err = "n_points must be 1..2^40"; return MI_ERR_RANGE; }
    if ((p.proc_mask & 0x3F) == 0 || (p.proc_mask & ~0x3F)) {
        ctx->err = "the adjoint covers the massless codes 111..116 (proc_mask bits 0..5)";
        return MI_ERR_ARG;
    }
    if (p.pdf_source != MI_PDF_TOY) { ctx->err = "the adjoint's PDF-shape derivatives need the toy PDFs"; return MI_ERR_ARG; }
    if (!(p.alphaS_value > 0.0) || !(p.renorm_fac > 0.0) || !(p.eCM > 2.0 * p.pTHatMin) || !(p.pTHatMin > 0.0) ||
        p.nQuarkIn < 2 || p.nQuarkIn > 5 || !(p.pdf_g[0] > -1.0) || !(p.pdf_g[1] > -1.0)) {
        ctx->err = "physics configuration out of range";
        return MI_ERR_ARG;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { ctx->err = "hipSetDevice failed"; return MI_ERR_HIP; }
    const Geom g = geom_for(run->n_points, 16, (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * occ_adjoint(run->precision));
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
    if (run->precision == MI_PREC_FP32) {
        HadConst<float> k;
        fill_had_consts_f32(p, k);
        AdjConst<float> ac{(float)dA_da, (float)dA_db, (float)as_fac, (float)p.renorm_fac};
        hipLaunchKernelGGL(k_adjoint<float>, dim3(g.blocks), dim3(kBlock), 0, ctx->stream, k, ac, p.proc_mask, run->key,
                           run->counter0, run->n_points, g.ppt, ctx->partials);
    } else {
        HadConst<double> k;
        fill_had_consts_f64(p, k);
        AdjConst<double> ac{dA_da, dA_db, as_fac, p.renorm_fac};
        hipLaunchKernelGGL(k_adjoint<double>, dim3(g.blocks), dim3(kBlock), 0, ctx->stream, k, ac, p.proc_mask, run->key,
                           run->counter0, run->n_points, g.ppt, ctx->partials);
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
