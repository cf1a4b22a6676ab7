// mi_solve.hip -- time-to-precision driver behind mi_solve: sigma to a relative-error target with
// plain VEGAS, or with the flow on top of VEGAS and plain VEGAS as the fallback (the decisions are
// mi_solve_policy.h's).  Host code only: every pass is one of the library's own entry points
// (mi_integrate, mi_flow_init, mi_flow_train), and the GPU time of the solve is the sum of the
// device times those passes report.
#include "mi_internal.h"
#include "mi_solve_policy.h"

#include <cstring>
#include <string>

using namespace mi;

namespace {

int serr(mi_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

mi_run pass(const mi_run &base, int sampler, uint64_t counter0, uint64_t n)
{
    mi_run r = base;
    r.sampler = sampler;
    r.counter0 = counter0;
    r.n_points = n;
    return r;
}

}  // namespace

extern "C" {

int mi_solve(mi_ctx *ctx, const mi_run *run, const mi_solve_opts *opts_in, mi_sigma_est *out, mi_solve_info *info)
{
    if (!ctx) return MI_ERR_ARG;
    if (!run || !opts_in || !out) return serr(ctx, MI_ERR_ARG, "run, opts and out must not be NULL");
    mi_solve_opts o = *opts_in;
    if (!(o.rel_target > 0.0 && o.rel_target < 1.0)) return serr(ctx, MI_ERR_ARG, "rel_target must be in (0, 1)");
    if (o.sampler != MI_SAMPLER_VEGAS && o.sampler != MI_SAMPLER_FLOW)
        return serr(ctx, MI_ERR_ARG, "mi_solve samples with MI_SAMPLER_VEGAS or MI_SAMPLER_FLOW");
    if (run->phys.mode != MI_MODE_HADRONIC_3D) return serr(ctx, MI_ERR_ARG, "mi_solve integrates the hadronic mode");
    if (o.sampler == MI_SAMPLER_FLOW && run->precision == MI_PREC_FP64)
        return serr(ctx, MI_ERR_ARG, "the flow sampler evaluates FP32 weights: run->precision must be FP32");
    if (o.chunk_steps == 0) o.chunk_steps = 10;
    if (o.max_steps == 0) o.max_steps = 200;
    if (o.pilot_points == 0) o.pilot_points = 1ULL << 20;
    if (o.probe_points == 0) o.probe_points = 1ULL << 20;
    if (o.batch == 0) o.batch = 1ULL << 16;
    if (o.min_points == 0) o.min_points = 1ULL << 16;
    if (!(o.lr > 0.0)) o.lr = 2e-3;
    if (o.chunk_steps < 1 || o.max_steps < o.chunk_steps || o.max_steps > 100000)
        return serr(ctx, MI_ERR_ARG, "need 1 <= chunk_steps <= max_steps <= 100000");
    const uint64_t cap = 1ULL << 40;
    if (o.pilot_points > cap || o.probe_points > cap || o.batch > (1ULL << 24) || o.min_points > cap)
        return serr(ctx, MI_ERR_RANGE, "pass sizes above 2^40 (training batch above 2^24)");

    mi_solve_info I;
    std::memset(&I, 0, sizeof I);
    const double target = o.rel_target;
    uint64_t next = run->counter0;                  // passes take consecutive, disjoint counter ranges
    mi_sigma_est e;

    // the VEGAS pilot: plain VEGAS's own, and the fallback's projection in the flow arm
    mi_run r = pass(*run, MI_SAMPLER_VEGAS, next, o.pilot_points);
    int rc = mi_integrate(ctx, &r, &e);
    if (rc) return rc;
    next += o.pilot_points;
    I.pilot_ms = e.kernel_ms;
    I.vegas_ms_per_point = e.kernel_ms / (double)o.pilot_points;
    const double rel_v = e.sigmaGen[0] > 0.0 ? e.sigmaErr[0] / e.sigmaGen[0] : 0.0;
    I.cv_vegas = rel_v * std::sqrt((double)o.pilot_points);
    const uint64_t n_v = solve_final_points(rel_v, o.pilot_points, target, o.min_points);
    if (n_v > cap) return serr(ctx, MI_ERR_RANGE, "the target needs more than 2^40 VEGAS points");
    I.vegas_final_ms_proj = I.vegas_ms_per_point * (double)n_v;

    int final_sampler = MI_SAMPLER_VEGAS;
    uint64_t n_final = n_v;
    I.decision = MI_SOLVE_VEGAS;
    const bool forced = o.force_flow != 0;
    // A flow point evaluates the same integrand through the same VEGAS map as a VEGAS point, plus the
    // flow's layers: VEGAS's own cost per point bounds the flow's from below, so when even that bound
    // says no flow can pay back, plain VEGAS finishes without a probe.
    const bool hopeless_now = !forced && solve_flow_hopeless(I.vegas_final_ms_proj, I.vegas_ms_per_point,
                                                             o.chunk_steps, o.batch, o.pilot_points, o.min_points);
    if (o.sampler == MI_SAMPLER_FLOW && hopeless_now) I.decision = MI_SOLVE_FALLBACK_PROBE;
    if (o.sampler == MI_SAMPLER_FLOW && !hopeless_now) {
        rc = mi_flow_init(ctx, o.flow_seed);
        if (rc) return rc;
        // identity flow: its cost per point and starting variance
        r = pass(*run, MI_SAMPLER_FLOW, next, o.probe_points);
        rc = mi_integrate(ctx, &r, &e);
        if (rc) return rc;
        next += o.probe_points;
        I.probe_ms = e.kernel_ms;
        I.flow_ms_per_point = e.kernel_ms / (double)o.probe_points;
        const double cv0 = e.sigmaGen[0] > 0.0 ? e.sigmaErr[0] / e.sigmaGen[0] * std::sqrt((double)o.probe_points) : 0.0;
        I.cv_flow = cv0;
        double rest = solve_flow_rest_ms(cv0, I.flow_ms_per_point, o.pilot_points, target, o.min_points);
        if (!forced && solve_flow_hopeless(I.vegas_final_ms_proj, I.flow_ms_per_point, o.chunk_steps, o.batch,
                                           o.pilot_points, o.min_points)) {
            I.decision = MI_SOLVE_FALLBACK_PROBE;
        } else {
            // training in chunks on their own counters (a block of 2^40 above the passes)
            const uint64_t train0 = run->counter0 + (1ULL << 41);
            for (;;) {
                r = pass(*run, MI_SAMPLER_FLOW, train0 + (uint64_t)I.train_steps * o.batch, o.batch);
                mi_flow_info fi;
                rc = mi_flow_train(ctx, &r, o.chunk_steps, o.batch, o.lr, &fi);
                if (rc) return rc;
                I.train_steps += o.chunk_steps;
                I.train_ms += fi.kernel_ms;
                I.cv_flow = fi.cv_flow;
                const double now = solve_flow_rest_ms(fi.cv_flow, I.flow_ms_per_point, o.pilot_points, target,
                                                      o.min_points);
                const bool more = forced ? I.train_steps < o.max_steps
                                         : solve_keep_training(rest, now, fi.kernel_ms, I.train_steps, o.max_steps,
                                                               I.vegas_final_ms_proj);
                rest = now;
                if (!more) break;
            }
            I.flow_rest_ms_proj = rest;
            if (forced || solve_use_flow(rest, I.vegas_final_ms_proj)) {
                I.decision = forced ? MI_SOLVE_FLOW_FORCED : MI_SOLVE_FLOW;
                r = pass(*run, MI_SAMPLER_FLOW, next, o.pilot_points);
                rc = mi_integrate(ctx, &r, &e);
                if (rc) return rc;
                next += o.pilot_points;
                I.flow_pilot_ms = e.kernel_ms;
                const double rel_f = e.sigmaGen[0] > 0.0 ? e.sigmaErr[0] / e.sigmaGen[0] : 0.0;
                I.cv_flow = rel_f * std::sqrt((double)o.pilot_points);
                n_final = solve_final_points(rel_f, o.pilot_points, target, o.min_points);
                if (n_final > cap) return serr(ctx, MI_ERR_RANGE, "the target needs more than 2^40 flow points");
                final_sampler = MI_SAMPLER_FLOW;
            } else {
                I.decision = MI_SOLVE_FALLBACK_TRAINED;
            }
        }
    }
    r = pass(*run, final_sampler, next, n_final);
    rc = mi_integrate(ctx, &r, out);
    if (rc) return rc;
    I.final_ms = out->kernel_ms;
    I.final_points = n_final;
    I.sampler_used = final_sampler;
    I.gpu_ms = I.pilot_ms + I.probe_ms + I.train_ms + I.flow_pilot_ms + I.final_ms;
    if (info) *info = I;
    return MI_OK;
}

}  // extern "C"
