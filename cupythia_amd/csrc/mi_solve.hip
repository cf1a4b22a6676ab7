// mi_solve.hip -- time-to-precision driver behind mi_solve: sigma to a relative-error target with
// plain VEGAS, or with the flow on top of VEGAS and plain VEGAS as the fallback (the decisions are
// mi_solve_policy.h's).  Host code only: every pass is one of the library's own entry points
// (mi_integrate, mi_flow_init, mi_flow_train), and the GPU time of the solve is the sum of the
// device times those passes report.
#include "mi_internal.h"
#include "mi_solve_policy.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

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
    if (o.max_steps == 0) o.max_steps = 400;
    if (o.pilot_points == 0) o.pilot_points = 1ULL << 20;
    if (o.probe_points == 0) o.probe_points = 1ULL << 20;
    if (o.batch == 0) o.batch = 1ULL << 16;
    if (o.min_points == 0) o.min_points = 1ULL << 16;
    if (!(o.lr > 0.0)) o.lr = 1e-2;
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
    const double pilot_ms = e.kernel_ms;
    const double rel_v = e.sigmaGen[0] > 0.0 ? e.sigmaErr[0] / e.sigmaGen[0] : 0.0;
    I.cv_vegas = rel_v * std::sqrt((double)o.pilot_points);
    const uint64_t n_v = solve_final_points(rel_v, o.pilot_points, target, o.min_points);
    if (n_v > cap) return serr(ctx, MI_ERR_RANGE, "the target needs more than 2^40 VEGAS points");
    PassCost vc{0.0, pilot_ms / (double)o.pilot_points};     // the pilot's average, until the affine fit
    I.vegas_ms_per_point = vc.per_pt;
    I.vegas_final_ms_proj = pass_ms(vc, (double)n_v);

    int final_sampler = MI_SAMPLER_VEGAS;
    uint64_t n_final = n_v;
    I.decision = MI_SOLVE_VEGAS;
    const bool forced = o.force_flow != 0;
    const uint64_t P = o.pilot_points, nmin = o.min_points;
    // A flow point evaluates the same integrand through the same VEGAS map as a VEGAS point, plus the
    // flow's layers: VEGAS's own cost per point bounds the flow's from below, so when even that bound
    // says no flow can pay back, plain VEGAS finishes without another pass.
    bool fallback = !forced && solve_flow_hopeless(I.vegas_final_ms_proj, vc.per_pt, o.chunk_steps, o.batch, P, nmin);
    if (o.sampler == MI_SAMPLER_FLOW && fallback) I.decision = MI_SOLVE_FALLBACK_PROBE;
    if (o.sampler == MI_SAMPLER_FLOW && !fallback) {
        // VEGAS over one training batch: with the pilot, VEGAS's affine pass cost (the pilot's
        // average per point still carries the fixed cost of a pass) and the cost of a pass the size
        // of one training step's sample
        r = pass(*run, MI_SAMPLER_VEGAS, next, o.batch);
        rc = mi_integrate(ctx, &r, &e);
        if (rc) return rc;
        next += o.batch;
        I.pilot_ms += e.kernel_ms;
        vc = solve_affine(e.kernel_ms, o.batch, pilot_ms, P);
        I.vegas_ms_per_point = vc.per_pt;
        I.vegas_final_ms_proj = pass_ms(vc, (double)n_v);
        fallback = !forced && solve_flow_hopeless_passes(I.vegas_final_ms_proj, vc, o.chunk_steps, o.batch, P, nmin);
        if (fallback) I.decision = MI_SOLVE_FALLBACK_PROBE;
    }
    if (o.sampler == MI_SAMPLER_FLOW && !fallback) {
        const double vfinal = I.vegas_final_ms_proj;
        rc = mi_flow_init(ctx, o.flow_seed);
        if (rc) return rc;
        // identity flow: its cost per point and starting variance; a flow pass has at least VEGAS's
        // fixed cost, the rest of the probe is per point
        r = pass(*run, MI_SAMPLER_FLOW, next, o.probe_points);
        rc = mi_integrate(ctx, &r, &e);
        if (rc) return rc;
        next += o.probe_points;
        I.probe_ms = e.kernel_ms;
        const double per_f = (e.kernel_ms - vc.t0) / (double)o.probe_points;
        const PassCost fc{vc.t0, per_f > vc.per_pt ? per_f : vc.per_pt};
        I.flow_ms_per_point = fc.per_pt;
        const double cv0 = e.sigmaGen[0] > 0.0 ? e.sigmaErr[0] / e.sigmaGen[0] * std::sqrt((double)o.probe_points) : 0.0;
        I.cv_flow = cv0;
        std::vector<double> rest{solve_flow_rest_ms(cv0, fc, P, target, nmin)};
        I.flow_rest_ms_proj = rest[0];
        if (!forced && solve_flow_hopeless_passes(vfinal, fc, o.chunk_steps, o.batch, P, nmin)) {
            I.decision = MI_SOLVE_FALLBACK_PROBE;
        } else {
            // training in chunks on their own counters (a block of 2^40 above the passes)
            const uint64_t train0 = run->counter0 + (1ULL << 41);
            mi_flow_info fi;
            auto train = [&](int steps) -> int {
                mi_run t = pass(*run, MI_SAMPLER_FLOW, train0 + (uint64_t)I.train_steps * o.batch, o.batch);
                const int c = mi_flow_train(ctx, &t, steps, o.batch, o.lr, &fi);
                if (c) return c;
                I.train_steps += steps;
                I.train_ms += fi.kernel_ms;
                I.cv_flow = fi.cv_flow;
                return MI_OK;
            };
            // the first step alone: its cost bounds the rest of the first chunk
            rc = train(1);
            if (rc) return rc;
            double chunk_ms = fi.kernel_ms;
            bool go = forced || o.chunk_steps == 1 ||
                      (double)(o.chunk_steps - 1) * fi.kernel_ms + pass_ms(fc, (double)P) + pass_ms(fc, (double)nmin) < vfinal;
            int todo = o.chunk_steps - 1;
            while (go) {
                if (todo > 0) {
                    rc = train(todo);
                    if (rc) return rc;
                    chunk_ms += fi.kernel_ms;
                }
                rest.push_back(solve_flow_rest_ms(fi.cv_flow, fc, P, target, nmin));
                const int k = (int)rest.size() - 1;
                const bool more = forced ? I.train_steps < o.max_steps
                                         : solve_keep_training(rest.data(), k, chunk_ms, I.train_steps, o.max_steps,
                                                               o.chunk_steps, vfinal);
                if (!more) break;
                todo = std::min(o.chunk_steps, o.max_steps - I.train_steps);
                chunk_ms = 0.0;
            }
            I.flow_rest_ms_proj = rest.back();
            if (forced || (rest.size() > 1 && solve_use_flow(rest.back(), vfinal))) {
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
