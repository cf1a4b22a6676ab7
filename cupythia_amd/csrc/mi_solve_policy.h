// mi_solve_policy.h -- the decisions of the time-to-precision driver (mi_solve), host-only and
// free of HIP so that they can be exercised on any machine.  Costs are device milliseconds as the
// passes report them; a sampler's cost per point is measured on its own pass.
//
// Protocol of a solve: a variance pilot of P points fixes the points a final pass needs,
// N = max(N_min, ceil(1.1 P (rel / target)^2)) (the pilot's relative error rel; 10% margin); the
// final pass alone is the estimate.  Plain VEGAS: pilot + final.  Flow + VEGAS: the VEGAS pilot
// (the fallback's own projection), a VEGAS pass of one training batch (with the pilot, VEGAS's
// affine pass cost), an identity-flow probe for the flow's cost per point, training in chunks, a
// flow pilot, the final pass -- or, whenever the flow cannot beat it, the VEGAS final pass (plain
// VEGAS stays the fallback).
#pragma once

#include <cmath>
#include <cstdint>

namespace mi {

// Points of a final pass for relative error `target`, from a pilot of n_pilot points whose
// relative error was rel.
inline uint64_t solve_final_points(double rel, uint64_t n_pilot, double target, uint64_t n_min)
{
    if (!(rel > 0.0) || !(target > 0.0)) return n_min;
    const double n = std::ceil(1.1 * (double)n_pilot * (rel / target) * (rel / target));
    if (!(n < 1.8e19)) return UINT64_MAX;
    return n > (double)n_min ? (uint64_t)n : n_min;
}

// The same from a per-point coefficient of variation cv (rel = cv / sqrt(n_pilot)).
inline uint64_t solve_points_for_cv(double cv, uint64_t n_pilot, double target, uint64_t n_min)
{
    return solve_final_points(cv / std::sqrt((double)n_pilot), n_pilot, target, n_min);
}


// Affine cost of a pass of n points, t0 + per_pt n: the fixed part is the launches, the reduction
// and the ramp of the grid, which a pass of 2^20 points still spends as much time on as on its
// points (measured on MI355X: 12.7 us + 11.3 ps per VEGAS point).
struct PassCost {
    double t0, per_pt;
};
inline double pass_ms(const PassCost &c, double n) { return c.t0 + c.per_pt * n; }

// The affine cost from two passes of n_small < n_large points; with degenerate or noisy timings
// (no larger time for the larger pass) the large pass's average per point.
inline PassCost solve_affine(double t_small, uint64_t n_small, double t_large, uint64_t n_large)
{
    if (n_large > n_small && t_large > t_small) {
        const double c = (t_large - t_small) / (double)(n_large - n_small);
        const double t0 = t_small - c * (double)n_small;
        return {t0 > 0.0 ? t0 : 0.0, c};
    }
    return {0.0, n_large > 0 ? t_large / (double)n_large : 0.0};
}

// What the flow still costs once trained to the coefficient of variation cv: its pilot and a final
// pass of the points cv needs, at the flow's pass cost.
inline double solve_flow_rest_ms(double cv, const PassCost &flow, uint64_t n_pilot, double target, uint64_t n_min)
{
    return pass_ms(flow, (double)n_pilot) + pass_ms(flow, (double)solve_points_for_cv(cv, n_pilot, target, n_min));
}

// Before any training: the cheapest possible continuation of the flow -- one chunk of training,
// which samples at least chunk * batch points through the flow, then its pilot and the smallest
// final pass (a perfect flow) -- against the VEGAS final pass.  True: fall back now.  ms_pt: a
// cost per point no flow point can undercut (VEGAS's: a flow point evaluates the same integrand
// through the same map, plus the flow).
inline bool solve_flow_hopeless(double vegas_final_ms, double flow_ms_pt, int chunk, uint64_t batch,
                                uint64_t n_pilot, uint64_t n_min)
{
    const double lb = flow_ms_pt * ((double)chunk * (double)batch + (double)n_pilot + (double)n_min);
    return !(lb < vegas_final_ms);
}

// The same with an affine lower cost: each training step is at least one pass of batch points,
// then the flow's pilot and its smallest final pass.
inline bool solve_flow_hopeless_passes(double vegas_final_ms, const PassCost &lower, int chunk, uint64_t batch,
                                       uint64_t n_pilot, uint64_t n_min)
{
    const double lb = (double)chunk * pass_ms(lower, (double)batch) + pass_ms(lower, (double)n_pilot) +
                      pass_ms(lower, (double)n_min);
    return !(lb < vegas_final_ms);
}

// After k >= 1 training chunks: rest[0..k] is the flow's projected rest (pilot + final pass)
// before training and after each chunk.  Train another chunk?  The last (up to three) chunks show
// the average rate r = (rest[k] / rest[k-w])^(1/w) per chunk; a training run gains less with every
// chunk, so the j-th further chunk is taken to multiply the rest by 1 - (1 - r) d^j, the gain shrinking
// by d = kSolveDecel per chunk.  Train on while steps remain and some number j >= 1 of further
// chunks, chunk_ms each, is projected to end below the cheaper of the two ways to finish now (the
// flow's rest, VEGAS's final pass).  Training is judged against plain VEGAS, not only against the
// flow's own state, and the window keeps one noisy batch (the coefficients of variation are
// estimated on one batch each) from ending it.
// kSolveDecel: on MI355X the 5-process library's flow (2^16-point batches, lr 1e-2, chunks of 10
// steps) gains 1 - r = 0.68, 0.51, 0.43, 0.34, 0.27, 0.22, 0.14 in its first seven chunks, about
// 0.8 of the previous chunk's gain each time.
constexpr double kSolveDecel = 0.8;
inline bool solve_keep_training(const double *rest, int k, double chunk_ms, int steps, int max_steps,
                                int chunk_steps, double vegas_final_ms)
{
    if (k < 1 || steps >= max_steps || chunk_steps < 1) return false;
    const int w = k < 3 ? k : 3;
    const double now = rest[k], then = rest[k - w];
    if (!(now > 0.0) || !(then > 0.0)) return false;
    const double r = std::pow(now / then, 1.0 / (double)w);
    if (!(r < 1.0)) return false;
    const int left = (max_steps - steps) / chunk_steps;
    const double best_now = now < vegas_final_ms ? now : vegas_final_ms;
    double proj = now, gain = 1.0 - r;
    for (int j = 1; j <= left; ++j) {
        gain *= kSolveDecel;
        proj *= 1.0 - gain;
        if ((double)j * chunk_ms + proj < best_now) return true;
    }
    return false;
}

// The final pass: the flow (its pilot and final) only if that projection is below VEGAS's final.
inline bool solve_use_flow(double flow_rest_ms, double vegas_final_ms) { return flow_rest_ms < vegas_final_ms; }

}  // namespace mi
