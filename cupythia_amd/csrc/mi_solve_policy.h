// mi_solve_policy.h -- the decisions of the time-to-precision driver (mi_solve), host-only and
// free of HIP so that they can be exercised on any machine.  Costs are device milliseconds as the
// passes report them; a sampler's cost per point is measured on its own pass.
//
// Protocol of a solve: a variance pilot of P points fixes the points a final pass needs,
// N = max(N_min, ceil(1.1 P (rel / target)^2)) (the pilot's relative error rel; 10% margin); the
// final pass alone is the estimate.  Plain VEGAS: pilot + final.  Flow + VEGAS: the VEGAS pilot
// (the fallback's own projection), an identity-flow probe for the flow's cost per point, training
// in chunks, a flow pilot, the final pass -- or, whenever the flow cannot beat it, the VEGAS final
// pass (plain VEGAS stays the fallback).
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

// What the flow still costs at best once it has been trained enough: its pilot and a final pass
// of the points its coefficient of variation cv needs (ms_pt: the flow's cost per point).
inline double solve_flow_rest_ms(double cv, double ms_pt, uint64_t n_pilot, double target, uint64_t n_min)
{
    return ms_pt * ((double)n_pilot + (double)solve_points_for_cv(cv, n_pilot, target, n_min));
}

// Before any training: the cheapest possible continuation of the flow -- one chunk of training,
// which samples at least chunk * batch points through the flow, then its pilot and the smallest
// final pass (a perfect flow) -- against the VEGAS final pass.  True: fall back now.
inline bool solve_flow_hopeless(double vegas_final_ms, double flow_ms_pt, int chunk, uint64_t batch,
                                uint64_t n_pilot, uint64_t n_min)
{
    const double lb = flow_ms_pt * ((double)chunk * (double)batch + (double)n_pilot + (double)n_min);
    return !(lb < vegas_final_ms);
}

// After a training chunk that cost chunk_ms and moved the flow's projected rest from prev_rest to
// rest: train another chunk?  Only while steps remain, the last chunk saved more than it cost, and
// another chunk saving as much again would end below the cheaper of the two ways to finish now
// (the flow's rest, VEGAS's final pass): training is judged against plain VEGAS, not only against
// the flow's own previous state.
inline bool solve_keep_training(double prev_rest, double rest, double chunk_ms, int steps, int max_steps,
                                double vegas_final_ms)
{
    if (steps >= max_steps) return false;
    const double saving = prev_rest - rest;
    if (!(saving > chunk_ms)) return false;
    const double best_now = rest < vegas_final_ms ? rest : vegas_final_ms;
    return rest - saving + chunk_ms < best_now;
}

// The final pass: the flow (its pilot and final) only if that projection is below VEGAS's final.
inline bool solve_use_flow(double flow_rest_ms, double vegas_final_ms) { return flow_rest_ms < vegas_final_ms; }

}  // namespace mi
