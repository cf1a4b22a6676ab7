#!/usr/bin/env python3
"""Benchmark of the MI355X matrix-element integration path.

Workload (one step): the 5-process QCD 2->2 library at sqrt(S) = 13.6 TeV, pT > 20 GeV, toy
PDFs, 1-loop alphaS(MZ) = 0.118, 3-D (tau, y, cos theta) phase space, 2^28 weighted points per
step, FP32 matrix elements with FP64 sum w / sum w^2, Philox4x32-10 key 42, counters advancing
by 2^28 per step (every step integrates new events).  Inputs are generated on the device, so
nothing crosses PCIe inside the timed region except the 24-double result of each step.

Prints one JSON line: weighted events/s (value), the roofline of the integration kernel (its
duration measured with a hipEvent pair on its own stream, inside the library), the CPU baseline
(the Pythia-form FP64 oracle on the host cores, bounded sample), and, with --tts, the
time-to-solution at 0.1% relative error of plain VEGAS and flow + VEGAS.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("σ rel-err 0.1% time-to-solution (flow+VEGAS vs plain VEGAS); weighted events/s")
UNIT = "weighted events/s"
# Algorithmic flops per weighted point (FMA = 2; div, sqrt, exp, log = 1): 162 in the FP32
# integrand of the five processes + 18 FP64 accumulation flops (sum w, sum w^2 of total and
# five processes).  Philox's integer work is excluded (reported by rocprof instead).
FLOPS_PER_POINT = 180
PEAK_FP32_TFLOPS = 157.3          # MI355X vector FP32 (vendor figure, counts packed FP32)
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_integrate_had_2p28.txt")


def traffic_from_profile():
    """HBM bytes per launch of the integration kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by the copy-kernel calibration)."""
    if not os.path.exists(PMC_FILE):
        return None
    for line in open(PMC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(n_sample, threads):
    import ctypes as C
    from oracle import oracle as O
    cfg = O.hadcfg_default()
    s = (C.c_double * 2)()
    L = O.lib()
    L.orc_had_mc(C.byref(cfg), 42, 0, 1024, threads, s)            # warm
    t0 = time.perf_counter()
    L.orc_had_mc(C.byref(cfg), 42, 0, n_sample, threads, s)
    dt = time.perf_counter() - t0
    return n_sample / dt, dt, s[0] / n_sample


def time_to_solution(ctx, mi, target=1e-3, flow=False):
    """GPU time (sum of device-side kernel times of every launch) until sigma_err/sigma <= target:
    VEGAS adaptation (+ flow training) + the final integration with the frozen map."""
    phys = mi.physics_default(1)
    t0 = time.perf_counter()
    ctx.vegas_init(3, 64)
    info = ctx.vegas_adapt(mi.make_run(phys, 1 << 20, key=42), 8, 1.5)
    gpu_ms = info.kernel_ms
    sampler = mi.MI_SAMPLER_VEGAS
    if flow:
        ctx.flow_init(7)
        fi = ctx.flow_train(mi.make_run(phys, 1 << 20, key=43), 200, 1 << 16, 2e-3)
        gpu_ms += fi.kernel_ms
        sampler = mi.MI_SAMPLER_FLOW
    # pilot for the variance, then the final pass sized for the target
    pilot = ctx.integrate(mi.make_run(phys, 1 << 20, key=44, sampler=sampler))
    gpu_ms += pilot.kernel_ms
    rel = pilot.sigmaErr() / pilot.sigmaGen()
    n = int(math.ceil((1 << 20) * (rel / target) ** 2 * 1.1))
    n = max(n, 1 << 16)
    fin = ctx.integrate(mi.make_run(phys, n, key=45, sampler=sampler))
    gpu_ms += fin.kernel_ms
    wall = time.perf_counter() - t0
    return {"gpu_ms": gpu_ms, "wall_s": wall, "points_final": n, "sigma": fin.sigmaGen(),
            "rel_err": fin.sigmaErr() / fin.sigmaGen()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1 << 28)
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tts", action="store_true", help="also measure time-to-solution (VEGAS, flow+VEGAS)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 --pmc: 3 integration launches and one 1 GiB device copy (the "
                         "calibration of FETCH_SIZE / WRITE_SIZE), nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the integration shards by "
              "counter range and needs one all-reduce of a few hundred bytes, which this build does not "
              "launch.", file=sys.stderr)
        sys.exit(2)

    import torch
    import cupythia_amd as mi

    ctx = mi.Context(0)
    phys = mi.physics_default(1)
    n = args.points
    if args.pmc_probe:
        for k in range(3):
            ctx.integrate(mi.make_run(phys, n, key=42, counter0=k * n))
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")      # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "copy_bytes": src.numel() * 4}))
        ctx.close()
        return
    for w in range(args.warmup):
        ctx.integrate(mi.make_run(phys, n, key=42, counter0=w * n))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kern_ms = []
    est = None
    for k in range(args.steps):
        est = ctx.integrate(mi.make_run(phys, n, key=42, counter0=(args.warmup + k) * n))
        kern_ms.append(est.kernel_ms)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = t1 - t0
    value = n * args.steps / dt
    avg_kernel_s = sum(kern_ms) / len(kern_ms) / 1e3
    achieved = FLOPS_PER_POINT * n / avg_kernel_s / 1e12

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "5-process QCD 2->2 library (gg->gg, gg->qqbar, qg->qg, qq'->qq', qqbar->gg), "
                               "sqrt(S)=13.6 TeV, pT>20 GeV, toy PDFs, 1-loop alphaS(MZ)=0.118, 3-D (tau,y,cos), "
                               "FP32 ME + FP64 sums, Philox4x32-10 key 42",
                   "points_per_step": n, "precision": "fp32 ME / fp64 accumulation",
                   "sigma_mb": est.sigmaGen(), "sigma_err_mb": est.sigmaErr()},
        "roofline": {"bound": "mfma", "engine": "fp32 vector ALU (compute roof; the FP32 matrix and vector peaks are both 157.3)",
                     "kernel": "k_integrate_had<float,flat>", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic_from_profile(), "flops_per_point": FLOPS_PER_POINT,
                     "avg_kernel_ms": avg_kernel_s * 1e3},
    }
    if not args.no_cpu:
        rate1, dt1, _ = cpu_baseline(args.cpu_sample, 1)
        out["cpu_baseline"] = {"value": rate1, "unit": UNIT, "cores": 1, "kind": "port",
                               "sample": f"{args.cpu_sample} points of the same workload through the FP64 "
                                         f"Pythia-form oracle (flux lists, literal sigmaHat), flat map, "
                                         f"{dt1:.1f} s on one host core"}
        nthr = min(16, os.cpu_count() or 1)
        rateN, dtN, _ = cpu_baseline(args.cpu_sample * 4, nthr)
        out["cpu_baseline"]["all_cores"] = {"value": rateN, "cores": nthr, "seconds": dtN}
    if args.tts:
        tv = time_to_solution(ctx, mi, flow=False)
        out["time_to_solution"] = {"vegas": tv}
        try:
            tf = time_to_solution(ctx, mi, flow=True)
            out["time_to_solution"]["flow_vegas"] = tf
            out["time_to_solution"]["speedup"] = tv["gpu_ms"] / tf["gpu_ms"]
        except mi.MIError as e:
            out["time_to_solution"]["flow_vegas"] = {"error": str(e)}
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
