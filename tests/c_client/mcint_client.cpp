// mcint_client.cpp -- a compiled host program on the library's C ABI alone (include/mcint.h),
// with the HIP runtime only for its device buffers: the chain a Pythia-side integration would
// drive -- integrate, unweight with compaction, build the device event record, write LHEF.
// Prints one JSON line.  Usage: mcint_client <n_points> <key> <lhe path>
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mcint.h"

#define CK(x)                                                                                      \
    do {                                                                                           \
        const int rc_ = (x);                                                                       \
        if (rc_ != MI_OK) {                                                                        \
            std::fprintf(stderr, "%s: %d %s\n", #x, rc_, ctx ? mi_last_error(ctx) : "");            \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)
#define HK(x)                                                                                      \
    do {                                                                                           \
        const hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                           \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

template <typename T>
static hipError_t dalloc(T **p, uint64_t n) { return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (n ? n : 1)); }

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <n_points> <key> <lhe path>\n", argv[0]);
        return 2;
    }
    const uint64_t n = std::strtoull(argv[1], nullptr, 10), key = std::strtoull(argv[2], nullptr, 10);
    mi_ctx *ctx = nullptr;
    CK(mi_ctx_create(0, &ctx));
    mi_run run{};
    CK(mi_physics_default(MI_MODE_HADRONIC_3D, &run.phys));
    run.precision = MI_PREC_FP32;
    run.sampler = MI_SAMPLER_FLAT;
    run.rng = MI_RNG_PHILOX;
    run.key = key;
    run.counter0 = 0;
    run.n_points = n;
    mi_sigma_est est;
    CK(mi_integrate(ctx, &run, &est));

    mi_events ev{};
    HK(dalloc(&ev.sHat, n)); HK(dalloc(&ev.tHat, n)); HK(dalloc(&ev.x1, n)); HK(dalloc(&ev.x2, n));
    HK(dalloc(&ev.weight, n)); HK(dalloc(&ev.pid, n)); HK(dalloc(&ev.index, n));
    ev.capacity = n;
    uint64_t nacc = 0;
    double wmax = 0.0;
    mi_sigma_est uest;
    CK(mi_unweight_compact(ctx, &run, 0.0, &ev, &nacc, &wmax, &uest));

    mi_particles rec{};
    const uint64_t np = nacc * 4;
    HK(dalloc(&rec.px, np)); HK(dalloc(&rec.py, np)); HK(dalloc(&rec.pz, np)); HK(dalloc(&rec.e, np));
    HK(dalloc(&rec.m, np)); HK(dalloc(&rec.pdg, np)); HK(dalloc(&rec.status, np)); HK(dalloc(&rec.col, np));
    HK(dalloc(&rec.acol, np)); HK(dalloc(&rec.mo1, np)); HK(dalloc(&rec.mo2, np)); HK(dalloc(&rec.d1, np));
    HK(dalloc(&rec.d2, np)); HK(dalloc(&rec.n_part, nacc)); HK(dalloc(&rec.scale, nacc));
    rec.capacity = nacc;
    rec.max_part = 4;
    CK(mi_build_events(ctx, &run.phys, &ev, nacc, key, &rec));
    std::vector<int32_t> pdg(np);
    std::vector<double> e(np);
    if (np) {
        HK(hipMemcpy(pdg.data(), rec.pdg, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
        HK(hipMemcpy(e.data(), rec.e, sizeof(double) * np, hipMemcpyDeviceToHost));
    }
    long long pdg_sum = 0;
    double e_sum = 0.0;
    for (uint64_t i = 0; i < np; ++i) { pdg_sum += pdg[i]; e_sum += e[i]; }
    CK(mi_write_lhe(ctx, argv[3], &run.phys, &ev, nacc, key, &uest));

    std::printf("{\"sigma_mb\": %.17g, \"sigma_err_mb\": %.17g, \"n_acc\": %llu, \"wmax\": %.17g, "
                "\"pdg_sum\": %lld, \"e_sum\": %.17g}\n",
                est.sigmaGen[0], est.sigmaErr[0], (unsigned long long)nacc, wmax, pdg_sum, e_sum);
    void *bufs[] = {ev.sHat, ev.tHat, ev.x1, ev.x2, ev.weight, ev.pid, ev.index, rec.px, rec.py, rec.pz, rec.e,
                    rec.m, rec.pdg, rec.status, rec.col, rec.acol, rec.mo1, rec.mo2, rec.d1, rec.d2, rec.n_part,
                    rec.scale};
    for (void *b : bufs) HK(hipFree(b));
    mi_ctx_destroy(ctx);
    return 0;
}
