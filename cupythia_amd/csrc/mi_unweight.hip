// mi_unweight.hip -- hit-or-miss unweighting and order-preserving compaction on the device.
//
// Reference behaviour: accept iff u * wMax < w with u from a disjoint stream
// (cuPythia/pipeline/generate.cu:56-62, cuPythia/10_cub_compaction/cub_compaction.cu:36-46;
// Pythia ProcessContainer.cc:437), then a stable filter of the accepted candidates
// (cub::DeviceSelect::Flagged semantics, cub_compaction.cu:76-80, generate.cu:98-103).
//
// Three launches, no spin-waits:
//   count:  the integrand of every candidate, accept flags, per-block accepted counts and the
//           weighted-estimate partials;
//   scan:   exclusive scan of the block counts (one block);
//   write:  the integrand again (cheaper than storing and re-reading every candidate), per-step
//           wave ballot + popcount for lane ranks, LDS prefix over the 4 waves, and a
//           contiguous, coalesced SoA write of the accepted events at their global rank.
// Block b owns candidates [b*256*ppt, (b+1)*256*ppt) and walks them in 256-wide steps, so the
// global rank order is the candidate order.
#include "mi_kernel_common.h"

namespace mi {

namespace {

struct Cand {
    double w;          // total weight (mb), the FP32 value promoted exactly in FP32 mode
    double ws[kMaxProc];
    double sH, tH, x1, x2;
    bool acc;
    int pid;
};

// Accept stream: Philox stream 1 of the same event: u = 53-bit uniform from words (0, 1),
// process choice from the 24-bit uniform of word 2 against the cumulative FP32/FP64 weights.
template <typename T>
__device__ __forceinline__ void decide(uint64_t key, uint64_t e, T tot, const T *ws, int mask, double wmax, Cand &c)
{
    const U4 a = philox_event(key, e, 1u);
    const double u = u53(a.x, a.y);
    c.acc = u * wmax < (double)tot;
    const T target = (T)k24(a.z) * (T)(1.0 / 16777216.0) * tot;
    T cum = (T)0;
    int pid = 0;
    for (int i = 0; i < kMaxProc; ++i) {
        if (!((mask >> i) & 1) || !(ws[i] > (T)0)) continue;
        pid = proc_code(i);
        cum = cum + ws[i];
        if (cum > target) break;
    }
    c.pid = pid;
}

// HEAVY: the event's kinematics are those of the chosen process's family (each family has its
// own map at the same uniforms).
template <typename T, bool VEGAS, bool HEAVY>
__device__ __forceinline__ void had_cand(const HadConst<T> &k, int mask, uint64_t key, uint64_t e, const T *stab,
                                         int nbins, double wmax, const GridPdf<T> *gp, Cand &c)
{
    T ws[kMaxProc];
    int bins[3];
    constexpr int NF = HEAVY ? 3 : 1;
    T sH[NF], tH[NF], x1[NF], x2[NF];
    const T tot = had_candidate<T, VEGAS, HEAVY>(k, mask, key, e, stab, nbins, (T)nbins, ws, bins, sH, tH, x1, x2, gp);
    c.w = (double)tot;
#pragma unroll
    for (int i = 0; i < kMaxProc; ++i) c.ws[i] = (double)ws[i];
    decide<T>(key, e, tot, ws, mask, wmax, c);
    int f = 0;
    if (HEAVY) {
        f = c.pid >= 123 ? 2 : (c.pid >= 121 ? 1 : 0);
        if (c.pid == 0) f = (mask & kFamMask[0]) ? 0 : ((mask & kFamMask[1]) ? 1 : 2);
    }
    c.sH = (double)pick_fam<T, NF>(sH, f); c.tH = (double)pick_fam<T, NF>(tH, f);
    c.x1 = (double)pick_fam<T, NF>(x1, f); c.x2 = (double)pick_fam<T, NF>(x2, f);
}

// Partonic gg -> gg candidate.  Philox: stream 0 for cos(theta), stream 1 for u.  Compat:
// the reference kernel-10 stream, ctr = seed + e * 0x100000001B3, draws cos(theta) then u.
__device__ __forceinline__ void c1_cand(double pref, double cMax, double sHat, uint64_t key, uint64_t e, bool compat,
                                        double wmax, Cand &c)
{
    double u, uacc;
    if (compat) {
        const uint64_t ctr = key + e * 0x100000001B3ULL;
        u = u01(splitmix64(ctr));
        uacc = u01(splitmix64(ctr + 1));
    } else {
        const U4 r = philox_event(key, e, 0u);
        u = u53(r.x, r.y);
        const U4 a = philox_event(key, e, 1u);
        uacc = u53(a.x, a.y);
    }
    const double w = c1_weight(pref, cMax, u);
    const double cth = cMax * (2.0 * u - 1.0);
    c.w = w;
    for (int i = 0; i < kMaxProc; ++i) c.ws[i] = 0.0;
    c.ws[0] = w;
    c.sH = sHat; c.tH = -0.5 * sHat * (1.0 - cth); c.x1 = 1.0; c.x2 = 1.0;
    c.acc = uacc * wmax < w;
    c.pid = 111;
}

// KIND 0: hadronic flat, 1: hadronic VEGAS, 2: partonic, 3 / 4: hadronic flat / VEGAS with the
// heavy-quark codes selected.
constexpr bool kind_vegas(int kind) { return kind == 1 || kind == 4; }
constexpr bool kind_heavy(int kind) { return kind >= 3; }

template <int KIND, typename T>
struct Params {
    HadConst<T> k;
    int mask;
    uint64_t key, ctr0, n;
    uint32_t ppt;
    const T *gtab;
    int nbins;
    double wmax;
    double pref, cMax, sHat;
    int compat;
    GridPdf<T> gpdf;       // PDF grid (active = 0: toy PDFs)
};

template <int KIND, typename T>
__device__ __forceinline__ void candidate(const Params<KIND, T> &P, const T *stab, uint64_t e, Cand &c)
{
    const GridPdf<T> *gp = P.gpdf.active ? &P.gpdf : nullptr;
    if (KIND == 2) c1_cand(P.pref, P.cMax, P.sHat, P.key, e, P.compat != 0, P.wmax, c);
    else had_cand<T, kind_vegas(KIND), kind_heavy(KIND)>(P.k, P.mask, P.key, e, stab, P.nbins, P.wmax, gp, c);
}

template <int KIND, typename T>
__global__ __launch_bounds__(kBlock) void k_unw_count(Params<KIND, T> P, double *__restrict__ partials,
                                                      uint64_t *__restrict__ counts)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *stab = reinterpret_cast<T *>(smem);
    if (kind_vegas(KIND)) {
        for (int i = threadIdx.x; i < 9 * P.nbins; i += blockDim.x) stab[i] = P.gtab[i];
        __syncthreads();
    }
    double v[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) v[i] = 0.0;
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * P.ppt;
    for (uint32_t j = 0; j < P.ppt; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        if (i >= P.n) break;
        Cand c;
        candidate<KIND, T>(P, stab, P.ctr0 + i, c);
        v[kRS] += c.w;
        v[kRQ] = fma(c.w, c.w, v[kRQ]);
        v[kRM] = fmax(v[kRM], c.w);
        v[kRNz] += (c.w > 0.0) ? 1.0 : 0.0;
        v[kRAcc] += c.acc ? 1.0 : 0.0;
#pragma unroll
        for (int p = 0; p < (KIND < 2 ? kLightProc : kMaxProc); ++p) {
            v[kRS + 1 + p] += c.ws[p];
            v[kRQ + 1 + p] = fma(c.ws[p], c.ws[p], v[kRQ + 1 + p]);
            v[kRM + 1 + p] = fmax(v[kRM + 1 + p], c.ws[p]);
        }
    }
    block_record(v, partials + (size_t)blockIdx.x * kRec, counts + blockIdx.x);
}

// Exclusive scan of nblk counts in place; counts[nblk] receives the total.  One block of 1024.
__global__ __launch_bounds__(1024) void k_scan_counts(uint64_t *__restrict__ counts, uint32_t nblk)
{
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (nblk + 1023) / 1024;
    const uint32_t lo = t * chunk, hi = min(nblk, lo + chunk);
    uint64_t s = 0;
    for (uint32_t b = lo; b < hi; ++b) s += counts[b];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint64_t run = 0;
        for (int i = 0; i < 1024; ++i) { const uint64_t x = part[i]; part[i] = run; run += x; }
        counts[nblk] = run;
    }
    __syncthreads();
    uint64_t run = part[t];
    for (uint32_t b = lo; b < hi; ++b) { const uint64_t x = counts[b]; counts[b] = run; run += x; }
}

template <int KIND, typename T>
__global__ __launch_bounds__(kBlock) void k_unw_write(Params<KIND, T> P, const uint64_t *__restrict__ offsets,
                                                      mi_events out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *stab = reinterpret_cast<T *>(smem);
    __shared__ uint32_t wcnt[kBlock / 64];
    if (kind_vegas(KIND)) {
        for (int i = threadIdx.x; i < 9 * P.nbins; i += blockDim.x) stab[i] = P.gtab[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    uint64_t pos0 = offsets[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * P.ppt;
    for (uint32_t j = 0; j < P.ppt; ++j) {                 // block-uniform trip count (barriers inside)
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        Cand c;
        c.acc = false;
        if (i < P.n) candidate<KIND, T>(P, stab, P.ctr0 + i, c);
        const uint64_t m = __ballot(c.acc);
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) {
            const uint32_t x = wcnt[w];
            before += (w < wave) ? x : 0u;
            total += x;
        }
        if (c.acc) {
            const uint64_t pos = pos0 + before + (uint64_t)__popcll(m & lt);
            if (pos < out.capacity) {
                out.sHat[pos] = c.sH;
                out.tHat[pos] = c.tH;
                out.x1[pos] = c.x1;
                out.x2[pos] = c.x2;
                out.weight[pos] = fmax(c.w, P.wmax);
                out.pid[pos] = c.pid;
                out.index[pos] = P.ctr0 + i;
            }
        }
        pos0 += total;
        __syncthreads();
    }
}

inline const GridPdf<float> &grid_of(const HadLaunch &h, float *) { return h.gf; }
inline const GridPdf<double> &grid_of(const HadLaunch &h, double *) { return h.gd; }

template <int KIND, typename T>
Params<KIND, T> make_params(const UnwLaunch &L, const HadConst<T> &k, const T *tab)
{
    Params<KIND, T> P{};
    P.k = k;
    if (!L.partonic) P.gpdf = grid_of(L.h, (T *)nullptr);
    if (L.partonic) {
        P.key = L.c.key; P.ctr0 = L.c.ctr0; P.n = L.c.n; P.ppt = L.c.g.ppt;
    } else {
        P.mask = L.h.mask; P.key = L.h.key; P.ctr0 = L.h.ctr0; P.n = L.h.n; P.ppt = L.h.g.ppt;
    }
    P.gtab = tab;
    P.nbins = (L.partonic || !L.h.vegas) ? 0 : L.h.grid->nbins;
    P.wmax = L.wmax;
    P.pref = L.c.pref; P.cMax = L.c.cMax; P.sHat = L.sHat;
    P.compat = L.compat ? 1 : 0;
    return P;
}

template <int KIND, typename T>
hipError_t count_t(hipStream_t s, const UnwLaunch &L, const HadConst<T> &k, const T *tab, uint64_t *counts)
{
    const Params<KIND, T> P = make_params<KIND, T>(L, k, tab);
    const Geom g = L.partonic ? L.c.g : L.h.g;
    double *parts = L.partonic ? L.c.partials : L.h.partials;
    const size_t sm = kind_vegas(KIND) ? sizeof(T) * 9 * P.nbins : 0;
    hipLaunchKernelGGL((k_unw_count<KIND, T>), dim3(g.blocks), dim3(kBlock), sm, s, P, parts, counts);
    return hipGetLastError();
}

template <int KIND, typename T>
hipError_t write_t(hipStream_t s, const UnwLaunch &L, const HadConst<T> &k, const T *tab, const uint64_t *offsets,
                   const mi_events &out)
{
    const Params<KIND, T> P = make_params<KIND, T>(L, k, tab);
    const Geom g = L.partonic ? L.c.g : L.h.g;
    const size_t sm = kind_vegas(KIND) ? sizeof(T) * 9 * P.nbins : 0;
    hipLaunchKernelGGL((k_unw_write<KIND, T>), dim3(g.blocks), dim3(kBlock), sm, s, P, offsets, out);
    return hipGetLastError();
}

template <typename T>
hipError_t count_h(hipStream_t s, const UnwLaunch &L, const HadConst<T> &k, const T *tab, uint64_t *counts)
{
    const bool hv = has_heavy(L.h.mask);
    if (L.h.vegas) return hv ? count_t<4, T>(s, L, k, tab, counts) : count_t<1, T>(s, L, k, tab, counts);
    return hv ? count_t<3, T>(s, L, k, nullptr, counts) : count_t<0, T>(s, L, k, nullptr, counts);
}

template <typename T>
hipError_t write_h(hipStream_t s, const UnwLaunch &L, const HadConst<T> &k, const T *tab, const uint64_t *offsets,
                   const mi_events &out)
{
    const bool hv = has_heavy(L.h.mask);
    if (L.h.vegas) return hv ? write_t<4, T>(s, L, k, tab, offsets, out) : write_t<1, T>(s, L, k, tab, offsets, out);
    return hv ? write_t<3, T>(s, L, k, nullptr, offsets, out) : write_t<0, T>(s, L, k, nullptr, offsets, out);
}

}  // namespace

hipError_t launch_unweight_count(hipStream_t s, const UnwLaunch &L, uint64_t *counts)
{
    if (L.partonic) {
        HadConst<double> kd{};
        return count_t<2, double>(s, L, kd, nullptr, counts);
    }
    if (L.h.prec == MI_PREC_FP32) return count_h<float>(s, L, L.h.kf, L.h.vegas ? L.h.grid->tab32 : nullptr, counts);
    return count_h<double>(s, L, L.h.kd, L.h.vegas ? L.h.grid->tab64 : nullptr, counts);
}

hipError_t launch_scan_counts(hipStream_t s, uint64_t *counts, uint32_t nblk)
{
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, nblk);
    return hipGetLastError();
}

hipError_t launch_unweight_write(hipStream_t s, const UnwLaunch &L, const uint64_t *offsets, const mi_events &out)
{
    if (L.partonic) {
        HadConst<double> kd{};
        return write_t<2, double>(s, L, kd, nullptr, offsets, out);
    }
    if (L.h.prec == MI_PREC_FP32)
        return write_h<float>(s, L, L.h.kf, L.h.vegas ? L.h.grid->tab32 : nullptr, offsets, out);
    return write_h<double>(s, L, L.h.kd, L.h.vegas ? L.h.grid->tab64 : nullptr, offsets, out);
}

}  // namespace mi
