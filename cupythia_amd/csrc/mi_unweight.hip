// mi_unweight.hip -- hit-or-miss unweighting and order-preserving compaction on the device.
//
// Reference behaviour: accept iff u * wMax < w with u from a disjoint stream
// (cuPythia/pipeline/generate.cu:56-62, cuPythia/10_cub_compaction/cub_compaction.cu:36-46;
// Pythia ProcessContainer.cc:437), then a stable filter of the accepted candidates
// (cub::DeviceSelect::Flagged semantics, cub_compaction.cu:76-80, generate.cu:98-103).
//
// Six launches per chunk of candidates, no spin-waits or atomics; the integrand runs once per candidate
// and once more per accepted event:
//   eval:        the integrand of every candidate; its total weight is stored (FP32 or FP64, 4 or
//                8 B per candidate) and the weighted-estimate partials are written per block;
//   flags:       the accept uniform (Philox stream 1) against the stored weight: a 64-bit
//                wave ballot per 64 candidates; then the accepted count of each eval block from
//                its ballots;
//   scan:        exclusive scan of the block counts (one block);
//   compact:     the candidate index of every accepted event at its global rank (the block's
//                ballots as a bitset in candidate order, popcounts scanned over the block);
//   materialize: one thread per accepted event re-evaluates its candidate (same arithmetic, so the
//                same weights) for the kinematics and the process choice, and writes the SoA.
// Block b owns candidates [b*256*ppt, (b+1)*256*ppt) and walks them in 256-wide steps, so the
// global rank order is the candidate order.
#include "mi_kernel_common.h"

namespace mi {

namespace {

struct Cand {
    double w;          // total weight (mb), the FP32 value promoted exactly in FP32 mode
    double sH, tH, x1, x2;
    bool acc;
    int pid;
};

// Process choice from the 24-bit uniform of word 2 of the accept stream (Philox stream 1)
// against the cumulative FP32/FP64 weights; the accept decision from words 0, 1.
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

// KIND: kPartonic (partonic gg -> gg), or for the hadronic integrand the bits kVegas (VEGAS map),
// kHeavy (heavy-quark codes selected), kGrid (PDF grid instead of the toy PDFs).
constexpr int kVegas = 1, kHeavy = 2, kGrid = 4, kPartonic = 8;
constexpr bool kind_vegas(int kind) { return kind != kPartonic && (kind & kVegas); }
constexpr bool kind_heavy(int kind) { return kind != kPartonic && (kind & kHeavy); }
constexpr bool kind_grid(int kind) { return kind != kPartonic && (kind & kGrid); }

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

// Partonic gg -> gg: the cos(theta) uniform (Philox stream 0; compat: the reference kernel-10
// stream, ctr = seed + e * 0x100000001B3, first draw) and the accept uniform (stream 1; compat:
// the second draw).
__device__ __forceinline__ double c1_u(uint64_t key, uint64_t e, bool compat)
{
    if (compat) return u01(splitmix64(key + e * 0x100000001B3ULL));
    const U4 r = philox_event(key, e, 0u);
    return u53(r.x, r.y);
}
__device__ __forceinline__ double c1_uacc(uint64_t key, uint64_t e, bool compat)
{
    if (compat) return u01(splitmix64(key + e * 0x100000001B3ULL + 1));
    const U4 a = philox_event(key, e, 1u);
    return u53(a.x, a.y);
}

// Total weight (and per-process weights) of candidate e.
template <int KIND, typename T>
__device__ __forceinline__ T eval_cand(const Params<KIND, T> &P, const T *stab, uint64_t e, T *ws)
{
    if (KIND == kPartonic) {
        const double w = c1_weight(P.pref, P.cMax, c1_u(P.key, e, P.compat != 0));
#pragma unroll
        for (int i = 0; i < kMaxProc; ++i) ws[i] = (T)0;
        ws[0] = (T)w;
        return (T)w;
    }
    int bins[3];
    return had_candidate<T, kind_vegas(KIND), kind_heavy(KIND)>(P.k, P.mask, P.key, e, stab, P.nbins, (T)P.nbins, ws,
                                                                bins, (T *)nullptr, (T *)nullptr, (T *)nullptr,
                                                                (T *)nullptr, &P.gpdf, kind_grid(KIND));
}

// The full candidate for an accepted event: weights, kinematics of the chosen process's family
// (each family has its own map at the same uniforms), accept flag and process code.
template <int KIND, typename T>
__device__ __forceinline__ void full_cand(const Params<KIND, T> &P, const T *stab, uint64_t e, Cand &c)
{
    if (KIND == kPartonic) {
        const bool compat = P.compat != 0;
        const double u = c1_u(P.key, e, compat);
        const double w = c1_weight(P.pref, P.cMax, u);
        const double cth = P.cMax * (2.0 * u - 1.0);
        c.w = w;
        c.sH = P.sHat; c.tH = -0.5 * P.sHat * (1.0 - cth); c.x1 = 1.0; c.x2 = 1.0;
        c.acc = c1_uacc(P.key, e, compat) * P.wmax < w;
        c.pid = 111;
        return;
    }
    constexpr bool HEAVY = kind_heavy(KIND);
    T ws[kMaxProc];
    int bins[3];
    constexpr int NF = HEAVY ? 3 : 1;
    T sH[NF], tH[NF], x1[NF], x2[NF];
    const T tot = had_candidate<T, kind_vegas(KIND), HEAVY>(P.k, P.mask, P.key, e, stab, P.nbins, (T)P.nbins, ws,
                                                            bins, sH, tH, x1, x2, &P.gpdf, kind_grid(KIND));
    c.w = (double)tot;
    decide<T>(P.key, e, tot, ws, P.mask, P.wmax, c);
    int f = 0;
    if (HEAVY) {
        f = c.pid >= 123 ? 2 : (c.pid >= 121 ? 1 : 0);
        if (c.pid == 0) f = (P.mask & kFamMask[0]) ? 0 : ((P.mask & kFamMask[1]) ? 1 : 2);
    }
    c.sH = (double)pick_fam<T, NF>(sH, f); c.tH = (double)pick_fam<T, NF>(tH, f);
    c.x1 = (double)pick_fam<T, NF>(x1, f); c.x2 = (double)pick_fam<T, NF>(x2, f);
}

template <int KIND, typename T>
__device__ __forceinline__ void stage_tables(const Params<KIND, T> &P, T *stab)
{
    if (kind_vegas(KIND)) {
        for (int i = threadIdx.x; i < 9 * P.nbins; i += blockDim.x) stab[i] = P.gtab[i];
        __syncthreads();
    }
}

// Per-256-thread-group record of a block of NT threads: group g's four waves fold into
// partials[vb] (vb = blockIdx.x * NT / 256 + g) when live, in the order block_record uses.
template <int NT>
__device__ __forceinline__ void group_record(const double *v, double *partials, uint32_t vb, bool live)
{
    __shared__ double red[NT / 64][kRec];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kRec; ++i) {
        const double r = rec_is_max(i) ? wave_max(v[i]) : wave_sum(v[i]);
        if (lane == 0) red[wave][i] = r;
    }
    __syncthreads();
    const int t = threadIdx.x & (kBlock - 1), w0 = (threadIdx.x / kBlock) * (kBlock / 64);
    if (t < kRec && live) {
        const bool isMax = rec_is_max(t);
        double r = red[w0][t];
        for (int w = 1; w < kBlock / 64; ++w) r = isMax ? fmax(r, red[w0 + w][t]) : r + red[w0 + w][t];
        partials[(size_t)vb * kRec + t] = r;
    }
}

// eval: the integrand of every candidate, the stored total (wout may be null: estimate and
// maximum only) and one kRec record per 256-candidate-wide block of the layout ("virtual block"
// vb < nvb: candidates [vb*256*ppt, (vb+1)*256*ppt)).  NT = 256: one virtual block per block.
// WIN (FP32 grid): NT = 512 threads run two virtual blocks and first copy the window of the PDF
// grid's node values that P.gpdf describes into their dynamic LDS at P.gpdf.woff.
template <int KIND, typename T, int NT = kBlock, bool WIN = false>
__global__ __launch_bounds__(NT) void k_unw_eval(Params<KIND, T> P, T *__restrict__ wout,
                                                 double *__restrict__ partials, uint32_t nvb)
{
    static_assert(!WIN || (kind_grid(KIND) && sizeof(T) == 4), "the LDS window is for FP32 grid candidates");
    unsigned char *smem = mi_lds;
    T *stab = reinterpret_cast<T *>(smem);
    if (kind_vegas(KIND))
        for (int i = threadIdx.x; i < 9 * P.nbins; i += NT) stab[i] = P.gtab[i];
    if (WIN) {
        float4 *dst = reinterpret_cast<float4 *>(smem + P.gpdf.woff);
        const int rowq = P.gpdf.wstride * (kGridNF / 4);
        const int total = (P.gpdf.wq1 - P.gpdf.wq0 + 1) * rowq;
        for (int t = threadIdx.x; t < total; t += NT) {
            const int r = t / rowq, c = t - r * rowq;
            dst[t] = reinterpret_cast<const float4 *>(
                P.gpdf.val + ((size_t)(P.gpdf.wq0 + r) * P.gpdf.nx + P.gpdf.wx0) * kGridNF)[c];
        }
    }
    if (kind_vegas(KIND) || WIN) __syncthreads();
    const uint32_t vb = blockIdx.x * (NT / kBlock) + threadIdx.x / kBlock;
    const bool live = vb < nvb;
    const int tid = threadIdx.x & (kBlock - 1);
    constexpr int NP = (KIND == kPartonic) ? 1 : (kind_heavy(KIND) ? kMaxProc : kLightProc);
    double v[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) v[i] = 0.0;
    MaxOf<T> mx[NP + 1];
#pragma unroll
    for (int i = 0; i <= NP; ++i) mx[i] = 0;
    uint32_t nnz = 0;
    const uint64_t base = (uint64_t)vb * kBlock * P.ppt;
    for (uint32_t j = 0; live && j < P.ppt; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + tid;
        if (i >= P.n) break;
        T ws[kMaxProc];
        const T tot = eval_cand<KIND, T>(P, stab, P.ctr0 + i, ws);
        if (wout) wout[i] = tot;
        const double w = (double)tot;
        v[kRS] += w;
        v[kRQ] = fma(w, w, v[kRQ]);
        mx[0] = run_max(mx[0], tot);
        nnz += (tot > (T)0) ? 1u : 0u;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const double d = (double)ws[p];
            v[kRS + 1 + p] += d;
            v[kRQ + 1 + p] = fma(d, d, v[kRQ + 1 + p]);
            mx[1 + p] = run_max(mx[1 + p], ws[p]);
        }
    }
#pragma unroll
    for (int i = 0; i <= NP; ++i) v[kRM + i] = max_value<T>(mx[i]);
    v[kRNz] = (double)nnz;
    group_record<NT>(v, partials, vb, live);
}

// flags: accept decisions from the stored totals over the padded range of nwords x 64 candidates
// (blocks x ppt x 256 of the eval layout; candidates at or beyond n are rejected), so that
// masks[k] is the ballot of candidates 64 k ... 64 k + 63: the bitset of the whole chunk in
// candidate order.  Each wave takes 4 consecutive words: its 4 loads are issued before the
// Philox blocks, and lanes 0..3 store the 4 ballots.
template <int KIND, typename T>
__global__ __launch_bounds__(kBlock) void k_unw_flags(Params<KIND, T> P, const T *__restrict__ w, uint64_t nwords,
                                                      uint64_t *__restrict__ masks)
{
    const int lane = threadIdx.x & 63;
    const uint64_t word0 = ((uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 4;
    if (word0 >= nwords) return;                       // whole waves: nwords is a multiple of 4
    T wv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint64_t i = (word0 + r) * 64 + lane;
        wv[r] = (i < P.n) ? w[i] : (T)0;
    }
    uint64_t mine = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint64_t i = (word0 + r) * 64 + lane;
        bool acc = false;
        if (i < P.n) {
            const uint64_t e = P.ctr0 + i;
            double u;
            if (KIND == kPartonic) {
                u = c1_uacc(P.key, e, P.compat != 0);
            } else {
                const U4 a = philox_event(P.key, e, 1u);
                u = u53(a.x, a.y);
            }
            acc = u * P.wmax < (double)wv[r];
        }
        const uint64_t m = __ballot(acc);
        mine = (lane == r) ? m : mine;
    }
    if (lane < 4) masks[word0 + lane] = mine;
}

// counts[b] = accepted candidates of eval block b (its 4 ppt mask words), one block each.
__global__ __launch_bounds__(kBlock) void k_unw_count(const uint64_t *__restrict__ masks, uint32_t ppt,
                                                      uint64_t *__restrict__ counts)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t words = ppt * (kBlock / 64);
    const uint64_t *mb = masks + (size_t)blockIdx.x * words;
    uint32_t c = 0;
    for (uint32_t k = threadIdx.x; k < words; k += kBlock) c += (uint32_t)__popcll(mb[k]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int q = 0; q < kBlock / 64; ++q) t += wsum[q];
        counts[blockIdx.x] = t;
    }
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

// compact: the candidate index of every accepted event at rank base + offsets[b] + (rank in the
// block), written if below cap.  The block's ballots, in mask order, are a bitset over its
// candidates in candidate order (word k = candidates 64k .. 64k + 63 of the block); thread t
// takes W consecutive words, the block scans the threads' popcounts, and each thread writes the
// set bits of its words in order.
__global__ __launch_bounds__(kBlock) void k_unw_compact(const uint64_t *__restrict__ masks,
                                                        const uint64_t *__restrict__ offsets, uint64_t ctr0,
                                                        uint32_t ppt, uint64_t base, uint64_t cap,
                                                        uint64_t *__restrict__ index)
{
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t words = ppt * (kBlock / 64);
    const uint32_t W = (words + kBlock - 1) / kBlock;
    const uint32_t w0 = min(words, threadIdx.x * W), w1 = min(words, w0 + W);
    const uint64_t *mb = masks + (size_t)blockIdx.x * words;
    uint32_t c = 0;
    for (uint32_t k = w0; k < w1; ++k) c += (uint32_t)__popcll(mb[k]);
    // exclusive scan of c over the block: within the wave, then the wave totals
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = incl - c;
    for (int q = 0; q < wave; ++q) before += wsum[q];
    uint64_t pos = base + offsets[blockIdx.x] + before;
    const uint64_t first = ctr0 + (uint64_t)blockIdx.x * kBlock * ppt;
    for (uint32_t k = w0; k < w1; ++k) {
        uint64_t m = mb[k];
        while (m) {
            const int bit = __ffsll((unsigned long long)m) - 1;
            if (pos < cap) index[pos] = first + 64ull * k + (uint64_t)bit;
            ++pos;
            m &= m - 1;
        }
    }
}

// materialize: events [base, min(base + *chunk_total, cap)) from their candidate indices.
template <int KIND, typename T>
__global__ __launch_bounds__(kBlock) void k_unw_materialize(Params<KIND, T> P, const uint64_t *__restrict__ chunk_total,
                                                            uint64_t base, mi_events out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *stab = reinterpret_cast<T *>(smem);
    stage_tables<KIND, T>(P, stab);
    const uint64_t hi = min(base + *chunk_total, out.capacity);
    for (uint64_t j = base + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < hi;
         j += (uint64_t)gridDim.x * blockDim.x) {
        Cand c;
        full_cand<KIND, T>(P, stab, out.index[j], c);
        out.sHat[j] = c.sH;
        out.tHat[j] = c.tH;
        out.x1[j] = c.x1;
        out.x2[j] = c.x2;
        out.weight[j] = fmax(c.w, P.wmax);
        out.pid[j] = c.pid;
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

enum class Stage { Eval, Flags, Materialize };

template <int KIND, typename T>
hipError_t stage_t(hipStream_t s, Stage st, const UnwLaunch &L, const HadConst<T> &k, const T *tab)
{
    const Params<KIND, T> P = make_params<KIND, T>(L, k, tab);
    const Geom g = L.partonic ? L.c.g : L.h.g;
    const size_t sm = kind_vegas(KIND) ? sizeof(T) * 9 * P.nbins : 0;
    switch (st) {
    case Stage::Eval: {
        double *parts = L.partonic ? L.c.partials : L.h.partials;
        if constexpr (kind_grid(KIND) && sizeof(T) == 4) {
            if (L.win) {
                Params<KIND, T> PW = P;
                PW.gpdf.woff = L.wgrid.woff; PW.gpdf.wq0 = L.wgrid.wq0; PW.gpdf.wq1 = L.wgrid.wq1;
                PW.gpdf.wx0 = L.wgrid.wx0; PW.gpdf.wstride = L.wgrid.wstride;
                auto kern = k_unw_eval<KIND, T, kWinBlock, true>;
                (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.win_smem);
                (void)hipGetLastError();
                constexpr uint32_t per = kWinBlock / kBlock;
                hipLaunchKernelGGL(kern, dim3((g.blocks + per - 1) / per), dim3(kWinBlock), L.win_smem, s, PW,
                                   (T *)L.w, parts, g.blocks);
                break;
            }
        }
        hipLaunchKernelGGL((k_unw_eval<KIND, T>), dim3(g.blocks), dim3(kBlock), sm, s, P, (T *)L.w, parts, g.blocks);
        break;
    }
    case Stage::Flags: {
        const uint64_t nwords = (uint64_t)g.blocks * g.ppt * (kBlock / 64);     // a multiple of 4
        const uint64_t fb = (nwords + 4 * (kBlock / 64) - 1) / (4 * (kBlock / 64));
        hipLaunchKernelGGL((k_unw_flags<KIND, T>), dim3((uint32_t)fb), dim3(kBlock), 0, s, P, (const T *)L.w, nwords,
                           L.masks);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_unw_count, dim3(g.blocks), dim3(kBlock), 0, s, L.masks, g.ppt, L.counts);
        break;
    }
    case Stage::Materialize:
        hipLaunchKernelGGL((k_unw_materialize<KIND, T>), dim3(L.mat_blocks), dim3(kBlock), sm, s, P,
                           L.counts + g.blocks, L.base, *L.out);
        break;
    }
    return hipGetLastError();
}

// The hadronic KIND of a launch.
inline int had_kind(const UnwLaunch &L, bool grid)
{
    return (L.h.vegas ? kVegas : 0) | (has_heavy(L.h.mask) ? kHeavy : 0) | (grid ? kGrid : 0);
}

template <typename T>
hipError_t stage_h(hipStream_t s, Stage st, const UnwLaunch &L, const HadConst<T> &k, const T *tab)
{
    switch (had_kind(L, grid_of(L.h, (T *)nullptr).active != 0)) {
    case 0: return stage_t<0, T>(s, st, L, k, nullptr);
    case 1: return stage_t<1, T>(s, st, L, k, tab);
    case 2: return stage_t<2, T>(s, st, L, k, nullptr);
    case 3: return stage_t<3, T>(s, st, L, k, tab);
    case 4: return stage_t<4, T>(s, st, L, k, nullptr);
    case 5: return stage_t<5, T>(s, st, L, k, tab);
    case 6: return stage_t<6, T>(s, st, L, k, nullptr);
    default: return stage_t<7, T>(s, st, L, k, tab);
    }
}

hipError_t stage(hipStream_t s, Stage st, const UnwLaunch &L)
{
    if (L.partonic) {
        HadConst<double> kd{};
        return stage_t<kPartonic, double>(s, st, L, kd, nullptr);
    }
    if (L.h.prec == MI_PREC_FP32) return stage_h<float>(s, st, L, L.h.kf, L.h.vegas ? L.h.grid->tab32 : nullptr);
    return stage_h<double>(s, st, L, L.h.kd, L.h.vegas ? L.h.grid->tab64 : nullptr);
}

template <int KIND, typename T>
int occ_eval_t(size_t sm)
{
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_unw_eval<KIND, T>, kBlock, sm);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

template <typename T>
int occ_eval_h(const UnwLaunch &L)
{
    const size_t sm = L.h.vegas ? sizeof(T) * 9 * L.h.grid->nbins : 0;
    switch (had_kind(L, grid_of(L.h, (T *)nullptr).active != 0)) {
    case 0: return occ_eval_t<0, T>(sm);
    case 1: return occ_eval_t<1, T>(sm);
    case 2: return occ_eval_t<2, T>(sm);
    case 3: return occ_eval_t<3, T>(sm);
    case 4: return occ_eval_t<4, T>(sm);
    case 5: return occ_eval_t<5, T>(sm);
    case 6: return occ_eval_t<6, T>(sm);
    default: return occ_eval_t<7, T>(sm);
    }
}

}  // namespace

// Resident blocks per CU of the eval kernel the launch will use (virtual 256-thread blocks).
int occ_unweight_eval(const UnwLaunch &L)
{
    if (L.partonic) return occ_eval_t<kPartonic, double>(0);
    if (L.win) return kWinBlock / kBlock;                // one window block per CU
    return L.h.prec == MI_PREC_FP32 ? occ_eval_h<float>(L) : occ_eval_h<double>(L);
}

hipError_t launch_unweight_eval(hipStream_t s, const UnwLaunch &L) { return stage(s, Stage::Eval, L); }
hipError_t launch_unweight_flags(hipStream_t s, const UnwLaunch &L) { return stage(s, Stage::Flags, L); }
hipError_t launch_unweight_materialize(hipStream_t s, const UnwLaunch &L) { return stage(s, Stage::Materialize, L); }

hipError_t launch_scan_counts(hipStream_t s, uint64_t *counts, uint32_t nblk)
{
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, nblk);
    return hipGetLastError();
}

hipError_t launch_unweight_compact(hipStream_t s, const UnwLaunch &L)
{
    const Geom g = L.partonic ? L.c.g : L.h.g;
    const uint64_t ctr0 = L.partonic ? L.c.ctr0 : L.h.ctr0;
    hipLaunchKernelGGL(k_unw_compact, dim3(g.blocks), dim3(kBlock), 0, s, L.masks, L.counts, ctr0, g.ppt, L.base,
                       L.out->capacity, L.out->index);
    return hipGetLastError();
}

}  // namespace mi
