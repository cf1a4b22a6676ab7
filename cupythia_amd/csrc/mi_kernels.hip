// mi_kernels.hip -- gfx950 kernels of the integration path: Philox fill, batched sigmaHat,
// the fused hadronic / partonic integration kernels (RNG in registers -> map -> PDFs x
// sigmaHat -> FP64 sum w, sum w^2, max), per-point dumps, deterministic reductions and the
// on-device VEGAS refinement.
//
// Reduction scheme (deterministic, no floating-point atomics except the per-wave LDS VEGAS
// histogram): each thread accumulates its points sequentially in FP64; the 64 lanes of a wave
// combine by an xor butterfly of __shfl_xor; the 4 waves of a block combine in wave order
// through LDS; one record per block goes to global memory; a single-block kernel folds the
// block records in a fixed order.  The result is bit-identical from run to run.
#include "mi_kernel_common.h"
#include <type_traits>

namespace mi {

// ------------------------------------------------------------------------------------ Philox fill
__global__ __launch_bounds__(kBlock) void k_philox_fill(uint64_t key, uint64_t ctr0, uint32_t sid,
                                                        uint64_t n, uint32_t *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const U4 r = philox_event(key, ctr0 + i, sid);
    uint4 v; v.x = r.x; v.y = r.y; v.z = r.z; v.w = r.w;
    reinterpret_cast<uint4 *>(out)[i] = v;
}

hipError_t launch_philox_fill(hipStream_t s, uint64_t key, uint64_t ctr0, uint32_t sid, uint64_t n,
                              uint32_t *out)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_philox_fill, dim3((uint32_t)blocks), dim3(kBlock), 0, s, key, ctr0, sid, n, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ sigmaHat
// Kfactor * sigmaHatWrap for one code, FP64.  Brackets in terms of the ratios t/s, u/s with
// three reciprocals; algebraically identical to SigmaQCD.cc:115-378.  The heavy pairs 121..124
// (Sigma2gg2QQbar, Sigma2qqbar2QQbar::sigmaKin with m3 = m4) take m^2 = (s + t + u) / 2 from
// the three invariants and Pythia's tHQ = -(s - t + u)/2, uHQ = -(s + t - u)/2.
__global__ __launch_bounds__(kBlock) void k_sigma_hat(int code, uint64_t n, const double *__restrict__ sH,
                                                      const double *__restrict__ tH, const double *__restrict__ uH,
                                                      const double *__restrict__ alpS,
                                                      const int32_t *__restrict__ id1,
                                                      const int32_t *__restrict__ id2, double nQN,
                                                      double kfac, double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = sH[i], t = tH[i];
    const double u = uH ? uH[i] : -(s + t);
    const double as = alpS[i];
    const double is = 1.0 / s, it = 1.0 / t, iu = 1.0 / u;
    const double ts = t * is, us = u * is, st = s * it, su = s * iu, tu = t * iu, ut = u * it;
    double br;
    switch (code) {
    case MI_GG2GG:
        br = 2.25 * (3.0 - ts * us - us * st * st - ts * su * su);
        break;
    case MI_GG2QQBAR:
        br = nQN * ((1.0 / 6.0) * (ut + tu) - 0.375 * (ts * ts + us * us));
        break;
    case MI_QG2QG:
        br = (1.0 + us * us) * (st * st - (4.0 / 9.0) * su);
        break;
    case MI_QQ2QQ: {
        const int a = id1 ? id1[i] : 1, b = id2 ? id2[i] : 2;
        const double sigT = (4.0 / 9.0) * (1.0 + us * us) * (st * st);
        if (b == a) {
            const double sigU = (4.0 / 9.0) * (1.0 + ts * ts) * (su * su);
            const double sigTU = -(8.0 / 27.0) * (st * su);
            br = 0.5 * (sigT + sigU + sigTU);
        } else if (b == -a) {
            br = sigT - (8.0 / 27.0) * (us * us) * st;
        } else {
            br = sigT;
        }
        break;
    }
    case MI_QQBAR2GG:
        br = (16.0 / 27.0) * (ut + tu) - (4.0 / 3.0) * (ts * ts + us * us);
        break;
    case MI_QQBAR2QQBARNEW:
        br = nQN * (4.0 / 9.0) * (ts * ts + us * us);
        break;
    case MI_GG2CCBAR:
    case MI_GG2BBBAR:
    case MI_QQBAR2CCBAR:
    case MI_QQBAR2BBBAR: {
        const double m2 = 0.5 * (s + t + u);
        const double tq = -0.5 * (s - t + u), uq = -0.5 * (s + t - u);
        const double ta = -tq * is, ub = -uq * is, mu = m2 * is;       // a, b, m^2 / s
        if (code == MI_GG2CCBAR || code == MI_GG2BBBAR) {
            const double ia = 1.0 / ta, ib = 1.0 / ub;
            const double sTS = ub * ia - 2.25 * ub * ub + mu * ia * ia * (ta * (4.5 * ub - 0.5 + mu) - 4.0 * mu);
            const double sUS = ta * ib - 2.25 * ta * ta + mu * ib * ib * (ub * (4.5 * ta - 0.5 + mu) - 4.0 * mu);
            br = (sTS + sUS) / 6.0;
        } else {
            br = (4.0 / 9.0) * (ta * ta + ub * ub + 2.0 * mu);
        }
        break;
    }
    default:
        br = 0.0;
    }
    out[i] = kfac * kConvert2mb * (kPi * is * is) * (as * as) * br;
}

hipError_t launch_sigma_hat(hipStream_t s, int code, uint64_t n, const double *sH, const double *tH,
                            const double *uH, const double *alpS, const int32_t *id1, const int32_t *id2,
                            int nQuarkNew, double kfactor, double *out)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sigma_hat, dim3((uint32_t)blocks), dim3(kBlock), 0, s, code, n, sH, tH, uH, alpS,
                       id1, id2, (double)nQuarkNew, kfactor, out);
    return hipGetLastError();
}


// ------------------------------------------------------------------------------------ hadronic integration
// Block b owns points [b*256*ppt, (b+1)*256*ppt); in step j thread t takes b*256*ppt + j*256 + t.
// GRID: parton densities from the PDF grid gpdf instead of the toy PDFs.  HEAVY: mask may
// select the heavy-quark codes 121..124 (their families are evaluated); otherwise only the
// massless codes are computed and accumulated.  NT threads per block.  WIN (FP32 grid): the
// block first copies the window of the grid's node values that gpdf describes (rows wq0..wq1,
// columns wx0..nx-1) into its dynamic LDS at gpdf.woff (after the VEGAS table and histograms),
// so the interpolation stencils read LDS instead of gathering from L2.  HW (FP32): the hardware
// transcendentals (MI_PREC_FP32_HW).  Only the slots of the processes mask selects are summed.
// One point of k_integrate_had: the candidate at counter e, its histogram entry (HIST) and, if
// live, its sums.  A function of its own rather than a lambda: a closure that refers to the
// kernel's GridPdf argument made the compiler copy that argument to scratch.
template <typename T, bool VEGAS, bool HIST, bool GRID, bool HEAVY, bool HW, int NP, typename MaxT>
__device__ __forceinline__ void integrate_point(const HadConst<T> &k, int mask, uint64_t key, uint64_t e, bool live,
                                                const T *stab, int nbins, T nbf, const GridPdf<T> *gp,
                                                unsigned char *hwave, double *sw, double *sw2, MaxT *mx, uint32_t &nnz)
{
    T ws[kMaxProc];
    int bins[3];
    const T tot = had_candidate<T, VEGAS, HEAVY, HW>(k, mask, key, e, stab, nbins, nbf, ws, bins, (T *)nullptr,
                                                     (T *)nullptr, (T *)nullptr, (T *)nullptr, gp, GRID);
    if (HIST) WaveHist<3>::add(live ? (double)tot * (double)tot : 0.0, bins, hwave, nbins);
    if (!live) return;
    const double dt = (double)tot;
    sw[0] += dt;
    sw2[0] = fma(dt, dt, sw2[0]);
    mx[0] = run_max(mx[0], tot);
    nnz += (tot > (T)0) ? 1u : 0u;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        if (!((mask >> p) & 1)) continue;                       // uniform: the slot stays 0
        // an empty volatile asm keeps this a branch: if-converted, each of the slot's updates
        // became a select per register (5 per slot, for every slot).  Not for the ten slots of
        // HEAVY, where the branches cost registers (115 -> 146 VGPRs)
        if (!HEAVY) __asm__ volatile("");
        const double d = (double)ws[p];
        sw[1 + p] += d;
        sw2[1 + p] = fma(d, d, sw2[1 + p]);
        mx[1 + p] = run_max(mx[1 + p], ws[p]);
    }
}

template <typename T, bool VEGAS, bool HIST, bool GRID, bool HEAVY, int NT = kBlock, bool WIN = false, bool HW = false>
__global__ __launch_bounds__(NT) void k_integrate_had(HadConst<T> k, int mask, uint64_t key,
                                                      uint64_t ctr0, uint64_t n, uint32_t ppt,
                                                      const T *__restrict__ gtab, int nbins,
                                                      double *__restrict__ partials,
                                                      double *__restrict__ hist, GridPdf<T> gpdf)
{
    static_assert(!WIN || (GRID && sizeof(T) == 4), "the LDS window is for FP32 grid integration");
    unsigned char *smem = mi_lds;
    T *stab = reinterpret_cast<T *>(smem);                                   // 3 dims x nbins x {lo, wd, 1-lo}
    const size_t tabBytes = (sizeof(T) * 9 * (VEGAS ? nbins : 0) + 7) & ~(size_t)7;
    const size_t areaBytes = WaveHist<3>::area_bytes(nbins);                 // per wave: its histogram
    unsigned char *harea = smem + tabBytes;
    const int wave = threadIdx.x >> 6;
    if (VEGAS) {
        for (int i = threadIdx.x; i < 9 * nbins; i += NT) stab[i] = gtab[i];
    }
    if (WIN) {
        float4 *dst = reinterpret_cast<float4 *>(smem + gpdf.woff);
        const int rowq = gpdf.wstride * (kGridNF / 4);                       // 16-byte words per row
        const int total = (gpdf.wq1 - gpdf.wq0 + 1) * rowq;
        for (int t = threadIdx.x; t < total; t += NT) {
            const int r = t / rowq, c = t - r * rowq;
            dst[t] = reinterpret_cast<const float4 *>(
                gpdf.val + ((size_t)(gpdf.wq0 + r) * gpdf.nx + gpdf.wx0) * kGridNF)[c];
        }
    }
    if (VEGAS || WIN) __syncthreads();
    const T nbf = (T)nbins;
    // Running maxima as the weights' own type: the weights are >= 0 (never NaN), so for FP32 the
    // maximum is an integer max of the bit patterns.  Converted to FP64 once at the end, which is
    // exact and order-preserving.  The nonzero count is an integer.
    using MaxT = typename std::conditional<sizeof(T) == 4, uint32_t, double>::type;
    // VEGAS adaptation (HIST) reads only the totals of its records: the per-process slots stay 0
    constexpr int NP = HIST ? 0 : (HEAVY ? kMaxProc : kLightProc);
    MaxT mx[NP + 1];
    double sw[NP + 1], sw2[NP + 1];                                          // sum w, sum w^2 per slot
#pragma unroll
    for (int i = 0; i <= NP; ++i) { mx[i] = 0; sw[i] = 0.0; sw2[i] = 0.0; }
    uint32_t nnz = 0;

    if (HIST) WaveHist<3>::init(harea + wave * areaBytes, nbins);
    const uint64_t base = (uint64_t)blockIdx.x * NT * ppt;
    if (!HIST && base + (uint64_t)NT * ppt <= n) {
        // every point of the block is live (all blocks but the last): a uniform loop, no tests
        for (uint32_t j = 0; j < ppt; ++j)
            integrate_point<T, VEGAS, HIST, GRID, HEAVY, HW, NP>(k, mask, key, ctr0 + base + (uint64_t)j * NT + threadIdx.x,
                                                                 true, stab, nbins, nbf, &gpdf, harea + wave * areaBytes,
                                                                 sw, sw2, mx, nnz);
    } else {
        for (uint32_t j = 0; j < ppt; ++j) {
            const uint64_t i = base + (uint64_t)j * NT + threadIdx.x;
            const bool live = i < n;
            // the histogram's batches need the whole wave: it leaves the loop together
            if (HIST ? __ballot(live) == 0 : !live) break;
            // a lane past the end (HIST only) evaluates the last point again and adds nothing
            integrate_point<T, VEGAS, HIST, GRID, HEAVY, HW, NP>(k, mask, key, ctr0 + (live ? i : n - 1), live, stab, nbins,
                                                                 nbf, &gpdf, harea + wave * areaBytes, sw, sw2, mx, nnz);
        }
    }
    double mxd[NP + 1];
#pragma unroll
    for (int i = 0; i <= NP; ++i) mxd[i] = max_value<T>(mx[i]);
    block_record_sums<NT, NP>(sw, sw2, mxd, (double)nnz, partials + (size_t)blockIdx.x * kRec);
    if (HIST) {
        __syncthreads();
        for (int c = threadIdx.x; c < 3 * nbins; c += NT) {
            double s = WaveHist<3>::hist(harea)[c];
            for (int w = 1; w < NT / 64; ++w) s += WaveHist<3>::hist(harea + w * areaBytes)[c];
            hist[(size_t)blockIdx.x * 3 * nbins + c] = s;
        }
    }
}

static size_t had_smem(const HadLaunch &L)
{
    if (!L.vegas) return 0;
    const size_t tb = ((L.prec == MI_PREC_FP32 ? sizeof(float) : sizeof(double)) * 9 * L.grid->nbins + 7) & ~(size_t)7;
    return tb + (L.hist ? (kBlock / 64) * WaveHist<3>::area_bytes(L.grid->nbins) : 0);
}

template <typename T, bool GRID, bool HEAVY, bool HW = false>
static hipError_t launch_had_tg(hipStream_t s, const HadLaunch &L, const HadConst<T> &k, const T *tab,
                                const GridPdf<T> &gp)
{
    const size_t sm = had_smem(L);
    const int nb = L.vegas ? L.grid->nbins : 0;
    dim3 grid(L.g.blocks), blk(kBlock);
    if (!L.vegas)
        hipLaunchKernelGGL((k_integrate_had<T, false, false, GRID, HEAVY, kBlock, false, HW>), grid, blk, sm, s, k,
                           L.mask, L.key, L.ctr0, L.n, L.g.ppt, tab, nb, L.partials, L.hist, gp);
    else if (!L.hist)
        hipLaunchKernelGGL((k_integrate_had<T, true, false, GRID, HEAVY, kBlock, false, HW>), grid, blk, sm, s, k,
                           L.mask, L.key, L.ctr0, L.n, L.g.ppt, tab, nb, L.partials, L.hist, gp);
    else
        hipLaunchKernelGGL((k_integrate_had<T, true, true, GRID, HEAVY, kBlock, false, HW>), grid, blk, sm, s, k,
                           L.mask, L.key, L.ctr0, L.n, L.g.ppt, tab, nb, L.partials, L.hist, gp);
    return hipGetLastError();
}

// FP32 grid integration with the LDS window: one kWinBlock-thread block per CU.
template <bool VEGAS, bool HIST, bool HEAVY, bool HW>
static hipError_t launch_had_win(hipStream_t s, const HadLaunch &L, const float *tab)
{
    auto kern = k_integrate_had<float, VEGAS, HIST, true, HEAVY, kWinBlock, true, HW>;
    // allows dynamic LDS above the default limit where the runtime asks for it; a runtime that
    // has no such limit may refuse the call, which the launch below then does not depend on
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)L.win_smem);
    (void)hipGetLastError();
    hipLaunchKernelGGL(kern, dim3(L.g.blocks), dim3(kWinBlock), L.win_smem, s, L.kf, L.mask, L.key, L.ctr0, L.n,
                       L.g.ppt, tab, VEGAS ? L.grid->nbins : 0, L.partials, L.hist, L.gf);
    return hipGetLastError();
}

template <typename T, bool HW = false>
static hipError_t launch_had_t(hipStream_t s, const HadLaunch &L, const HadConst<T> &k, const T *tab,
                               const GridPdf<T> &gp)
{
    if constexpr (sizeof(T) == 4) {
        if (L.win && gp.active) {
            const bool hv = has_heavy(L.mask);
            if (L.vegas && L.hist)
                return hv ? launch_had_win<true, true, true, HW>(s, L, tab) : launch_had_win<true, true, false, HW>(s, L, tab);
            if (L.vegas)
                return hv ? launch_had_win<true, false, true, HW>(s, L, tab) : launch_had_win<true, false, false, HW>(s, L, tab);
            return hv ? launch_had_win<false, false, true, HW>(s, L, tab) : launch_had_win<false, false, false, HW>(s, L, tab);
        }
    }
    if (has_heavy(L.mask))
        return gp.active ? launch_had_tg<T, true, true, HW>(s, L, k, tab, gp) : launch_had_tg<T, false, true, HW>(s, L, k, tab, gp);
    return gp.active ? launch_had_tg<T, true, false, HW>(s, L, k, tab, gp) : launch_had_tg<T, false, false, HW>(s, L, k, tab, gp);
}

hipError_t launch_integrate_had(hipStream_t s, const HadLaunch &L)
{
    if (L.prec == MI_PREC_FP32)
        return L.hw ? launch_had_t<float, true>(s, L, L.kf, L.vegas ? L.grid->tab32 : nullptr, L.gf)
                    : launch_had_t<float, false>(s, L, L.kf, L.vegas ? L.grid->tab32 : nullptr, L.gf);
    return launch_had_t<double>(s, L, L.kd, L.vegas ? L.grid->tab64 : nullptr, L.gd);
}

// ------------------------------------------------------------------------------------ partonic integration
template <bool VEGAS, bool HIST>
__global__ __launch_bounds__(kBlock) void k_integrate_c1(double pref, double cMax, uint64_t key, uint64_t ctr0,
                                                         uint64_t n, uint32_t ppt, const double *__restrict__ gtab,
                                                         int nbins, double *__restrict__ partials,
                                                         double *__restrict__ hist)
{
    extern __shared__ __align__(16) unsigned char smem[];
    double *stab = reinterpret_cast<double *>(smem);
    unsigned char *harea = smem + sizeof(double) * 3 * (VEGAS ? nbins : 0);
    const size_t areaBytes = WaveHist<1>::area_bytes(nbins);
    const int wave = threadIdx.x >> 6;
    if (VEGAS) {
        for (int i = threadIdx.x; i < 3 * nbins; i += blockDim.x) stab[i] = gtab[i];
        __syncthreads();
    }
    if (HIST) WaveHist<1>::init(harea + wave * areaBytes, nbins);
    double v[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) v[i] = 0.0;
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * ppt;
    for (uint32_t j = 0; j < ppt; ++j) {
        const uint64_t i = base + (uint64_t)j * kBlock + threadIdx.x;
        const bool live = i < n;
        if (HIST ? __ballot(live) == 0 : !live) break;
        // a lane past the end (HIST only) evaluates the last point again and adds nothing
        int bin;
        const double w = c1_candidate(pref, cMax, key, ctr0 + (live ? i : n - 1), VEGAS, stab, nbins, &bin);
        if (HIST) WaveHist<1>::add(live ? w * w : 0.0, &bin, harea + wave * areaBytes, nbins);
        if (!live) continue;
        v[kRS] += w;
        v[kRQ] = fma(w, w, v[kRQ]);
        v[kRM] = fmax(v[kRM], w);
        v[kRNz] += (w > 0.0) ? 1.0 : 0.0;
    }
    // gg -> gg is the only process: copy total into slot of code 111
    v[kRS + 1] = v[kRS]; v[kRQ + 1] = v[kRQ]; v[kRM + 1] = v[kRM];
    block_record(v, partials + (size_t)blockIdx.x * kRec);
    if (HIST) {
        __syncthreads();
        for (int c = threadIdx.x; c < nbins; c += blockDim.x) {
            double s = WaveHist<1>::hist(harea)[c];
            for (int w = 1; w < kBlock / 64; ++w) s += WaveHist<1>::hist(harea + w * areaBytes)[c];
            hist[(size_t)blockIdx.x * nbins + c] = s;
        }
    }
}

hipError_t launch_integrate_c1(hipStream_t s, const C1Launch &L)
{
    const int nb = L.vegas ? L.grid->nbins : 0;
    const size_t sm = L.vegas ? sizeof(double) * 3 * nb + (L.hist ? (kBlock / 64) * WaveHist<1>::area_bytes(nb) : 0) : 0;
    const double *tab = L.vegas ? L.grid->tab64 : nullptr;
    dim3 grid(L.g.blocks), blk(kBlock);
    if (!L.vegas)
        hipLaunchKernelGGL((k_integrate_c1<false, false>), grid, blk, sm, s, L.pref, L.cMax, L.key, L.ctr0, L.n,
                           L.g.ppt, tab, nb, L.partials, L.hist);
    else if (!L.hist)
        hipLaunchKernelGGL((k_integrate_c1<true, false>), grid, blk, sm, s, L.pref, L.cMax, L.key, L.ctr0, L.n,
                           L.g.ppt, tab, nb, L.partials, L.hist);
    else
        hipLaunchKernelGGL((k_integrate_c1<true, true>), grid, blk, sm, s, L.pref, L.cMax, L.key, L.ctr0, L.n,
                           L.g.ppt, tab, nb, L.partials, L.hist);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ per-point dumps
// Kinematics of the first family that mask selects (massless, c cbar, b bbar).
template <typename T, bool VEGAS, bool HEAVY, bool HW = false>
__global__ __launch_bounds__(kBlock) void k_points_had(HadConst<T> k, int mask, uint64_t key, uint64_t ctr0,
                                                       uint64_t n, const T *__restrict__ gtab, int nbins,
                                                       double *__restrict__ w, double *__restrict__ kin,
                                                       GridPdf<T> gpdf)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T ws[kMaxProc];
    int bins[3];
    constexpr int NF = HEAVY ? 3 : 1;
    T sH[NF], tH[NF], x1[NF], x2[NF];
    const T tot = had_candidate<T, VEGAS, HEAVY, HW>(k, mask, key, ctr0 + i, gtab, nbins, (T)nbins, ws, bins, sH,
                                                     tH, x1, x2, &gpdf, gpdf.active);
    w[i * kRecN] = (double)tot;
#pragma unroll
    for (int p = 0; p < kMaxProc; ++p) w[i * kRecN + 1 + p] = (double)ws[p];
    if (kin) {
        const int f = !HEAVY ? 0 : ((mask & kFamMask[0]) ? 0 : ((mask & kFamMask[1]) ? 1 : 2));
        kin[4 * i + 0] = (double)pick_fam<T, NF>(sH, f); kin[4 * i + 1] = (double)pick_fam<T, NF>(tH, f);
        kin[4 * i + 2] = (double)pick_fam<T, NF>(x1, f); kin[4 * i + 3] = (double)pick_fam<T, NF>(x2, f);
    }
}

template <typename T, bool HEAVY, bool HW = false>
static void launch_points_had_t(hipStream_t s, const HadLaunch &L, const HadConst<T> &k, const T *tab,
                                const GridPdf<T> &gp, double *w, double *kin)
{
    dim3 grid((uint32_t)((L.n + kBlock - 1) / kBlock)), blk(kBlock);
    const int nb = L.vegas ? L.grid->nbins : 0;
    if (L.vegas) hipLaunchKernelGGL((k_points_had<T, true, HEAVY, HW>), grid, blk, 0, s, k, L.mask, L.key, L.ctr0, L.n, tab, nb, w, kin, gp);
    else         hipLaunchKernelGGL((k_points_had<T, false, HEAVY, HW>), grid, blk, 0, s, k, L.mask, L.key, L.ctr0, L.n, tab, nb, w, kin, gp);
}

hipError_t launch_points_had(hipStream_t s, const HadLaunch &L, double *w, double *kin)
{
    if (L.n == 0) return hipSuccess;
    const bool hv = has_heavy(L.mask);
    if (L.prec == MI_PREC_FP32) {
        const float *tab = L.vegas ? L.grid->tab32 : nullptr;
        if (L.hw) {
            if (hv) launch_points_had_t<float, true, true>(s, L, L.kf, tab, L.gf, w, kin);
            else    launch_points_had_t<float, false, true>(s, L, L.kf, tab, L.gf, w, kin);
        } else {
            if (hv) launch_points_had_t<float, true>(s, L, L.kf, tab, L.gf, w, kin);
            else    launch_points_had_t<float, false>(s, L, L.kf, tab, L.gf, w, kin);
        }
    } else {
        const double *tab = L.vegas ? L.grid->tab64 : nullptr;
        if (hv) launch_points_had_t<double, true>(s, L, L.kd, tab, L.gd, w, kin);
        else    launch_points_had_t<double, false>(s, L, L.kd, tab, L.gd, w, kin);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(kBlock) void k_points_c1(double pref, double cMax, double sHat, uint64_t key,
                                                      uint64_t ctr0, uint64_t n, int vegas,
                                                      const double *__restrict__ gtab, int nbins,
                                                      double *__restrict__ w, double *__restrict__ kin)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int bin;
    const double wt = c1_candidate(pref, cMax, key, ctr0 + i, vegas != 0, gtab, nbins, &bin);
    w[i * kRecN] = wt;
    w[i * kRecN + 1] = wt;
    for (int p = 1; p < kMaxProc; ++p) w[i * kRecN + 1 + p] = 0.0;
    if (kin) {
        // recompute c for the kinematics record
        const U4 r = philox_event(key, ctr0 + i, 0u);
        const uint64_t x53 = ((((uint64_t)r.y) << 32) | r.x) >> 11;
        double u = (double)x53 * (1.0 / 9007199254740992.0);
        if (vegas) {
            const uint64_t kb = x53 * (uint64_t)nbins;
            const int b = (int)(kb >> 53);
            const double frac = (double)(kb & ((1ULL << 53) - 1)) * (1.0 / 9007199254740992.0);
            u = fma(frac, gtab[3 * b + 1], gtab[3 * b]);
        }
        const double c = cMax * (2.0 * u - 1.0);
        kin[4 * i + 0] = sHat; kin[4 * i + 1] = -0.5 * sHat * (1.0 - c);
        kin[4 * i + 2] = 1.0; kin[4 * i + 3] = 1.0;
    }
}

hipError_t launch_points_c1(hipStream_t s, const C1Launch &L, double sHat, double *w, double *kin)
{
    if (L.n == 0) return hipSuccess;
    dim3 grid((uint32_t)((L.n + kBlock - 1) / kBlock)), blk(kBlock);
    hipLaunchKernelGGL(k_points_c1, grid, blk, 0, s, L.pref, L.cMax, sHat, L.key, L.ctr0, L.n, L.vegas ? 1 : 0,
                       L.vegas ? L.grid->tab64 : nullptr, L.vegas ? L.grid->nbins : 0, w, kin);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ reductions
// Column c of an nrow x stride row-major matrix folded in a fixed order: thread t takes rows
// t, t+256, ... sequentially, then a fixed LDS tree.  Sum, or max for the max slots.  One block
// per column, so the columns fold in parallel (each one's order is that of a single block).
__device__ __forceinline__ void fold_column(const double *__restrict__ parts, uint32_t nrow, uint32_t stride,
                                            uint32_t c, bool isMax, double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nrow; b += kBlock) {
        const double x = parts[(size_t)b * stride + c];
        acc = isMax ? fmax(acc, x) : acc + x;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = kBlock / 2; off >= 1; off >>= 1) {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] = isMax ? fmax(sh[threadIdx.x], sh[threadIdx.x + off]) : sh[threadIdx.x] + sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Fold nblk block records into one: block i folds slot i.
__global__ __launch_bounds__(kBlock) void k_reduce_records(const double *__restrict__ parts, uint32_t nblk,
                                                           double *__restrict__ out)
{
    const uint32_t i = blockIdx.x;
    fold_column(parts, nblk, kRec, i, rec_is_max((int)i), out + i);
}

hipError_t launch_reduce_records(hipStream_t s, const double *partials, uint32_t nblk, double *out)
{
    hipLaunchKernelGGL(k_reduce_records, dim3(kRec), dim3(kBlock), 0, s, partials, nblk, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ VEGAS refinement
// Lepage, J. Comput. Phys. 27 (1978) 192: per dimension, d_i = sum of w^2 in bin i, smoothed
// over neighbours (ends averaged with one neighbour), normalised, damped
// m_i = ((1 - d_i) / -ln d_i)^alpha, then the edges are moved so every new bin holds an equal
// share of sum m_i, interpolating linearly inside the old bins.  One block, one thread per bin:
// the smoothed total is a fixed-shape block sum, sum m an inclusive block scan C_i (fixed shape,
// so the result does not depend on thread timing), and new edge j takes the first old bin
// k < nbins - 1 with C_k >= j * C_last / nbins (else the last bin) by binary search over C: the
// bin that the sequential walk of the old edges stops in.  The per-bin tables {lo, width, 1 - lo}
// of the new edges (FP64 and FP32) are written as well.
static_assert(kBlock == kMaxBins, "the refinement runs one thread per bin in a kBlock-thread block");

__device__ __forceinline__ double block_sum_bins(double v, double *wpart)
{
    const double s = wave_sum(v);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) wpart[wave] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kMaxBins / 64; ++w) t += wpart[w];
    __syncthreads();
    return t;
}

// Refines one dimension: hd = its nbins histogram sums, xd = its nbins + 1 edges (updated in
// place), t64d / t32d = its 3 * nbins table entries.  Called by every thread of the block.
__device__ void refine_dim(const double *__restrict__ hd, double *__restrict__ xd, int nbins, double alpha,
                           double *__restrict__ t64d, float *__restrict__ t32d)
{
    __shared__ double h[kMaxBins], c[kMaxBins], x[kMaxBins + 1], ne[kMaxBins + 1], wpart[kMaxBins / 64];
    const int i = threadIdx.x, wave = i >> 6, lane = i & 63;
    if (i < nbins) h[i] = hd[i];
    for (int j = i; j <= nbins; j += kMaxBins) { x[j] = xd[j]; ne[j] = x[j]; }
    __syncthreads();
    double s = 0.0;
    if (i < nbins) {
        if (nbins == 1) s = h[0];
        else if (i == 0) s = 0.5 * (h[0] + h[1]);
        else if (i == nbins - 1) s = 0.5 * (h[nbins - 2] + h[nbins - 1]);
        else s = (h[i - 1] + h[i] + h[i + 1]) / 3.0;
    }
    const double tot = block_sum_bins(s, wpart);
    double mi = 0.0;
    if (i < nbins && tot > 0.0) {
        const double r = s / tot;
        if (r > 0.0) mi = (r < 1.0) ? pow((1.0 - r) / -log(r), alpha) : 1.0;
    }
    if (i < nbins) h[i] = mi;                                        // the smoothing reads are done
    // inclusive scan: within the wave (shifts 1..32), then the totals of the lower waves
    double sc = mi;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double up = __shfl_up(sc, off, 64);
        if (lane >= off) sc += up;
    }
    if (lane == 63) wpart[wave] = sc;
    __syncthreads();
    for (int w = 0; w < wave; ++w) sc += wpart[w];
    c[i] = sc;
    __syncthreads();
    const double msum = c[nbins - 1];
    if (tot > 0.0 && msum > 0.0 && i >= 1 && i < nbins) {
        const double target = (msum / nbins) * i;
        int lo = 0, hi = nbins - 1;                                  // first k < nbins - 1 with c[k] >= target
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (c[mid] >= target) hi = mid; else lo = mid + 1;
        }
        const double acc = lo > 0 ? c[lo - 1] : 0.0;
        double frac = (h[lo] > 0.0) ? (target - acc) / h[lo] : 0.0;
        frac = fmin(fmax(frac, 0.0), 1.0);
        ne[i] = x[lo] + frac * (x[lo + 1] - x[lo]);
    }
    __syncthreads();
    for (int j = i; j <= nbins; j += kMaxBins) xd[j] = ne[j];
    if (i < nbins) {
        const double lo = ne[i], hi = ne[i + 1];
        t64d[3 * i + 0] = lo; t64d[3 * i + 1] = hi - lo; t64d[3 * i + 2] = 1.0 - lo;
        t32d[3 * i + 0] = (float)lo; t32d[3 * i + 1] = (float)(hi - lo); t32d[3 * i + 2] = (float)(1.0 - lo);
    }
    __syncthreads();                                                 // LDS reused by the next dimension
}

// The reductions of one VEGAS iteration in one launch: blocks 0 .. kRec - 1 fold the block records
// into rec_out (as k_reduce_records), blocks kRec .. kRec + ncol - 1 fold the histogram columns
// into hsum (one block per column).
__global__ __launch_bounds__(kBlock) void k_vegas_fold(const double *__restrict__ parts, uint32_t nblk,
                                                       double *__restrict__ rec_out,
                                                       const double *__restrict__ hparts, uint32_t ncol,
                                                       double *__restrict__ hsum)
{
    const uint32_t b = blockIdx.x;
    if (b < (uint32_t)kRec) fold_column(parts, nblk, kRec, b, rec_is_max((int)b), rec_out + b);
    else fold_column(hparts, nblk, ncol, b - kRec, false, hsum + (b - kRec));
}

// Block d refines dimension d of the grid from hsum (the next launch on the stream, so hsum is
// complete).
__global__ __launch_bounds__(kBlock) void k_vegas_refine(const double *__restrict__ hsum, double *__restrict__ edges,
                                                         int nbins, double alpha, double *__restrict__ t64,
                                                         float *__restrict__ t32)
{
    const int d = blockIdx.x;
    refine_dim(hsum + (size_t)d * nbins, edges + (size_t)d * (nbins + 1), nbins, alpha, t64 + (size_t)3 * d * nbins,
               t32 + (size_t)3 * d * nbins);
}

hipError_t launch_vegas_step(hipStream_t s, const double *partials, uint32_t nblk, double *rec_out,
                             const double *hist_parts, Grid &g, double *hsum, double alpha)
{
    if (g.ndim < 1 || g.nbins < 1 || g.nbins > kMaxBins) return hipErrorInvalidValue;
    const uint32_t ncol = (uint32_t)(g.ndim * g.nbins);
    hipLaunchKernelGGL(k_vegas_fold, dim3(kRec + ncol), dim3(kBlock), 0, s, partials, nblk, rec_out, hist_parts,
                       ncol, hsum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_vegas_refine, dim3(g.ndim), dim3(kBlock), 0, s, hsum, g.edges, g.nbins, alpha, g.tab64,
                       g.tab32);
    return hipGetLastError();
}

// Per-bin tables {lo, width, 1 - lo} in FP64 and FP32 from the FP64 edges (after an explicit
// grid set or reset; the refinement writes them itself).
__global__ void k_vegas_tables(const double *__restrict__ edges, int ndim, int nbins, double *__restrict__ t64,
                               float *__restrict__ t32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndim * nbins) return;
    const int d = i / nbins, b = i % nbins;
    const double lo = edges[d * (nbins + 1) + b], hi = edges[d * (nbins + 1) + b + 1];
    t64[3 * i + 0] = lo; t64[3 * i + 1] = hi - lo; t64[3 * i + 2] = 1.0 - lo;
    t32[3 * i + 0] = (float)lo; t32[3 * i + 1] = (float)(hi - lo); t32[3 * i + 2] = (float)(1.0 - lo);
}

hipError_t launch_vegas_tables(hipStream_t s, Grid &g)
{
    const int n = g.ndim * g.nbins;
    hipLaunchKernelGGL(k_vegas_tables, dim3((n + 255) / 256), dim3(256), 0, s, g.edges, g.ndim, g.nbins, g.tab64,
                       g.tab32);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ occupancy
// Resident blocks per CU of each integration kernel instance (for one-wave, equal-work grids).
template <typename T, bool GRID, bool HEAVY, bool HW>
static int occ_had_t(bool vegas, bool hist, size_t sm)
{
    int nb = 0;
    hipError_t e;
    if (!vegas)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_integrate_had<T, false, false, GRID, HEAVY, kBlock, false, HW>, kBlock, sm);
    else if (!hist)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_integrate_had<T, true, false, GRID, HEAVY, kBlock, false, HW>, kBlock, sm);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_integrate_had<T, true, true, GRID, HEAVY, kBlock, false, HW>, kBlock, sm);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

template <typename T, bool HW = false>
static int occ_had_tt(bool vegas, bool hist, size_t sm, bool grid, bool heavy)
{
    if (heavy) return grid ? occ_had_t<T, true, true, HW>(vegas, hist, sm) : occ_had_t<T, false, true, HW>(vegas, hist, sm);
    return grid ? occ_had_t<T, true, false, HW>(vegas, hist, sm) : occ_had_t<T, false, false, HW>(vegas, hist, sm);
}

int occ_integrate_had(int prec, bool vegas, bool hist, int nbins, bool grid, bool heavy)
{
    HadLaunch L{};
    L.prec = prec == MI_PREC_FP32_HW ? MI_PREC_FP32 : prec; L.vegas = vegas;
    Grid g; g.nbins = nbins; g.ndim = 3;
    L.grid = &g; L.hist = hist ? (double *)1 : nullptr;
    const size_t sm = had_smem(L);
    if (prec == MI_PREC_FP32_HW) return occ_had_tt<float, true>(vegas, hist, sm, grid, heavy);
    if (prec == MI_PREC_FP32) return occ_had_tt<float>(vegas, hist, sm, grid, heavy);
    return occ_had_tt<double>(vegas, hist, sm, grid, heavy);
}

int occ_integrate_c1(bool vegas, bool hist, int nbins)
{
    const int nb = vegas ? nbins : 0;
    const size_t sm = vegas ? sizeof(double) * 3 * nb + (hist ? (kBlock / 64) * WaveHist<1>::area_bytes(nb) : 0) : 0;
    int b = 0;
    hipError_t e;
    if (!vegas) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_integrate_c1<false, false>, kBlock, sm);
    else if (!hist) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_integrate_c1<true, false>, kBlock, sm);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_integrate_c1<true, true>, kBlock, sm);
    return (e == hipSuccess && b > 0) ? b : 1;
}

// General variations: the event's own process weight recomputed at (xi_R^2 Q^2, xi_F^2 Q^2) with
// the PDFs of a slot, Q^2 = sH a b (pT^2, or pT^2 + m^2 for the heavy pairs), from the event's
// (sHat, tHat, x1, x2, code); out = w * W(variation) / W(the run's own scales, slot 0), where
// W = alphaS^2 * bracket * luminosity of the code (the phase-space Jacobian cancels).
__device__ __forceinline__ double proc_weight(const HadConst<double> &k, const GridSet64 &gs, bool grid, int fam, int j,
                                              double a, double b, double mu, double Q2, double x1, double lx1,
                                              double x2, double lx2, const Lumi<double> &toyL, double xiR2, double xiF2,
                                              int slot)
{
    const double alpS = alphas_f64(k, xiR2 * Q2);
    Lumi<double> L = toyL;
    if (grid) {
        const double q = sqrt(xiF2 * Q2);
        L = grid_lumi<double>(gs.g[slot], (int)(0.5 * k.nSea), x1, lx1, x2, lx2, q, log(q));
    }
    double wb[6];
    if (fam == 0) brackets_f64<0>(k, a, b, mu, L, wb);
    else if (fam == 1) brackets_f64<1>(k, a, b, mu, L, wb);
    else brackets_f64<2>(k, a, b, mu, L, wb);
    const double br = wb[j] > 0.0 ? wb[j] : 0.0;
    return alpS * alpS * br;
}

__global__ __launch_bounds__(kBlock) void k_reweight_vars(HadConst<double> k, GridSet64 gs, int grid, uint64_t n,
                                                          const double *__restrict__ sH, const double *__restrict__ tH,
                                                          const double *__restrict__ x1v, const double *__restrict__ x2v,
                                                          const int32_t *__restrict__ pid, const double *__restrict__ w,
                                                          int nvar, const RwVar *__restrict__ vars, double xiR2_0,
                                                          double xiF2_0, double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int code = pid[i];
    int fam = -1, j = 0;
    if (code >= 111 && code <= 116) { fam = 0; j = code - 111; }
    else if (code == 121 || code == 122) { fam = 1; j = code - 121; }
    else if (code == 123 || code == 124) { fam = 2; j = code - 123; }
    if (fam < 0) {                                     // no process (a zero-weight candidate)
        for (int v = 0; v < nvar; ++v) out[i * nvar + v] = 0.0;
        return;
    }
    const double m2 = fam == 0 ? 0.0 : k.fam[fam - 1].m2;
    const double s = sH[i], t = tH[i];
    const double a = (m2 - t) / s, b = (s - m2 + t) / s;  // (m^2 - tH)/sH, (m^2 - uH)/sH
    const double Q2 = s * (a * b), mu = m2 / s;
    const double x1 = x1v[i], x2 = x2v[i], lx1 = log(x1), lx2 = log(x2);
    double l1a, l1b;
    const Lumi<double> toyL = grid ? Lumi<double>{} : toy_lumi_f64(k, lx1, lx2, l1a, l1b);
    const double W0 = proc_weight(k, gs, grid != 0, fam, j, a, b, mu, Q2, x1, lx1, x2, lx2, toyL, xiR2_0, xiF2_0, 0);
    for (int v = 0; v < nvar; ++v) {
        const RwVar q = vars[v];
        const double Wv = proc_weight(k, gs, grid != 0, fam, j, a, b, mu, Q2, x1, lx1, x2, lx2, toyL, q.xiR2, q.xiF2,
                                      q.slot);
        out[i * nvar + v] = W0 > 0.0 ? w[i] * (Wv / W0) : 0.0;
    }
}

hipError_t launch_reweight_vars(hipStream_t s, const HadConst<double> &k, const GridSet64 &gs, bool grid, uint64_t n,
                                const double *sH, const double *tH, const double *x1, const double *x2,
                                const int32_t *pid, const double *w, int nvar, const RwVar *vars_dev,
                                double xiR2_0, double xiF2_0, double *out)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reweight_vars, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, k, gs,
                       grid ? 1 : 0, n, sH, tH, x1, x2, pid, w, nvar, vars_dev, xiR2_0, xiF2_0, out);
    return hipGetLastError();
}

}  // namespace mi

namespace mi {

// ------------------------------------------------------------------------------------ scale reweighting
// Same-sample renormalisation-scale variations of weighted events (cuPythia/pipeline/reweight.cu:36-51):
// w_k = w * (alphaS(xi_k^2 Q^2) / alphaS(xi0^2 Q^2))^2 with Pythia's Q^2 of the event's process
// (SigmaProcess.cc:876-922): pT^2 = tHat uHat / sHat for the massless codes, and for the heavy
// pairs (m3 = m4 = m) sqrt((pT^2 + m^2)^2) = (m^2 - tHat)(m^2 - uHat) / sHat with
// uHat = 2 m^2 - sHat - tHat; alphaS in Pythia's order-0/1 form (FP64).
__device__ __forceinline__ double alphas64(const HadConst<double> &k, double Q2)
{
    if (k.as_order == 0.0) return k.as_fixed;
    Q2 = (Q2 < k.scale2Min) ? k.scale2Min : Q2;
    const double lq = log(Q2);
    const bool t6 = Q2 > k.mt2 && k.nf6 != 0.0, t5 = Q2 > k.mb2, t4 = Q2 > k.mc2;
    const double lam = t6 ? k.lnL6s : (t5 ? k.lnL5s : (t4 ? k.lnL4s : k.lnL3s));
    const double cf = t6 ? k.c6 : (t5 ? k.c5 : (t4 ? k.c4 : k.c3));
    return cf / (lq - lam);
}

__global__ __launch_bounds__(kBlock) void k_reweight(HadConst<double> k, uint64_t n, const double *__restrict__ sH,
                                                     const double *__restrict__ tH, const int32_t *__restrict__ pid,
                                                     const double *__restrict__ w, int nvar,
                                                     const double *__restrict__ fac, double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int code = pid ? pid[i] : 111;
    const double m2 = (code == 121 || code == 122) ? k.fam[0].m2 : ((code == 123 || code == 124) ? k.fam[1].m2 : 0.0);
    const double s = sH[i], t = tH[i], u = 2.0 * m2 - (s + t);
    const double pT2 = (m2 - t) * (m2 - u) / s;
    const double a0 = alphas64(k, k.xiR2 * pT2);
    for (int v = 0; v < nvar; ++v) {
        const double r = alphas64(k, fac[v] * fac[v] * pT2) / a0;
        out[i * nvar + v] = w[i] * (r * r);
    }
}

hipError_t launch_reweight(hipStream_t s, const HadConst<double> &k, uint64_t n, const double *sH, const double *tH,
                           const int32_t *pid, const double *w, int nvar, const double *fac, double *out)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reweight, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, k, n, sH, tH, pid,
                       w, nvar, fac, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------ regenerated variations
// The reference's pinned-seed reweighting (cuPythia/pipeline/reweight.cu:36-67: each event's weights
// recomputed from its own counter seed): event i is the point of counter idx[i] (Philox stream 0
// under key, the flat or VEGAS map), evaluated in FP64 at every variation v with that variation's
// physics constants and PDF view -- the same device function as the FP64 per-point dump, so W_v is
// bit for bit the dump of a run at those scales.  W_v is the weight of the event's own process
// code.  vars[0] holds the run's own constants (W_0) when w is given: out = w * W_v / W_0; with w
// null, out = W_v.  Codes outside mask (or no process) give 0.
template <bool VEGAS, bool HEAVY>
__global__ __launch_bounds__(kBlock) void k_reweight_regen(const RwRegen *__restrict__ vars, int nvar, int mask,
                                                           uint64_t key, uint64_t n, const uint64_t *__restrict__ idx,
                                                           const int32_t *__restrict__ pid,
                                                           const double *__restrict__ w,
                                                           const double *__restrict__ gtab, int nbins,
                                                           double *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int code = pid[i];
    const int slot = (code >= 111 && code <= 116) ? code - 111 : ((code >= 121 && code <= 124) ? code - 115 : -1);
    if (slot < 0 || !((mask >> slot) & 1)) {
        for (int v = 0; v < nvar; ++v) out[i * nvar + v] = 0.0;
        return;
    }
    const uint64_t e = idx[i];
    const int v0 = w ? 0 : 1;               // vars[0]: the run's own scales (W_0), only with w
    double W0 = 0.0;
    for (int v = v0; v <= nvar; ++v) {
        const RwRegen &q = vars[v];
        double ws[kMaxProc];
        int bins[3];
        had_candidate<double, VEGAS, HEAVY>(q.k, mask, key, e, gtab, nbins, (double)nbins, ws, bins, (double *)nullptr,
                                            (double *)nullptr, (double *)nullptr, (double *)nullptr,
                                            &q.g, q.g.active);
        double Wv = 0.0;
#pragma unroll
        for (int p = 0; p < kMaxProc; ++p) Wv = (p == slot) ? ws[p] : Wv;   // selects: no dynamic register index
        if (v == 0) { W0 = Wv; continue; }
        double r = Wv;
        if (w) r = W0 > 0.0 ? w[i] * (Wv / W0) : 0.0;
        out[i * nvar + (v - 1)] = r;
    }
}

hipError_t launch_reweight_regen(hipStream_t s, const RwRegen *vars_dev, int nvar, int mask, bool vegas,
                                 uint64_t key, uint64_t n, const uint64_t *idx, const int32_t *pid, const double *w,
                                 const double *gtab, int nbins, double *out)
{
    if (n == 0) return hipSuccess;
    const dim3 grid((uint32_t)((n + kBlock - 1) / kBlock)), blk(kBlock);
    const bool hv = has_heavy(mask);
    if (vegas) {
        if (hv) hipLaunchKernelGGL((k_reweight_regen<true, true>), grid, blk, 0, s, vars_dev, nvar, mask, key, n, idx, pid, w, gtab, nbins, out);
        else    hipLaunchKernelGGL((k_reweight_regen<true, false>), grid, blk, 0, s, vars_dev, nvar, mask, key, n, idx, pid, w, gtab, nbins, out);
    } else {
        if (hv) hipLaunchKernelGGL((k_reweight_regen<false, true>), grid, blk, 0, s, vars_dev, nvar, mask, key, n, idx, pid, w, gtab, nbins, out);
        else    hipLaunchKernelGGL((k_reweight_regen<false, false>), grid, blk, 0, s, vars_dev, nvar, mask, key, n, idx, pid, w, gtab, nbins, out);
    }
    return hipGetLastError();
}

}  // namespace mi
