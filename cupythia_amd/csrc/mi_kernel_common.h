// mi_kernel_common.h -- device helpers shared by the kernel translation units: wave / block
// reductions, the VEGAS grid map and the per-candidate integrand evaluation.
#pragma once

#include "mi_internal.h"

#include <type_traits>

namespace mi {

// ------------------------------------------------------------------------------------ utils
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ double wave_max(double x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}

// Per-wave histogram of w^2 over D dimensions in which every addition happens in an order the
// code fixes.  A batch is one point per lane (w2 = 0 for a lane without a point: it adds nothing).
// For each dimension, every lane with w2 != 0 ORs its lane bit into the group mask of its bin (an
// OR gives the same mask in any order), reads the mask back and counts the lanes below it: its
// rank in the bin.  Round t then adds the w2 of the rank-t lanes, which lie in distinct bins, so
// each bin takes its batch's values in increasing lane order, batches in sequence, and the sums
// do not depend on the order in which the hardware applies atomics to one address.  A batch takes
// as many rounds per dimension as its largest number of points in one bin.  The masks are cleared
// after each dimension, so one mask per bin serves every dimension.
//
// The wave's LDS area holds the histogram h[d][bin] (FP64), then the masks m[bin] (u64).
template <int D>
struct WaveHist {
    __host__ __device__ static constexpr size_t area_bytes(int nbins) { return wave_hist_bytes(D, nbins); }

    // every lane of the wave, before the first batch
    __device__ static void init(unsigned char *area, int nbins)
    {
        double *h = reinterpret_cast<double *>(area);
        uint64_t *m = reinterpret_cast<uint64_t *>(area + (size_t)D * nbins * 8);
        for (int i = threadIdx.x & 63; i < D * nbins; i += 64) h[i] = 0.0;
        for (int i = threadIdx.x & 63; i < nbins; i += 64) m[i] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // every lane of the wave, once per batch
    __device__ static void add(double w2, const int *bins, unsigned char *area, int nbins)
    {
        double *h = reinterpret_cast<double *>(area);
        uint64_t *m = reinterpret_cast<uint64_t *>(area + (size_t)D * nbins * 8);
        const uint64_t bit = 1ull << (threadIdx.x & 63);
        const bool has = w2 != 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            uint64_t *mb = m + bins[d];
            if (has) __hip_atomic_fetch_or(mb, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            const uint64_t grp = has ? __hip_atomic_load(mb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) : 0ull;
            // lanes below this one in its bin (< 64, so the rounds end); 64 for a lane that adds nothing
            const uint32_t rank =
                has ? __builtin_amdgcn_mbcnt_hi((uint32_t)(grp >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)grp, 0u))
                    : 64u;
            double *hb = h + d * nbins + bins[d];
            // The first kUnrolled rounds without a test (a round with no lane of its rank skips its
            // add).  The fence after each round keeps the rounds' adds in separate instructions: a
            // thread takes at most one of them, so without it they may be merged into one add that
            // all lanes of rank < kUnrolled issue together.
            constexpr uint32_t kUnrolled = 4;
#pragma unroll
            for (uint32_t t = 0; t < kUnrolled; ++t) {
                if (rank == t) __hip_atomic_fetch_add(hb, w2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
            for (uint32_t t = kUnrolled; __ballot(rank >= t && rank < 64u) != 0; ++t) {
                if (rank == t) __hip_atomic_fetch_add(hb, w2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (has) __hip_atomic_store(mb, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        }
    }

    // the wave's histogram as [d][bin], once the block has synchronised
    __device__ static const double *hist(const unsigned char *area) { return reinterpret_cast<const double *>(area); }
};

// Combine kRec per-thread values into one block record (sum except the max slots).
// If cnt is given, it receives slot kRAcc (the unweighting accept count) as an integer.
template <int NT = kBlock>
__device__ __forceinline__ void block_record(const double *v, double *dst, uint64_t *cnt = nullptr)
{
    __shared__ double red[NT / 64][kRec];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kRec; ++i) {
        const bool isMax = rec_is_max(i);
        const double r = isMax ? wave_max(v[i]) : wave_sum(v[i]);
        if (lane == 0) red[wave][i] = r;
    }
    __syncthreads();
    if (threadIdx.x < kRec) {
        const int i = threadIdx.x;
        const bool isMax = rec_is_max(i);
        double r = red[0][i];
        for (int w = 1; w < NT / 64; ++w) r = isMax ? fmax(r, red[w][i]) : r + red[w][i];
        dst[i] = r;
        if (cnt && i == kRAcc) *cnt = (uint64_t)r;
    }
}

// block_record for a record whose only non-zero slots are the NP + 1 sums, sums of squares and
// maxima and the non-zero count (the integration kernels): the same wave reductions and fold
// over waves per slot, zeros elsewhere, from small per-thread arrays instead of a full record.
template <int NT, int NP>
__device__ __forceinline__ void block_record_sums(const double *sw, const double *sw2, const double *mx,
                                                  double nz, double *dst)
{
    __shared__ double red[NT / 64][kRec];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p <= NP; ++p) {
        const double a = wave_sum(sw[p]), b = wave_sum(sw2[p]), c = wave_max(mx[p]);
        if (lane == 0) { red[wave][kRS + p] = a; red[wave][kRQ + p] = b; red[wave][kRM + p] = c; }
    }
    const double z = wave_sum(nz);
    if (lane == 0) red[wave][kRNz] = z;
    __syncthreads();
    if (threadIdx.x < kRec) {
        const int i = threadIdx.x;
        const int slot = i < kRQ ? i - kRS : (i < kRM ? i - kRQ : i - kRM);
        const bool used = (i < kRNz && slot <= NP) || i == kRNz;
        double r = 0.0;
        if (used) {
            const bool isMax = rec_is_max(i);
            r = red[0][i];
            for (int w = 1; w < NT / 64; ++w) r = isMax ? fmax(r, red[w][i]) : r + red[w][i];
        }
        dst[i] = r;
    }
}

// Running maximum of non-negative weights: FP32 as the integer order of the bit patterns, FP64 as
// fmax; MaxOf<T> is the running value's type.  Converted to FP64 once, exactly.
__device__ __forceinline__ uint32_t run_max(uint32_t m, float w) { return max(m, __float_as_uint(w)); }
__device__ __forceinline__ double run_max(double m, double w) { return fmax(m, w); }
template <typename T> __device__ __forceinline__ double max_value(uint32_t m) { return (double)__uint_as_float(m); }
template <typename T> __device__ __forceinline__ double max_value(double m) { return m; }
template <typename T> using MaxOf = typename std::conditional<sizeof(T) == 4, uint32_t, double>::type;

// a[f] of a family array (f < N) as selects, so a register array never needs a dynamic index.
template <typename T, int N>
__device__ __forceinline__ T pick_fam(const T *a, int f)
{
    if (N == 1) return a[0];
    return f == 0 ? a[0] : (f == 1 ? a[1] : a[N - 1]);
}

// ------------------------------------------------------------------------------------ map
// 24-bit uniform k -> (x, 1 - x, bin, jacobian factor) through the grid table of one dim.
template <typename T>
struct MapOut { T x, omx, jac; int bin; };

template <typename T>
__device__ __forceinline__ MapOut<T> map_flat(uint32_t k)
{
    MapOut<T> m;
    m.x = (T)k * (T)(1.0 / 16777216.0);
    m.omx = (T)(16777216u - k) * (T)(1.0 / 16777216.0);
    m.jac = (T)1;
    m.bin = 0;
    return m;
}

// The flat map's integers k and 2^24 - k (exact in FP32): the base-2 FP32_HW point takes these
// and scales inside the products that use them.
template <typename T>
__device__ __forceinline__ MapOut<T> map_flat_u24(uint32_t k)
{
    MapOut<T> m;
    m.x = (T)k;
    m.omx = (T)16777216.0f - m.x;
    m.jac = (T)1;
    m.bin = 0;
    return m;
}

template <typename T>
__device__ __forceinline__ MapOut<T> map_grid(uint32_t k, const T *tab, int nbins, T nbf)
{
    const uint32_t kb = k * (uint32_t)nbins;              // < 2^32 for nbins <= 256
    const int b = (int)(kb >> 24);
    const T frac = (T)(kb & 0xFFFFFFu) * (T)(1.0 / 16777216.0);
    const T lo = tab[3 * b + 0], wd = tab[3 * b + 1], omlo = tab[3 * b + 2];
    MapOut<T> m;
    m.x = fma(frac, wd, lo);
    // FP32: 1 - x from the complement table keeps ~1e-7 relative accuracy next to x = 1.
    // FP64: 1 - x of the mapped point itself (exact for x >= 1/2), so the point is exactly x.
    m.omx = (sizeof(T) == 4) ? fma(-frac, wd, omlo) : (T)1 - m.x;
    m.jac = nbf * wd;
    m.bin = b;
    return m;
}

// grid: the PDFs of the grid *gp, else the toy PDFs (gp is then not read).  The choice is a flag
// rather than a null test of gp, which keeps a kernel argument's view out of scratch memory.  HEAVY: see had_point_f32.  HW: the FP32 point
// with the hardware transcendentals (MI_PREC_FP32_HW); the FP64 point ignores it.
// U24: the uniforms are map_flat_u24's integers (FP32 HW only).
template <typename T, bool HEAVY, bool HW = false, bool U24 = false> struct HadPoint;
template <bool HEAVY, bool HW, bool U24> struct HadPoint<float, HEAVY, HW, U24> {
    static __device__ __forceinline__ float eval(const HadConst<float> &k, int mask, float u0, float o0,
                                                 float u1, float o1, float u2, float o2, float *ws,
                                                 float *sH, float *tH, float *x1, float *x2,
                                                 const GridPdf<float> *gp, bool grid)
    { return had_point_f32<HEAVY, HW, U24>(k, mask, u0, o0, u1, o1, u2, o2, ws, sH, tH, x1, x2, nullptr, gp, grid); }
};
template <bool HEAVY, bool HW, bool U24> struct HadPoint<double, HEAVY, HW, U24> {
    static __device__ __forceinline__ double eval(const HadConst<double> &k, int mask, double u0, double o0,
                                                  double u1, double o1, double u2, double o2, double *ws,
                                                  double *sH, double *tH, double *x1, double *x2,
                                                  const GridPdf<double> *gp, bool grid)
    { return had_point_f64<HEAVY>(k, mask, u0, o0, u1, o1, u2, o2, ws, sH, tH, x1, x2, nullptr, gp, grid); }
};

// One candidate of the hadronic integrand: Philox stream 0 of event e -> three 24-bit
// uniforms -> (optional VEGAS map) -> weights (mapped by the grid Jacobian).  HEAVY = false:
// the massless codes only (ws[6..9] = 0); HEAVY = true: every family of mask, kinematics
// outputs (if given) one per family.
template <typename T, bool VEGAS, bool HEAVY = false, bool HW = false>
__device__ __forceinline__ T had_candidate(const HadConst<T> &k, int mask, uint64_t key, uint64_t e,
                                           const T *stab, int nbins, T nbf, T *ws, int *bins,
                                           T *sH, T *tH, T *x1, T *x2, const GridPdf<T> *gp, bool grid)
{
    const U4 r = philox_event(key, e, 0u);
    MapOut<T> m0, m1, m2;
    constexpr bool U24 = HW && !VEGAS && sizeof(T) == 4;
    if (VEGAS) {
        m0 = map_grid<T>(k24(r.x), stab, nbins, nbf);
        m1 = map_grid<T>(k24(r.y), stab + 3 * nbins, nbins, nbf);
        m2 = map_grid<T>(k24(r.z), stab + 6 * nbins, nbins, nbf);
    } else if (U24) {
        m0 = map_flat_u24<T>(k24(r.x)); m1 = map_flat_u24<T>(k24(r.y)); m2 = map_flat_u24<T>(k24(r.z));
    } else {
        m0 = map_flat<T>(k24(r.x)); m1 = map_flat<T>(k24(r.y)); m2 = map_flat<T>(k24(r.z));
    }
    T tot = HadPoint<T, HEAVY, HW, U24>::eval(k, mask, m0.x, m0.omx, m1.x, m1.omx, m2.x, m2.omx, ws, sH, tH, x1, x2, gp, grid);
    if (VEGAS) {
        T jac = (T)1;
        jac = jac * m0.jac;
        jac = jac * m1.jac;
        jac = jac * m2.jac;
#pragma unroll
        for (int i = 0; i < (HEAVY ? kMaxProc : kLightProc); ++i) ws[i] = ws[i] * jac;
        tot = tot * jac;
        bins[0] = m0.bin; bins[1] = m1.bin; bins[2] = m2.bin;
    }
    return tot;
}

// ------------------------------------------------------------------------------------ partonic candidate
// Event e: Philox stream 0, u = 53-bit uniform from words (0, 1); optional 1-D grid
// (FP64 edges, bin from the top bits of the 53-bit integer).
__device__ __forceinline__ double c1_candidate(double pref, double cMax, uint64_t key, uint64_t e, bool vegas,
                                               const double *tab, int nbins, int *bin)
{
    const U4 r = philox_event(key, e, 0u);
    const uint64_t x53 = ((((uint64_t)r.y) << 32) | r.x) >> 11;
    double u, jac = 1.0;
    if (vegas) {
        // bin = floor(u * nbins) exactly: the top bits of x53 * nbins (fits in 61 bits)
        const uint64_t kb = x53 * (uint64_t)nbins;
        const int b = (int)(kb >> 53);
        const double frac = (double)(kb & ((1ULL << 53) - 1)) * (1.0 / 9007199254740992.0);
        const double lo = tab[3 * b], wd = tab[3 * b + 1];
        u = fma(frac, wd, lo);
        jac = (double)nbins * wd;
        *bin = b;
    } else {
        u = (double)x53 * (1.0 / 9007199254740992.0);
        *bin = 0;
    }
    return c1_weight(pref, cMax, u) * jac;
}

}  // namespace mi
